"""The committed fixtures of tests/golden (written by the reference's own
CPU code, tests/golden/gen_fixtures.py) against the oracle restatement and
the host-side parser and geometry.  Runs without the reference checkout."""
import ctypes
import os
import tempfile

import numpy as np
import pytest

import oracle_lib as orc
from oracle_lib import P
from seismic_tomography_3d_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden")
HOST = os.path.join(os.path.dirname(HERE), "seismic_tomography_3d_amd", "libst_host.so")


def load(name):
    return dict(np.load(os.path.join(G, name), allow_pickle=False))


@pytest.mark.parametrize("fx", ["ref_params.npz", "ref_params_recip.npz"])
def test_fsm_picks_and_volume_bit_exact(fx):
    f = load(fx)
    vel = f["vel"]
    ny, nx, nz = vel.shape
    h = 25.0
    recip = "reciprocity = True" in str(f["params_text"])
    shots, nodes = f["shots_xyz"], f["nodes_xyz"]
    if recip:
        shots, nodes = nodes, shots
    o = orc.oracle()
    s = np.ascontiguousarray((1.0 / vel).astype(np.float32))
    for i, src in enumerate(shots):
        T = np.zeros(vel.shape, np.float32)
        o.or_fsm(P(s), nx, ny, nz, h, h, h, *(float(v) for v in src), P(T))
        if i == 0:
            assert np.array_equal(T, f["fsm_T_shot1"])
        assert np.array_equal(orc.picks(T, (h, h, h), nodes), f["fsm_picks"][i])


@pytest.mark.parametrize("fx", ["ref_params.npz", "ref_params_recip.npz"])
def test_regular_geometry_bit_exact(fx):
    f = load(fx)
    x, y, z = synth.regular_geometry((100, 120), (100, 600), (700, 120), 3, 2, 10.0)
    assert np.array_equal(np.stack([x, y, z], 1), f["shots_xyz"])
    x, y, z = synth.regular_geometry((50, 40), (50, 650), (775, 40), 5, 4, 0.0)
    assert np.array_equal(np.stack([x, y, z], 1), f["nodes_xyz"])


@pytest.mark.parametrize("fx", ["ref_params.npz", "ref_params_recip.npz"])
def test_podvin_picks_bit_exact(fx):
    """The Podvin & Lecomte restatement reproduces the picks the reference's
    own classic.cpp wrote for every shot of the parameter file, and stays
    within discretisation error of the FSM."""
    f = load(fx)
    if "pod_picks" not in f:
        pytest.skip("fixture without Podvin picks")
    vel = f["vel"]
    ny, nx, nz = vel.shape
    h = 25.0
    recip = "reciprocity = True" in str(f["params_text"])
    shots, nodes = f["shots_xyz"], f["nodes_xyz"]
    if recip:
        shots, nodes = nodes, shots
    o = orc.oracle()
    s = np.ascontiguousarray((1.0 / vel).astype(np.float32))
    for i, src in enumerate(shots):
        T = np.zeros(vel.shape, np.float32)
        o.or_podvin(P(s), nx, ny, nz, h, h, h, *(float(v) for v in src), P(T))
        if i == 0:
            assert np.array_equal(T, f["pod_T_shot1"])
        assert np.array_equal(orc.picks(T, (h, h, h), nodes), f["pod_picks"][i]), i
    assert np.abs(f["pod_picks"] - f["fsm_picks"]).max() < 0.05


def test_trilinear_and_gaussian_fixture():
    f = load("ref_kernels.npz")
    o = orc.oracle()
    for a, r in zip(f["tri_args"], f["tri_out"]):
        assert np.float32(o.or_trilinear(*[float(v) for v in a])) == r
    g = f["gauss_in"]
    out = np.zeros_like(g)
    o.or_gaussian(P(g), P(out), 11, 12, 10, 5, 2.0)
    assert np.array_equal(out, f["gauss_out"])


@pytest.mark.skipif(not os.path.exists(HOST), reason="libst_host.so not built")
@pytest.mark.parametrize("fx", ["ref_params.npz", "ref_params_recip.npz"])
def test_host_parser_and_geometry(fx):
    f = load(fx)
    lib = ctypes.CDLL(HOST)
    lib.sth_catch_parameter.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    lib.sth_regular.argtypes = [ctypes.c_char_p, ctypes.c_char_p, orc.fp, orc.fp, orc.fp, ctypes.c_int]
    d = tempfile.mkdtemp()
    pf = os.path.join(d, "params.txt")
    with open(pf, "w") as fh:
        fh.write(str(f["params_text"]).replace("{dir}", d))
    buf = ctypes.create_string_buffer(512)
    for k, v in zip(f["keys"], f["values"]):
        lib.sth_catch_parameter(str(k).encode(), pf.encode(), buf, 512)
        assert buf.value.decode().replace(d, "{dir}") == str(v), k
    for name in ("shots", "nodes"):
        x = np.zeros(64, np.float32); y = np.zeros(64, np.float32); z = np.zeros(64, np.float32)
        n = lib.sth_regular(pf.encode(), name.encode(), P(x), P(y), P(z), 64)
        assert np.array_equal(np.stack([x[:n], y[:n], z[:n]], 1), f[f"{name}_xyz"])
