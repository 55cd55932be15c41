"""The Podvin & Lecomte eikonal (classic.cpp, eikonal_type 0) on the
device against the oracle's restatement (itself bit-exact against the
reference's own build): whole volumes and picks, bit for bit, on
heterogeneous models, sources on and off nodes and at the model edges,
unequal spacings (the reference uses h = dx), several shots per batch and
batches split."""
import numpy as np
import pytest

import oracle_lib as orc
from oracle_lib import P

pytestmark = pytest.mark.gpu


def _model(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    vel = np.empty((ny, nx, nz), np.float32)
    vel[:] = np.linspace(1500.0, 3500.0, nz, dtype=np.float32)[None, None, :]
    vel *= rng.uniform(0.8, 1.2, vel.shape).astype(np.float32)
    return np.ascontiguousarray((1.0 / vel).astype(np.float32))


def _oracle(s, d, src):
    ny, nx, nz = s.shape
    T = np.zeros(s.shape, np.float32)
    orc.oracle().or_podvin(P(s), nx, ny, nz, *d, *(float(v) for v in src), P(T))
    return T


@pytest.mark.parametrize("shape,d,seed", [
    ((21, 17, 13), (10.0, 10.0, 10.0), 1),
    ((19, 23, 11), (12.5, 12.5, 12.5), 3),
    ((15, 12, 14), (10.0, 8.0, 7.0), 4),
])
def test_podvin_bit_exact(ctx, shape, d, seed):
    nx, ny, nz = shape
    s = _model(nx, ny, nz, seed)
    ctx.set_model(s, *d)
    ext = np.array([(nx - 1) * d[0], (ny - 1) * d[1], (nz - 1) * d[2]], np.float32)
    rng = np.random.default_rng(seed)
    src = (rng.uniform(0, 1, (5, 3)).astype(np.float32) * ext).astype(np.float32)
    src[0] = 0.0                                  # corner node
    src[1] = ext                                  # far corner
    src[2] = [ext[0], 0.0, ext[2] * 0.5]          # on a face
    rcv = (rng.uniform(0, 1, (12, 3)).astype(np.float32) * ext).astype(np.float32)
    print(f"shape {shape} d {d}", flush=True)
    picks, T, ms = ctx.podvin_solve(src, rcv, return_T=True, batch=3)
    for k in range(len(src)):
        ref = _oracle(s, d, src[k])
        bad = np.argwhere(T[k] != ref)
        assert bad.size == 0, (k, bad[:5])
        assert np.array_equal(picks[k], orc.picks(ref, d, rcv))
    p2, _, _ = ctx.podvin_solve(src[3:], rcv)
    assert np.array_equal(p2, picks[3:])


def test_podvin_fixture_files(ctx):
    """The reference's own picks for the fixture parameter file."""
    import os
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_params.npz"),
                allow_pickle=False)
    vel = f["vel"]
    h = 25.0
    ctx.set_model((1.0 / vel).astype(np.float32), h)
    picks, T, _ = ctx.podvin_solve(f["shots_xyz"], f["nodes_xyz"], return_T=True)
    assert np.array_equal(picks, f["pod_picks"])
    assert np.array_equal(T[0], f["pod_T_shot1"])
