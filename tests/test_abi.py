"""The C-ABI library loads and exports every entry point include/*.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "seismic_tomography_3d_amd", "libst_hip.so")


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(st_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared()
    assert "st_fim_solve" in names and "st_forward" in names and "st_optimize" in names
    assert len(names) >= 19


@pytest.mark.skipif(not os.path.exists(LIB), reason="libst_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libst_hip.so not built")
def test_python_binding_covers_header():
    from seismic_tomography_3d_amd import _lib
    assert set(declared()) <= set(_lib.exported_names())


@pytest.mark.skipif(not os.path.exists(LIB), reason="libst_hip.so not built")
@pytest.mark.parametrize("nb,factor", [(8, 0.028), (50, 0.0045), (4, 0.1)])
def test_acoustic_dampers_host_entry_matches_oracle(nb, factor):
    """st_acoustic_dampers runs on the host only: its tables equal the
    oracle's (and the reference build's, where present) bit for bit."""
    import numpy as np
    import oracle_lib as orc
    from oracle_lib import P
    from seismic_tomography_3d_amd.api import acoustic_dampers
    d1, d2, d3 = acoustic_dampers(factor, nb)
    o = [np.zeros(n, np.float32) for n in (nb, nb * nb, nb ** 3)]
    orc.oracle().or_acoustic_dampers(*(P(x) for x in o), factor, nb)
    assert np.array_equal(d1, o[0]) and np.array_equal(d2.ravel(), o[1])
    assert np.array_equal(d3.ravel(), o[2])
    if orc.ref_available():
        r = [np.zeros(n, np.float32) for n in (nb, nb * nb, nb ** 3)]
        orc.ref().ref_dampers(*(P(x) for x in r), factor, nb)
        assert np.array_equal(d3.ravel(), r[2])


@pytest.mark.skipif(not os.path.exists(LIB), reason="libst_hip.so not built")
@pytest.mark.parametrize("nt,dt,fmax,delay", [(3001, 1e-3, 40.0, 0.1), (500, 5e-4, 25.0, 0.05)])
def test_ricker_host_entry(nt, dt, fmax, delay):
    """st_ricker (host only) equals the reference's rickerGeneration bit for
    bit where the reference build is present; in any case it peaks at 1 at
    t = delay and is symmetric about it."""
    import numpy as np
    import oracle_lib as orc
    from oracle_lib import P
    from seismic_tomography_3d_amd.api import ricker
    w = ricker(nt, dt, fmax, delay)
    k = int(round(delay / dt))
    assert abs(w[k] - 1.0) < 1e-6 and w.max() == w[k]
    assert np.allclose(w[k - 20:k], w[k + 20:k:-1], atol=1e-5)
    if orc.ref_available():
        r = np.zeros(nt, np.float32)
        orc.ref().ref_ricker(delay, fmax, dt, nt, P(r))
        assert np.array_equal(w, r)
