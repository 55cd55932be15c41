"""The Podvin & Lecomte restatement (oracle/oracle.c or_podvin) against the
reference's own classic.cpp build: whole travel-time volumes, bit for bit,
on heterogeneous models, sources on and off nodes and next to the model
edges, and unequal spacings (the reference uses h = dx for every axis)."""
import numpy as np
import pytest

import oracle_lib as orc
from oracle_lib import P


def _case(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    vel = np.empty((ny, nx, nz), np.float32)
    vel[:] = np.linspace(1500.0, 3500.0, nz, dtype=np.float32)[None, None, :]
    vel *= rng.uniform(0.8, 1.2, vel.shape).astype(np.float32)
    return np.ascontiguousarray((1.0 / vel).astype(np.float32))


def _run(lib, s, d, src):
    ny, nx, nz = s.shape
    if hasattr(lib, "or_podvin"):
        T = np.zeros(s.shape, np.float32)
        lib.or_podvin(P(s), nx, ny, nz, *d, *(float(v) for v in src), P(T))
        return T
    T = np.zeros(s.shape, np.float32)
    sxyz = np.asarray(src, np.float32).reshape(1, 3)
    r = np.zeros(1, np.float32)
    lib.ref_eikonal(0, P(s), nx, ny, nz, *d, P(sxyz), 1, P(r), P(r), P(r), 0, P(T), None, None)
    return T


@pytest.mark.skipif(not orc.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("shape,d,src,seed", [
    ((21, 17, 13), (10.0, 10.0, 10.0), (103.7, 81.2, 44.9), 1),
    ((21, 17, 13), (10.0, 10.0, 10.0), (0.0, 0.0, 0.0), 2),            # corner node
    ((19, 23, 11), (12.5, 12.5, 12.5), (225.0, 0.0, 125.0), 3),        # far x face, on a node
    ((15, 12, 14), (10.0, 8.0, 7.0), (70.0, 40.0, 30.0), 4),           # h = dx for every axis
])
def test_podvin_bit_exact_vs_reference(shape, d, src, seed):
    nx, ny, nz = shape
    s = _case(nx, ny, nz, seed)
    a = _run(orc.oracle(), s, d, src)
    b = _run(orc.ref(), s, d, src)
    assert a.max() < 1e5                                  # every node reached
    assert np.array_equal(a, b)


def test_podvin_homogeneous_properties():
    """Homogeneous medium, source on a node: the source node is 0, the
    axial neighbours are h s (head wave), and times grow with distance."""
    nx = ny = nz = 15
    h, v = 10.0, 2000.0
    s = np.full((ny, nx, nz), 1.0 / v, np.float32)
    T = _run(orc.oracle(), s, (h, h, h), (70.0, 70.0, 70.0))
    assert T[7, 7, 7] == 0.0
    hs = np.float32(h) * np.float32(1.0 / v)
    for a, b, c in ((7, 7, 8), (7, 8, 7), (8, 7, 7), (7, 7, 6)):
        assert abs(T[a, b, c] - hs) <= 1e-7
    ii, jj, kk = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    dist = h * np.sqrt((ii - 7) ** 2 + (jj - 7) ** 2 + (kk - 7) ** 2)
    assert np.abs(T - dist / v).max() < 0.05 * (dist.max() / v)
