"""Pins of the tomography part of the oracle (CPU only).

* or_ray_trace_shot -- homogeneous medium: rays are straight, every step
  moves rayStep along the receiver->source line, the row's sum reproduces
  the path length (+ the reference's extra step on the first entry and the
  finalDist entry); CSR rows have unique ascending columns; the source cell
  carries finalDist; brute force against a direct Python transcription of
  least_squares.cpp:192-280 on a small volume;
* or_tk_stencil     -- binomial closed form (-1)^(k-j) C(k, j);
* or_cgls_f64       -- equals the least-squares solution (numpy lstsq) when
  iterations >= columns on small full-rank systems;
* or_cgls_ref_f32   -- the reference's FP32 recurrence, close to the FP64 one;
* or_gaussian       -- bit-exact against the reference gaussian() built in
  oracle/_ref; kernel sums to one; constant fields invariant;
* or_model_update   -- constant xdm gives constant dm inside the reference's
  margins and zero outside; slowness updated only inside its own margins.
"""
import math

import numpy as np
import pytest

import oracle_lib as orc
from oracle_lib import P


def _trace(T, sp, tg, src, rcv, max_steps=100000):
    o = orc.oracle()
    ny, nx, nz = T.shape
    Tp = np.zeros((ny + 2, nx + 2, nz + 2), np.float32)
    o.or_expand_fdm(P(np.ascontiguousarray(T, np.float32)), nx, ny, nz, P(Tp))
    rcv = np.ascontiguousarray(rcv, np.float32).reshape(-1, 3)
    rx, ry, rz = (np.ascontiguousarray(rcv[:, i]) for i in range(3))
    nr = rcv.shape[0]
    cap = 200000
    rn = np.zeros(nr, np.int32)
    cols = np.zeros(cap, np.int32)
    vals = np.zeros(cap, np.float32)
    ill = np.zeros(T.shape, np.float32)
    (nxt, nyt, nzt), (dxt, dyt, dzt) = tg
    nnz = o.or_ray_trace_shot(P(Tp), nx, ny, nz, *sp, nxt, nyt, nzt, dxt, dyt, dzt,
                              *(float(v) for v in src), P(rx), P(ry), P(rz), nr, max_steps,
                              P(rn), P(cols), P(vals), cap, P(ill))
    return nnz, rn, cols[:max(nnz, 0)], vals[:max(nnz, 0)], ill, Tp


def _trace_literal(Tp, nx, ny, nz, sp, tg, src, rcv):
    """least_squares.cpp:192-280 transcribed in numpy float32 scalars."""
    f = np.float32
    dx, dy, dz = (f(v) for v in sp)
    (nxt, nyt, nzt), (dxt, dyt, dzt) = tg
    dxt, dyt, dzt = f(dxt), f(dyt), f(dzt)
    nxx, nzz = nx + 2, nz + 2
    Tf = Tp.ravel()
    sx, sy, sz = (f(v) for v in src)
    sId = int(sz / dzt) + int(sx / dxt) * nzt + int(sy / dyt) * nxt * nzt
    step = f(0.2) * dz
    rows = []
    for r in rcv:
        xi, yi, zi = f(r[0]), f(r[1]), f(r[2])
        idx = []
        while True:
            i = int(zi / dz) + 1
            j = int(xi / dx) + 1
            k = int(yi / dy) + 1
            dTz = f(Tf[(i + 1) + j * nzz + k * nxx * nzz] - Tf[(i - 1) + j * nzz + k * nxx * nzz]) / (f(2.0) * dz)
            dTx = f(Tf[i + (j + 1) * nzz + k * nxx * nzz] - Tf[i + (j - 1) * nzz + k * nxx * nzz]) / (f(2.0) * dx)
            dTy = f(Tf[i + j * nzz + (k + 1) * nxx * nzz] - Tf[i + j * nzz + (k - 1) * nxx * nzz]) / (f(2.0) * dy)
            norm = np.sqrt(f(f(dTx * dTx) + f(dTy * dTy)) + f(dTz * dTz), dtype=np.float32)
            zi = f(zi - f(f(step * dTz) / norm))
            xi = f(xi - f(f(step * dTx) / norm))
            yi = f(yi - f(f(step * dTy) / norm))
            idx.append(int(zi / dzt) + int(xi / dxt) * nzt + int(yi / dyt) * nxt * nzt)
            if idx[-1] == sId:
                break
        fd = np.sqrt(f(f(f((zi - sz) * (zi - sz)) + f((xi - sx) * (xi - sx))) + f((yi - sy) * (yi - sy))),
                     dtype=np.float32)
        idx.sort()
        cur, dist, row = idx[0], step, []
        for v in idx:
            if v == cur:
                dist = f(dist + step)
            else:
                row.append((cur, fd if cur == sId else dist))
                dist, cur = step, v
        row.append((cur, fd if cur == sId else dist))
        rows.append(row)
    return rows


def _homog(nx=25, ny=23, nz=15, h=10.0, v=2000.0, src=(120.0, 110.0, 0.0)):
    y, x, z = np.meshgrid(np.arange(ny) * h, np.arange(nx) * h, np.arange(nz) * h, indexing="ij")
    T = (np.sqrt((x - src[0]) ** 2 + (y - src[1]) ** 2 + (z - src[2]) ** 2) / v).astype(np.float32)
    return T


def test_raytrace_matches_literal_transcription():
    h = 10.0
    src = (120.0, 110.0, 0.0)
    T = _homog(src=src)
    ny, nx, nz = T.shape
    tg = ((int((nx - 1) * h / 40) + 1, int((ny - 1) * h / 40) + 1, int((nz - 1) * h / 40) + 1),
          (40.0, 40.0, 40.0))
    rng = np.random.default_rng(2)
    rcv = np.stack([rng.uniform(5, 235, 12), rng.uniform(5, 215, 12), rng.uniform(20, 130, 12)], 1)
    nnz, rn, cols, vals, _, Tp = _trace(T, (h, h, h), tg, src, rcv)
    assert nnz > 0
    lit = _trace_literal(Tp, nx, ny, nz, (h, h, h), tg, src, rcv.astype(np.float32))
    o = 0
    for r, row in enumerate(lit):
        assert rn[r] == len(row)
        for c, v in row:
            assert cols[o] == c and vals[o] == np.float32(v)
            o += 1


def test_raytrace_homogeneous_straight_rays():
    h, v = 10.0, 2000.0
    src = (120.0, 110.0, 0.0)
    T = _homog(src=src, v=v)
    ny, nx, nz = T.shape
    tg = ((int((nx - 1) * h / 40) + 1, int((ny - 1) * h / 40) + 1, int((nz - 1) * h / 40) + 1),
          (40.0, 40.0, 40.0))
    rcv = np.array([[10.0, 20.0, 100.0], [230.0, 200.0, 120.0], [120.0, 110.0, 140.0]], np.float32)
    nnz, rn, cols, vals, ill, _ = _trace(T, (h, h, h), tg, src, rcv)
    step = 0.2 * h
    nzt, nxt = tg[0][2], tg[0][0]
    sId = int(src[2] / 40) + int(src[0] / 40) * nzt + int(src[1] / 40) * nxt * nzt
    o = 0
    for r in range(3):
        c = cols[o:o + rn[r]]
        w = vals[o:o + rn[r]]
        assert np.all(np.diff(c) > 0)  # unique, ascending: CSR order
        assert sId in c
        L = math.dist(rcv[r], (src[0], src[1], src[2]))
        # path length: every entry is a multiple of the step except the
        # source cell (finalDist) and the first entry carries one extra step
        assert abs(w.sum() - L) < 3 * step + 0.5 * 40
        o += rn[r]
    assert ill.sum() > 0


def test_raytrace_flat_field_diverges():
    T = np.ones((8, 9, 7), np.float32)
    tg = ((3, 3, 2), (40.0, 40.0, 40.0))
    nnz, *_ = _trace(T, (10.0, 10.0, 10.0), tg, (40.0, 40.0, 0.0), np.array([[10, 10, 30]], np.float32))
    assert nnz == -1


def test_tk_stencil_binomial():
    o = orc.oracle()
    for k in range(0, 6):
        df = np.zeros(k + 1, np.int32)
        o.or_tk_stencil(k, P(df))
        if k == 0:
            assert df[0] == 1
        else:
            assert list(df) == [(-1) ** (k - j) * math.comb(k, j) for j in range(k + 1)]


def _dense_to_csr(A):
    rp = [0]
    c, v = [], []
    for row in A:
        nzs = np.nonzero(row)[0]
        c.extend(nzs.tolist())
        v.extend(row[nzs].tolist())
        rp.append(len(c))
    return np.array(rp, np.int64), np.array(c, np.int32), np.array(v, np.float32)


def test_cgls_f64_equals_lstsq():
    o = orc.oracle()
    rng = np.random.default_rng(0)
    N, M = 40, 12
    A = rng.normal(size=(N, M)).astype(np.float32)
    A[rng.random((N, M)) < 0.3] = 0
    b = rng.normal(size=N).astype(np.float32)
    rp, c, v = _dense_to_csr(A)
    x = np.zeros(M, np.float64)
    o.or_cgls_f64(P(rp), P(c), P(v), N, M, P(b), M + 5, P(x))
    ref = np.linalg.lstsq(A.astype(np.float64), b.astype(np.float64), rcond=None)[0]
    assert np.allclose(x, ref, rtol=1e-8, atol=1e-10)
    xf = np.zeros(M, np.float32)
    o.or_cgls_ref_f32(P(rp), P(c), P(v), N, M, P(b), 8, P(xf))
    x8 = np.zeros(M, np.float64)
    o.or_cgls_f64(P(rp), P(c), P(v), N, M, P(b), 8, P(x8))
    assert np.allclose(xf, x8, rtol=1e-3, atol=1e-4)


@pytest.mark.skipif(not orc.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("samples,stdv", [(5, 2.0), (3, 0.7), (7, 1.5), (5, 0.0)])
def test_gaussian_bit_exact_vs_reference(samples, stdv):
    rng = np.random.default_rng(4)
    a = rng.normal(size=(14, 13, 12)).astype(np.float32)
    out_o = np.zeros_like(a)
    out_r = np.zeros_like(a)
    orc.oracle().or_gaussian(P(a), P(out_o), 13, 14, 12, samples, stdv)
    orc.ref().ref_gaussian(P(a.copy()), P(out_r), 13, 14, 12, samples, stdv)
    assert np.array_equal(out_o, out_r)


def test_gaussian_kernel_and_constant_field():
    o = orc.oracle()
    k = np.zeros(125, np.float32)
    o.or_gaussian_kernel(5, 2.0, P(k))
    assert abs(k.sum() - 1.0) < 1e-6
    a = np.full((10, 11, 9), 3.25, np.float32)
    out = np.zeros_like(a)
    o.or_gaussian(P(a), P(out), 11, 10, 9, 5, 2.0)
    assert np.allclose(out, 3.25, atol=1e-5)


def test_model_update_margins():
    o = orc.oracle()
    nx, ny, nz, h = 40, 38, 20, 10.0
    dxt = 50.0
    nxt, nyt, nzt = (int((n - 1) * h / dxt) + 1 for n in (nx, ny, nz))
    slow = np.full((ny, nx, nz), 1 / 2000.0, np.float32)
    xdm = np.full(nxt * nyt * nzt, 1e-5, np.float32)
    dm = np.zeros_like(slow)
    vel = np.zeros_like(slow)
    s0 = slow.copy()
    o.or_model_update(P(slow), P(xdm), nx, ny, nz, h, h, h, nxt, nyt, nzt, dxt, dxt, dxt, 0, 5, 2.0,
                      P(dm), P(vel))
    y, x, z = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    inside = (z >= 11) & (z < nz - 2) & (x >= 15) & (x < nx - 16) & (y >= 15) & (y < ny - 16)
    assert np.allclose(dm[inside], 1e-5, rtol=1e-6) and np.all(dm[~inside] == 0)
    upd = (z >= 11) & (z < nz - 2) & (x >= 15) & (x < nx - 15) & (y >= 15) & (y < ny - 15)
    assert np.all(slow[~upd] == s0[~upd])
    assert np.allclose(slow[inside], s0[inside] + 1e-5, rtol=1e-6)
    assert np.array_equal(vel, (1.0 / slow).astype(np.float32))


def test_residual_norm():
    o = orc.oracle()
    a = np.array([1.0, 2.0, 3.0], np.float32)
    b = np.array([1.5, 2.0, 1.0], np.float32)
    assert o.or_residual_norm(P(a), P(b), 3) == pytest.approx(math.sqrt(0.25 + 4.0), rel=1e-6)
