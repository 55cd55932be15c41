"""st_fim_solve_noalloc (no device allocation: all state in the library's
static scratch array) against st_fim_solve on the same inputs, through the
C ABI: picks within the 1e-4 s gate of each other and of the FP64 fixed
point, on two calls in a row (the second with fewer shots and receivers,
so stale state from the first would show), with both visit orderings, and
the size limit reported as an argument error without touching the GPU."""
import numpy as np
import pytest

import oracle_lib as orc
from seismic_tomography_3d_amd import Context, synth
from seismic_tomography_3d_amd.api import StError

pytestmark = pytest.mark.gpu
TOL = 1e-4


def test_noalloc_matches_fim_solve_twice():
    vel = synth.layered_model(45, 41, 29, 25.0, n_layers=4, n_anom=2, seed=9)
    h = 25.0
    srcs = np.array([[500.0, 437.3, 110.0], [0.0, 0.0, 0.0], [1100.0, 1000.0, 700.0],
                     [12.5, 880.0, 460.0], [600.0, 20.0, 0.0], [300.0, 300.0, 300.0]], np.float32)
    rcv = np.stack(synth.surface_grid(6, 25.0, 1000.0, 0.0), axis=1)
    with Context(0) as ctx:
        ctx.set_model(1.0 / vel, h)
        for call, (ns, nr, order) in enumerate([(6, 36, 0), (3, 20, 1)]):
            s, r = srcs[:ns], rcv[:nr]
            ref, _, _ = ctx.fim_solve(s, r, order=order)
            got, st = ctx.fim_solve_noalloc(s, r, order=order, batch=4)
            print(f"call {call}: {ns} shots, {nr} receivers, order {order}, stats {st}")
            assert got.shape == (ns, nr)
            assert np.abs(got - ref).max() < TOL
            for i in range(ns):
                T64 = orc.eikonal_fp64(vel, h, s[i])
                p64 = orc.picks(T64.astype(np.float32), (h, h, h), r)
                assert np.abs(got[i].astype(np.float64) - p64).max() < TOL


def test_noalloc_rejects_oversized_model_and_warm():
    with Context(0) as ctx:
        vel = np.full((401, 401, 201), 2000.0, np.float32)
        ctx.set_model(1.0 / vel, 10.0)
        src = np.array([[100.0, 100.0, 0.0]], np.float32)
        with pytest.raises(StError) as e:
            ctx.fim_solve_noalloc(src, src)
        assert e.value.code == 1 and "scratch" in str(e.value)
        with pytest.raises(StError) as e:
            ctx.fim_solve_noalloc(src, src, warm=1)
        assert e.value.code == 1
