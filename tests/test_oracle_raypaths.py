"""The oracle's ray tracer (or_ray_trace_shot, the restatement of
least_squares.cpp:192-280) against sensitivity rows known in closed form
(tests/raypaths.py): axis-aligned rays along z (source cell sorted first
and last), x (dx = 20 m) and y (dy = 12.5 m), and a U-shaped path that
visits one tomography cell twice.  Every column, every value and every
illumination cell must equal the literal.

Plausible misreadings of the reference each fail one case: dropping the
first run's extra step (z_down), keeping the source cell's own count
instead of finalDist (z_up, z_down), stepping by 0.2 dx (x_axis), the
unpadded/padded illumination index (all), not merging a cell's two visits
(u_turn)."""
import numpy as np
import pytest

import oracle_lib as orc
import raypaths
from oracle_lib import P


def oracle_trace(c):
    T = np.ascontiguousarray(c["T"])
    ny, nx, nz = T.shape
    o = orc.oracle()
    Tp = np.zeros((ny + 2, nx + 2, nz + 2), np.float32)
    o.or_expand_fdm(P(T), nx, ny, nz, P(Tp))
    rx, ry, rz = (np.array([v], np.float32) for v in c["rcv"])
    rn = np.zeros(1, np.int32)
    cols = np.zeros(4096, np.int32)
    vals = np.zeros(4096, np.float32)
    ill = np.zeros(T.shape, np.float32)
    nxt, nyt, nzt = c["nt"]
    dt = c["dt"]
    nnz = o.or_ray_trace_shot(P(Tp), nx, ny, nz, *c["d"], nxt, nyt, nzt, dt, dt, dt,
                              *c["src"], P(rx), P(ry), P(rz), 1, 100000, P(rn), P(cols),
                              P(vals), 4096, P(ill))
    assert nnz == rn[0]
    return cols[:nnz], vals[:nnz], ill


@pytest.mark.parametrize("name", ["z_up", "z_down", "x_axis", "y_axis", "u_turn"])
def test_oracle_rows_match_closed_form(name):
    c = raypaths.cases()[name]
    cols, vals, ill = oracle_trace(c)
    assert cols.tolist() == c["cols"], (name, cols.tolist())
    assert np.array_equal(vals, np.array(c["vals"], np.float32)), (name, vals.tolist())
    assert np.array_equal(ill, c["illum"]), name
