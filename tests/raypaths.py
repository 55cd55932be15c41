"""Ray paths whose sensitivity rows are known in closed form.

Each case is a travel-time field T on a grid with dz = 10 m, so the
reference's ray step (0.2 * dz, least_squares.cpp:203) is exactly 2.0 m,
built from small integers so that every central difference is exact and
the off-path components of the gradient are exactly zero: each step then
moves the ray by exactly 2 m along one axis and every position is an
integer.  From the geometry alone the visited tomography cells, their
counts, the sorted run-length values (the first sorted run n + 1 steps,
later runs n steps, the source cell the straight remainder finalDist,
least_squares.cpp:241-276) and the illumination additions (rayStep at the
pre-step node, indexed with the unpadded strides, :234) follow; the cases
below write them down as literals.

  z_up      receiver 260 m deeper than the source on one vertical line (the
            ray climbs): the source cell sorts first and its run becomes
            finalDist
  z_down    receiver 260 m shallower: the receiver's cell sorts first and its
            run carries the extra step; the source cell is the last run
  x_axis    dx = 20 m: the step still follows dz (2 m), nodes change every
            10 steps
  y_axis    dy = 12.5 m
  u_turn    a path that leaves a tomography cell and comes back: the two
            visits merge into one run after the sort
"""
import numpy as np

STEP = 2.0


def _grid(nx, ny, nz):
    z, x, y = np.meshgrid(np.arange(nz), np.arange(nx), np.arange(ny), indexing="ij")
    # arrays in the volume layout (ny, nx, nz), z fastest
    return x.transpose(2, 1, 0), y.transpose(2, 1, 0), z.transpose(2, 1, 0)


def _illum(shape, pre_steps, d):
    """rayStep at the padded pre-step node (i, j, k) = (z/dz, x/dx, y/dy) + 1,
    at the flat index i + j nz + k nx nz of the unpadded volume."""
    ny, nx, nz = shape
    ill = np.zeros(nx * ny * nz, np.float32)
    dx, dy, dz = d
    for (x, y, z) in pre_steps:
        i, j, k = int(z // dz) + 1, int(x // dx) + 1, int(y // dy) + 1
        ill[i + j * nz + k * nx * nz] += np.float32(STEP)
    return ill.reshape(shape)


def _line(start, stop, axis, fixed):
    """Pre-step positions from start towards stop in steps of 2 m along one
    axis (stop excluded), the other coordinates fixed."""
    s = -STEP if stop < start else STEP
    out = []
    for v in np.arange(start, stop, s):
        p = list(fixed)
        p[axis] = float(v)
        out.append(tuple(p))
    return out


def cases():
    out = {}
    # ---- vertical line, receiver deeper than the source (z is depth)
    nx, ny, nz, d = 11, 11, 31, (10.0, 10.0, 10.0)
    x, y, z = _grid(nx, ny, nz)
    T = np.abs(z * 10.0 - 20.0).astype(np.float32)
    # tomography 50 m: nzt 7, nxt 3, nyt 3; cell = im + 7 jm + 21 km; jm = km = 1
    # path z = 280 -> 48: im 5 x15, 4..1 x25, 0 (the source cell) x1
    out["z_up"] = dict(T=T, d=d, dt=50.0, src=(50.0, 50.0, 20.0), rcv=(50.0, 50.0, 280.0),
                       cols=[28, 29, 30, 31, 32, 33],
                       vals=[28.0, 50.0, 50.0, 50.0, 50.0, 30.0],
                       illum=_illum(T.shape, _line(280.0, 48.0, 2, (50.0, 50.0, 0.0)), d))
    # ---- vertical line, receiver at 20 m, source at 280 m
    T = np.abs(z * 10.0 - 280.0).astype(np.float32)
    # path z = 20 -> 250: im 0 x14 (sorts first: (14 + 1) steps), 1..4 x25,
    # 5 (source cell) x1 with finalDist |250 - 280|
    out["z_down"] = dict(T=T, d=d, dt=50.0, src=(50.0, 50.0, 280.0), rcv=(50.0, 50.0, 20.0),
                         cols=[28, 29, 30, 31, 32, 33],
                         vals=[30.0, 50.0, 50.0, 50.0, 50.0, 30.0],
                         illum=_illum(T.shape, _line(20.0, 250.0, 2, (50.0, 50.0, 0.0)), d))
    # ---- x line, dx = 20 m
    nx, ny, nz, d = 16, 11, 11, (20.0, 10.0, 10.0)
    x, y, z = _grid(nx, ny, nz)
    T = np.abs(x * 20.0 - 20.0).astype(np.float32)
    # tomography 50 m: nzt 3, nxt 7, nyt 3; cell = im + 3 jm + 21 km; im = km = 1
    out["x_axis"] = dict(T=T, d=d, dt=50.0, src=(20.0, 50.0, 50.0), rcv=(280.0, 50.0, 50.0),
                         cols=[22, 25, 28, 31, 34, 37],
                         vals=[28.0, 50.0, 50.0, 50.0, 50.0, 30.0],
                         illum=_illum(T.shape, _line(280.0, 48.0, 0, (0.0, 50.0, 50.0)), d))
    # ---- y line, dy = 12.5 m
    nx, ny, nz, d = 11, 25, 11, (10.0, 12.5, 10.0)
    x, y, z = _grid(nx, ny, nz)
    T = np.abs(y * 12.5 - 20.0).astype(np.float32)
    # tomography 50 m: nzt 3, nxt 3, nyt 7; cell = im + 3 jm + 9 km; im = jm = 1
    out["y_axis"] = dict(T=T, d=d, dt=50.0, src=(50.0, 20.0, 50.0), rcv=(50.0, 280.0, 50.0),
                         cols=[4, 13, 22, 31, 40, 49],
                         vals=[28.0, 50.0, 50.0, 50.0, 50.0, 30.0],
                         illum=_illum(T.shape, _line(280.0, 48.0, 1, (50.0, 0.0, 50.0)), d))
    # ---- U turn: +x along z = 282 to x = 100, -z to 278, -x to 98, then -z
    # down the column x in [90, 100) to the source cell.  T is independent of
    # y and integer valued.  At each visited node (a = x / 10, b = z / 10)
    # exactly one central difference is non-zero: f[a+1, 28] < f[a-1, 28]
    # for a = 6..9 (the ray moves +x), f[10, 29] > f[10, 27] (-z),
    # f[11, 27] > f[9, 27] (-x) and f[9, b+1] > f[9, b-1] down column 9 (-z),
    # while the other axis' neighbours are equal
    nx, ny, nz, d = 16, 11, 31, (10.0, 10.0, 10.0)
    f = np.zeros((nx, nz), np.float64)       # f[a, b]
    f[8, 4:28] = 1000.0
    f[10, 4:28] = 1000.0
    f[9, :28] = 10.0 * np.arange(28)         # rises with depth index: ray goes up
    f[9, 28] = 300.0
    f[9, 29] = 270.0
    f[8, 28], f[8, 29] = 2000.0, 1000.0
    f[6, 28], f[5, 28], f[7, 28] = 3000.0, 4000.0, 400.0
    f[10, 28], f[10, 29] = 1000.0, 1500.0
    f[11, 28], f[11, 27] = 300.0, 500.0
    T = np.ascontiguousarray(np.broadcast_to(f[None, :, :], (ny, nx, nz)).astype(np.float32))
    pre = (_line(62.0, 100.0, 0, (0.0, 50.0, 282.0)) + [(100.0, 50.0, 282.0), (100.0, 50.0, 280.0),
                                                       (100.0, 50.0, 278.0)]
           + _line(278.0, 48.0, 2, (98.0, 50.0, 0.0)))
    # tomography 50 m: nzt 7, nxt 4, nyt 3; cell = im + 7 jm + 28 km; km = 1.
    # (im 5, jm 1) is visited 18 times on the way out and 15 on the way down
    # (33 steps, one run after the sort); (im 5, jm 2) 3 times; the column
    # im 4..1 25 each; the source cell (im 0, jm 1) once, finalDist =
    # sqrt((48 - 20)^2 + (98 - 95)^2) = sqrt(793)
    out["u_turn"] = dict(T=T, d=d, dt=50.0, src=(95.0, 50.0, 20.0), rcv=(62.0, 50.0, 282.0),
                         cols=[35, 36, 37, 38, 39, 40, 47],
                         vals=[float(np.sqrt(np.float32(793.0), dtype=np.float32)), 50.0, 50.0,
                               50.0, 50.0, 66.0, 6.0],
                         illum=_illum(T.shape, pre, d))
    for c in out.values():
        ny, nx, nz = c["T"].shape
        dx, dy, dz = c["d"]
        c["nt"] = (int((nx - 1) * dx / c["dt"]) + 1, int((ny - 1) * dy / c["dt"]) + 1,
                   int((nz - 1) * dz / c["dt"]) + 1)
    return out
