"""Seeded synthetic models and acquisition geometry for the benchmark and
test workloads.

Volumes are float32 numpy arrays of shape (ny, nx, nz) in C order, so the
flat index is ``z + x*nz + y*nx*nz`` -- the reference's z-fastest layout
(src/eikonal/block_FIM/block_FIM.cpp:17, src/utils/py_codes/functions.py:23-25
reads them as Fortran-order (nz, nx, ny)).

Workloads (grid, spacing h):

* ``gradient_model`` -- velocity rising linearly with depth from v0 at z=0 to
  v1 at the bottom, times (1 + a*exp(-r^2/(2 sigma^2))) for one Gaussian
  anomaly whose centre is drawn uniformly in the middle 60% of the volume.
  101x101x51 at h=25 m is the single-shot case; 201x201x101 at h=25 m with
  256 shots at depths U(0, 2 km) is the batched case.
* ``layered_model`` -- 3-7 flat layers with velocities drawn in
  [1500, 4500] m/s, sorted increasing with depth, plus 1-4 Gaussian
  anomalies of +-8% (the tomography and surrogate-training recipe).
"""
from __future__ import annotations

import numpy as np


def grid_coords(nx, ny, nz, h):
    z = np.arange(nz, dtype=np.float64) * h
    x = np.arange(nx, dtype=np.float64) * h
    y = np.arange(ny, dtype=np.float64) * h
    return x, y, z


def gaussian_bump(nx, ny, nz, h, centre, sigma):
    x, y, z = grid_coords(nx, ny, nz, h)
    Y, X, Z = np.meshgrid(y, x, z, indexing="ij")
    r2 = (X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2
    return np.exp(-r2 / (2.0 * sigma * sigma))


def gradient_model(nx, ny, nz, h, v0=1500.0, v1=4000.0, amp=-0.10,
                   sigma=250.0, seed=1):
    """Linear-gradient velocity with one seeded Gaussian anomaly (m/s)."""
    rng = np.random.default_rng(seed)
    Lx, Ly, Lz = (nx - 1) * h, (ny - 1) * h, (nz - 1) * h
    centre = (rng.uniform(0.2, 0.8) * Lx, rng.uniform(0.2, 0.8) * Ly,
              rng.uniform(0.2, 0.8) * Lz)
    _, _, z = grid_coords(nx, ny, nz, h)
    v = v0 + (v1 - v0) * z / Lz
    vel = np.broadcast_to(v[None, None, :], (ny, nx, nz)).astype(np.float64)
    if amp != 0.0:
        vel = vel * (1.0 + amp * gaussian_bump(nx, ny, nz, h, centre, sigma))
    return np.ascontiguousarray(vel, np.float32)


def layered_model(nx, ny, nz, h, n_layers=5, n_anom=3, amp=0.08,
                  vmin=1500.0, vmax=4500.0, seed=2):
    """Flat layers with sorted velocities plus +-amp Gaussian anomalies."""
    rng = np.random.default_rng(seed)
    Lx, Ly, Lz = (nx - 1) * h, (ny - 1) * h, (nz - 1) * h
    vels = np.sort(rng.uniform(vmin, vmax, n_layers))
    tops = np.sort(rng.uniform(0.1, 0.9, n_layers - 1)) * Lz
    _, _, z = grid_coords(nx, ny, nz, h)
    v = np.full(nz, vels[0])
    for i, t in enumerate(tops):
        v[z >= t] = vels[i + 1]
    vel = np.broadcast_to(v[None, None, :], (ny, nx, nz)).astype(np.float64)
    for _ in range(n_anom):
        c = (rng.uniform(0.2, 0.8) * Lx, rng.uniform(0.2, 0.8) * Ly,
             rng.uniform(0.2, 0.8) * Lz)
        sig = rng.uniform(0.08, 0.15) * min(Lx, Ly, max(Lz, h))
        sign = rng.choice([-1.0, 1.0])
        vel = vel * (1.0 + sign * amp * gaussian_bump(nx, ny, nz, h, c, sig))
    return np.ascontiguousarray(vel, np.float32)


def random_layered_model(nx, ny, nz, h, seed):
    """Surrogate-training recipe: 3-7 layers, 1-4 anomalies of +-8%."""
    rng = np.random.default_rng(seed)
    n_layers = int(rng.integers(3, 8))
    n_anom = int(rng.integers(1, 5))
    return layered_model(nx, ny, nz, h, n_layers=n_layers, n_anom=n_anom,
                         seed=int(rng.integers(0, 2**31 - 1)))


def lateral_average(vel):
    """1D start model: the mean over x and y of each depth sample."""
    v = vel.astype(np.float64).mean(axis=(0, 1))
    return np.ascontiguousarray(np.broadcast_to(v[None, None, :], vel.shape), np.float32)


def linspace_f32(xi, xf, n):
    """Regular::linspace (src/geometry/regular/regular.cpp:6-27) in float32."""
    xi = np.float32(xi)
    xf = np.float32(xf)
    if n == 0:
        return np.zeros(0, np.float32)
    if n == 1:
        return np.array([xi], np.float32)
    delta = np.float32((xf - xi) / np.float32(n - 1))
    i = np.arange(n, dtype=np.float32)
    return (xi + (delta * i).astype(np.float32)).astype(np.float32)


def regular_geometry(sw, nw, se, n_xline, n_yline, z):
    """Regular::build_geometry (regular.cpp:93-100): x[k + j*n_yline]."""
    xs = linspace_f32(sw[0], se[0], n_xline)
    ys = linspace_f32(sw[1], nw[1], n_yline)
    n = n_xline * n_yline
    x = np.zeros(n, np.float32)
    y = np.zeros(n, np.float32)
    for j in range(n_xline):
        for k in range(n_yline):
            x[k + j * n_yline] = xs[j]
            y[k + j * n_yline] = ys[k]
    zz = np.full(n, np.float32(z), np.float32)
    return x, y, zz


def surface_grid(n_side, lo, hi, z):
    """n_side x n_side regular surface grid over [lo, hi] in x and y."""
    return regular_geometry((lo, lo), (lo, hi), (hi, lo), n_side, n_side, z)


def random_sources(n, nx, ny, h, zmin, zmax, seed):
    """Sources with x, y uniform inside the grid and depth U(zmin, zmax)."""
    rng = np.random.default_rng(seed)
    Lx, Ly = (nx - 1) * h, (ny - 1) * h
    s = np.empty((n, 3), np.float32)
    s[:, 0] = rng.uniform(0.0, Lx, n)
    s[:, 1] = rng.uniform(0.0, Ly, n)
    s[:, 2] = rng.uniform(zmin, zmax, n)
    return s


# --------------------------------------------------------------------------
# named workloads
# --------------------------------------------------------------------------

def single_shot_case():
    """101x101x51, h=25 m, linear 1500->4000 m/s + one -10% anomaly (seed 1),
    one source at the surface centre, 20x20 receivers over 50..2450 m."""
    nx, ny, nz, h = 101, 101, 51, 25.0
    vel = gradient_model(nx, ny, nz, h, seed=1)
    src = np.array([[1250.0, 1250.0, 0.0]], np.float32)
    rx, ry, rz = surface_grid(20, 50.0, 2450.0, 0.0)
    return dict(nx=nx, ny=ny, nz=nz, h=h, vel=vel, src=src, rx=rx, ry=ry, rz=rz)


def batched_case(n_shots=256):
    """201x201x101, h=25 m, the single-shot recipe over a 5 km x 5 km x
    2.5 km volume; n_shots sources at x, y uniform and depth U(0, 2 km)
    (seed 4); 20x20 surface receivers over 50..4950 m."""
    nx, ny, nz, h = 201, 201, 101, 25.0
    vel = gradient_model(nx, ny, nz, h, seed=1)
    src = random_sources(n_shots, nx, ny, h, 0.0, 2000.0, seed=4)
    rx, ry, rz = surface_grid(20, 50.0, 4950.0, 0.0)
    return dict(nx=nx, ny=ny, nz=nz, h=h, vel=vel, src=src, rx=rx, ry=ry, rz=rz)


def tomography_case(nx=201, ny=201, nz=101, h=12.5, n_shot_side=4,
                    n_rec_side=50, seafloor=100.0):
    """One tomography iteration: 5 seeded layers (1500-4500 m/s) plus 3
    anomalies of +-8% (seed 2); start model = lateral average; 4x4 surface
    shots and 50x50 receivers at the seafloor depth; 41x41x21 inversion
    grid at 62.5 m (for the 201x201x101 / 12.5 m grid)."""
    vel_true = layered_model(nx, ny, nz, h, n_layers=5, n_anom=3, seed=2)
    vel0 = lateral_average(vel_true)
    Lx, Ly = (nx - 1) * h, (ny - 1) * h
    sx, sy, sz = regular_geometry((0.1 * Lx, 0.1 * Ly), (0.1 * Lx, 0.9 * Ly),
                                  (0.9 * Lx, 0.1 * Ly), n_shot_side,
                                  n_shot_side, 0.0)
    src = np.stack([sx, sy, sz], axis=1).astype(np.float32)
    rx, ry, rz = regular_geometry((0.05 * Lx, 0.05 * Ly), (0.05 * Lx, 0.95 * Ly),
                                  (0.95 * Lx, 0.05 * Ly), n_rec_side,
                                  n_rec_side, seafloor)
    d_tomo = 5.0 * h
    return dict(nx=nx, ny=ny, nz=nz, h=h, vel_true=vel_true, vel0=vel0,
                src=src, rx=rx, ry=ry, rz=rz, d_tomo=d_tomo)
