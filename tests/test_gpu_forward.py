"""The tomography forward pass (st_forward / st_forward_times) with each of
the reference's eikonal types, held bit for bit against the oracle on the
volumes the pass itself solved.

Least_squares::forward_modeling (least_squares.cpp:163-190) solves every
shot with the eikonal the parameter file selects (eikonal_type,
least_squares.cpp:85-104; the reference's run/parameters.txt uses 2,
Accurate_FSM), picks the receivers and traces the rays on that shot's
travel times.  Bars:
  * block FIM (not reproducible run to run: its visit order depends on
    device timing): dcal and every CSR row bit-exact against the oracle's
    picks and ray tracer on the travel times the same call returned;
  * Accurate_FSM and Podvin (deterministic, bit-identical to the CPU
    code): travel times and dcal equal to the oracle's solver, and every
    CSR row equal to the oracle's rays on the oracle's own volumes -- the
    reference's result for the same inputs;
  * illumination: a sum of FP32 additions made by many rays in a
    device-dependent order, within 1e-5 of its largest value.
Also: the solver cache follows the model through updates, the round-graph
cache follows the geometry, and the failure paths return their codes.
"""
import numpy as np
import pytest

import oracle_lib as orc
from oracle_lib import P
from seismic_tomography_3d_amd import synth
from seismic_tomography_3d_amd.api import StError

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _fim_forward_after(ctx):
    """The session context goes back to the block FIM forward after each
    test, whatever the test selected."""
    yield
    ctx.tomo_set_eikonal(1)


def _case(nx=49, ny=45, nz=25, h=12.5, seed=5, dy=None, dz=None):
    vel = synth.layered_model(nx, ny, nz, h, n_layers=4, n_anom=2, seed=seed)
    dy = h if dy is None else dy
    dz = h if dz is None else dz
    Lx, Ly = (nx - 1) * h, (ny - 1) * dy
    src = np.array([[0.3 * Lx, 0.4 * Ly, 0.0], [0.7 * Lx, 0.6 * Ly, 0.0],
                    [0.5 * Lx, 0.5 * Ly, 0.4 * (nz - 1) * dz]], np.float32)
    rx, ry, rz = synth.regular_geometry((25.0, 25.0), (25.0, Ly - 25), (Lx - 25, 25.0), 7, 6, 50.0)
    return vel, (h, dy, dz), src, np.stack([rx, ry, rz], 1)


def _oracle_rows(T, d, tg, src, rcv):
    """CSR rows (lengths, columns, values) and illumination of one shot by
    the oracle's ray tracer on the unpadded volume T."""
    o = orc.oracle()
    ny, nx, nz = T.shape
    Tp = np.zeros((ny + 2, nx + 2, nz + 2), np.float32)
    o.or_expand_fdm(P(np.ascontiguousarray(T)), nx, ny, nz, P(Tp))
    rx, ry, rz = (np.ascontiguousarray(rcv[:, i]) for i in range(3))
    cap = 1 << 21
    rn = np.zeros(len(rcv), np.int32)
    cols = np.zeros(cap, np.int32)
    vals = np.zeros(cap, np.float32)
    ill = np.zeros(T.shape, np.float32)
    (nxt, nyt, nzt), dt = tg
    nnz = o.or_ray_trace_shot(P(Tp), nx, ny, nz, *d, nxt, nyt, nzt, dt, dt, dt,
                              *(float(v) for v in src), P(rx), P(ry), P(rz), len(rcv), 1 << 20,
                              P(rn), P(cols), P(vals), cap, P(ill))
    assert nnz > 0
    return rn, cols[:nnz], vals[:nnz], ill


def _check_G(ctx, Ts, d, tg, src, rcv, ill):
    """Every row of the device G equals the oracle's row on Ts[shot]."""
    rp, c, v = ctx.csr()
    nr = len(rcv)
    assert len(rp) == len(src) * nr + 1
    assert ctx.csr_size() == (len(rp) - 1, int(rp[-1]))
    oill = np.zeros(Ts.shape[1:], np.float32)
    for i in range(len(src)):
        rn, oc, ov, oi = _oracle_rows(Ts[i], d, tg, src[i], rcv)
        oill += oi
        assert np.array_equal(np.diff(rp[i * nr:(i + 1) * nr + 1]), rn), i
        lo, hi = rp[i * nr], rp[(i + 1) * nr]
        assert np.array_equal(c[lo:hi], oc), i
        assert np.array_equal(v[lo:hi], ov), i
    assert np.abs(ill - oill).max() <= 1e-5 * oill.max()


def test_forward_fim_rows_bit_exact_on_own_volumes(ctx):
    vel, d, src, rcv = _case()
    ctx.set_model(1.0 / vel, d[0])
    ctx.tomo_set_eikonal(1)
    dt = 5 * d[0]
    nt = ctx.tomo_set_grid(dt)
    dcal, ill, st, T = ctx.forward(src, rcv, illum=True, return_T=True, batch=2)
    print(f"FIM forward: {st['visits']} visits, {st['batches']} batches", flush=True)
    assert st["batches"] == 2
    for i in range(len(src)):
        assert np.array_equal(dcal[i], orc.picks(T[i], d, rcv)), i
    _check_G(ctx, T, d, (nt, dt), src, rcv, ill)
    # the travel times reach the FP64 fixed point of the same scheme
    T64 = orc.eikonal_fp64(vel, d[0], src[2])
    assert np.abs(T[2] - T64).max() < 1e-4


@pytest.mark.parametrize("spacing", [(12.5, 12.5, 12.5), (12.5, 15.0, 10.0)])
def test_forward_accurate_fsm_is_the_reference_path(ctx, spacing):
    """eikonal_type 2 (the reference's default): picks, travel times and G
    equal what the reference's Accurate_FSM + ray tracing give."""
    vel, d, src, rcv = _case(h=spacing[0], dy=spacing[1], dz=spacing[2], seed=7)
    s = np.ascontiguousarray((1.0 / vel).astype(np.float32))
    ctx.set_model(s, *d)
    ctx.tomo_set_eikonal(2)
    dt = 62.5
    nt = ctx.tomo_set_grid(dt)
    dcal, ill, st, T = ctx.forward(src, rcv, illum=True, return_T=True)
    print(f"FSM forward {spacing}: device {st['solve_ms']:.2f} ms", flush=True)
    ny, nx, nz = s.shape
    To = np.empty((len(src),) + s.shape, np.float32)
    for i, p in enumerate(src):
        orc.oracle().or_fsm(P(s), nx, ny, nz, *d, *(float(v) for v in p), P(To[i]))
        assert np.array_equal(T[i], To[i]), i
        assert np.array_equal(dcal[i], orc.picks(To[i], d, rcv)), i
    _check_G(ctx, To, d, (nt, dt), src, rcv, ill)
    # the same pass through a second call: deterministic, so identical
    dcal2, _, _ = ctx.forward(src, rcv)
    assert np.array_equal(dcal2, dcal)


def test_forward_podvin_is_the_reference_path(ctx):
    vel, d, src, rcv = _case(nx=31, ny=27, nz=17, h=20.0, seed=3)
    s = np.ascontiguousarray((1.0 / vel).astype(np.float32))
    ctx.set_model(s, *d)
    ctx.tomo_set_eikonal(0)
    dt = 60.0
    nt = ctx.tomo_set_grid(dt)
    dcal, ill, st, T = ctx.forward(src, rcv, illum=True, return_T=True)
    ny, nx, nz = s.shape
    To = np.empty((len(src),) + s.shape, np.float32)
    for i, p in enumerate(src):
        orc.oracle().or_podvin(P(s), nx, ny, nz, *d, *(float(v) for v in p), P(To[i]))
        assert np.array_equal(T[i], To[i]), i
        assert np.array_equal(dcal[i], orc.picks(To[i], d, rcv)), i
    _check_G(ctx, To, d, (nt, dt), src, rcv, ill)


def test_eikonal_type_default_branch_and_warm_refusal(ctx):
    """An eikonal_type outside 0-2 is Accurate_FSM (least_squares.cpp:97-103);
    warm starts belong to the block FIM."""
    vel, d, src, rcv = _case(nx=33, ny=29, nz=17, h=20.0, seed=4)
    ctx.set_model(1.0 / vel, d[0])
    ctx.tomo_set_grid(60.0)
    ctx.tomo_set_eikonal(2)
    d2, _, _ = ctx.forward(src, rcv)
    ctx.tomo_set_eikonal(7)
    d7, _, _ = ctx.forward(src, rcv)
    assert np.array_equal(d2, d7)
    with pytest.raises(StError) as e:
        ctx.forward(src, rcv, warm=1)
    assert e.value.code == 1
    ctx.tomo_set_eikonal(1)


def test_model_update_refreshes_every_solver(ctx):
    """After st_model_update / st_optimize_update, the sweeping solvers and
    the FIM solve the updated model: their results equal solves on a fresh
    context loaded with st_get_model."""
    from seismic_tomography_3d_amd import Context
    vel, d, src, rcv = _case(nx=41, ny=37, nz=21, h=20.0, seed=6)
    # the update is confined to 15 <= x, y cells < n - 16 and 11 <= z < nz - 2
    # (least_squares.cpp:520,559): a source inside that region sees it
    src[0] = [20 * d[0], 18 * d[0], 15 * d[0]]
    ctx.set_model(1.0 / vel, d[0])
    ctx.tomo_set_eikonal(2)
    ctx.tomo_set_grid(60.0)
    before, _, _ = ctx.fsm_solve(src, rcv)      # caches the padded slowness
    pod_before, _, _ = ctx.podvin_solve(src[:1], rcv)
    dcal, _, _ = ctx.forward(src, rcv)
    dobs = (dcal + np.float32(0.02)).astype(np.float32)   # every ray 20 ms late
    ctx.optimize_update(dobs, dcal, 2, 1e-2, 10, smooth=True, samples=5, stdv=2.0, graph=True,
                        fetch=False)
    s_new = ctx.get_model()
    assert not np.array_equal(s_new, (1.0 / vel).astype(np.float32))
    after, _, _ = ctx.fsm_solve(src, rcv)
    pod_after, _, _ = ctx.podvin_solve(src[:1], rcv)
    fim_after, _, _ = ctx.fim_solve(src, rcv, eps=0.0)
    with Context(0) as fresh:
        fresh.set_model(s_new, d[0])
        ref, _, _ = fresh.fsm_solve(src, rcv)
        pod_ref, _, _ = fresh.podvin_solve(src[:1], rcv)
        fim_ref, _, _ = fresh.fim_solve(src, rcv, eps=0.0)
    assert not np.array_equal(after, before)
    assert np.array_equal(after, ref)
    assert not np.array_equal(pod_after, pod_before)
    assert np.array_equal(pod_after, pod_ref)
    assert np.abs(fim_after - fim_ref).max() < 1e-5
    # st_model_update with a caller solution too
    nt = ctx.tomo_dims()
    ctx.model_update(np.full(nt[0] * nt[1] * nt[2], 1e-6, np.float32))
    s2 = ctx.get_model()
    after2, _, _ = ctx.fsm_solve(src, rcv)
    with Context(0) as fresh:
        fresh.set_model(s2, d[0])
        ref2, _, _ = fresh.fsm_solve(src, rcv)
    assert np.array_equal(after2, ref2)
    ctx.tomo_set_eikonal(1)


def test_round_graph_follows_geometry(ctx):
    """With use_graph = 1 the cached round graph is re-captured when the
    geometry changes, even where the block count stays: nx 41 and 45 both
    give 6 blocks of 8 along x but pad-to-4 extents of 44 and 48."""
    h = 25.0
    for nx in (41, 45, 41):
        vel = synth.gradient_model(nx, 33, 21, h, seed=1)
        ctx.set_model(1.0 / vel, h)
        L = (nx - 1) * h
        src = np.array([[0.9 * L, 400.0, 0.0], [L, 10.0, 300.0]], np.float32)
        rcv = np.array([[L, 800.0, 0.0], [0.5 * L, 100.0, 0.0], [L, 0.0, 500.0]], np.float32)
        pe, _, _ = ctx.fim_solve(src, rcv, eps=0.0, use_graph=0)
        pg, _, _ = ctx.fim_solve(src, rcv, eps=0.0, use_graph=1)
        print(f"nx {nx}: eager {pe.ravel()} graph {pg.ravel()}", flush=True)
        assert np.abs(pg - pe).max() < 1e-5, nx
        assert pg.max() < 1e5, nx


def test_failure_paths_return_codes(ctx):
    """The C boundary's error paths, forced from the host: every kernel
    exits normally."""
    nx, ny, nz, h = 21, 19, 13, 25.0
    ok = np.full((ny, nx, nz), 1.0 / 2000.0, np.float32)
    for bad in (np.nan, np.inf, 0.0, -1e-4):
        s = ok.copy()
        s[3, 4, 5] = bad
        with pytest.raises(StError) as e:
            ctx.set_model(s, h)
        assert e.value.code == 1, bad
    ctx.set_model(ok, h)
    ctx.tomo_set_grid(50.0)
    src = np.array([250.0, 225.0, 0.0], np.float32)
    rcv = np.array([[400.0, 300.0, 0.0], [100.0, 50.0, 100.0]], np.float32)
    # a flat travel-time field has no gradient: the tracer's own check ends
    # the ray (the reference's loop would never end, least_squares.cpp:213-237)
    flat = np.full((ny, nx, nz), 0.5, np.float32)
    ctx.csr_reset()
    with pytest.raises(StError) as e:
        ctx.raytrace_volume(flat, src, rcv)
    assert e.value.code == 4
    # a ray that cannot reach the source cell within the step cap: a field
    # rising away from the source's opposite corner leads rays away from it
    ctx.tomo_set_grid(50.0, max_steps=40)
    z, x, y = np.meshgrid(np.arange(nz), np.arange(nx), np.arange(ny), indexing="ij")
    ramp = np.ascontiguousarray(((x + y + z) * h * 5e-4).transpose(2, 1, 0).astype(np.float32))
    far = np.array([(nx - 1) * h, (ny - 1) * h, (nz - 1) * h], np.float32)
    ctx.csr_reset()
    with pytest.raises(StError) as e:
        ctx.raytrace_volume(ramp, far, np.array([[0.0, 0.0, 0.0]], np.float32))
    assert e.value.code == 4
    ctx.tomo_set_grid(50.0)
    # an update that drives the slowness negative is refused at the next
    # solve (the update region is 15 <= x, y < n - 16, 11 <= z < nz - 2)
    big = np.full((40, 40, 20), 1.0 / 2000.0, np.float32)
    ctx.set_model(big, h)
    nt = ctx.tomo_set_grid(50.0)
    ctx.tomo_set_eikonal(1)
    ctx.model_update(np.full(nt[0] * nt[1] * nt[2], -1.0, np.float32))
    assert ctx.get_model().min() < 0
    with pytest.raises(StError) as e:
        ctx.fim_solve(src[None], rcv)
    assert e.value.code == 1
    with pytest.raises(StError) as e:
        ctx.fsm_solve(src[None], rcv)
    assert e.value.code == 1
    ctx.set_model(ok, h)
    ctx.tomo_set_grid(50.0)
    # Tikhonov orders beyond the stencil cap
    dcal, _, _ = ctx.forward(src[None], rcv)
    with pytest.raises(StError) as e:
        ctx.optimize(dcal, dcal, 17, 1e-2, 3)
    assert e.value.code == 1
    p, _, _ = ctx.fim_solve(src[None], rcv)
    assert np.all(np.isfinite(p))
