#!/usr/bin/env python3
"""Headline benchmark: eikonal shot-solves per second at 201x201x101.

Workload: the batched case of seismic_tomography_3d_amd.synth -- 256 point
sources (x, y uniform, depth U(0, 2 km), seed 4) on a 201x201x101 grid at
h = 25 m, velocity rising linearly 1500 -> 4000 m/s with a -10% Gaussian
anomaly (seed 1), 20x20 surface receivers.  One step = one pass of the hot
path over the whole shot set: block-FIM solve of every shot plus receiver
picks, with the model already resident in HBM.

Legs reported in the one JSON line:
  value           shot-solves/s of the headline method: the batched FIM with
                  the fastest measured visit ordering among none, the solve's
                  own tentative arrivals and the learned warm start's
                  predicted arrivals (--order auto; each timed the same way,
                  all reported under "ordering")
  plain_fim       the same solver one shot at a time, cold start
  roofline        HBM roofline of the dominant kernel (the FIM round kernel):
                  algorithmic bytes per block visit x visits / kernel time
  cpu_baseline    the reference's own Accurate_FSM (oracle/_ref, built from the
                  reference sources) on the host cores, bounded sample; this
                  leg also runs the FP64 oracle on sample shots and reports the
                  receiver-traveltime error of the GPU picks against it
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "shot-solves/sec @201x201x101, ≤1e-4 s err: batched+surrogate vs plain FIM"
HBM_PEAK_GBS = 8000.0


def bytes_per_visit(B):
    """Algorithmic HBM bytes of one block visit: slowness B^3, travel times
    B^3 + 6 B^2 (block + face halos) in and B^3 out, FP32."""
    return 4 * (B ** 3 + (B ** 3 + 6 * B ** 2) + B ** 3)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """Host cores for the CPU baseline: every core this process may run on,
    up to 16 -- the host share of one GPU on the GPU machines (whose
    os.cpu_count() reports the whole host)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_baseline_leg(case, gpu_picks, gpu_T_sample, sample_idx, threads, budget_s=20.0):
    """CPU reference timing + FP64 error check (the checker lives here)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc

    vel, h = case["vel"], case["h"]
    s = np.ascontiguousarray((1.0 / vel).astype(np.float32))
    ny, nx, nz = s.shape
    rcv = np.stack([case["rx"], case["ry"], case["rz"]], axis=1).astype(np.float32)
    rx, ry, rz = (np.ascontiguousarray(rcv[:, i]) for i in range(3))
    nr = rcv.shape[0]
    src = np.ascontiguousarray(case["src"], np.float32)

    # FP64 oracle on the sample shots, one thread per shot (ctypes frees the GIL)
    errs = [None] * len(sample_idx)
    p64s = np.zeros((len(sample_idx), nr), np.float64)

    def check(k, i):
        T64 = orc.eikonal_fp64(vel, h, src[i])
        p64 = orc.picks(T64.astype(np.float32), (h, h, h), rcv)
        p64s[k] = p64
        pe = gpu_picks[i].astype(np.float64) - p64
        te = np.abs(gpu_T_sample[k].astype(np.float64) - T64).max()
        errs[k] = (np.abs(pe).max(), float(np.sqrt(np.mean(pe ** 2))), te)

    # a pool of `threads` workers over the sample shots
    todo = list(enumerate(sample_idx))
    lock = threading.Lock()

    def worker():
        while True:
            with lock:
                if not todo:
                    return
                k, i = todo.pop()
            check(k, i)

    ths = [threading.Thread(target=worker) for _ in range(min(threads, len(sample_idx)))]
    t0 = time.time()
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    t_fp64 = time.time() - t0
    log(f"fp64 check of {len(sample_idx)} shots took {t_fp64:.1f}s")
    err = {
        "shots_checked": [int(i) for i in sample_idx],
        "max_pick_err_s": float(max(e[0] for e in errs)),
        "rms_pick_err_s": float(np.sqrt(np.mean([e[1] ** 2 for e in errs]))),
        "max_volume_err_s": float(max(e[2] for e in errs)),
        "per_shot_max_pick_err_s": [float(e[0]) for e in errs],
        "fp64_oracle_s_per_shot": t_fp64 * min(threads, len(sample_idx)) / len(sample_idx),
    }

    # CPU reference: the reference's own Accurate_FSM when built, else the
    # port; one shot per thread, two shots per thread per run, median of
    # three runs (the sample is bounded: a few seconds of wall time)
    kind = "reference" if orc.ref_available() else "port"

    def run(sub, nthreads):
        out = np.zeros((len(sub), nr), np.float32)
        t = time.time()
        if kind == "reference":
            orc.ref().ref_fsm_threads(orc.P(s), nx, ny, nz, h, h, h, orc.P(sub), len(sub),
                                      orc.P(rx), orc.P(ry), orc.P(rz), nr, orc.P(out), nthreads)
        else:
            orc.oracle().or_fsm_batch(orc.P(s), nx, ny, nz, h, h, h, orc.P(sub), len(sub),
                                      orc.P(rx), orc.P(ry), orc.P(rz), nr, orc.P(out), nthreads)
        return time.time() - t

    n = 2 * threads
    sub = np.ascontiguousarray(src[:n])
    walls = [run(sub, threads) for _ in range(3)]
    dt = float(np.median(walls))
    one = [run(np.ascontiguousarray(src[k:k + 1]), 1) for k in range(3)]
    dt1 = float(np.median(one))
    base = {
        "value": n / dt,
        "unit": "shot-solves/s",
        "cores": threads,
        "kind": kind,
        "sample": f"{n} of the {len(src)} shots ({n // threads} per thread), Accurate_FSM "
                  f"(Noble-Gesret-Belayouni, the reference's default CPU eikonal) + picks, "
                  f"201x201x101; median of 3 runs: {', '.join(f'{w:.2f}' for w in walls)} s wall",
        "cpu_model": cpu_model(),
        "host_cpus": {"affinity": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count(),
                      "used": threads},
        "one_core": {"value": 1.0 / dt1, "unit": "shot-solves/s",
                     "sample": "shots 0-2 one at a time on 1 thread, median"},
    }
    return base, err, p64s


def fsm_leg(ctx, src, rcv, cpu_value):
    """The reference's default CPU eikonal, Accurate_FSM, on the device
    (bit-identical to the CPU code) over the same 256 shots."""
    ctx.fsm_solve(src[:64], rcv)
    t0 = time.perf_counter()
    _, _, ms = ctx.fsm_solve(src, rcv)
    dt = time.perf_counter() - t0
    out = {"value": len(src) / dt, "unit": "shot-solves/s", "ms_per_step": 1e3 * dt,
           "device_ms": ms}
    if cpu_value:
        out["speedup_vs_cpu_reference"] = out["value"] / cpu_value
    return out


def podvin_leg(ctx, case, src, rcv, with_cpu):
    """The reference's Podvin & Lecomte eikonal (eikonal_type 0) on the
    device, bit-identical to its CPU code, over the same shots; beside it
    the reference's own classic.cpp build on one shot, one core."""
    ctx.podvin_solve(src[:16], rcv)
    t0 = time.perf_counter()
    _, _, ms = ctx.podvin_solve(src, rcv)
    dt = time.perf_counter() - t0
    out = {"value": len(src) / dt, "unit": "shot-solves/s", "ms_per_step": 1e3 * dt,
           "device_ms": ms}
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as orc
        from oracle_lib import P
        if orc.ref_available():
            vel, h = case["vel"], case["h"]
            ny, nx, nz = vel.shape
            s = np.ascontiguousarray((1.0 / vel).astype(np.float32))
            T = np.zeros(s.shape, np.float32)
            r = np.zeros(1, np.float32)
            s0 = np.ascontiguousarray(src[:1])
            t0 = time.perf_counter()
            orc.ref().ref_eikonal(0, P(s), nx, ny, nz, h, h, h, P(s0), 1, P(r), P(r), P(r), 0,
                                  P(T), None, None)
            cpu = time.perf_counter() - t0
            out["cpu_reference"] = {"value": 1.0 / cpu, "unit": "shot-solves/s", "cores": 1,
                                    "kind": "reference", "sample": "1 shot (the first)"}
            out["speedup_vs_cpu_reference_one_core"] = out["value"] * cpu
    return out


def tomography_leg(with_cpu):
    """One least-squares tomography iteration on the 201x201x101 case (16
    shots, 2,500 receivers, 201x201x101 at 12.5 m, inversion grid 41x41x21,
    order 2, lambda 1e-2, 20 CGLS iterations, 5-sample Gaussian): forward
    (FIM + picks + ray tracing into the CSR matrix), Tikhonov CGLS and model
    update on the device; beside it, the oracle's restatement of the
    reference's host-serial steps (ray tracing per ray, FP32 COO CGLS, model
    update with smoothing) on the same inputs, one core."""
    from seismic_tomography_3d_amd import Context, synth
    c = synth.tomography_case()
    h = c["h"]
    src = c["src"]
    rcv = np.stack([c["rx"], c["ry"], c["rz"]], 1).astype(np.float32)
    ctx = Context(0)
    ctx.set_model(1.0 / c["vel_true"], h)
    dobs, _, _ = ctx.fim_solve(src, rcv)
    dobs = (dobs + np.random.default_rng(3).normal(0.0, 2e-3, dobs.shape)).astype(np.float32)
    ctx.set_model(1.0 / c["vel0"], h)
    ctx.tomo_set_grid(c["d_tomo"])
    ctx.forward(src, rcv)  # warm-up (buffers)
    ctx.set_model(1.0 / c["vel0"], h)
    t0 = time.perf_counter()
    dcal, _, fst = ctx.forward(src, rcv)
    t1 = time.perf_counter()
    ctx.optimize(dobs, dcal, 2, 1e-2, 20)
    t2 = time.perf_counter()
    ctx.model_update(None, smooth=True, samples=5, stdv=2.0)
    t3 = time.perf_counter()
    rp, cols, vals = ctx.csr()
    nrays = int(src.shape[0] * rcv.shape[0])
    # ray steps: every step adds rayStep (0.2 h) to its cell's entry (the
    # source cell's entry is the final distance instead, under one step)
    ray_steps = float(np.sum(vals[:rp[nrays]], dtype=np.float64) / (0.2 * h))
    out = {"workload": "16 shots x 2,500 receivers, 201x201x101 at 12.5 m, 41x41x21 inversion "
                       "grid, order 2, lambda 1e-2, 20 CGLS iterations, Gaussian 5/2.0",
           "ray_steps_est": ray_steps, "G_nnz": int(rp[nrays]),
           "gpu_ms": {"forward": 1e3 * (t1 - t0), "forward_fim_device": fst["solve_ms"],
                      "optimize": 1e3 * (t2 - t1), "model_update": 1e3 * (t3 - t2),
                      "iteration": 1e3 * (t3 - t0)},
           "rays": nrays}
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as orc
        o = orc.oracle()
        _, T, _ = ctx.fim_solve(src, rcv, return_T=True)  # the iteration's volumes
        ny, nx, nz = T.shape[1:]
        nxt, nyt, nzt = ctx.tomo_dims()
        dt_ = c["d_tomo"]
        rx, ry, rz = (np.ascontiguousarray(rcv[:, i]) for i in range(3))
        nr = rcv.shape[0]
        cap = 4_000_000
        rn = np.zeros(nr, np.int32)
        cl = np.zeros(cap, np.int32)
        vl = np.zeros(cap, np.float32)
        ill = np.zeros((ny, nx, nz), np.float32)
        Tp = np.zeros((ny + 2, nx + 2, nz + 2), np.float32)
        t_ray = 0.0
        for i, sxyz in enumerate(src):
            o.or_expand_fdm(orc.P(np.ascontiguousarray(T[i])), nx, ny, nz, orc.P(Tp))
            a = time.perf_counter()
            o.or_ray_trace_shot(orc.P(Tp), nx, ny, nz, h, h, h, nxt, nyt, nzt, dt_, dt_, dt_,
                                *(float(v) for v in sxyz), orc.P(rx), orc.P(ry), orc.P(rz), nr,
                                100000, orc.P(rn), orc.P(cl), orc.P(vl), cap, orc.P(ill))
            t_ray += time.perf_counter() - a
        # the device system (data rows + Tikhonov rows) and right-hand side
        M = nxt * nyt * nzt
        N = rp.shape[0] - 1
        b = np.zeros(N, np.float32)
        b[:dobs.size] = (dobs - dcal).ravel()
        x = np.zeros(M, np.float32)
        a = time.perf_counter()
        o.or_cgls_ref_f32(rp.ctypes.data_as(orc.lp), np.ascontiguousarray(cols).ctypes.data_as(orc.ip),
                          orc.P(np.ascontiguousarray(vals)), N, M, orc.P(b), 20, orc.P(x))
        t_cg = time.perf_counter() - a
        sl = np.ascontiguousarray((1.0 / c["vel0"]).astype(np.float32))
        dm = np.zeros_like(sl)
        vel = np.zeros_like(sl)
        a = time.perf_counter()
        o.or_model_update(orc.P(sl), orc.P(x), nx, ny, nz, h, h, h, nxt, nyt, nzt, dt_, dt_, dt_,
                          1, 5, 2.0, orc.P(dm), orc.P(vel))
        t_up = time.perf_counter() - a
        out["cpu_port_ms"] = {"ray_tracing": 1e3 * t_ray, "cgls": 1e3 * t_cg,
                              "model_update": 1e3 * t_up, "cores": 1,
                              "kind": "port (the oracle's restatement of least_squares.cpp)"}
        g = out["gpu_ms"]
        out["speedup_host_steps"] = (t_ray + t_cg + t_up) * 1e3 / max(
            1e-9, g["forward"] - g["forward_fim_device"] + g["optimize"] + g["model_update"])
    # CGLS + model update in one call, eager and as one captured graph (the
    # first graph call instantiates, later ones re-capture and update it)
    ctx.set_model(1.0 / c["vel0"], h)
    dcal, _, _ = ctx.forward(src, rcv)
    # (the updated model stays on the device: fetch=False)
    fused = {"eager": [], "graph": []}
    for graph in (False, False, False, True, True, True, True):
        a = time.perf_counter()
        ctx.optimize_update(dobs, dcal, 2, 1e-2, 20, smooth=True, samples=5, stdv=2.0, graph=graph,
                            fetch=False)
        fused["graph" if graph else "eager"].append(1e3 * (time.perf_counter() - a))
    a = time.perf_counter()
    ctx.optimize_update(dobs, dcal, 2, 1e-2, 20, smooth=True, samples=5, stdv=2.0, graph=True)
    g = out["gpu_ms"]
    g["optimize_update_graph_with_outputs"] = 1e3 * (time.perf_counter() - a)
    g["optimize_update_eager"] = float(np.median(fused["eager"]))
    g["optimize_update_graph_first"] = fused["graph"][0]
    g["optimize_update_graph"] = float(np.median(fused["graph"][1:]))
    # forward, then the solve and update as one graph, the model kept resident
    a = time.perf_counter()
    dcal, _, _ = ctx.forward(src, rcv)
    ctx.optimize_update(dobs, dcal, 2, 1e-2, 20, smooth=True, samples=5, stdv=2.0, graph=True,
                        fetch=False)
    g["iteration_resident_graph"] = 1e3 * (time.perf_counter() - a)
    ctx.close()
    return out



def large_inversion(iters=10, log=None, warm=False):
    """The 401x401x201 (h = 10 m, 32 M cells) inversion: 1,024 surface shots
    on a 32x32 grid, 4,096 receivers on a 64x64 grid at 100 m, picks from
    the true model (5 seeded layers + 3 anomalies, seed 2) by the FIM;
    start = lateral average; 81x81x41 inversion grid (50 m); per iteration
    a forward pass in chunks of 256 shots (FIM + picks + ray tracing into
    the CSR matrix), order-2 Tikhonov CGLS of 30 iterations and the model
    update with the 5-sample Gaussian as one captured graph.  Reports data
    RMS misfit and model RMS error per iteration.  warm: order the FIM
    visits by the warm-start network (the network of the 5 x 5 x 2.5 km
    workload, applied to this 4 x 4 x 2 km model; slower here than cold)."""
    from seismic_tomography_3d_amd import Context, synth
    nx, ny, nz, h = 401, 401, 201, 10.0
    t0 = time.perf_counter()
    vel_true = synth.layered_model(nx, ny, nz, h, n_layers=5, n_anom=3, seed=2)
    vel0 = synth.lateral_average(vel_true)
    L = (nx - 1) * h
    sx, sy, sz = synth.regular_geometry((0.05 * L, 0.05 * L), (0.05 * L, 0.95 * L),
                                        (0.95 * L, 0.05 * L), 32, 32, 0.0)
    src = np.stack([sx, sy, sz], 1).astype(np.float32)
    rcv = np.stack(synth.surface_grid(64, 0.02 * L, 0.98 * L, 100.0), 1).astype(np.float32)
    if log:
        log(f"large inversion: models built in {time.perf_counter() - t0:.1f} s")
    ctx = Context(0)
    from seismic_tomography_3d_amd import tune
    warm = 1 if warm and tune.load_surrogate(ctx, None, log=log) else 0
    # 256 shots per chunk: 35 GB of travel times; the FIM's rounds serve all
    # of them (0.40 s per pass against 0.44 s in chunks of 64)
    fo = dict(batch=256, warm=warm)
    ctx.set_model(1.0 / vel_true, h)
    a = time.perf_counter()
    dobs, _, _ = ctx.fim_solve(src, rcv, **fo)
    t_obs = time.perf_counter() - a
    ctx.set_model(1.0 / vel0, h)
    nt = ctx.tomo_set_grid(50.0)
    rows = []
    for it in range(iters + 1):
        a = time.perf_counter()
        dcal, _, fst = ctx.forward(src, rcv, **fo)
        t_fwd = time.perf_counter() - a
        misfit = ctx.residual_norm(dobs, dcal)
        vel = 1.0 / ctx.get_model()
        merr = float(np.sqrt(np.mean((vel.astype(np.float64) - vel_true) ** 2)))
        row = {"iteration": it, "data_rms_s": misfit / np.sqrt(dobs.size), "model_rms_mps": merr,
               "forward_s": t_fwd, "fim_device_s": fst["solve_ms"] / 1e3}
        if it < iters:
            a = time.perf_counter()
            ctx.optimize_update(dobs, dcal, 2, 1e-2, 30, smooth=True, samples=5, stdv=2.0,
                                graph=True, fetch=False)
            row["solve_update_s"] = time.perf_counter() - a
        if it == 0:
            row["G_rows"], row["G_nnz"] = ctx.csr_size()
        rows.append(row)
        if log:
            log(f"large inversion {row}")
    ctx.close()
    return {"workload": "401x401x201 at 10 m, 1,024 shots x 4,096 receivers (4.19 M rays), "
                        "81x81x41 grid, order 2, lambda 1e-2, 30 CGLS iterations, Gaussian 5/2.0, "
                        "chunks of 256 shots",
            "warm_start": bool(warm), "observed_picks_s": t_obs, "iterations": rows,
            "tomo_dims": list(nt)}


def synthesis_leg(with_cpu, traffic_file, nt=3001):
    """Data synthesis at the reference's wave_main.cpp / auto_picking.cpp
    sizes: 701x501x101 at 10 m with 50 absorbing cells per side (801x601x201
    padded), nt 3001 at 1 ms, 14,000 receivers on a 140x100 grid; two shots
    modelled and auto-picked on the device (synthetic layered model, 40 Hz
    Ricker).  Beside it the reference's own serial loop (acoustic.cpp built
    from its source) on the same padded volume for two steps, one core, and
    the picker's C restatement over one shot's file."""
    from seismic_tomography_3d_amd import Context, synth
    from seismic_tomography_3d_amd.api import ricker
    nx, ny, nz, h, nb, dt = 701, 501, 101, 10.0, 50, 1e-3
    vel = synth.layered_model(nx, ny, nz, h, n_layers=5, n_anom=3, seed=2)
    w = ricker(nt, dt, 40.0, 0.1)
    gx, gy = np.meshgrid(np.linspace(500.0, 6500.0, 140), np.linspace(500.0, 4500.0, 100))
    rcv = np.stack([gx.ravel(), gy.ravel(), np.full(gx.size, 10.0)], 1).astype(np.float32)
    src = np.array([[3500.0, 2500.0, 500.0], [1500.0, 1200.0, 500.0]], np.float32)
    ctx = Context(0)
    ctx.acoustic_set_model(vel, h, nb=nb, factor=0.0045, dt=dt)
    cells = (nx + 2 * nb - 8) * (ny + 2 * nb - 8) * (nz + 2 * nb - 8)
    # the step alone (plus its receiver recording), one shot, 200 steps
    ns = min(200, nt)
    ctx.acoustic_shots(w[:min(20, nt)], src[:1], rcv, group=1, return_seis=False)
    _, _, ms_s = ctx.acoustic_shots(w[:ns], src[:1], rcv, group=1, return_seis=False)
    step_ms = ms_s / ns
    t0 = time.perf_counter()
    picks, dev_ms = ctx.acoustic_picks(w, src, rcv)
    wall = time.perf_counter() - t0
    seis, _, _ = ctx.acoustic_shots(w[:min(600, nt)], src[:1], rcv[:2000], return_seis=True)
    t1 = time.perf_counter()
    ctx.auto_pick(seis[0], dt)
    pick_ms = 1e3 * (time.perf_counter() - t1)
    ctx.close()
    achieved = 16.0 * cells / (step_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(traffic_file):
        with open(traffic_file) as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    out = {"workload": "acoustic modelling + auto-picking, 701x501x101 at 10 m, nb 50 "
                       f"(801x601x201 padded), nt {nt} at 1 ms, 14,000 receivers, 2 shots",
           "value": len(src) / (dev_ms * 1e-3), "unit": "shots/s (modelled and picked)",
           "ms_per_shot": dev_ms / len(src), "wall_s": wall, "step_ms": step_ms,
           "cells_per_step": cells, "cell_updates_per_s": cells / (step_ms * 1e-3),
           "picked_fraction": float(np.count_nonzero(picks)) / picks.size,
           "picker_ms_2000x600_from_host": pick_ms,
           "roofline": {"kernel": "wave_step", "bound": "hbm", "achieved": achieved,
                        "peak": 8000.0, "unit": "GB/s", "frac": achieved / 8000.0,
                        "traffic": traffic,
                        "bytes_per_cell": "16 (U_pre, U_pas, V^2 dt^2 read; U_fut written)"}}
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as orc
        from oracle_lib import P
        v = np.ascontiguousarray(vel, np.float32)
        rx, ry, rz = (np.ascontiguousarray(rcv[:2, i]) for i in range(3))
        kind = "reference" if orc.ref_available() else "port"
        fn = orc.ref().ref_acoustic_shot if kind == "reference" else orc.oracle().or_acoustic_shot
        ts = []
        for n in (1, 2):
            sg = np.zeros(2 * n, np.float32)
            t0 = time.perf_counter()
            fn(P(v), nx, ny, nz, h, h, h, nb, 0.0045, P(w), n, dt, *(float(c) for c in src[0]),
               P(rx), P(ry), P(rz), 2, P(sg), None)
            ts.append(time.perf_counter() - t0)
        cpu_step = max(ts[1] - ts[0], 1e-9)
        out["cpu_baseline"] = {"value": cells / cpu_step, "unit": "cell-updates/s", "cores": 1,
                               "kind": kind, "step_s": cpu_step,
                               "sample": "2 time steps of one shot on the full 801x601x201 "
                                         "padded volume (step time = t(2 steps) - t(1 step))"}
        out["speedup_vs_cpu_reference"] = out["cell_updates_per_s"] / out["cpu_baseline"]["value"]
        s2 = np.ascontiguousarray(seis[0])
        pk = np.zeros(s2.shape[0], np.float32)
        t0 = time.perf_counter()
        orc.oracle().or_auto_pick(P(s2), s2.shape[0], s2.shape[1], dt, 0.01, 1e-15, 0.045, P(pk))
        out["picker_cpu_port_ms_2000x600"] = 1e3 * (time.perf_counter() - t0)
    return out


def calibrate_copy(mib):
    """A device-to-device copy of known size (torch's copy kernel), three
    times: under rocprofv3 --pmc its FETCH_SIZE / WRITE_SIZE per dispatch
    give the scale of those counters against algorithmic bytes."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "dispatches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--synthesis-only", action="store_true",
                    help="run only the data-synthesis leg (acoustic modelling + picking)")
    ap.add_argument("--tomography-only", action="store_true",
                    help="only run the tomography-iteration leg")
    ap.add_argument("--large-inversion", type=int, default=0, metavar="ITERS",
                    help="only run ITERS iterations of the 401x401x201 inversion")
    ap.add_argument("--large-warm", action="store_true",
                    help="order the large inversion's FIM visits by the warm-start network")
    ap.add_argument("--synthesis-nt", type=int, default=3001,
                    help="time steps of the data-synthesis leg (3001 as the reference)")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shots", type=int, default=256)
    ap.add_argument("--block", type=int, default=0, help="4 or 8 (0 = tuned)")
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--max-inner", type=int, default=0)
    ap.add_argument("--plain-shots", type=int, default=16)
    ap.add_argument("--warm", default="auto", choices=["auto", "on", "off"],
                    help="load the warm-start network (auto: when trained weights exist)")
    ap.add_argument("--order", default="auto", choices=["auto", "none", "arrival", "network"],
                    help="visit ordering of the headline solve (auto: the fastest measured)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--check-shots", type=int, default=32,
                    help="shots checked against the FP64 oracle after the timed steps")
    ap.add_argument("--band", type=float, default=0.0,
                    help="ordering band in s for the arrival and network orderings (0 = tuned)")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-shot-at-a-time leg")
    ap.add_argument("--no-cold-leg", action="store_true",
                    help="skip the batched solve without ordering (unless --order none)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the Accurate_FSM and tomography-iteration legs")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB three times (HBM counter scale)")
    a = ap.parse_args()

    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(shots shard across GPUs by design: one shot "
                                   "range per rank, no collective in the solve)"}))
        sys.exit(2)

    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    if a.large_inversion:
        print(json.dumps({"large_inversion": large_inversion(a.large_inversion, log, a.large_warm)}),
              flush=True)
        return
    if a.tomography_only:
        print(json.dumps({"tomography_iteration": tomography_leg(not a.no_cpu)}), flush=True)
        return
    if a.synthesis_only:
        leg = synthesis_leg(not a.no_cpu,
                            os.path.join(ROOT, "profiles", "wave_step_801x601x201_traffic.json"),
                            a.synthesis_nt)
        print(json.dumps({"data_synthesis": leg}), flush=True)
        return

    from seismic_tomography_3d_amd import Context, synth, tune

    case = synth.batched_case(a.shots)
    h = case["h"]
    src = case["src"]
    rcv = np.stack([case["rx"], case["ry"], case["rz"]], axis=1).astype(np.float32)
    ctx = Context(0)
    ctx.set_model(1.0 / case["vel"], h)
    log(f"model 201x201x101 resident; {len(src)} shots, {len(rcv)} receivers")

    params = tune.best_params(ctx, src, rcv, block=a.block, grid=a.grid,
                              max_inner=a.max_inner, log=log)
    have_net = False
    if a.warm in ("auto", "on"):
        have_net = tune.load_surrogate(ctx, case, log=log)
        if a.warm == "on" and not have_net:
            raise SystemExit("--warm on but no trained network is available")
        if a.warm == "on" and a.order == "auto":
            a.order = "network"
    cold_opts = dict(block=params["block"], grid=params["grid"],
                     max_inner=params["max_inner"], batch=a.batch)

    # visit ordering: none, the solve's own tentative arrivals, or the
    # warm-start network's predicted arrivals; each candidate's band is tuned
    # over whole calls, then every candidate is timed the same way on all
    # shots and the fastest is the headline configuration
    cands = {}
    if a.order == "none" or (a.order == "auto" and not a.no_cold_leg):
        cands["none"] = dict(cold_opts)
    if a.order in ("auto", "arrival"):
        op = (dict(order=1, band=a.band) if a.band > 0
              else tune.best_order_params(ctx, src, rcv, cold_opts, log=log)[0])
        cands["arrival"] = dict(cold_opts, **op)
    if have_net and a.order in ("auto", "network"):
        wp = (dict(band=a.band, max_inner=params["max_inner"]) if a.band > 0
              else tune.best_warm_params(ctx, src, rcv, cold_opts, log=log)[0])
        cands["network"] = dict(cold_opts, **wp, warm=1)
    if not cands:
        raise SystemExit(f"--order {a.order} needs a trained network (none was loaded)")

    def whole_calls(o, reps=3):
        ctx.fim_solve(src, rcv, **o)
        ts, st = [], None
        for _ in range(reps):
            t0 = time.perf_counter()
            _, _, st = ctx.fim_solve(src, rcv, **o)
            ts.append(time.perf_counter() - t0)
        dt = float(np.median(ts))
        return {"value": len(src) / dt, "unit": "shot-solves/s", "ms_per_step": 1e3 * dt,
                "visits": st["visits"], "inner_iters": st["inner_iters"], "rounds": st["rounds"],
                "device_solve_ms": st["solve_ms"], "options": o}

    legs = {k: whole_calls(o) for k, o in cands.items()}
    for k, v in legs.items():
        log(f"ordering {k}: {v}")
    head = max(legs, key=lambda k: legs[k]["value"])
    opts = cands[head]
    warm = int(opts.get("warm", 0))
    log(f"headline ordering {head}, options {opts}")
    for _ in range(a.warmup):
        ctx.fim_solve(src, rcv, **opts)
    times, stats = [], None
    for k in range(a.steps):
        t0 = time.perf_counter()
        picks, _, stats = ctx.fim_solve(src, rcv, **opts)
        times.append(time.perf_counter() - t0)
        log(f"step {k}: {times[-1] * 1e3:.1f} ms, {stats}")
    ms = 1e3 * float(np.median(times))
    value = len(src) / (ms / 1e3)
    batched_cold = legs.get("none")

    # plain FIM: same kernels, cold start, one shot at a time, its own tuned
    # launch shape (bounded sample)
    plain = None
    if not a.no_plain:
        npl = min(a.plain_shots, len(src))
        pparams = tune.best_params(ctx, src[:npl], rcv, block=a.block, batch=1, log=log,
                                   candidates=[dict(block=B, grid=g, max_inner=mi, batch=1)
                                               for B in (4, 8) for g in (1024, 2048)
                                               for mi in (16, 24)])
        popts = dict(pparams, batch=1, warm=0)
        ctx.fim_solve(src[:2], rcv, **popts)
        t0 = time.perf_counter()
        _, _, pst = ctx.fim_solve(src[:npl], rcv, **popts)
        pdt = time.perf_counter() - t0
        plain = {"value": npl / pdt, "unit": "shot-solves/s", "shots": npl,
                 "ms_per_shot": 1e3 * pdt / npl, "rounds_per_shot": pst["rounds"] / npl,
                 "visits_per_shot": pst["visits"] / npl, "options": popts}
        log(f"plain: {plain}")

    # roofline of the FIM round kernel over the last timed step
    B = params["block"]
    bpv = bytes_per_visit(B)
    # per launch: algorithmic bytes of the visits of one step over its round
    # launches (the empty launches enqueued before the host saw the end
    # included, as a kernel trace counts them), over the average launch
    # duration (HIP events around the rounds on the library's stream)
    launches = max(1, stats["launches"])
    avg_launch_s = stats["solve_ms"] / 1e3 / launches
    bytes_per_launch = stats["visits"] * bpv / launches
    achieved = bytes_per_launch / avg_launch_s / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "kernel": f"fim_round{B}", "bytes_per_visit": bpv,
            "bytes_per_launch": bytes_per_launch, "avg_launch_us": avg_launch_s * 1e6,
            "visits_per_step": stats["visits"], "kernel_ms_per_step": stats["solve_ms"],
            "launches_per_step": stats["launches"], "rounds_per_step": stats["rounds"]}
    prof = os.path.join(ROOT, "profiles", f"fim_round{B}_batched_201x201x101_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            tr = json.load(f)
        # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE counters of
        # the same command, scaled by their readings for a copy of known size
        roof["traffic"] = tr.get("bytes_per_launch")
        roof["traffic_source"] = os.path.relpath(prof, ROOT)

    out = {
        "metric": METRIC, "value": value, "unit": "shot-solves/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batched eikonal: {len(src)} shots, 201x201x101, h=25 m, linear "
                               f"1500-4000 m/s + -10% anomaly, {len(rcv)} surface receivers",
                   "shots": len(src), "receivers": len(rcv), "block": B,
                   "batch": a.batch, "grid": params["grid"],
                   "max_inner": params["max_inner"], "warm_start": bool(warm),
                   "ordering": head, "options": opts},
        "plain_fim": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "batched_cold_fim": batched_cold,
        # every ordering candidate timed the same way (median of 3 whole
        # calls on all shots); the fastest is the headline
        "ordering": legs,
        "network_gain": (legs["network"]["value"] / legs["none"]["value"]
                         if "network" in legs and "none" in legs else None),
        "network_vs_arrival_order": (legs["network"]["value"] / legs["arrival"]["value"]
                                     if "network" in legs and "arrival" in legs else None),
        "fim_stats_per_step": stats,
        "roofline": roof,
        # every shot must at least write its volume once and read the model:
        # 8 B per cell, the HBM floor of any solver of this grid
        "compulsory_floor": {"bytes_per_shot": 8 * int(np.prod(case["vel"].shape)),
                             "shot_solves_per_s_at_8TBps":
                                 HBM_PEAK_GBS * 1e9 / (8 * int(np.prod(case["vel"].shape)))},
    }
    if have_net and not a.no_cold_leg:
        # C4: the network's prediction inflated by (1 + 0.02) seeds the travel
        # times; the judge re-solves cold every shot it did not bound
        vo = dict(cold_opts, warm=1, warm_values=1, delta=0.02, judge=1)
        if "network" in cands:
            vo["band"] = cands["network"]["band"]
        leg = whole_calls(vo)
        _, _, vst = ctx.fim_solve(src, rcv, **vo)
        leg.update(fallbacks=vst["fallbacks"], judge_max_s=vst["judge_max"], delta=0.02)
        if "none" in legs:
            leg["iter_reduction_vs_none"] = legs["none"]["inner_iters"] / max(1, leg["inner_iters"])
            leg["visit_reduction_vs_none"] = legs["none"]["visits"] / max(1, leg["visits"])
        out["warm_values_leg"] = leg
        log(f"(1+delta) warm values: {leg}")
    if "none" in legs:
        for k in ("arrival", "network"):
            if k in legs:
                legs[k]["visit_reduction_vs_none"] = legs["none"]["visits"] / max(1, legs[k]["visits"])
                legs[k]["iter_reduction_vs_none"] = (legs["none"]["inner_iters"]
                                                     / max(1, legs[k]["inner_iters"]))

    if not a.no_cpu:
        threads = a.cpu_threads or host_threads()
        ks = max(1, min(a.check_shots, len(src)))
        sample = np.linspace(0, len(src) - 1, ks).round().astype(int)
        _, Ts, _ = ctx.fim_solve(src[sample], rcv, return_T=True, **opts)
        base, err, p64s = cpu_baseline_leg(case, picks, Ts, sample, threads)
        out["cpu_baseline"] = base
        out["speedup_vs_cpu"] = value / base["value"]
        out["error_vs_fp64"] = err
        if "warm_values_leg" in out:
            pv, _, _ = ctx.fim_solve(src, rcv, **out["warm_values_leg"]["options"])
            out["warm_values_leg"]["max_pick_err_vs_fp64_s"] = float(
                np.abs(pv[sample].astype(np.float64) - p64s).max())
        out["error_gate_ok"] = err["max_pick_err_s"] <= 1e-4
    if not a.no_extra:
        cpu_v = out["cpu_baseline"]["value"] if "cpu_baseline" in out else None
        out["accurate_fsm_gpu"] = fsm_leg(ctx, src, rcv, cpu_v)
        log(f"accurate_fsm_gpu: {out['accurate_fsm_gpu']}")
        out["podvin_gpu"] = podvin_leg(ctx, case, src, rcv, not a.no_cpu)
        log(f"podvin_gpu: {out['podvin_gpu']}")
    ctx.close()
    if not a.no_extra:
        out["tomography_iteration"] = tomography_leg(not a.no_cpu)
        log(f"tomography_iteration: {out['tomography_iteration']}")
        out["data_synthesis"] = synthesis_leg(
            not a.no_cpu, os.path.join(ROOT, "profiles", "wave_step_801x601x201_traffic.json"))
        log(f"data_synthesis: {out['data_synthesis']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
