"""Autotuner for the batched FIM.

Chooses the block edge (4 or 8), the number of waves per round launch
(resident waves per CU; each wave walks its own share of the active list)
and the inner-iteration cap by timing the device solve (HIP events inside the
library) on a sample of the caller's shots, then the visit ordering (none,
the solve's own tentative arrivals, the warm-start network's predictions) and
its band by timing whole calls.  Launch-shape results are cached per model
shape and shot count for the life of the process.
"""
from __future__ import annotations

import os

import numpy as np

_CACHE = {}

CUS = 256  # MI355X compute units


def best_params(ctx, src, rcv, block=0, grid=0, max_inner=0, sample=256, log=None,
                candidates=None, batch=None):
    key = (ctx.shape, len(src), block, grid, max_inner, batch,
           None if candidates is None else tuple(tuple(sorted(c.items())) for c in candidates))
    if key in _CACHE:
        return _CACHE[key]
    if block and grid and max_inner:
        p = dict(block=block, grid=grid, max_inner=max_inner)
        _CACHE[key] = p
        return p
    n = min(sample, len(src))
    idx = np.linspace(0, len(src) - 1, n).round().astype(int)
    sub = np.ascontiguousarray(src[idx])
    if candidates is None:
        candidates = []
        for B in ([block] if block else [4, 8]):
            # more waves than fit at once (24 per CU) run in generations:
            # a CU's slot then works through several waves' shares of the
            # list, which evens out a round's tail
            for wpc in ([grid // CUS] if grid else ([16] if B == 4 else [32, 48, 96, 192])):
                for mi in ([max_inner] if max_inner else ([16] if B == 4 else [16, 24])):
                    candidates.append(dict(block=B, grid=wpc * CUS, max_inner=mi))
    best, best_ms = None, float("inf")
    for c in candidates:
        c = dict(c)
        b = c.pop("batch", batch or n)
        ctx.fim_solve(sub[:2], rcv, batch=b, **c)  # first-touch of buffers
        _, _, st = ctx.fim_solve(sub, rcv, batch=b, **c)
        c["batch"] = b
        ms = st["solve_ms"] if b > 1 else st["total_ms"]
        if log:
            log(f"tune {c}: {ms:.2f} ms for {n} shots, {st['visits']} visits")
        if ms < best_ms:
            best, best_ms = c, ms
    _CACHE[key] = best
    return best


def load_surrogate(ctx, case, log=None):
    """Loads the trained warm-start network into the context when one is
    available (see seismic_tomography_3d_amd.surrogate); False otherwise."""
    try:
        from . import surrogate
    except ImportError:
        return False
    return surrogate.attach(ctx, case, log=log)


def best_warm_params(ctx, src, rcv, cold, sample=256, log=None):
    """Band of the predicted-arrival ordering and inner cap for warm solves,
    timed over the whole call (network inference included).  Returns
    (options, ms)."""
    n = min(sample, len(src))
    idx = np.linspace(0, len(src) - 1, n).round().astype(int)
    sub = np.ascontiguousarray(src[idx])
    h = ctx.spacing[0]
    # one block crossing at 1500 m/s as the unit of the band
    unit = cold["block"] * h / 1500.0
    best, best_ms = None, float("inf")
    grids = sorted({cold["grid"], 48 * CUS}) if cold["block"] == 8 else [cold["grid"]]
    caps = sorted({cold["max_inner"], 24})
    for g, mi, f in ((g, mi, f) for g in grids for mi in caps for f in (1.0, 1.5, 2.0)):
        o = dict(cold, batch=n, warm=1, band=f * unit, max_inner=mi, grid=g)
        ctx.fim_solve(sub[:2], rcv, **dict(o, batch=2))
        _, _, st = ctx.fim_solve(sub, rcv, **o)
        ms = st["total_ms"]
        if log:
            log(f"tune warm grid {g} max_inner {mi} band {f * unit:.3f}s: {ms:.2f} ms "
                f"(solve {st['solve_ms']:.2f}) for {n} shots, {st['visits']} visits, "
                f"{st['fallbacks']} fallbacks")
        if ms < best_ms:
            best, best_ms = dict(band=f * unit, max_inner=mi, grid=g), ms
    return best, best_ms


def best_order_params(ctx, src, rcv, cold, sample=256, log=None, factors=(0.6, 0.9, 1.5)):
    """Band of the tentative-arrival ordering (order = 1: each round visits,
    per shot, the scheduled blocks whose arrival lower bound lies within the
    band of the shot's earliest), timed over the whole call; the band unit is
    one block crossing at 1500 m/s.  Returns (options, ms)."""
    n = min(sample, len(src))
    idx = np.linspace(0, len(src) - 1, n).round().astype(int)
    sub = np.ascontiguousarray(src[idx])
    unit = cold["block"] * ctx.spacing[0] / 1500.0
    best, best_ms = None, float("inf")
    for f in factors:
        o = dict(cold, batch=n, order=1, band=f * unit)
        ctx.fim_solve(sub[:2], rcv, **dict(o, batch=2))
        _, _, st = ctx.fim_solve(sub, rcv, **o)
        ms = st["total_ms"]
        if log:
            log(f"tune order 1 band {f * unit:.3f}s: {ms:.2f} ms (solve {st['solve_ms']:.2f}) "
                f"for {n} shots, {st['visits']} visits, {st['inner_iters']} inner iterations")
        if ms < best_ms:
            best, best_ms = dict(order=1, band=f * unit), ms
    return best, best_ms
