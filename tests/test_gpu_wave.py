"""Data synthesis on the device against the oracle, bit for bit: the
acoustic modelling (src/utils/wave) and the auto-picker
(src/utils/auto_picking).

The oracle restates the reference's serial loop (pinned bit-exact against
the reference's own build in test_oracle_wave.py).  Cases cover several
absorbing widths (nb = 4, 5, 8), unequal spacings, shots advanced one, two
and four per launch (an odd shot count leaves a group part-empty), a
receiver on the source cell, and the final wavefield.
"""
import numpy as np
import pytest

import oracle_lib as orc
from oracle_lib import P

pytestmark = pytest.mark.gpu


def _model(ctx, c):
    ctx.acoustic_set_model(c["vel"], c["d"], nb=c["nb"], factor=c["factor"], dt=c["dt"])


def _sources(c, n, seed=0):
    rng = np.random.default_rng(seed)
    ext = np.array([(c["nx"] - 1) * c["d"][0], (c["ny"] - 1) * c["d"][1],
                    (c["nz"] - 1) * c["d"][2]], np.float32)
    s = (rng.uniform(0.05, 0.95, (n, 3)).astype(np.float32) * ext).astype(np.float32)
    s[0] = c["src"]
    return s


@pytest.mark.parametrize("case,group", [
    (dict(seed=3), 1),
    (dict(seed=4, nb=5, spacing=(10.0, 12.5, 7.5)), 2),
    (dict(seed=5, nb=4, nt=80), 4),
    (dict(seed=6, nx=70, ny=9, nz=90, nt=90), 2),          # several z and x tiles
    (dict(seed=7, nx=40, ny=11, nz=112, nt=70), 0),        # 64-cell z tiles
])
def test_seismograms_bit_exact(ctx, case, group):
    c = orc.acoustic_case(**case)
    src = _sources(c, 3, seed=case["seed"])
    c["rcv"] = np.concatenate([c["rcv"], src[:1]])          # a receiver on the first source
    _model(ctx, c)
    print(f"case {case} group {group}", flush=True)
    seis, uf, ms = ctx.acoustic_shots(c["wavelet"], src, c["rcv"], group=group, return_ufut=True)
    for s in range(src.shape[0]):
        if s < src.shape[0] - 1:
            ref = orc.acoustic_shot(c, src=src[s])
        else:
            ref, ref_uf = orc.acoustic_shot(c, src=src[s], want_ufut=True)
            assert np.array_equal(uf, ref_uf)
        assert np.abs(ref).max() > 0
        bad = np.argwhere(seis[s] != ref)
        assert bad.size == 0, (s, bad[:5], seis[s][tuple(bad[0])], ref[tuple(bad[0])])


def test_ychunks_and_long_axis(ctx):
    """A model long in y: the planes are split over several workgroups,
    each re-reading the queue at its start."""
    c = orc.acoustic_case(nx=12, ny=150, nz=12, nb=6, nt=70, seed=8)
    _model(ctx, c)
    seis, _, _ = ctx.acoustic_shots(c["wavelet"], c["src"][None], c["rcv"], group=1)
    assert np.array_equal(seis[0], orc.acoustic_shot(c))


def _pick_oracle(seis, dt, window=0.01, amp_cut=1e-15, shift=0.045):
    seis = np.ascontiguousarray(seis, np.float32)
    out = np.zeros(seis.shape[0], np.float32)
    orc.oracle().or_auto_pick(P(seis), seis.shape[0], seis.shape[1], dt, window, amp_cut, shift,
                              P(out))
    return out


def test_auto_pick_bit_exact(ctx):
    rng = np.random.default_rng(1)
    nt, dt = 900, 0.001
    ntr = 300
    seis = (rng.normal(0, 1e-3, (ntr, nt))).astype(np.float32)
    on = rng.integers(50, 800, ntr)
    for i, n0 in enumerate(on):
        seis[i, n0:] += np.sin(np.arange(nt - n0) * 0.3).astype(np.float32) * rng.uniform(0.1, 2)
    seis[5] = 0.0                                   # a dead trace
    seis[6, :] = -1.0                               # a negative constant
    for window, cut in ((0.01, 1e-15), (0.005, 0.3), (0.02, 1e-3)):
        got = ctx.auto_pick(seis, dt, window=window, amp_cut=cut)
        ref = _pick_oracle(seis, dt, window, cut)
        assert np.array_equal(got, ref), (window, cut, np.argwhere(got != ref)[:5])
    # closed form: a unit step at n0, iw = 10 (0.01 / 0.001 is 9.9999995 in float)
    step = np.zeros((3, 600), np.float32)
    for i, n0 in enumerate((100, 250, 480)):
        step[i, n0:] = 1.0
    got = ctx.auto_pick(step, 0.001)
    assert np.array_equal(got, (np.array([91, 241, 471]).astype(np.float32) * np.float32(0.001)
                                + np.float32(0.045)))
    # all zero file, and traces shorter than two windows: no pick
    assert np.all(ctx.auto_pick(np.zeros((4, 300), np.float32), 0.001) == 0)
    short = rng.normal(0, 1, (3, 15)).astype(np.float32)
    assert np.array_equal(ctx.auto_pick(short, 0.001), _pick_oracle(short, 0.001))


def test_modelled_picks_fused_equal_two_step(ctx):
    """st_acoustic_picks (modelling + picking without leaving the device)
    equals st_acoustic_shots followed by the oracle picker per shot file."""
    c = orc.acoustic_case(nx=30, ny=26, nz=20, nb=8, nt=300, seed=9, nrec=40)
    _model(ctx, c)
    src = _sources(c, 3, seed=2)
    seis, _, _ = ctx.acoustic_shots(c["wavelet"], src, c["rcv"], group=2)
    picks, _ = ctx.acoustic_picks(c["wavelet"], src, c["rcv"], group=2)
    for s in range(3):
        assert np.array_equal(seis[s], orc.acoustic_shot(c, src=src[s]))
        ref = _pick_oracle(seis[s], c["dt"])
        assert np.array_equal(picks[s], ref)
        assert np.array_equal(ctx.auto_pick(seis[s], c["dt"]), ref)
    assert np.count_nonzero(picks) > 0


def test_bad_arguments(ctx):
    from seismic_tomography_3d_amd.api import StError
    c = orc.acoustic_case()
    with pytest.raises(StError):
        ctx.acoustic_set_model(c["vel"], c["d"], nb=3, dt=c["dt"])
    _model(ctx, c)
    with pytest.raises(StError):
        ctx.acoustic_shots(c["wavelet"], np.array([[-5.0, 10.0, 10.0]]), c["rcv"])
    with pytest.raises(StError):
        ctx.acoustic_shots(c["wavelet"], c["src"][None], c["rcv"], group=-1)
