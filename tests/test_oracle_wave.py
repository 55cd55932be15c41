"""Oracle for data synthesis: the acoustic modelling loop (src/utils/wave)
and the auto-picker (src/utils/auto_picking), pinned against the
reference's own build and closed forms."""
import numpy as np
import pytest

import oracle_lib as orc
from oracle_lib import P


@pytest.mark.skipif(not orc.ref_available(), reason="reference build absent")
def test_dampers_bit_exact_vs_reference():
    for nb, factor in ((8, 0.028), (50, 0.0045)):
        a = [np.zeros(n, np.float32) for n in (nb, nb * nb, nb ** 3)]
        b = [np.zeros(n, np.float32) for n in (nb, nb * nb, nb ** 3)]
        orc.oracle().or_acoustic_dampers(*(P(x) for x in a), factor, nb)
        orc.ref().ref_dampers(*(P(x) for x in b), factor, nb)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


@pytest.mark.skipif(not orc.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("case", [dict(seed=3), dict(seed=4, nb=5, spacing=(10.0, 12.5, 7.5)),
                                  dict(seed=5, nb=4, nt=60)])
def test_acoustic_shot_bit_exact_vs_reference(case):
    """Every sample of the seismogram and every cell of the final U_fut
    equal the reference's own loop (its in-place damping, where later cells
    read already damped neighbours, included)."""
    c = orc.acoustic_case(**case)
    a, ua = orc.acoustic_shot(c, want_ufut=True)
    b, ub = orc.acoustic_shot(c, orc.ref(), want_ufut=True)
    assert np.abs(a).max() > 0
    assert np.array_equal(a, b)
    assert np.array_equal(ua, ub)
    nb = c["nb"]
    assert np.all(ua[:4] == 0) and np.all(ua[-4:] == 0)          # the shell is never written
    assert np.count_nonzero(ua[nb:-nb, nb:-nb, nb:-nb]) > 0


def test_dampers_closed_form():
    nb, f = 6, 0.05
    d1, d2, d3 = (np.zeros(n, np.float32) for n in (nb, nb * nb, nb ** 3))
    orc.oracle().or_acoustic_dampers(P(d1), P(d2), P(d3), f, nb)
    i = np.arange(nb)
    assert np.allclose(d1, np.exp(-(f * (nb - i)) ** 2), rtol=1e-6)
    D2 = d2.reshape(nb, nb)            # [b][a] = d1[a] + d1[b] - 1
    assert np.allclose(D2, d1[None, :] + d1[:, None] - 1.0, atol=1e-6)
    assert d1[-1] < 1.0 and d1[0] < d1[-1]


def test_auto_pick_step_onset_closed_form():
    """A step of height 1 at sample n0 (zero before): the first window
    whose forward sum sees the step is time = n0 - iw + 1, so the pick is
    (n0 - iw + 1) dt + shift; scaling the traces leaves it unchanged."""
    nt, dt, window = 600, 0.001, 0.01
    iw = int(np.float32(window) / np.float32(dt)) + 1     # 10: the quotient is 9.9999995 in float
    n0 = np.array([100, 250, 333, 480])
    seis = np.zeros((4, nt), np.float32)
    for t, n in enumerate(n0):
        seis[t, n:] = 1.0
    picks = np.zeros(4, np.float32)
    orc.oracle().or_auto_pick(P(seis), 4, nt, dt, window, 1e-15, 0.045, P(picks))
    expect = (n0 - iw + 1).astype(np.float32) * np.float32(dt) + np.float32(0.045)
    assert np.array_equal(picks, expect)
    p2 = np.zeros(4, np.float32)
    orc.oracle().or_auto_pick(P((seis * 7.5).astype(np.float32)), 4, nt, dt, window, 1e-15, 0.045,
                              P(p2))
    assert np.array_equal(p2, picks)


def test_powf_square_folds_to_the_product():
    """The reference writes powf(x, 2.0f) (V^2, dt^2, d^2, the picker's
    (iw + t)^2).  Optimised builds fold it to x * x, while the libm call of
    an unoptimised build rounds some exact ties the other way (4109^2 =
    16883881 lies halfway between two floats): the restatement and the
    device code square, as the reference's optimised build (O2, the build
    the seismogram tests compare against) does."""
    import ctypes
    import ctypes.util
    libm = ctypes.CDLL(ctypes.util.find_library("m"))
    libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]
    libm.powf.restype = ctypes.c_float
    for w in range(0, 4097):
        f = np.float32(w)
        assert np.float32(libm.powf(float(f), 2.0)) == f * f, w
    assert np.float32(libm.powf(4109.0, 2.0)) != np.float32(4109.0) * np.float32(4109.0)
    # the picker restatement squares: a step onset late in a long trace,
    # where (iw + t)^2 passes 2^24, still picks at the closed-form sample
    nt, dt = 6000, np.float32(0.001)
    seis = np.zeros((1, nt), np.float32)
    seis[0, 5000:] = 1.0
    picks = np.zeros(1, np.float32)
    orc.oracle().or_auto_pick(P(seis), 1, nt, float(dt), 0.01, 1e-15, 0.0, P(picks))
    assert picks[0] == np.float32(5000 - 10 + 1) * dt


def _damper_formula(d1, lz, lx, ly):
    """The damper of a cell from the 1D profile alone, in the reference's
    accumulation order (acoustic.cpp:108-136): one layered axis d1[l]; two
    (d1[a] + d1[b]) - 1; three ((d1z + d1x) + (d1x + d1y)) + (d1z + d1y) - 5."""
    f = np.float32
    ins = [l >= 0 for l in (lz, lx, ly)]
    v = [d1[l] if l >= 0 else f(0) for l in (lz, lx, ly)]
    n = sum(ins)
    if n == 0:
        return f(1)
    if n == 1:
        return v[ins.index(True)]
    if n == 2:
        a, b = (v[i] for i in range(3) if ins[i])
        return f(f(a + b) - f(1))
    dz, dx, dy = v
    return f(f(f(f(dz + dx) + f(dx + dy)) + f(dz + dy)) - f(5))


@pytest.mark.parametrize("nb,factor", [(8, 0.028), (13, 0.02), (50, 0.0045)])
def test_damper_from_profile_equals_tables(nb, factor):
    """The device computes a cell's damper from d1 alone; every entry of
    the 1D, 2D and 3D tables must come out the same, bit for bit."""
    d1, d2, d3 = (np.zeros(n, np.float32) for n in (nb, nb * nb, nb ** 3))
    orc.oracle().or_acoustic_dampers(P(d1), P(d2), P(d3), factor, nb)
    for l in range(nb):
        assert _damper_formula(d1, l, -1, -1) == d1[l]
    for a in range(nb):
        for b in range(nb):
            assert _damper_formula(d1, -1, a, b) == d2[a + b * nb]
            assert _damper_formula(d1, a, -1, b) == d2[a + b * nb]
            assert _damper_formula(d1, a, b, -1) == d2[a + b * nb]
    rng = np.random.default_rng(nb)
    idx = rng.integers(0, nb, (4000, 3)) if nb > 16 else np.array(
        [(i, j, k) for i in range(nb) for j in range(nb) for k in range(nb)])
    for i, j, k in idx:
        assert _damper_formula(d1, i, j, k) == d3[i + j * nb + k * nb * nb], (i, j, k)
