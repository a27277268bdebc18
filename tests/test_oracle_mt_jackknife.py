"""Pins of the multitaper jackknife errors (calculate_mt_error,
src/preproc/adjoint_sources/mt_tt_misfit.f90:302-409) in the oracle's
orc_mt_adjoint with use_mt_error.

  estimates    every leave-one-taper-out estimate is rebuilt here from
               independent parts (scipy's DPSS tapers, numpy's FFT, numpy's
               arctan2 / log): the errors the oracle reports per frequency
               equal the jackknife spread of those estimates
  exact model  data = a whole-sample shift and gain of the synthetic: every
               taper subset measures the same dtau and dlnA, so the errors
               vanish to rounding
  selection    with use_cc_error the errors change only the acceptance test:
               misfits equal those without use_mt_error; an err_fac that no
               error can pass rejects the window to the cross-correlation
               measurement (imeas 11), bit for bit
  weighting    without use_cc_error the errors (with their water level) weight
               the misfit, which then differs from the unweighted one
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

signal = pytest.importorskip("scipy.signal")

DT, N = 1 / 64, 3000
WIN = [10.0, 32.0]


def _case(tau=0.2345, noise=0.02, seed=3):
    rng = np.random.default_rng(seed)
    t = np.arange(N) * DT
    w = lambda t0: np.exp(-(t - t0) ** 2) * np.sin(2 * np.pi * (t - t0))
    d = 1.1 * w(20.0 + tau) + noise * rng.standard_normal(N)
    return d, w(20.0)


def _leave_one_out(d, s, o, ntaper=5, nw=4.0, wl=1e-10, taper_pct=0.05):
    """(dtau, dlnA) of every leave-one-out taper set, [ntaper, 2, nf], from
    scipy's DPSS and numpy's FFT; phases unwrapped by nothing (the test case
    keeps them smooth in the band)."""
    nb, ne = int(WIN[0] / DT) + 1, int(WIN[1] / DT)
    nl = ne - nb + 1
    sv = O.window_taper(s[nb - 1:ne], taper_pct, 1)
    ish = int(o["tshift"][0] / DT)
    nbd = max(1, nb + ish)
    dm = O.window_taper(d[nbd - 1:nbd - 1 + nl] * np.exp(-o["dlna"][0]), taper_pct, 1)
    tap = signal.windows.dpss(nl, nw, ntaper) * np.sqrt(nl)
    nf = 2 ** int(np.frexp(np.float32(N))[1])
    D = np.fft.fft(dm * tap, nf, axis=1) * DT
    S = np.fft.fft(sv * tap, nf, axis=1) * DT
    P, B = D * np.conj(S), (S * np.conj(S)).real
    f = np.fft.fftfreq(nf, DT)
    out = np.zeros((ntaper, 2, nf))
    for q in range(ntaper):
        keep = [j for j in range(ntaper) if j != q]
        top, bot = P[keep].sum(0), B[keep].sum(0)
        lev = np.abs(bot[:nf // 2 + 1]).max() * wl ** 2
        tf = top / np.where(np.abs(bot) >= lev, bot, bot + lev)
        phi = np.arctan2(tf.imag, tf.real)
        with np.errstate(divide="ignore", invalid="ignore"):  # DC and frequencies with tf = 0: outside the band
            out[q, 0] = -phi / (2 * np.pi * f) + o["tshift"][0]
            out[q, 1] = np.log(np.abs(tf)) + o["dlna"][0]
        out[q, 0, 0] = o["tshift"][0]
    return out, nf


def _jackknife(full, loo):
    n = loo.shape[0]
    ave = sum(n * full - (n - 1) * loo[q] for q in range(n)) / n
    return np.sqrt(sum((loo[q] - ave) ** 2 for q in range(n)) / (n * (n - 1)))


def test_errors_are_the_spread_of_independent_leave_one_out_estimates():
    d, s = _case()
    _, _, o, err = O.mt_adjoint(d, s, DT, [WIN], 0.5, 10.0, imeas=13, errors=True, use_mt_error=True,
                                use_cc_error=True)
    assert o["imeas"][0] == 13
    loo, nf = _leave_one_out(d, s, o)
    band = np.nonzero(err[0, 0])[0]
    lo, hi = band[0], band[-1]
    assert hi - lo > 20  # a real band, not a stray frequency
    # the phases of every subset stay clear of the cycle-skipping threshold in the band
    for q in range(5):
        assert np.abs(np.diff(loo[q, 0, lo:hi + 1] * 2 * np.pi * np.fft.fftfreq(nf, DT)[lo:hi + 1])).max() < 1.0
    # pseudo-values from the five-taper estimate
    full = _all_taper_estimate(d, s, o, nf)
    exp = np.stack([_jackknife(full[c], loo[:, c]) for c in range(2)])
    for c in range(2):
        ref = exp[c, lo:hi + 1]
        np.testing.assert_allclose(err[0, c, lo:hi + 1], ref, rtol=2e-6, atol=1e-9 * ref.max())
    assert np.all(err[0, :, :lo] == 0) and np.all(err[0, :, hi + 1:] == 0)


def _all_taper_estimate(d, s, o, nf, ntaper=5, nw=4.0, wl=1e-10, taper_pct=0.05):
    nb, ne = int(WIN[0] / DT) + 1, int(WIN[1] / DT)
    nl = ne - nb + 1
    sv = O.window_taper(s[nb - 1:ne], taper_pct, 1)
    nbd = max(1, nb + int(o["tshift"][0] / DT))
    dm = O.window_taper(d[nbd - 1:nbd - 1 + nl] * np.exp(-o["dlna"][0]), taper_pct, 1)
    tap = signal.windows.dpss(nl, nw, ntaper) * np.sqrt(nl)
    D = np.fft.fft(dm * tap, nf, axis=1) * DT
    S = np.fft.fft(sv * tap, nf, axis=1) * DT
    top, bot = (D * np.conj(S)).sum(0), (S * np.conj(S)).real.sum(0)
    lev = np.abs(bot[:nf // 2 + 1]).max() * wl ** 2
    tf = top / np.where(np.abs(bot) >= lev, bot, bot + lev)
    f = np.fft.fftfreq(nf, DT)
    with np.errstate(divide="ignore", invalid="ignore"):
        dtau = -np.arctan2(tf.imag, tf.real) / (2 * np.pi * f) + o["tshift"][0]
        dlna = np.log(np.abs(tf)) + o["dlna"][0]
    dtau[0] = o["tshift"][0]
    return np.stack([dtau, dlna])


def test_errors_vanish_for_an_exact_shift_and_gain():
    _, s = _case(noise=0.0)
    d = 1.3 * np.roll(s, 9)
    _, _, o, err = O.mt_adjoint(d, s, DT, [WIN], 0.5, 10.0, imeas=14, taper_pct=0.0, errors=True,
                                use_mt_error=True, use_cc_error=True)
    assert o["imeas"][0] == 14
    assert np.abs(err[0]).max() < 1e-9


def test_errors_only_gate_acceptance_with_cc_error_weights():
    d, s = _case()
    t0, a0, o0 = O.mt_adjoint(d, s, DT, [WIN], 0.5, 10.0, imeas=13)
    t1, a1, o1 = O.mt_adjoint(d, s, DT, [WIN], 0.5, 10.0, imeas=13, use_mt_error=True)
    assert o1["imeas"][0] == 13 and t1 == t0
    np.testing.assert_array_equal(a1, a0)
    # no error can be below a period / 1e12: the window falls back to cross-correlation
    t2, a2, o2 = O.mt_adjoint(d, s, DT, [WIN], 0.5, 10.0, imeas=13, use_mt_error=True, err_fac=1e12)
    tc, ac, _ = O.cc_adjoint(d, s, DT, [WIN], imeas=11)
    assert o2["imeas"][0] == 11 and t2 == tc
    np.testing.assert_array_equal(a2, ac)


def test_errors_weight_the_misfit_without_cc_error():
    d, s = _case()
    t0, _, o0 = O.mt_adjoint(d, s, DT, [WIN], 0.5, 10.0, imeas=13, use_cc_error=False)
    t1, a1, o1, err = O.mt_adjoint(d, s, DT, [WIN], 0.5, 10.0, imeas=13, use_cc_error=False, use_mt_error=True,
                                   errors=True)
    assert o0["imeas"][0] == 13 and o1["imeas"][0] == 13
    assert np.isfinite(t1) and np.all(np.isfinite(a1)) and t1 > 0
    assert abs(t1 - t0) > 1e-3 * t0
