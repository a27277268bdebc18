"""Pins of the adjoint-source filter oracle (oracle/measure_oracle.cpp).

  bandpass design   frequency response of the double-precision sections equals
                    scipy.signal's Butterworth bandpass (an independent
                    implementation of the same published design: tangent
                    prewarping, lowpass-to-bandpass, bilinear) to 1e-8
  bandpass apply    forward + reverse direct-form passes equal scipy's sosfilt
                    applied forward and to the reversed output, zero initial
                    state; zero phase (a centred impulse stays symmetric)
  filter_adj        zero outside [min window start, max window end], tapered
                    filtered trace inside
  fft               the discrete Fourier transform (numpy) in both directions
  dpss              scipy.signal.windows.dpss (independent implementation): same
                    tapers and sign convention to 1e-9
  MT measurement    a sub-sample delay is recovered (misfit 0.5 (tau/sigma)^2
                    where the cross-correlation one is quantised); with a pure
                    gain and whole-sample shift dtau = tshift, dlnA = ln A and
                    Parseval closes dt sum(adj_q s) = -ln A / sigma^2; the
                    cross-correlation fall back is bit-identical to orc_cc_adjoint
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

signal = pytest.importorskip("scipy.signal")


@pytest.mark.parametrize("dt,f1,f2", [(0.004, 0.02, 0.5), (0.004, 0.5, 5.0), (0.02, 0.05, 1.0), (0.005, 0.1, 2.0)])
def test_bandpass_design_is_the_butterworth_bandpass(dt, f1, f2):
    sn, sd = O.bandpass_design_d(dt, f1, f2)
    ref = signal.butter(2, [f1, f2], btype="bandpass", fs=1.0 / dt, output="sos")
    w = np.linspace(1e-3, 0.45 / dt, 3000)
    _, h = signal.sosfreqz(np.hstack([sn, sd]), worN=w, fs=1.0 / dt)
    _, hr = signal.sosfreqz(ref, worN=w, fs=1.0 / dt)
    assert np.abs(h - hr).max() < 1e-8
    # unit gain at the centre of the (prewarped) band
    fc = np.arctan(np.sqrt(np.tan(np.pi * f1 * dt) * np.tan(np.pi * f2 * dt))) / (np.pi * dt)
    _, hc = signal.sosfreqz(np.hstack([sn, sd]), worN=[fc], fs=1.0 / dt)
    assert abs(abs(hc[0]) - 1.0) < 1e-9


def test_bandpass_apply_is_forward_backward_filtering():
    dt, f1, f2 = 0.004, 0.5, 5.0
    sn, sd = O.bandpass_design(dt, np.float32(f1), np.float32(f2))
    x = np.random.default_rng(1).standard_normal(4000).astype(np.float32)
    y = O.bandpass_apply(x, sn, sd, zero_phase=True)
    sos = np.hstack([sn, sd]).astype(np.float64)
    ref = signal.sosfilt(sos, x.astype(np.float64))
    ref = signal.sosfilt(sos, ref[::-1])[::-1]
    # single-precision recursion: rounding grows ~1/(1 - pole radius) ~ 50-fold
    assert np.abs(y - ref).max() < 3e-5 * np.abs(ref).max()
    one = O.bandpass_apply(x, sn, sd, zero_phase=False)
    assert np.abs(one - signal.sosfilt(sos, x.astype(np.float64))).max() < 3e-5 * np.abs(ref).max()
    imp = np.zeros(2001, np.float32)
    imp[1000] = 1.0
    r = O.bandpass_apply(imp, sn, sd)
    assert np.abs(r[1000 - 300:1000] - r[1001:1301][::-1]).max() < 2e-4 * np.abs(r).max()


def test_filter_adj_window_and_taper():
    dt = 0.01
    n = 3000
    x = np.random.default_rng(2).standard_normal(n)
    wins = [[5.0, 12.0], [8.0, 20.0]]
    out = O.filter_adj(x, dt, np.float32(1 / 20.0), np.float32(1 / 0.5), wins, 0.05, 1)
    nb, ne = int(5.0 / dt) + 1, int(20.0 / dt)
    assert np.all(out[:nb - 1] == 0) and np.all(out[ne:] == 0)
    sn, sd = O.bandpass_design(dt, np.float32(1 / 20.0), np.float32(1 / 0.5))
    f = O.bandpass_apply(x.astype(np.float32), sn, sd).astype(np.float64)
    np.testing.assert_array_equal(out[nb - 1:ne], O.window_taper(f[nb - 1:ne], 0.05, 1))


def _wavelet(n, dt, t0, f):
    t = np.arange(n) * dt
    return np.exp(-((t - t0) * f) ** 2) * np.sin(2 * np.pi * f * (t - t0))


@pytest.mark.parametrize("k,amp", [(7, 0.8), (-12, 1.3), (0, 1.0)])
def test_cc_measures_a_pure_shift_and_gain(k, amp):
    dt, n = 0.01, 1200
    d = _wavelet(n, dt, 5.0, 1.5)
    s = np.roll(d, k) * amp  # synthetic delayed by k samples
    tot, adj, o = O.cc_adjoint(d, s, dt, [[0.5, 10.0]], imeas=11)
    assert abs(o["tshift"][0] + k * dt) < 1e-12
    assert abs(o["dlna"][0] + np.log(amp)) < 1e-12
    assert abs(o["cc_max"][0] - 1.0) < 1e-12
    assert o["error"][0] == 1.0  # a perfect fit floors sigma_dt at dt_sigma_min
    assert abs(tot - 0.5 * (k * dt) ** 2) < 1e-15
    _, _, q = O.cc_adjoint(d, s, dt, [[0.5, 10.0]], imeas=12)
    assert abs(q["misfit"][0] - 0.5 * (np.log(amp) / 0.5) ** 2) < 1e-12


def test_cc_adjoint_sources_are_normalised_by_the_synthetic_energy():
    """Without taper: simpson(adj_p * ds/dt) = dT / sigma^2 and
    simpson(adj_q * s) = -dlnA / sigma^2 (ds/dt by numpy's gradient)."""
    from specfwat_amd.misfit import simpson
    dt, n = 0.01, 1200
    d = _wavelet(n, dt, 5.0, 1.2)
    s = np.roll(d, 9) * 0.9 + 0.05 * _wavelet(n, dt, 5.5, 2.0)
    win = [[1.0, 9.0]]
    nb, ne = int(1.0 / dt) + 1, int(9.0 / dt)
    _, adj, o = O.cc_adjoint(d, s, dt, win, imeas=11, taper_pct=0.0)
    sw = s[nb - 1:ne]
    ds = np.gradient(sw, dt, edge_order=1)
    sig = o["error"][0]
    assert abs(simpson(adj[nb - 1:ne] * ds, dt) - o["tshift"][0] / sig ** 2) < 1e-10
    assert np.all(adj[:nb - 1] == 0) and np.all(adj[ne:] == 0)
    _, adj, o = O.cc_adjoint(d, s, dt, win, imeas=12, taper_pct=0.0)
    sig = o["error"][0]
    assert abs(simpson(adj[nb - 1:ne] * sw, dt) + o["dlna"][0] / sig ** 2) < 1e-10


def test_cc_selection_thresholds():
    dt, n = 0.01, 1500
    rng = np.random.default_rng(4)
    d = _wavelet(n, dt, 6.0, 1.5)
    noise = rng.standard_normal(n)
    tot, adj, o = O.cc_adjoint(d, noise, dt, [[0.5, 12.0]])
    assert o["cc_max"][0] < 0.7 and not o["selected"][0] and tot == 0.0 and np.all(adj == 0)
    d2 = _wavelet(n, dt, 9.0, 1.5)
    s2 = np.roll(d2, -550)  # synthetic 5.5 s early, both inside the window
    tot, adj, o = O.cc_adjoint(d2, s2, dt, [[0.5, 14.5]])  # 5.5 s > TSHIFT_MAX
    assert abs(o["tshift"][0] - 5.5) < 1e-12 and not o["selected"][0] and tot == 0.0
    tot, adj, o = O.cc_adjoint(d2, s2, dt, [[0.5, 14.5]], tshift_max=7.0)
    assert o["selected"][0] and tot > 0


# ---- multitaper family: FFT, DPSS and the MT measurement of the oracle

def test_fft_is_the_discrete_fourier_transform():
    x = np.random.default_rng(5).standard_normal(256) + 1j * np.random.default_rng(6).standard_normal(256)
    assert np.abs(O.fft(x) - np.fft.fft(x)).max() < 1e-12
    assert np.abs(O.fft(x, inverse=True) - np.fft.ifft(x) * 256).max() < 1e-12
    # single precision (the reference's inverse of the adjoint source)
    assert np.abs(O.fft(x, inverse=True, single=True) - np.fft.ifft(x) * 256).max() < 2e-5


@pytest.mark.parametrize("n,nw", [(500, 4.0), (1537, 4.0), (2048, 2.5)])
def test_dpss_equals_scipy(n, nw):
    """Independent implementation (scipy.signal.windows.dpss, unit norm): same
    tapers, same sign convention (symmetric tapers sum positive,
    antisymmetric ones start positive)."""
    ref = signal.windows.dpss(n, nw, 5)
    v = O.dpss(n, nw, 5)
    assert np.abs(v - ref).max() < 1e-9
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-12)


def _mt_case(tau, amp=1.0, f=1.0, dt=1 / 64, n=3000):
    t = np.arange(n) * dt
    w = lambda t0: np.exp(-((t - t0) * f) ** 2) * np.sin(2 * np.pi * f * (t - t0))
    return w(20.0 + tau) * amp, w(20.0), dt


@pytest.mark.parametrize("tau", [0.2345, 0.05, -0.0999])
def test_mt_traveltime_recovers_a_sub_sample_shift(tau):
    """Data = synthetic delayed by tau (not a whole number of samples): the
    cross-correlation measurement is quantised to the sample, the multitaper one
    measures the residual shift per frequency, misfit 0.5 (tau / sigma)^2."""
    d, s, dt = _mt_case(tau)
    tot, adj, o = O.mt_adjoint(d, s, dt, [[10.0, 32.0]], 0.5, 10.0, imeas=13)
    assert o["imeas"][0] == 13 and o["selected"][0] and o["error"][0] == 1.0
    assert abs(o["misfit"][0] - 0.5 * tau ** 2) < 5e-3 * 0.5 * tau ** 2  # taper edges bias the estimate slightly
    assert abs(o["tshift"][0] - round(tau / dt) * dt) < 1e-12  # the CC stage's shift
    _, _, c = O.cc_adjoint(d, s, dt, [[10.0, 32.0]], imeas=11)
    assert abs(c["misfit"][0] - 0.5 * tau ** 2) > 10 * abs(o["misfit"][0] - 0.5 * tau ** 2)


def test_mt_adjoint_sources_close_the_parseval_identities():
    """Pure gain A and whole-sample shift, no taper: dtau(f) = tshift and
    dlnA(f) = ln A in the band, and Parseval gives dt * sum(adj_q * s) =
    -ln A / sigma^2 (single-precision inverse transform, as the reference);
    dt * sum(adj_p * ds/dt) = tshift / sigma^2 up to the tapers' own slope."""
    dt, n, k, A = 1 / 64, 3000, 9, 1.3
    _, s, _ = _mt_case(0.0)
    d = np.roll(s, k) * A
    nb, ne = int(10.0 / dt) + 1, int(32.0 / dt)
    sw = s[nb - 1:ne]
    _, adj, o = O.mt_adjoint(d, s, dt, [[10.0, 32.0]], 0.5, 10.0, imeas=14, taper_pct=0.0)
    assert o["imeas"][0] == 14 and abs(o["dlna"][0] - np.log(A)) < 1e-12
    assert abs(dt * np.sum(adj[nb - 1:ne] * sw) + np.log(A) / o["error"][0] ** 2) < 1e-6
    assert np.all(adj[:nb - 1] == 0) and np.all(adj[ne:] == 0)
    assert abs(o["misfit"][0] - 0.5 * (np.log(A) / 0.5) ** 2) < 1e-9
    _, adj, o = O.mt_adjoint(d, s, dt, [[10.0, 32.0]], 0.5, 10.0, imeas=13, taper_pct=0.0)
    assert o["imeas"][0] == 13 and abs(o["misfit"][0] - 0.5 * (k * dt) ** 2) < 1e-12
    ds = np.gradient(sw, dt, edge_order=1)
    assert abs(dt * np.sum(adj[nb - 1:ne] * ds) - k * dt) < 0.05 * k * dt


def test_mt_falls_back_to_the_cross_correlation_measurement():
    """|tshift| <= dt fails check_time_series_acceptability: the window takes the
    cross-correlation misfit and adjoint source (imeas 11), bit for bit."""
    d, s, dt = _mt_case(0.3 / 64)  # under one sample
    tot, adj, o = O.mt_adjoint(d, s, dt, [[10.0, 32.0]], 0.5, 10.0, imeas=13)
    ct, cadj, c = O.cc_adjoint(d, s, dt, [[10.0, 32.0]], imeas=11)
    assert o["imeas"][0] == 11 and tot == ct
    np.testing.assert_array_equal(adj, cadj)
    # a shift beyond half the shortest period fails the per-frequency test: fall back as well
    d, s, dt = _mt_case(-0.3111)
    _, _, o = O.mt_adjoint(d, s, dt, [[10.0, 32.0]], 0.5, 10.0, imeas=13)
    assert o["imeas"][0] == 11


def test_ccc_closed_forms():
    """CCCMisfit: s = A d gives coef 1, misfit 0 and a zero adjoint source; for
    s = d + e the coefficient is the normalised inner product (numpy) and the
    adjoint source -(d - (p/E_s) s)/sqrt(E_d E_s) is orthogonal to s."""
    dt, n = 0.01, 1200
    d = _wavelet(n, dt, 5.0, 1.5)
    tot, adj, o = O.cc_adjoint(d, 1.7 * d, dt, [[0.5, 10.0]], imeas=5, taper_pct=0.0)
    assert abs(o["cc_max"][0] - 1.0) < 1e-14 and abs(tot) < 1e-14 and np.abs(adj).max() < 1e-12
    e = 0.3 * _wavelet(n, dt, 5.4, 2.5)
    s = d + e
    nb, ne = int(0.5 / dt) + 1, int(10.0 / dt)
    dw, sw = d[nb - 1:ne], s[nb - 1:ne]
    coef = np.dot(dw, sw) / np.sqrt(np.dot(dw, dw) * np.dot(sw, sw))
    tot, adj, o = O.cc_adjoint(d, s, dt, [[0.5, 10.0]], imeas=5, taper_pct=0.0)
    assert abs(o["residual"][0] - coef) < 1e-13 and abs(tot - (1 - coef)) < 1e-13
    assert abs(np.dot(adj[nb - 1:ne], sw)) < 1e-10 * np.abs(adj).max() * np.abs(sw).sum()
    tot, adj, o = O.cc_adjoint(d, -s, dt, [[0.5, 10.0]], imeas=5)
    assert not o["selected"][0] and tot == 0.0 and np.all(adj == 0)


# ---- exponentiated phase

def test_hilbert_is_the_analytic_signal():
    """Even length: scipy.signal.hilbert on the same padded length; odd length:
    the same imaginary part, and a real part without the Nyquist component (the
    reference's weights follow the parity of the input length)."""
    x = np.random.default_rng(7).standard_normal(1000)
    assert np.abs(O.hilbert(x) - signal.hilbert(x, N=1024)[:1000]).max() < 1e-12
    x = np.random.default_rng(8).standard_normal(1001)
    h, r = O.hilbert(x), signal.hilbert(x, N=1024)[:1001]
    assert np.abs(h.imag - r.imag).max() < 1e-12
    xny = np.fft.fft(x, 1024)[512].real / 1024 * (-1.0) ** np.arange(1001)
    assert np.abs(h.real - (x - xny)).max() < 1e-12


def test_exponentiated_phase_gain_invariance_and_gradient():
    """The measurement ignores amplitude (d = A s: misfit and adjoint source
    vanish); with the envelope well above the water level the adjoint source is
    the gradient of the misfit: sum(adj * ds) dt equals the finite difference."""
    dt, n = 0.01, 2000
    t = np.arange(n) * dt
    s = (1 + 0.3 * np.sin(0.5 * t)) * np.sin(2 * np.pi * t)
    tot, adj, o = O.ep_adjoint(1.7 * s, s, dt, [[2.0, 14.0]])
    assert tot < 1e-25 and np.abs(adj).max() < 1e-13
    d = (1.2 + 0.2 * np.sin(0.4 * t)) * np.sin(2 * np.pi * (t - 0.07))
    tot, adj, o = O.ep_adjoint(d, s, dt, [[2.0, 14.0]], taper_pct=0.0, wtr_env=1e-6)
    assert abs(o["misfit"][0] - o["misfit_real"][0] - o["misfit_imag"][0]) < 1e-15
    ds = np.sin(2 * np.pi * 0.8 * t + 0.3) * np.exp(-((t - 8) / 3) ** 2)
    eps = 1e-6
    mp, _, _ = O.ep_adjoint(d, s + eps * ds, dt, [[2.0, 14.0]], taper_pct=0.0, wtr_env=1e-6)
    mm, _, _ = O.ep_adjoint(d, s - eps * ds, dt, [[2.0, 14.0]], taper_pct=0.0, wtr_env=1e-6)
    fd = (mp - mm) / (2 * eps)
    assert abs(np.sum(adj * ds) * dt - fd) < 5e-3 * abs(fd)


# ---- window selection (win_sel_type, no phase splitting)

def _hann_full(n):
    """window_taper(., 1.0, 1) weights written out: frac = n/2, Hann halves."""
    frac = n // 2
    w = np.ones(n)
    r = np.arange(frac)
    w[:frac] = 0.5 - 0.5 * np.cos(2 * np.pi * r / (2 * frac - 1))
    w[n - frac:] = 0.5 - 0.5 * np.cos(2 * np.pi * (frac + r) / (2 * frac - 1))
    return w


def test_sliding_correlation_is_numpy_correlate():
    """Every window start: lag of np.correlate's first maximum on the tapered
    windows and the coefficient of d with s shifted by it."""
    dt, n = 0.05, 320
    rng = np.random.default_rng(9)
    syn = np.convolve(rng.standard_normal(n), np.hanning(15), "same")
    dat = np.roll(syn, 2) + 0.1 * rng.standard_normal(n)
    tw, cc, ts = O.win_sel(dat, syn, dt, 0.0, 2.0, 40.0, 1.0)
    nlen = int(2.0 * 1.0 / dt) + 1
    w = _hann_full(nlen)
    scale = np.abs(syn).max()
    for i in range(0, n - nlen, 7):
        d = dat[i:i + nlen] / scale * w
        s = syn[i:i + nlen] / scale * w
        ish = int(np.argmax(np.correlate(d, s, "full"))) - (nlen - 1)
        assert ts[i] == ish * dt
        ss = np.zeros(nlen)
        src = np.arange(nlen) - ish
        ok = (src >= 0) & (src < nlen)
        ss[ok] = s[src[ok]]
        dd = d.copy()
        if ish > 0:
            dd[:ish] = 0
        elif ish < 0:
            dd[nlen + ish:] = 0
        assert abs(cc[i] - np.dot(dd, ss) / np.sqrt(np.dot(dd, dd) * np.dot(ss, ss))) < 1e-12


def test_windows_follow_the_selection_rules():
    """One clean arrival selected inside [tp - 1.5 T + t0, dist / v + 1.5 T + t0];
    a time-shift jump between two arrivals splits them into two windows that
    keep nint(jump_fac T / dt) samples away from the jump."""
    dt, n, T = 0.02, 3000, 2.0
    t = np.arange(n) * dt
    rng = np.random.default_rng(10)
    w = lambda t0, f: np.exp(-((t - t0) * f * 0.6) ** 2) * np.sin(2 * np.pi * f * (t - t0))
    syn = w(20, 0.5)
    dat = syn + 0.002 * rng.standard_normal(n)
    tw, cc, ts = O.win_sel(dat, syn, dt, 0.0, 15.0, 72.0, T)
    assert len(tw) == 1 and 15.0 - 1.5 * T <= tw[0, 0] < 20.0 < tw[0, 1] <= 72.0 / 2.4 + 1.5 * T
    syn2 = w(20, 0.5) + w(38, 0.5)
    dat2 = w(20, 0.5) + w(38.4, 0.5) + 0.002 * rng.standard_normal(n)
    tw2, cc2, ts2 = O.win_sel(dat2, syn2, dt, 0.0, 15.0, 100.0, T)
    assert len(tw2) == 2 and tw2[0, 1] < tw2[1, 0]
    nlen = int(2.0 * T / dt) + 1
    jumps = np.nonzero(np.abs(np.diff(ts2)) > 0.1 * T)[0] + 1
    assert len(jumps) > 0
    tc = (np.arange(len(ts2)) + nlen / 2.0) * dt
    nj = int(round(0.1 * T / dt))
    for a, b in tw2:
        ia, ib = int(round(a / dt - nlen / 2.0)), int(round(b / dt - nlen / 2.0))
        assert np.all(np.abs(jumps[:, None] - np.arange(ia, ib + 1)[None, :]) > nj)
        assert abs(tc[ia] - a) < 1e-9 and abs(tc[ib] - b) < 1e-9


def test_phase_splitting_cuts_at_the_sta_lta_trough():
    """Two arrivals that merge into one window are split (is_split_phases) into
    two touching windows whose common edge lies between the arrivals; a single
    arrival is left whole."""
    dt, n, T = 0.02, 3000, 2.0
    t = np.arange(n) * dt
    w = lambda t0, f: np.exp(-((t - t0) * f * 0.6) ** 2) * np.sin(2 * np.pi * f * (t - t0))
    rng = np.random.default_rng(3)
    syn = w(20, 0.5) + 0.8 * w(27, 0.5)
    dat = syn + 0.002 * rng.standard_normal(n)
    one, _, _ = O.win_sel(dat, syn, dt, 0.0, 15.0, 90.0, T)
    two, _, _ = O.win_sel(dat, syn, dt, 0.0, 15.0, 90.0, T, is_split_phases=1)
    assert len(one) == 1 and len(two) == 2
    assert two[0, 0] == one[0, 0] and two[1, 1] == one[0, 1] and two[0, 1] == two[1, 0]
    assert 20.0 < two[0, 1] < 27.0
    syn1 = w(20, 0.5)
    dat1 = syn1 + 0.002 * rng.standard_normal(n)
    a, _, _ = O.win_sel(dat1, syn1, dt, 0.0, 15.0, 90.0, T)
    b, _, _ = O.win_sel(dat1, syn1, dt, 0.0, 15.0, 90.0, T, is_split_phases=1)
    np.testing.assert_array_equal(a, b)


# ---- cross-convolution (waveform_conv_misfit)

def _conv_case(n=600, dt=0.05):
    t = np.arange(n) * dt
    w = lambda t0, f: np.exp(-((t - t0) * f) ** 2) * np.sin(2 * np.pi * f * (t - t0))
    return w(8, 0.5), 0.7 * w(8.3, 0.5), w(8.4, 0.5), 0.7 * w(8.6, 0.5), t, dt


def test_cross_convolution_is_source_independent():
    """Data = the synthetic convolved with any common source wavelet on both
    components: c1 = c2, misfit ~ 0; a different model gives a finite misfit."""
    sr, sz, dr, dz, t, dt = _conv_case()
    n = len(t)
    k = np.exp(-((t - 2) / 0.3) ** 2)
    cr, cz = np.convolve(sr, k)[:n], np.convolve(sz, k)[:n]
    m0, _, _, _ = O.conv_adjoint(cr, cz, sr, sz, dt, [2.0, 25.0])
    m1, _, _, _ = O.conv_adjoint(dr, dz, sr, sz, dt, [2.0, 25.0])
    assert m0 < 1e-15 * m1 and m1 > 1e-3
    m2, r2, a2, b2 = O.conv_adjoint(sr, sz, sr, sz, dt, [2.0, 25.0])
    assert m2 == 0.0 and r2 == 0.0 and not a2.any() and not b2.any()


@pytest.mark.parametrize("n", [600, 601])
def test_cross_convolution_matches_numpy_convolve(n):
    """myconvolution_dp part 0 (samples from n/2 + 1, or n/2 for odd n, 1-based)
    and the misfit / adjoint construction written with numpy's convolve."""
    sr, sz, dr, dz, t, dt = _conv_case(n)
    from specfwat_amd.misfit import simpson, window_taper
    ind = n // 2 + 1 if n % 2 == 0 else n // 2
    mid = lambda x, y: np.convolve(x, y)[ind - 1:ind - 1 + n]
    cd = (mid(sr, dz) - mid(dr, sz)) * dt
    nb, ne = int(2.0 / dt) + 1, int(25.0 / dt)
    ar = (mid(dz[::-1], cd) * dt)[nb - 1:ne]
    az = (mid(-dr[::-1], cd) * dt)[nb - 1:ne]
    c = window_taper(cd[nb - 1:ne], 0.05, 1)
    m, r, a, b = O.conv_adjoint(dr, dz, sr, sz, dt, [2.0, 25.0])
    assert abs(m - 0.5 * simpson(c ** 2, dt)) < 1e-12 * m and abs(r - c.sum() / len(c)) < 1e-12
    assert np.abs(a[nb - 1:ne] - window_taper(ar, 0.05, 1)).max() < 1e-12 * np.abs(ar).max()
    assert np.abs(b[nb - 1:ne] - window_taper(az, 0.05, 1)).max() < 1e-12 * np.abs(az).max()
