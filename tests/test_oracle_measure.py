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
