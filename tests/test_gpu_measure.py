"""GPU parity of the batched waveform measurement (fwat_waveform_adjoint_batch_)
against the oracle's per-trace restatement (orc_waveform_adjoint, orc_filter_adj).

Misfits are double precision with the reference's summation loops on both
sides (bar 1e-12 relative); adjoint sources are single precision (bar 1e-6 of
the trace maximum).
"""
import numpy as np
import pytest

from oracle import oracle as orc
from specfwat_amd import misfit

pytestmark = pytest.mark.gpu


def _traces(nstep=1501, nrec=7, dt=0.01, seed=0):
    rng = np.random.default_rng(seed)
    t = np.arange(nstep) * dt
    obs = np.zeros((nstep, nrec, 3), np.float32)
    syn = np.zeros_like(obs)
    for r in range(nrec):
        for c in range(3):
            f, t0 = rng.uniform(0.3, 1.5), rng.uniform(3, 10)
            w = np.exp(-((t - t0) * f) ** 2) * np.sin(2 * np.pi * f * (t - t0))
            obs[:, r, c] = w
            syn[:, r, c] = np.roll(w, rng.integers(-20, 20)) * rng.uniform(0.8, 1.2)
    return obs, syn, dt


@pytest.mark.parametrize("windows", [None, [[1.0, 8.0], [6.0, 13.01]], [[0.5, 14.5]]])
def test_batch_misfit_and_adjoint_match_oracle(windows):
    obs, syn, dt = _traces()
    nstep = obs.shape[0]
    total, adj = misfit.event_adjoint_sources(obs, syn, dt, windows)
    win = [[0.0, nstep * dt]] if windows is None else windows
    tot = 0.0
    for r in range(obs.shape[1]):
        for c in range(3):
            m, a, _, _ = orc.waveform_adjoint(obs[:, r, c], syn[:, r, c], dt, win)
            tot += m
            assert np.abs(adj[:, r, c] - a).max() <= 1e-6 * max(np.abs(a).max(), 1e-30)
    assert abs(total - tot) <= 1e-12 * tot
    print(f"windows {windows}: misfit {total:.6e}")


def test_filtered_adjoint_sources_match_oracle():
    obs, syn, dt = _traces(seed=3)
    win = [[1.0, 8.0], [6.0, 13.0]]
    total, adj = misfit.event_adjoint_sources(obs, syn, dt, win, min_period=0.4, max_period=12.0)
    for r in range(obs.shape[1]):
        for c in range(3):
            _, a, _, _ = orc.waveform_adjoint(obs[:, r, c], syn[:, r, c], dt, win)
            a = orc.filter_adj(a, dt, np.float32(1 / 12.0), np.float32(1 / 0.4), win)
            assert np.abs(adj[:, r, c] - a).max() <= 1e-6 * np.abs(a).max()
    assert np.all(adj[:int(1.0 / dt)] == 0) and np.all(adj[int(13.0 / dt):] == 0)


def test_batch_matches_the_host_statement():
    obs, syn, dt = _traces(nstep=1000, seed=5)
    total, adj = misfit.event_adjoint_sources(obs, syn, dt, [[0.2, 9.0]])
    m, a, _, _ = misfit.waveform_adjoint_source(obs[:, 2, 1], syn[:, 2, 1], dt, [[0.2, 9.0]])
    np.testing.assert_allclose(adj[:, 2, 1], a, rtol=0, atol=1e-6 * np.abs(a).max())


@pytest.mark.parametrize("imeas", [11, 12, 5])
def test_cc_batch_matches_oracle(imeas):
    rng = np.random.default_rng(imeas)
    nstep, nrec, dt = 1400, 6, 0.01
    t = np.arange(nstep) * dt
    obs = np.zeros((nstep, nrec, 3), np.float32)
    syn = np.zeros_like(obs)
    for r in range(nrec):
        for c in range(3):
            f, t0 = rng.uniform(0.6, 1.5), rng.uniform(4, 8)
            w = np.exp(-((t - t0) * f) ** 2) * np.sin(2 * np.pi * f * (t - t0))
            obs[:, r, c] = w + 0.02 * rng.standard_normal(nstep)
            syn[:, r, c] = np.roll(w, rng.integers(-40, 40)) * rng.uniform(0.6, 1.4)
    syn[:, 0, 0] = rng.standard_normal(nstep)  # one trace below CC_MIN
    win = [[1.0, 7.01], [3.0, 12.5]]  # 601 (odd) and 950 (even) samples
    total, adj, meas = misfit.event_cc_adjoint_sources(obs, syn, dt, win, imeas=imeas)
    tot = 0.0
    for r in range(nrec):
        for c in range(3):
            m, a, o = orc.cc_adjoint(obs[:, r, c], syn[:, r, c], dt, win, imeas=imeas)
            tot += m
            q = meas[3 * r + c]
            np.testing.assert_array_equal(q[:, 6].astype(bool), o["selected"])
            np.testing.assert_allclose(q[:, 4], o["tshift"], rtol=0, atol=1e-12)
            np.testing.assert_allclose(q[:, 5], o["dlna"], rtol=1e-12, atol=1e-14)
            np.testing.assert_allclose(q[:, 2], o["error"], rtol=1e-12)
            np.testing.assert_allclose(q[:, 3], o["cc_max"], rtol=1e-12)
            assert np.abs(adj[:, r, c] - a).max() <= 1e-6 * max(np.abs(a).max(), 1e-30)
    assert not meas[0, :, 6].any() and meas[1:, :, 6].any()
    assert abs(total - tot) <= 1e-12 * max(tot, 1e-30)
    print(f"imeas {imeas}: misfit {total:.6e}, selected {int(meas[:, :, 6].sum())} of {meas.shape[0] * 2}")


def _mt_event(seed, nstep=3000, dt=1 / 64):
    """Traces whose synthetics are shifted by sub-sample, whole-sample, sub-one-sample
    and too-large amounts, plus a noise trace below CC_MIN: every branch of the
    multitaper measurement (MT result, both fall backs, not selected)."""
    rng = np.random.default_rng(seed)
    t = np.arange(nstep) * dt
    shifts = [0.2345, 0.05, -0.0999, 0.3 / 64, -0.3111, 9 / 64, 0.171, -0.14]
    nrec = len(shifts)
    obs = np.zeros((nstep, nrec, 3), np.float32)
    syn = np.zeros_like(obs)
    for r, tau in enumerate(shifts):
        for c in range(3):
            f, t0 = rng.uniform(0.7, 1.3), rng.uniform(18, 24)
            w = lambda x: np.exp(-((t - x) * f) ** 2) * np.sin(2 * np.pi * f * (t - x))
            obs[:, r, c] = w(t0 + tau) * rng.uniform(0.8, 1.25) + 0.005 * rng.standard_normal(nstep)
            syn[:, r, c] = w(t0)
    syn[:, 0, 2] = rng.standard_normal(nstep)  # below CC_MIN
    # two narrow bands (0.6 and 1.6 Hz): the synthetic's spectrum falls below the
    # water level between them and rises past ten times it again, so the
    # frequency-limit search stops, restarts and stops
    nb2 = lambda x: (np.exp(-((t - x) / 2.0) ** 2) * np.sin(2 * np.pi * 0.6 * (t - x))
                     + 0.5 * np.exp(-((t - x - 1.0) / 2.0) ** 2) * np.sin(2 * np.pi * 1.6 * (t - x - 1.0)))
    obs[:, nrec - 1, 2] = nb2(20.0 + shifts[-1]) * 1.1
    syn[:, nrec - 1, 2] = nb2(20.0)
    return obs, syn, dt


@pytest.mark.parametrize("imeas,filt", [(13, False), (14, False), (13, True)])
def test_mt_batch_matches_oracle(imeas, filt):
    obs, syn, dt = _mt_event(imeas)
    win = [[10.0, 32.0], [12.0, 40.03]]  # 1408 (even) and 1922 samples
    total, adj, meas = misfit.event_mt_adjoint_sources(obs, syn, dt, win, 0.5, 10.0, imeas=imeas,
                                                       filter_adj=filt)
    tot = 0.0
    used = []
    for r in range(obs.shape[1]):
        for c in range(3):
            m, a, o = orc.mt_adjoint(obs[:, r, c], syn[:, r, c], dt, win, 0.5, 10.0, imeas=imeas)
            if filt:
                a = orc.filter_adj(a, dt, np.float32(1 / 10.0), np.float32(1 / 0.5), win, 0.05, 1)
            tot += m
            q = meas[3 * r + c]
            np.testing.assert_array_equal(q[:, 8], o["imeas"])
            np.testing.assert_array_equal(q[:, 6].astype(bool), o["selected"])
            np.testing.assert_allclose(q[:, 4], o["tshift"], rtol=0, atol=1e-12)
            np.testing.assert_allclose(q[:, 0], o["misfit"], rtol=1e-8, atol=1e-300)
            np.testing.assert_allclose(q[:, 9], o["misfit_other"], rtol=1e-8, atol=1e-300)
            np.testing.assert_allclose(q[:, 2], o["error"], rtol=1e-12)
            # the reference inverts the adjoint spectrum in single precision, the device in double
            assert np.abs(adj[:, r, c] - a).max() <= 2e-5 * max(np.abs(a).max(), 1e-30)
            used += list(o["imeas"])
    assert abs(total - tot) <= 1e-8 * max(tot, 1e-30)
    assert imeas in used and (imeas - 2) in used and 0 in used  # MT, fall back and unselected windows
    print(f"imeas {imeas} filter {filt}: misfit {total:.6e}, windows by measurement "
          f"{ {int(k): used.count(k) for k in set(used)} }")


@pytest.mark.parametrize("imeas,use_cc_error", [(13, True), (13, False), (14, False)])
def test_mt_batch_with_jackknife_errors_matches_oracle(imeas, use_cc_error):
    """use_mt_error: the leave-one-taper-out errors of calculate_mt_error gate
    acceptance and, without use_cc_error, weight the misfit and adjoint source;
    every trace against the oracle."""
    obs, syn, dt = _mt_event(imeas)
    win = [[10.0, 32.0], [12.0, 40.03]]
    kw = dict(use_mt_error=True, use_cc_error=use_cc_error)
    total, adj, meas = misfit.event_mt_adjoint_sources(obs, syn, dt, win, 0.5, 10.0, imeas=imeas,
                                                       filter_adj=False, **kw)
    tot = 0.0
    used = []
    for r in range(obs.shape[1]):
        for c in range(3):
            m, a, o = orc.mt_adjoint(obs[:, r, c], syn[:, r, c], dt, win, 0.5, 10.0, imeas=imeas, **kw)
            tot += m
            q = meas[3 * r + c]
            np.testing.assert_array_equal(q[:, 8], o["imeas"])
            np.testing.assert_allclose(q[:, 0], o["misfit"], rtol=1e-7, atol=1e-300)
            np.testing.assert_allclose(q[:, 9], o["misfit_other"], rtol=1e-7, atol=1e-300)
            assert np.abs(adj[:, r, c] - a).max() <= 2e-5 * max(np.abs(a).max(), 1e-30)
            used += list(o["imeas"])
    assert abs(total - tot) <= 1e-7 * max(tot, 1e-30)
    assert imeas in used
    print(f"jackknife errors, imeas {imeas}, use_cc_error {use_cc_error}: misfit {total:.6e}, windows by "
          f"measurement { {int(k): used.count(k) for k in set(used)} }")


@pytest.mark.parametrize("filt", [False, True])
def test_ep_batch_matches_oracle(filt):
    """Exponentiated phase over an even and an odd window (the two Nyquist
    weights of the reference's hilbert), every trace against the oracle."""
    rng = np.random.default_rng(41)
    nstep, nrec, dt = 1400, 6, 0.01
    t = np.arange(nstep) * dt
    obs = np.zeros((nstep, nrec, 3), np.float32)
    syn = np.zeros_like(obs)
    for r in range(nrec):
        for c in range(3):
            f, t0 = rng.uniform(0.6, 1.5), rng.uniform(4, 8)
            w = np.exp(-((t - t0) * f) ** 2) * np.sin(2 * np.pi * f * (t - t0))
            obs[:, r, c] = w + 0.02 * rng.standard_normal(nstep)
            syn[:, r, c] = np.roll(w, rng.integers(-40, 40)) * rng.uniform(0.6, 1.4)
    win = [[1.0, 7.01], [3.0, 12.5]]  # 601 (odd) and 950 (even) samples
    kw = dict(min_period=0.5, max_period=10.0) if filt else {}
    total, adj, meas = misfit.event_ep_adjoint_sources(obs, syn, dt, win, **kw)
    tot = 0.0
    for r in range(nrec):
        for c in range(3):
            m, a, o = orc.ep_adjoint(obs[:, r, c], syn[:, r, c], dt, win)
            if filt:
                a = orc.filter_adj(a, dt, np.float32(1 / 10.0), np.float32(1 / 0.5), win, 0.05, 1)
            tot += m
            q = meas[3 * r + c]
            for i, k in enumerate(orc.EP_FIELDS[:4]):
                np.testing.assert_allclose(q[:, i], o[k], rtol=1e-9)
            for i, k in ((4, "residual_real"), (5, "residual_imag")):
                np.testing.assert_allclose(q[:, i], o[k], rtol=0, atol=1e-12)
            # the single-precision bandpass recursion amplifies input rounding ~50-fold
            bar = 3e-5 if filt else 1e-6
            assert np.abs(adj[:, r, c] - a).max() <= bar * max(np.abs(a).max(), 1e-30)
    assert abs(total - tot) <= 1e-9 * tot
    print(f"exponentiated phase, filter {filt}: misfit {total:.6e}")


def test_calculate_adjoint_sources_dispatch():
    """The calculate_adjoint_source mirror runs every type the reference
    dispatches, always with filter_adj; each result is the direct batched call's."""
    obs, syn, dt = _mt_event(3)
    win = [[10.0, 32.0]]
    for im in (1, 4, 5, 11, 12, 13, 14):
        tot, adj, meas = misfit.calculate_adjoint_sources(obs, syn, dt, win, 0.5, 10.0, im)
        assert np.isfinite(tot) and np.isfinite(adj).all() and np.abs(adj).max() > 0
        if im == 13:
            t2, a2, m2 = misfit.event_mt_adjoint_sources(obs, syn, dt, win, 0.5, 10.0, imeas=13)
            assert t2 == tot and np.array_equal(a2, adj)
        print(f"imeasure_type {im}: misfit {tot:.6e}")


@pytest.mark.parametrize("split,len_fac", [(0, 2.0), (1, 2.0), (0, 4.5)])
def test_window_selection_matches_oracle(split, len_fac):
    """win_sel_type%select_windows for every trace: the same windows as the
    oracle (bit-identical correlation sums and decisions), including traces
    split by a time-shift jump and a trace with nothing to select; with phase
    splitting the groups are also cut at the STA/LTA troughs."""
    rng = np.random.default_rng(17)
    nstep, nrec, dt, T = 3000, 5, 0.02, 2.0
    t = np.arange(nstep) * dt
    w = lambda t0, f: np.exp(-((t - t0) * f * 0.6) ** 2) * np.sin(2 * np.pi * f * (t - t0))
    obs = np.zeros((nstep, nrec, 3), np.float32)
    syn = np.zeros_like(obs)
    tp = rng.uniform(8, 14, nrec)
    dist = rng.uniform(60, 120, nrec)
    for r in range(nrec):
        for c in range(3):
            a1, a2 = tp[r] + 4 + rng.uniform(0, 3), tp[r] + 20 + rng.uniform(0, 5)
            syn[:, r, c] = w(a1, 0.5) + 0.7 * w(a2, 0.45)
            jump = rng.choice([0.0, 0.0, 0.4])
            obs[:, r, c] = w(a1 + rng.uniform(-0.1, 0.1), 0.5) + 0.7 * w(a2 + jump, 0.45) \
                + 0.003 * rng.standard_normal(nstep)
    obs[:, 4, 1] = 0.003 * rng.standard_normal(nstep)  # nothing to select
    # sliding_win_len_fac 4.5: windows of 451 samples, 901 lags, two rounds of lag-block pairs per lane
    got = misfit.select_windows(obs, syn, dt, 0.0, tp, dist, T, is_split_phases=bool(split),
                                sliding_win_len_fac=len_fac)
    counts = []
    for r in range(nrec):
        for c in range(3):
            ref, _, _ = orc.win_sel(obs[:, r, c], syn[:, r, c], dt, 0.0, tp[r], dist[r], T, is_split_phases=split,
                                    sliding_win_len_fac=len_fac)
            np.testing.assert_array_equal(got[3 * r + c], ref)
            counts.append(len(ref))
    assert max(counts) >= (2 if len_fac < 3 else 1) and min(counts) == 0
    print(f"split {split} len_fac {len_fac}: windows per trace: {counts}, selection {misfit.last_time_ms():.2f} ms")


@pytest.mark.parametrize("split,nstep", [(0.0, 4000), (1.0, 4000), (1.0, 3999)])
def test_flexwin_matches_oracle(split, nstep):
    """FLEXWIN (WIN_TYPE 2) for every trace: the same windows, weights, cc,
    shifts and dlnA as the oracle, with merging and with interval scheduling,
    on even and odd trace lengths; a noise-only trace selects nothing."""
    rng = np.random.default_rng(23)
    nrec, dt = 4, 0.05
    t = np.arange(nstep) * dt
    w = lambda t0, f, a: a * np.exp(-((t - t0) * f / 2) ** 2) * np.sin(2 * np.pi * f * (t - t0))
    obs = np.zeros((nstep, nrec, 3), np.float32)
    syn = np.zeros_like(obs)
    tp = rng.uniform(20, 30, nrec)
    dist = rng.uniform(200, 260, nrec)
    for r in range(nrec):
        for c in range(3):
            arr = tp[r] + np.array([5.0, 28.0, 62.0]) + rng.uniform(-2, 2, 3)
            f = rng.uniform(0.18, 0.32, 3)
            a = rng.uniform(0.6, 1.6, 3)
            syn[:, r, c] = sum(w(arr[k], f[k], a[k]) for k in range(3))
            obs[:, r, c] = sum(w(arr[k] + rng.uniform(-0.6, 0.6), f[k], a[k] * rng.uniform(0.8, 1.25))
                               for k in range(3)) + 0.01 * rng.standard_normal(nstep)
    obs[:, 3, 2] = 0.01 * rng.standard_normal(nstep)  # nothing fits
    got, ginfo = misfit.flexwin_windows(obs, syn, dt, 0.0, tp, dist, 2.0, 20.0, is_split_phases=bool(split))
    counts = []
    for r in range(nrec):
        for c in range(3):
            ref, rinfo = orc.flexwin(obs[:, r, c], syn[:, r, c], dt, 0.0, tp[r], dist[r], 2.0, 20.0,
                                     is_split_phases=split)
            np.testing.assert_allclose(got[3 * r + c], ref, rtol=0, atol=1e-9)
            np.testing.assert_allclose(ginfo[3 * r + c], rinfo, rtol=1e-9, atol=1e-12)
            counts.append(len(ref))
    assert max(counts) >= (2 if split else 1) and counts[-1] == 0
    print(f"flexwin split {split} nstep {nstep}: windows per trace {counts}, {misfit.last_time_ms():.2f} ms")


@pytest.mark.parametrize("nstep", [3000, 2999])
def test_conv_batch_matches_oracle(nstep):
    """Cross-convolution of every receiver (even and odd trace lengths: the two
    part-0 offsets of myconvolution_dp), per-receiver windows, against the oracle."""
    rng = np.random.default_rng(nstep)
    nrec, dt = 7, 0.02
    t = np.arange(nstep) * dt
    w = lambda t0, f: np.exp(-((t - t0) * f) ** 2) * np.sin(2 * np.pi * f * (t - t0))
    tr = [np.zeros((nstep, nrec), np.float32) for _ in range(4)]
    win = np.zeros((nrec, 2))
    for r in range(nrec):
        t0, f = rng.uniform(15, 30), rng.uniform(0.3, 0.8)
        a, b = rng.uniform(0.5, 1.0), rng.uniform(0.3, 0.8)
        tr[2][:, r] = a * w(t0, f)
        tr[3][:, r] = b * w(t0 + 0.4, f)
        tr[0][:, r] = a * w(t0 + rng.uniform(-0.3, 0.3), f) + 0.01 * rng.standard_normal(nstep)
        tr[1][:, r] = b * w(t0 + 0.5, f) + 0.01 * rng.standard_normal(nstep)
        win[r] = [t0 - 8 - rng.uniform(0, 3), t0 + 10 + rng.uniform(0, 5)]
    total, mf, rs, ar, az = misfit.event_conv_adjoint_sources(*tr, dt, win)
    tot = 0.0
    for r in range(nrec):
        m, res, a, b = orc.conv_adjoint(tr[0][:, r], tr[1][:, r], tr[2][:, r], tr[3][:, r], dt, win[r])
        tot += m
        assert abs(mf[r] - m) <= 1e-9 * m and abs(rs[r] - res) <= 1e-9 * max(abs(res), 1e-12)
        assert np.abs(ar[:, r] - a).max() <= 1e-6 * np.abs(a).max()
        assert np.abs(az[:, r] - b).max() <= 1e-6 * np.abs(b).max()
    assert abs(total - tot) <= 1e-9 * tot
    print(f"cross-convolution nstep {nstep}: misfit {total:.6e}, {misfit.last_time_ms():.2f} ms")
