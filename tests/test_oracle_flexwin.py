"""Pins of the FLEXWIN window selection oracle (orc_flexwin,
src/preproc/window_selector/flexwin.f90:75-734 with interval_scheduling.f90).

  scheduling   schedule_weighted_intervals returns a set of pairwise compatible
               intervals (one ends at or before the next starts) of the largest
               total weight: brute force over every subset of small random sets
  STA/LTA      the two recursive averages are first-order IIR filters:
               scipy.signal.lfilter of the noise-extended envelope, bit for bit
  criteria     each selected window's weight, cc, time shift and dlnA equal a
               numpy computation on the window's samples (np.correlate's first
               maximum, the energy ratio, the zero-edge correlation coefficient)
  structure    merged windows are sorted and separated by at least one sample;
               scheduled windows do not overlap and each passes the data-fit
               thresholds and the minimum length
  known answer three isolated arrivals, data a slightly shifted and scaled copy:
               scheduling keeps one window per arrival, each holding its peak
"""
import itertools
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

signal = pytest.importorskip("scipy.signal")

DT, N = 0.05, 4000
T = np.arange(N) * DT
ARRIVALS = (30.0, 55.0, 90.0)


def _wav(t0, f, a):
    return a * np.exp(-((T - t0) * f / 2) ** 2) * np.sin(2 * np.pi * f * (T - t0))


def _event(seed=5, noise=0.01):
    rng = np.random.default_rng(seed)
    syn = _wav(30, 0.3, 1.0) + _wav(55, 0.2, 1.6) + _wav(90, 0.25, 0.7)
    dat = _wav(30.4, 0.3, 1.1) + _wav(55.9, 0.2, 1.4) + _wav(90.2, 0.25, 0.8) + noise * rng.standard_normal(N)
    return dat, syn


@pytest.mark.parametrize("seed", range(6))
def test_scheduling_is_the_best_compatible_subset(seed):
    rng = np.random.default_rng(seed)
    n = 10
    left = np.round(rng.uniform(0, 50, n), 1)
    right = left + np.round(rng.uniform(1, 15, n), 1)
    w = rng.uniform(0.1, 3.0, n)
    sel = O.schedule_intervals(left, right, w)
    for a, b in zip(sel[:-1], sel[1:]):
        assert right[a] <= left[b]  # time order and compatibility
    best = 0.0
    for k in range(1, n + 1):
        for sub in itertools.combinations(range(n), k):
            o = sorted(sub, key=lambda i: right[i])
            if all(right[o[i]] <= left[o[i + 1]] for i in range(k - 1)):
                best = max(best, w[list(sub)].sum())
    assert abs(w[sel].sum() - best) <= 1e-12 * best


def test_sta_lta_is_two_first_order_recursions():
    rng = np.random.default_rng(2)
    env = np.abs(rng.standard_normal(3000)) * np.exp(-np.linspace(0, 3, 3000))
    dt, mp = 0.05, 2.0
    out = O.sta_lta(env, dt, mp)
    noise = env.max() / 1e5
    x = np.concatenate([np.full(1000, noise), env + noise])
    sta = signal.lfilter([1.0], [1.0, -10.0 ** (-dt / mp)], x)[1000:]
    lta = signal.lfilter([1.0], [1.0, -10.0 ** (-dt / (12.0 * mp))], x)[1000:]
    np.testing.assert_array_equal(out, np.where(lta > 1e-9, sta / lta, noise))


def _numpy_criteria(dat, syn, i0, i1, min_period):
    smax = np.abs(syn).max()
    d, s = dat[i0 - 1:i1] / smax, syn[i0 - 1:i1] / smax
    n = len(d)
    cc = np.correlate(d, s, "full")  # cc[k] = sum d[i + k - (n - 1)] s[i]
    ish = int(np.argmax(cc)) - (n - 1)
    dlna = 0.5 * np.log(np.sum(d * d) / np.sum(s * s))
    ss = np.zeros(n)
    ds = d.copy()
    if ish >= 0:
        ss[ish:] = s[:n - ish]
        ds[:ish] = 0.0
    else:
        ss[:n + ish] = s[-ish:]
        ds[n + ish:] = 0.0
    coef = np.sum(ds * ss) / np.sqrt(np.sum(ds * ds) * np.sum(ss * ss))
    return (i1 - i0) * DT / min_period * coef, coef, ish * DT, dlna


@pytest.mark.parametrize("split", [0.0, 1.0])
def test_window_criteria_equal_numpy_on_the_window(split):
    dat, syn = _event()
    w, info = O.flexwin(dat, syn, DT, 0.0, 25.0, 250.0, 2.0, 20.0, is_split_phases=split)
    assert len(w) >= 1
    for (a, b), inf in zip(w, info):
        i0, i1 = int(round(a / DT)) + 1, int(round(b / DT)) + 1
        ref = _numpy_criteria(dat, syn, i0, i1, 2.0)
        np.testing.assert_allclose(inf, ref, rtol=1e-9, atol=1e-12)


def test_merged_windows_are_sorted_and_separated():
    dat, syn = _event(seed=9, noise=0.03)
    w, _ = O.flexwin(dat, syn, DT, 0.0, 25.0, 250.0, 2.0, 20.0)
    idx = np.round(w / DT).astype(int)
    assert np.all(idx[:, 1] > idx[:, 0])
    assert np.all(idx[1:, 0] > idx[:-1, 1] + 1)


def test_scheduled_windows_pass_every_threshold_and_do_not_overlap():
    dat, syn = _event(seed=11, noise=0.02)
    w, info = O.flexwin(dat, syn, DT, 0.0, 25.0, 250.0, 2.0, 20.0, is_split_phases=1.0)
    c = O.FLEXWIN_DEFAULTS
    assert len(w) >= 2
    assert np.all(w[1:, 0] >= w[:-1, 1] - 1e-9)
    assert np.all(info[:, 1] >= c["threshold_corr"])
    assert np.all(np.abs(info[:, 2]) <= c["threshold_shift_fac"] * 2.0)
    assert np.all(np.abs(info[:, 3]) <= c["threshold_dlna"])
    assert np.all(w[:, 1] - w[:, 0] >= c["min_win_len_fac"] * 2.0 - 1e-9)


def test_one_scheduled_window_per_isolated_arrival():
    dat, syn = _event()
    w, _ = O.flexwin(dat, syn, DT, 0.0, 25.0, 250.0, 2.0, 20.0, is_split_phases=1.0)
    assert len(w) == len(ARRIVALS)
    for (a, b), t in zip(w, ARRIVALS):
        assert a < t < b
        assert all(not (a < u < b) for u in ARRIVALS if u != t)


def test_no_windows_before_the_first_arrival_range():
    """A first arrival after the whole trace: every candidate fails the travel-time test."""
    dat, syn = _event()
    w, _ = O.flexwin(dat, syn, DT, 0.0, 400.0, 250.0, 2.0, 20.0)
    assert len(w) == 0
