"""Waveform-difference misfit and adjoint sources.

WaveformMisfit%calc_adjoint_source (src/preproc/adjoint_sources/waveform_misfit.f90:
15-64) with get_window_info / window_taper (src/shared/signal.f90:419-508) and
simpson (src/shared/utils.f90:1630-1663), then filter_adj
(src/preproc/adjoint_sources/adjoint_source.f90:92-114: bandpass_dp's zero-phase
Butterworth bandpass, signal.f90:7-36, and a taper over the windows' span).

event_adjoint_sources measures every trace of an event in one device pass
(fwat_waveform_adjoint_batch_; double-precision misfits as the reference,
single-precision filter as SAC's xapiir).  The single-trace functions below are
the host statement of the same arithmetic.
"""
import numpy as np

from . import _lib


def simpson(y, dx):
    y = np.asarray(y, np.float64)
    n = len(y)
    if n < 2:
        return 0.0
    if n % 2 == 0:
        s = y[0] + y[n - 2] + np.sum(4.0 * y[1:n - 2:2] + 2.0 * y[2:n - 1:2])
        return (dx / 3.0) * s + 0.5 * dx * (y[n - 2] + y[n - 1])
    s = y[0] + y[n - 1] + np.sum(4.0 * y[1:n - 1:2]) + np.sum(2.0 * y[2:n - 2:2])
    return (dx / 3.0) * s


def window_taper(sig, taper_percentage, itaper_type):
    sig = np.array(sig, np.float64)
    npts = len(sig)
    if taper_percentage < 0.0 or taper_percentage > 1.0:
        raise ValueError("taper_percentage must be 0 < % < 1")
    if taper_percentage in (0.0, 1.0):
        frac = int(npts * taper_percentage / 2.0)
    else:
        frac = int(npts * taper_percentage / 2.0 + 0.5)
    if frac <= 0:
        return sig
    idx2 = npts - frac

    def taper(off):
        r = np.arange(off, off + frac, dtype=np.float64)
        den = 2 * frac - 1
        if itaper_type == 1:
            return 0.5 - 0.5 * np.cos(2.0 * np.pi * r / den)
        if itaper_type == 2:
            return 0.54 - 0.46 * np.cos(2.0 * np.pi * r / den)
        if itaper_type == 3:
            return np.cos(np.pi * r / den - np.pi / 2.0)
        if itaper_type == 4:
            return 1.0 - np.cos(np.pi * r / den) ** 10
        raise ValueError("Taper type not supported!")

    sig[:frac] *= taper(0)
    sig[idx2:] *= taper(frac)
    return sig


def get_window_info(window, dt):
    nb = int(window[0] / dt) + 1
    ne = int(window[1] / dt)
    nlen = ne - nb + 1
    if nlen <= 0:
        raise ValueError("Invalid window length.")
    return nb, ne, nlen


def waveform_adjoint_source(dat, syn, dt, windows, taper_percentage=0.05, itaper_type=1):
    """Returns (total_misfit, adj_src[n], misfits[nwin], residuals[nwin])."""
    dat = np.asarray(dat, np.float64)
    syn = np.asarray(syn, np.float64)
    windows = np.reshape(np.asarray(windows, np.float64), (-1, 2))
    adj = np.zeros(len(dat))
    misfits = np.zeros(len(windows))
    residuals = np.zeros(len(windows))
    total = 0.0
    for w, win in enumerate(windows):
        nb, ne, nl = get_window_info(win, dt)
        s = window_taper(syn[nb - 1:ne], taper_percentage, itaper_type)
        d = window_taper(dat[nb - 1:ne], taper_percentage, itaper_type)
        tw = window_taper(s - d, taper_percentage, itaper_type)
        mf = 0.5 * simpson((s - d) ** 2, dt)
        misfits[w] = mf
        residuals[w] = np.sum(s - d) / nl
        total += mf
        adj[nb - 1:ne] = tw
    return total, adj, misfits, residuals


IORD = 2  # config.f90:76


def bandpass_sos(dt, f1, f2):
    """Second-order sections [nsect, 6] (b0 b1 b2 1 a1 a2, single precision) of
    the filter bandpass_dp asks of SAC's xapiir: Butterworth prototype of
    order IORD, band edges prewarped to w = tan(pi f dt), lowpass-to-bandpass
    s -> (s^2 + w0^2) / (B s), bilinear s = (1 - 1/z) / (1 + 1/z); one section
    per bandpass pole pair.  Coefficients are formed in double."""
    wl, wh = np.tan(np.pi * float(f1) * dt), np.tan(np.pi * float(f2) * dt)
    if not 0.0 < wl < wh:
        raise ValueError("bandpass edges must satisfy 0 < f1 < f2 < Nyquist")
    w0sq, B = wl * wh, wh - wl
    rows = []
    for k in range(IORD // 2):
        p = np.exp(1j * np.pi * (2 * k + IORD + 1) / (2 * IORD))  # upper prototype pole
        disc = np.sqrt(p * p * B * B - 4.0 * w0sq + 0j)
        for q in ((p * B + disc) / 2.0, (p * B - disc) / 2.0):
            a1, a0 = -2.0 * q.real, abs(q) ** 2
            D = 1.0 + a1 + a0
            rows.append([B / D, 0.0, -B / D, 1.0, (2.0 * a0 - 2.0) / D, (1.0 - a1 + a0) / D])
    return np.array(rows, np.float32)


def event_adjoint_sources(obs, syn, dt, windows=None, taper_percentage=0.05, itaper_type=1,
                          min_period=None, max_period=None):
    """obs/syn (NSTEP, nrec, 3) -> (misfit, adjoint traces float32 (NSTEP, nrec, 3)),
    every trace in one device pass.  Default window: the whole trace.  With
    min_period and max_period the adjoint sources go through filter_adj (band
    1/max_period .. 1/min_period, the reference's cfg%min_period / max_period)."""
    nstep, ntrace, win, o, s, sos = _event_inputs(obs, syn, dt, windows, min_period, max_period)
    adj = np.zeros(s.shape, np.float32)
    mf = np.zeros(ntrace)
    _lib.call("fwat_waveform_adjoint_batch_", nstep, ntrace, float(dt), o, s, win.shape[0], win,
              float(taper_percentage), int(itaper_type), sos.shape[0], sos if sos.size else np.zeros(6, np.float32),
              mf, None, adj)
    total = 0.0
    for v in mf:  # trace order, as the reference's receiver/component loop
        total += v
    return total, adj


IMEAS_CCC, IMEAS_CC_TT, IMEAS_CC_DLNA = 5, 11, 12  # adj_config
CC_DEFAULTS = dict(dt_sigma_min=1.0, dlna_sigma_min=0.5, tshift_min=-5.0, tshift_max=5.0, dlna_min=-1.5,
                   dlna_max=1.5, cc_min=0.7)  # ADJConfig defaults, adj_config.f90:18-23


def _event_inputs(obs, syn, dt, windows, min_period, max_period):
    nstep = syn.shape[0]
    if windows is None:
        windows = [[0.0, nstep * dt]]
    win = np.ascontiguousarray(np.reshape(np.asarray(windows, np.float64), (-1, 2)))
    o = np.ascontiguousarray(obs, np.float32)
    s = np.ascontiguousarray(syn, np.float32)
    if o.shape != s.shape:
        raise ValueError("observed and synthetic traces differ in shape")
    for w0, w1 in win:
        nb, ne, _ = get_window_info((w0, w1), dt)
        if nb < 1 or ne > nstep:
            raise ValueError(f"measurement window [{w0}, {w1}] s lies outside the {nstep}-sample trace")
    if (min_period is None) != (max_period is None):
        raise ValueError("give both min_period and max_period, or neither")
    sos = (bandpass_sos(dt, np.float32(1.0 / max_period), np.float32(1.0 / min_period))
           if min_period is not None else np.zeros((0, 6), np.float32))
    return nstep, s.size // nstep, win, o, s, sos


def event_cc_adjoint_sources(obs, syn, dt, windows=None, imeas=IMEAS_CC_TT, taper_percentage=0.05,
                             itaper_type=1, min_period=None, max_period=None, **cfg):
    """Cross-correlation traveltime (imeas 11) or amplitude (12) measurement of
    every trace (CCTTMisfit%calc_adjoint_source), or the zero-lag correlation
    coefficient (imeas 5, CCCMisfit: misfit 1 - coef, residual and cc_max
    = coef), in one device pass.  Returns
    (misfit, adjoint traces float32 like syn, per-window measurements
    [nrec*3, nwin, 8]: misfit, residual, error, cc_max, tshift, dlna, selected,
    ishift).  cfg overrides CC_DEFAULTS."""
    nstep, ntrace, win, o, s, sos = _event_inputs(obs, syn, dt, windows, min_period, max_period)
    unknown = set(cfg) - set(CC_DEFAULTS)
    if unknown:
        raise KeyError(f"unknown CC settings {sorted(unknown)}")
    c = dict(CC_DEFAULTS, **cfg)
    cv = np.array([c[k] for k in ("dt_sigma_min", "dlna_sigma_min", "tshift_min", "tshift_max", "dlna_min",
                                  "dlna_max", "cc_min")], np.float64)
    if imeas not in (IMEAS_CC_TT, IMEAS_CC_DLNA, IMEAS_CCC):
        raise ValueError("imeas must be 11 (CC-TT), 12 (CC-DLNA) or 5 (CCC)")
    adj = np.zeros(s.shape, np.float32)
    mf = np.zeros(ntrace)
    meas = np.zeros((ntrace, win.shape[0], 8))
    _lib.call("fwat_cc_adjoint_batch_", nstep, ntrace, float(dt), o, s, win.shape[0], win, float(taper_percentage),
              int(itaper_type), int(imeas), cv, sos.shape[0], sos if sos.size else np.zeros(6, np.float32), mf, meas,
              adj)
    total = 0.0
    for v in mf:
        total += v
    return total, adj, meas


IMEAS_CC_TT_MT, IMEAS_CC_DLNA_MT = 13, 14  # adj_config
MT_DEFAULTS = dict(transfunc_waterlevel=1e-10, water_threshold=0.02, mt_nw=4.0, num_taper=5, dt_fac=2.0,
                   err_fac=2.5, dt_max_scale=3.5, min_cycle_in_window=3, use_cc_error=True,
                   use_mt_error=False)  # adj_config.f90:18-26
MT_FIELDS = ("misfit", "residual", "error", "cc_max", "tshift", "dlna", "selected", "ishift", "imeas",
             "misfit_other")


def event_mt_adjoint_sources(obs, syn, dt, windows, min_period, max_period, imeas=IMEAS_CC_TT_MT,
                             taper_percentage=0.05, itaper_type=1, filter_adj=True, **cfg):
    """Multitaper traveltime (imeas 13) or amplitude (14) measurement of every
    trace (MTTTMisfit%calc_adjoint_source, mt_tt_misfit.f90:29-209) in one
    device pass, with the reference's fall back to the cross-correlation
    measurement for windows the multitaper tests reject.  min_period and
    max_period bound the measured band (cfg%min_period / max_period) and, with
    filter_adj, the adjoint-source bandpass.  Returns (misfit, adjoint traces
    float32 like syn, per-window measurements [nrec*3, nwin, 10] in MT_FIELDS
    order; "imeas" is the measurement used: 13/14, 11/12 after a fall back, 0
    when the window is not selected).  cfg overrides CC_DEFAULTS and
    MT_DEFAULTS (use_mt_error = .false. as the reference's default; with
    use_mt_error=True and use_cc_error=False the jackknife errors of
    calculate_mt_error weight the misfit and adjoint source, and with either
    setting they enter the acceptance test)."""
    nstep, ntrace, win, o, s, sos = _event_inputs(obs, syn, dt, windows, min_period, max_period)
    unknown = set(cfg) - set(CC_DEFAULTS) - set(MT_DEFAULTS)
    if unknown:
        raise KeyError(f"unknown MT settings {sorted(unknown)}")
    if imeas not in (IMEAS_CC_TT_MT, IMEAS_CC_DLNA_MT):
        raise ValueError("imeas must be 13 (MT-TT) or 14 (MT-DLNA)")
    if not 32 <= nstep <= 16383:
        raise ValueError("the multitaper measurement takes traces of 32 to 16383 samples")
    c = dict(CC_DEFAULTS, **MT_DEFAULTS)
    c.update(cfg)
    cv = np.array([c["dt_sigma_min"], c["dlna_sigma_min"], c["tshift_min"], c["tshift_max"], c["dlna_min"],
                   c["dlna_max"], c["cc_min"], c["transfunc_waterlevel"], c["water_threshold"], c["mt_nw"],
                   c["num_taper"], c["dt_fac"], c["err_fac"], c["dt_max_scale"], c["min_cycle_in_window"],
                   1.0 if c["use_cc_error"] else 0.0, min_period, max_period,
                   1.0 if c["use_mt_error"] else 0.0], np.float64)
    if not filter_adj:
        sos = np.zeros((0, 6), np.float32)
    adj = np.zeros(s.shape, np.float32)
    mf = np.zeros(ntrace)
    meas = np.zeros((ntrace, win.shape[0], 10))
    _lib.call("fwat_mt_adjoint_batch_", nstep, ntrace, float(dt), o, s, win.shape[0], win, float(taper_percentage),
              int(itaper_type), int(imeas), cv, sos.shape[0], sos if sos.size else np.zeros(6, np.float32), mf, meas,
              adj)
    total = 0.0
    for v in mf:
        total += v
    return total, adj, meas


IMEAS_EXP_PHASE = 4  # adj_config
EP_FIELDS = ("misfit", "residual", "misfit_real", "misfit_imag", "residual_real", "residual_imag", "selected")


def event_ep_adjoint_sources(obs, syn, dt, windows=None, taper_percentage=0.05, itaper_type=1, wtr_env=0.2,
                             min_period=None, max_period=None):
    """Exponentiated phase measurement (imeas 4, ExponentiatedPhaseMisfit%
    calc_adjoint_source, exponentiated_phase_misfit.f90:21-104) of every trace in
    one device pass: envelope-normalised real and Hilbert parts of data and
    synthetic, their differences' Simpson misfits and mean residuals, the
    adjoint source with its two Hilbert-transform terms.  wtr_env is the
    envelope water level (adj_config).  Returns (misfit, adjoint traces float32
    like syn, per-window measurements [nrec*3, nwin, 8] in EP_FIELDS order)."""
    nstep, ntrace, win, o, s, sos = _event_inputs(obs, syn, dt, windows, min_period, max_period)
    adj = np.zeros(s.shape, np.float32)
    mf = np.zeros(ntrace)
    meas = np.zeros((ntrace, win.shape[0], 8))
    _lib.call("fwat_ep_adjoint_batch_", nstep, ntrace, float(dt), o, s, win.shape[0], win, float(taper_percentage),
              int(itaper_type), float(wtr_env), sos.shape[0], sos if sos.size else np.zeros(6, np.float32), mf, meas,
              adj)
    total = 0.0
    for v in mf:
        total += v
    return total, adj, meas


IMEAS_WAVEFORM, IMEAS_WAVEFORM_CONV, IMEAS_RF = 1, 2, 3  # adj_config


def calculate_adjoint_sources(obs, syn, dt, windows, short_p, long_p, imeasure_type, taper_percentage=0.05,
                              itaper_type=1, **cfg):
    """calculate_adjoint_source (adjoint_source.f90:14-89) for every trace of an
    event: dispatch on imeasure_type to the batched measurement, then filter_adj
    (bandpass 1/long_p .. 1/short_p and the taper over the windows) as the
    reference always does.  Types 2 (waveform convolution) and 3 (receiver
    function) stop with an error there and raise here.  Returns (misfit,
    adjoint traces, per-window measurements or None for the waveform type)."""
    kw = dict(taper_percentage=taper_percentage, itaper_type=itaper_type)
    if imeasure_type == IMEAS_WAVEFORM:
        if cfg:
            raise KeyError(f"settings {sorted(cfg)} do not apply to the waveform misfit")
        total, adj = event_adjoint_sources(obs, syn, dt, windows, min_period=short_p, max_period=long_p, **kw)
        return total, adj, None
    if imeasure_type == IMEAS_WAVEFORM_CONV:
        raise ValueError("WAVEFORM_CONV misfit requires radial/vertical components "
                         "(event_conv_adjoint_sources; calc_adjoint_source_conv in the reference)")
    if imeasure_type == IMEAS_RF:
        raise ValueError("RF misfit requires additional parameters (calc_adjoint_source_rf in the reference)")
    if imeasure_type == IMEAS_EXP_PHASE:
        return event_ep_adjoint_sources(obs, syn, dt, windows, min_period=short_p, max_period=long_p, **kw, **cfg)
    if imeasure_type in (IMEAS_CCC, IMEAS_CC_TT, IMEAS_CC_DLNA):
        return event_cc_adjoint_sources(obs, syn, dt, windows, imeas=imeasure_type, min_period=short_p,
                                        max_period=long_p, **kw, **cfg)
    if imeasure_type in (IMEAS_CC_TT_MT, IMEAS_CC_DLNA_MT):
        return event_mt_adjoint_sources(obs, syn, dt, windows, short_p, long_p, imeas=imeasure_type, **kw, **cfg)
    raise ValueError(f"Unknown measurement type: {imeasure_type}")


WIN_DEFAULTS = dict(sliding_win_len_fac=2.0, min_velocity=2.4, threshold_corr=0.8, threshold_shift_fac=0.3,
                    jump_fac=0.1, min_win_len_fac=1.5, min_snr_window=5.0, min_energy_ratio=5.0,
                    min_peaks_troughs=3, is_split_phases=False)  # WIN_TYPE_1 defaults, input_params.f90:389-398


def select_windows(obs, syn, dt, t0, tp, dist, min_period, maxwin=64, **cfg):
    """win_sel_type%select_windows (window_selector/win_sel.f90:35-277) for every
    trace of an event in one device pass: the sliding cross-correlation (one full
    correlation of Hann-tapered windows per sample), good-sample mask with the
    time-shift jump rule, contiguous groups and the length / peak / SNR / energy
    checks.  obs, syn (NSTEP, nrec, 3), already filtered as the caller does;
    tp (nrec,) first-arrival times and dist (nrec,) distances (the receiver's
    three components share them), t0 the simulation start.  With
    is_split_phases the groups are split at the STA/LTA troughs between phases
    (split_window_by_phases).  Returns a list per trace (nrec*3, receiver-major)
    of [n, 2] window arrays in seconds."""
    o = np.ascontiguousarray(obs, np.float32)
    s = np.ascontiguousarray(syn, np.float32)
    if o.shape != s.shape:
        raise ValueError("observed and synthetic traces differ in shape")
    unknown = set(cfg) - set(WIN_DEFAULTS)
    if unknown:
        raise KeyError(f"unknown window settings {sorted(unknown)}")
    c = dict(WIN_DEFAULTS, **cfg)
    nstep = s.shape[0]
    ntrace = s.size // nstep
    ncomp = ntrace // np.size(tp)
    tpv = np.ascontiguousarray(np.repeat(np.asarray(tp, np.float64), ncomp))
    dv = np.ascontiguousarray(np.repeat(np.asarray(dist, np.float64), ncomp))
    if tpv.size != ntrace or dv.size != ntrace:
        raise ValueError("tp and dist need one value per receiver")
    cv = np.array([c[k] for k in WIN_DEFAULTS], np.float64)
    cv[-1] = 1.0 if c["is_split_phases"] else 0.0
    nwin = np.zeros(ntrace, np.int32)
    win = np.zeros((ntrace, maxwin, 2))
    _lib.call("fwat_win_sel_batch_", nstep, ntrace, float(dt), o, s, float(t0), tpv, dv, float(min_period), cv,
              int(maxwin), nwin, win)
    if (nwin > maxwin).any():
        raise RuntimeError(f"a trace has more than maxwin={maxwin} windows; raise maxwin")
    return [win[t, :nwin[t]].copy() for t in range(ntrace)]


FLEXWIN_DEFAULTS = dict(min_velocity=2.4, min_win_len_fac=1.5, threshold_corr=0.8, threshold_shift_fac=2.0,
                        min_peaks_troughs=3, threshold_dlna=1.0, min_snr_window=5.0, is_split_phases=False,
                        max_time_before_first_arrival=10.0, stalta_water_level=0.08, c_0=1.0, c_2=0.0,
                        c_3a=4.0, c_3b=2.5, c_4a=2.0, c_4b=6.0)  # input_params.f90:400-418 (WIN_TYPE_2)


def flexwin_windows(obs, syn, dt, t0, tp, dist, min_period, max_period, maxwin=64, **cfg):
    """FLEXWIN window selection (WIN_TYPE 2: flexwin_type%select_windows,
    window_selector/flexwin.f90:75-734, as leq_data.f90:116-121 calls it) for
    every trace of an event in one device pass: STA/LTA of the synthetic's
    envelope, candidate windows between STA/LTA troughs around each peak above the
    water level, the rejection passes (travel times, length, minima water level,
    prominence, phase separation, curtailing, duplicates), the data-fit criteria
    (cross-correlation shift, cc_max_coef, dlnA, SNR, peak count), then merging
    or, with is_split_phases, weighted interval scheduling.  obs, syn (NSTEP,
    nrec, 3) filtered as the caller does; tp, dist (nrec,) first-arrival times
    and distances; min/max_period the band.  Returns (windows, info): per trace
    (nrec*3, receiver-major) an [n, 2] array of window times (s, relative to t0)
    and an [n, 4] array of weight, cc, time shift, dlnA."""
    o = np.ascontiguousarray(obs, np.float32)
    s = np.ascontiguousarray(syn, np.float32)
    if o.shape != s.shape:
        raise ValueError("observed and synthetic traces differ in shape")
    unknown = set(cfg) - set(FLEXWIN_DEFAULTS)
    if unknown:
        raise KeyError(f"unknown FLEXWIN settings {sorted(unknown)}")
    c = dict(FLEXWIN_DEFAULTS, **cfg)
    nstep = s.shape[0]
    ntrace = s.size // nstep
    ncomp = ntrace // np.size(tp)
    tpv = np.ascontiguousarray(np.repeat(np.asarray(tp, np.float64), ncomp))
    dv = np.ascontiguousarray(np.repeat(np.asarray(dist, np.float64), ncomp))
    if tpv.size != ntrace or dv.size != ntrace:
        raise ValueError("tp and dist need one value per receiver")
    cv = np.array([float(c[k]) for k in FLEXWIN_DEFAULTS], np.float64)
    nwin = np.zeros(ntrace, np.int32)
    win = np.zeros((ntrace, maxwin, 2))
    info = np.zeros((ntrace, maxwin, 4))
    _lib.call("fwat_flexwin_batch_", nstep, ntrace, float(dt), o, s, float(t0), tpv, dv, float(min_period),
              float(max_period), cv, int(maxwin), nwin, win, info)
    if (nwin > maxwin).any():
        raise RuntimeError(f"a trace has more than maxwin={maxwin} windows; raise maxwin")
    return ([win[t, :nwin[t]].copy() for t in range(ntrace)], [info[t, :nwin[t]].copy() for t in range(ntrace)])


def event_conv_adjoint_sources(dat_r, dat_z, syn_r, syn_z, dt, windows, taper_percentage=0.05, itaper_type=1):
    """Cross-convolution misfit (imeas 2, WaveformConvMisfit%calc_adjoint_source,
    waveform_conv_misfit.f90:17-64, as telecc_data.f90:175-178 calls it) of every
    receiver in one device pass: dat_*, syn_* (NSTEP, nrec) radial and vertical
    traces, windows (nrec, 2) in seconds.  Returns (total misfit, misfits (nrec,),
    residuals (nrec,), adj_r, adj_z float32 (NSTEP, nrec)).  The caller rotates
    adj_r to N/E and applies the source weight as telecc_data does."""
    arrs = [np.ascontiguousarray(a, np.float32) for a in (dat_r, dat_z, syn_r, syn_z)]
    nstep, nrec = arrs[0].shape
    if any(a.shape != (nstep, nrec) for a in arrs):
        raise ValueError("the four traces must share their shape (NSTEP, nrec)")
    win = np.ascontiguousarray(np.reshape(np.asarray(windows, np.float64), (nrec, 2)))
    for w0, w1 in win:
        nb, ne, _ = get_window_info((w0, w1), dt)
        if nb < 1 or ne > nstep:
            raise ValueError(f"measurement window [{w0}, {w1}] s lies outside the {nstep}-sample trace")
    mf = np.zeros(nrec)
    rs = np.zeros(nrec)
    ar = np.zeros((nstep, nrec), np.float32)
    az = np.zeros((nstep, nrec), np.float32)
    _lib.call("fwat_conv_adjoint_batch_", nstep, nrec, float(dt), *arrs, win, float(taper_percentage),
              int(itaper_type), mf, rs, ar, az)
    total = 0.0
    for v in mf:
        total += v
    return total, mf, rs, ar, az


WIN_SELECTOR_TYPE, WIN_FLEX_TYPE, WIN_GROUPVEL_TYPE, WIN_ARRIVAL_TYPE = 1, 2, 3, 4  # config.f90:58-61


def fixed_windows(win_type, nstep, dt, t0, tp=None, dist=None, time_win=None, groupvel=None, long_p=None):
    """The one-window types of leq_data.f90:121-135 for every receiver: group
    velocity (3): [dist / vmax - long_p / 2, dist / vmin + long_p / 2]; arrival
    (4): [tp + time_win[0], tp + time_win[1]]; both clipped to
    [-t0, (NSTEP - 2) dt - t0].  Returns [nrec, 2] in seconds."""
    if win_type == WIN_GROUPVEL_TYPE:
        d = np.asarray(dist, np.float64)
        w0 = d / groupvel[1] - long_p / 2.0
        w1 = d / groupvel[0] + long_p / 2.0
    elif win_type == WIN_ARRIVAL_TYPE:
        p = np.asarray(tp, np.float64)
        w0 = p + time_win[0]
        w1 = p + time_win[1]
    else:
        raise ValueError(f"window type {win_type} is not a one-window type (3 or 4); use select_windows for 1")
    return np.stack([np.maximum(w0, -t0), np.minimum(w1, (nstep - 2) * dt - t0)], axis=1)


def last_time_ms():
    return float(_lib.call("fwat_waveform_adjoint_last_time_", 0.0)[0])
