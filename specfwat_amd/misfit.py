"""Waveform-difference misfit and adjoint source in double precision.

Restates WaveformMisfit%calc_adjoint_source
(src/preproc/adjoint_sources/waveform_misfit.f90:15-64) with its helpers
get_window_info / window_taper (src/shared/signal.f90:419-508) and simpson
(src/shared/utils.f90:1630-1663).  Host FP64 as in the reference.
"""
import numpy as np


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


def event_adjoint_sources(obs, syn, dt, windows=None, taper_percentage=0.05, itaper_type=1):
    """obs/syn (NSTEP, nrec, 3) -> (misfit, adjoint traces float32 (NSTEP, nrec, 3)).
    Default window: the whole trace."""
    nstep, nrec, _ = syn.shape
    if windows is None:
        windows = [[0.0, nstep * dt]]
    adj = np.zeros(syn.shape, np.float32)
    total = 0.0
    for r in range(nrec):
        for c in range(3):
            mf, a, _, _ = waveform_adjoint_source(obs[:, r, c], syn[:, r, c], dt, windows,
                                                  taper_percentage, itaper_type)
            total += mf
            adj[:, r, c] = a
    return total, adj
