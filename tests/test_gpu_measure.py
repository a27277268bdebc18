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
