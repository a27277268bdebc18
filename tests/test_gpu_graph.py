"""HIP-graph capture of the whole-run loop (fwat_set_graph_): forward and
adjoint runs captured into one graph each give the eager runs' seismograms and
kernels bit for bit, run after run (three runs, with eager runs of another
handle and host synchronisations in between), and the FP64 oracle's within the
parity bars."""
import numpy as np
import pytest

from specfwat_amd.solver import ElasticRun
from specfwat_amd.sources import Receivers, Sources
from test_gpu_parity import KERN_TOL, SEIS_TOL, make_case, oracle_of, rel

pytestmark = pytest.mark.gpu


def event(c, graph, adj, tuning=None):
    m = c["mesh"]
    src, rec = Sources(m, c["sxyz"], c["mt"], c["f0"]), Receivers(m, c["rxyz"])
    f = ElasticRun(m, c["rho"], c["vp"], c["vs"], src, rec, c["dt"], c["nstep"], 1, True)
    a = ElasticRun(m, c["rho"], c["vp"], c["vs"], src, rec, c["dt"], c["nstep"], 3, False)
    if tuning:
        f.set_tuning(**tuning)
        a.set_tuning(**tuning)
    f.set_graph(graph)
    a.set_graph(graph)
    f.iterate_time()
    syn = f.seismograms()
    a.take_forward(f)
    f.close()
    a.set_adjoint_sources(adj)
    a.iterate_time()
    k = a.kernels()
    a.close()
    return syn, k


def test_graph_runs_equal_eager_runs():
    c = make_case(seed=41, nstep=90, perturb=0.1)
    o = oracle_of(c, 64)
    obs = o.forward(save=True)
    adj = (0.5 * obs).astype(np.float32)
    ko = o.adjoint(adj)
    s0, k0 = event(c, False, adj)
    for rep in range(3):
        print(f"graph run {rep}", flush=True)
        s1, k1 = event(c, True, adj)
        assert np.array_equal(s1, s0), rep
        for n in k0:
            assert np.array_equal(k1[n], k0[n]), (rep, n)
        s2, _ = event(c, False, adj)  # eager work between the graph runs
        assert np.array_equal(s2, s0)
    assert rel(s1, obs) < SEIS_TOL
    for n in ("rho_kl", "mu_kl", "kappa_kl"):
        assert rel(k1[n], ko[n]) < KERN_TOL, n


def test_graph_runs_with_interior_points_equal_eager_runs():
    """The same with the pipelined pass finishing the element-interior points (a regular
    mesh, regular_path 3): captured runs equal eager runs bit for bit, twice."""
    c = make_case(nex=8, ney=8, nez=4, seed=47, nstep=96)
    adj = np.asarray(np.random.default_rng(9).standard_normal((c["nstep"], 5, 3)) * 1e-3, np.float32)
    tun = dict(regular_path=3, blocks_per_cu_fwd=2, blocks_per_cu_adj=2, interior=True)
    s0, k0 = event(c, False, adj, tun)
    for rep in range(2):
        print(f"graph run {rep} (interior points)", flush=True)
        s1, k1 = event(c, True, adj, tun)
        assert np.array_equal(s1, s0), rep
        for n in k0:
            assert np.array_equal(k1[n], k0[n]), (rep, n)
