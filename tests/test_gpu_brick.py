"""GPU parity of the brick-fused element + assembly pass (fwat_set_brick_).

Every brick shape, forward and adjoint, fp32 and bf16 statics, against the
FP64 oracle at the parity bars (seismograms 1e-5, kernels 1e-4 relative L2)
and against the gather path of the same handle.  The mesh puts sources and
receivers inside and on the edge of bricks, so both the brick-inner Newmark
and the halo pass (partials + Stacey + source + receiver terms) are covered.
"""
import numpy as np
import pytest

from test_gpu_parity import KERN_TOL, SEIS_TOL, make_case, oracle_of, rel
from specfwat_amd.solver import ElasticRun
from specfwat_amd.sources import Receivers, Sources

pytestmark = pytest.mark.gpu


def run_pair(c, adj, brick, static_precision=0, hess=False):
    m = c["mesh"]
    src = Sources(m, c["sxyz"], c["mt"], c["f0"])
    rec = Receivers(m, c["rxyz"])
    f = ElasticRun(m, c["rho"], c["vp"], c["vs"], src, rec, c["dt"], c["nstep"], 1, True,
                   static_precision=static_precision)
    f.set_tuning(1, 0, 8, brick=brick)
    f.iterate_time()
    syn = f.seismograms()
    a = ElasticRun(m, c["rho"], c["vp"], c["vs"], src, rec, c["dt"], c["nstep"], 3, False,
                   static_precision=static_precision, approximate_hess=hess)
    a.set_tuning(1, 0, 8, brick=brick)
    a.take_forward(f)
    f.close()
    a.set_adjoint_sources(adj)
    a.iterate_time()
    k = a.kernels()
    a.close()
    return syn, k


@pytest.fixture(scope="module")
def case():
    c = make_case(nex=8, ney=8, nez=4, seed=21, nstep=120)
    o = oracle_of(c, 64)
    obs = o.forward(save=True)
    adj = (0.5 * obs).astype(np.float32)
    ko = o.adjoint(adj)
    return c, obs, adj, ko


@pytest.mark.parametrize("brick", [(2, 2, 2), (4, 2, 2), (4, 4, 2)])
def test_brick_pass_matches_oracle_and_gather_path(case, brick):
    c, obs, adj, ko = case
    print(f"brick {brick}", flush=True)
    syn, k = run_pair(c, adj, brick)
    syn_g, k_g = run_pair(c, adj, (0, 0, 0))
    es = rel(syn, obs)
    print(f"brick {brick}: seismograms vs FP64 oracle {es:.2e}, vs gather path {rel(syn, syn_g):.2e}")
    assert es < SEIS_TOL
    assert rel(syn, syn_g) < 1e-6
    for n in ("rho_kl", "mu_kl", "kappa_kl"):
        e = rel(k[n], ko[n])
        print(f"  {n}: vs FP64 oracle {e:.2e}, vs gather path {rel(k[n], k_g[n]):.2e}")
        assert e < KERN_TOL, n
        assert rel(k[n], k_g[n]) < 1e-5, n


def test_brick_pass_reduced_precision_and_hessian(case):
    c, obs, adj, ko = case
    syn, k = run_pair(c, adj, (2, 2, 2), static_precision=1, hess=True)
    syn_g, k_g = run_pair(c, adj, (0, 0, 0), static_precision=1, hess=True)
    assert rel(syn, obs) < 1e-3
    assert rel(syn, syn_g) < 1e-6
    for n in ("rho_kl", "mu_kl", "kappa_kl"):
        assert rel(k[n], ko[n]) < 1e-3, n
        assert rel(k[n], k_g[n]) < 1e-5, n
