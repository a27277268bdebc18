"""The engine with phase lists in other visiting orders (colour by colour, as
SPECFEM's mesh colouring orders them, and a scrambled outer/inner split): its
assembly then sums every point's terms in that order, and seismograms and
kernels stay within the parity bars of the FP64 oracle assembling in the same
order."""
import numpy as np
import pytest

from oracle import oracle as orc
from specfwat_amd.solver import ElasticRun
from specfwat_amd.sources import Receivers, Sources
from test_gpu_parity import KERN_TOL, SEIS_TOL, make_case, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order", ["colour", "scrambled"])
def test_engine_in_phase_visiting_order_matches_oracle(order):
    c = make_case(seed=29, nstep=100, perturb=0.1)
    m = c["mesh"]
    if order == "colour":
        m.colour_sort_phases()
        elem_order = None  # the oracle's colour-by-colour assembly
    else:
        rng = np.random.default_rng(5)
        perm = rng.permutation(m.nspec) + 1
        nout = m.nspec // 4
        m.nspec_outer, m.nspec_inner = nout, m.nspec - nout
        m.phase_ispec_inner = np.zeros((2, m.nspec), np.int32)
        m.phase_ispec_inner[0, :nout] = perm[:nout]
        m.phase_ispec_inner[1, :m.nspec - nout] = perm[nout:]
        elem_order = perm - 1
    o = orc.Oracle(m.nex, m.ney, m.nez, m.lx, m.ly, m.lz, c["rho"].ravel(), c["vp"].ravel(), c["vs"].ravel(),
                   dt=c["dt"], nstep=c["nstep"], src_xyz=c["sxyz"], src_mt=c["mt"], src_f0=c["f0"],
                   rec_xyz=c["rxyz"], perturb=c["perturb"], elem_order=elem_order, precision=64)
    obs = o.forward(save=True)
    adj = (0.5 * obs).astype(np.float32)
    ko = o.adjoint(adj)
    src, rec = Sources(m, c["sxyz"], c["mt"], c["f0"]), Receivers(m, c["rxyz"])
    f = ElasticRun(m, c["rho"], c["vp"], c["vs"], src, rec, c["dt"], c["nstep"], 1, True)
    f.iterate_time()
    syn = f.seismograms()
    a = ElasticRun(m, c["rho"], c["vp"], c["vs"], src, rec, c["dt"], c["nstep"], 3, False)
    a.take_forward(f)
    f.close()
    a.set_adjoint_sources(adj)
    a.iterate_time()
    k = a.kernels()
    a.close()
    es = rel(syn, obs)
    ek = {n: rel(k[n], ko[n]) for n in ("rho_kl", "mu_kl", "kappa_kl")}
    print(f"{order}: seismograms {es:.2e}, " + ", ".join(f"{n} {v:.2e}" for n, v in ek.items()))
    assert es < SEIS_TOL and max(ek.values()) < KERN_TOL
