"""The engine run from a solver database (proc%06d_external_mesh.bin) against the
FP64 oracle of the same mesh and model: seismograms and sensitivity kernels."""
import numpy as np
import pytest

from specfwat_amd import databases as D
from specfwat_amd.sources import Receivers, Sources
from specfwat_amd.solver import ElasticRun
from test_gpu_parity import KERN_TOL, SEIS_TOL, make_case, oracle_of, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("aniso", [False, True])
def test_database_run_matches_oracle(tmp_path, aniso):
    c = make_case(seed=7, perturb=0.1, nstep=120, aniso=aniso)
    p = str(tmp_path / "proc000000_external_mesh.bin")
    D.write_external_mesh(p, c["mesh"], c["rho"], c["vp"], c["vs"], c["dt"], cij=c["cij"])
    dm = D.load_mesh(p, anisotropy=aniso)
    print(f"database {p}: nspec {dm.nspec} nglob {dm.nglob} irregular {dm.nspec_irregular} "
          f"faces {dm.num_abs_boundary_faces} aniso {aniso}", flush=True)
    mat = dm.materials()
    src = Sources(dm, c["sxyz"], c["mt"], c["f0"])
    rec = Receivers(dm, c["rxyz"])
    fwd = ElasticRun(dm, None, None, None, src, rec, c["dt"], c["nstep"], 1, True, cij=dm.cij(),
                     stacey=dm.stacey, materials=mat)
    fwd.iterate_time()
    syn = fwd.seismograms()
    o = oracle_of(c, 64)
    obs = o.forward(save=True)
    e = rel(syn, obs)
    print(f"seismograms rel L2 vs FP64 oracle {e:.2e}", flush=True)
    assert np.abs(obs).max() > 0 and e < SEIS_TOL
    if aniso:
        fwd.close()
        return
    adj = obs.astype(np.float32)
    a = ElasticRun(dm, None, None, None, src, rec, c["dt"], c["nstep"], 3, False, stacey=dm.stacey,
                   materials=mat)
    a.take_forward(fwd)
    fwd.close()
    a.set_adjoint_sources(adj)
    a.iterate_time()
    k = a.kernels()
    a.close()
    ko = o.adjoint(adj)
    for n in ("rho_kl", "mu_kl", "kappa_kl"):
        e = rel(k[n], ko[n])
        print(f"{n}: rel L2 vs FP64 oracle {e:.2e}", flush=True)
        assert e < KERN_TOL, n


def test_partitioned_database_run_matches_oracle(tmp_path):
    """Three slices of a partitioned database merged onto one GPU: seismograms
    and kernels vs the FP64 oracle of the whole mesh."""
    from test_databases import split_database
    c = make_case(seed=8, perturb=0.1, nstep=120)
    m = c["mesh"]
    p = str(tmp_path / "whole.bin")
    D.write_external_mesh(p, m, c["rho"], c["vp"], c["vs"], c["dt"])
    paths = split_database(p, m, c["rho"], [0, 17, 31, m.nspec], tmp_path)
    dm = D.load_partitioned_mesh(paths)
    print(f"merged {len(paths)} slices: nspec {dm.nspec} nglob {dm.nglob}", flush=True)
    mat = dm.materials()
    src = Sources(dm, c["sxyz"], c["mt"], c["f0"])
    rec = Receivers(dm, c["rxyz"])
    fwd = ElasticRun(dm, None, None, None, src, rec, c["dt"], c["nstep"], 1, True, stacey=dm.stacey, materials=mat)
    fwd.iterate_time()
    syn = fwd.seismograms()
    o = oracle_of(c, 64)
    obs = o.forward(save=True)
    e = rel(syn, obs)
    print(f"merged slices: seismograms rel L2 vs FP64 oracle {e:.2e}", flush=True)
    assert e < SEIS_TOL
    adj = obs.astype(np.float32)
    a = ElasticRun(dm, None, None, None, src, rec, c["dt"], c["nstep"], 3, False, stacey=dm.stacey, materials=mat)
    a.take_forward(fwd)
    fwd.close()
    a.set_adjoint_sources(adj)
    a.iterate_time()
    k = a.kernels()
    a.close()
    ko = o.adjoint(adj)
    # kernels are per local point: the merged mesh keeps the whole mesh's element order
    for n in ("rho_kl", "mu_kl", "kappa_kl"):
        ek = rel(k[n], ko[n])
        print(f"merged slices {n}: rel L2 vs FP64 oracle {ek:.2e}", flush=True)
        assert ek < KERN_TOL, n
