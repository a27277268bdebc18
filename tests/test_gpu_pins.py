"""The engine against the same independent closed forms that pin the oracle's
source and Stacey terms (tests/test_oracle_pins.py): the full-space
moment-tensor response of Aki & Richards (2002) eq. 4.29, and plain absorption
of normally incident P and S plane waves by the Stacey faces."""
import numpy as np
import pytest

from fk_plane_wave import ALPHA, BETA, RHO
from specfwat_amd.mesh import BoxMesh
from specfwat_amd.solver import ElasticRun
from specfwat_amd.sources import Receivers, Sources
from test_oracle_pins import (full_space_case, full_space_errors, plane_wave_injection, reflection_measure,
                              stacey_plane_wave_case)

pytestmark = pytest.mark.gpu


def test_engine_matches_full_space_moment_tensor_response():
    """667 m elements (the oracle's check uses 1 km): the tighter mesh leaves
    0.5-2.3 % discretisation error before the first boundary reflection."""
    c = full_space_case()
    c["ne"] = 24
    c["dt"] = 0.25 * (c["L"] / c["ne"] / 2 * (1 - np.sqrt(3 / 7))) / c["alpha"]
    c["nstep"] = int((c["t0"] + 2.2 * 3500.0 / c["beta"] + 1.2 / c["f0"]) / c["dt"]) + 1
    print(f"full space: {c['ne']}^3 elements, nstep {c['nstep']}", flush=True)
    mesh = BoxMesh(c["ne"], c["ne"], c["ne"], c["L"], c["L"], c["L"])
    n = (mesh.nspec, 125)
    run = ElasticRun(mesh, np.full(n, c["rho"], np.float32), np.full(n, c["alpha"], np.float32),
                     np.full(n, c["beta"], np.float32), Sources(mesh, [c["xs"]], [c["m6"]], [c["f0"]]),
                     Receivers(mesh, c["xr"]), c["dt"], c["nstep"], 1, False)
    run.iterate_time()
    seis = run.seismograms().astype(np.float64)
    run.close()
    e = full_space_errors(c, seis)
    print("A&R 4.29 relative L2 per receiver:", np.array2string(e, precision=4))
    assert e.max() < 0.03


@pytest.mark.parametrize("wave", ["P", "S"])
def test_engine_stacey_absorbs_normally_incident_plane_wave(wave):
    c = stacey_plane_wave_case(wave)
    print(f"Stacey plane wave {wave}: nstep {c['nstep']}", flush=True)
    mesh = BoxMesh(c["nex"], c["nex"], c["nez"], c["nex"] * c["h"], c["nex"] * c["h"], c["L"])
    e = mesh.abs_boundary_ispec - 1
    ijk = mesh.abs_boundary_ijk - 1
    loc = ijk[..., 0] + 5 * ijk[..., 1] + 25 * ijk[..., 2]
    z = mesh.xyz[mesh.ibool0[e[:, None], loc], 2].ravel()
    V, T = plane_wave_injection(c, z, mesh.abs_boundary_normal)
    n = (mesh.nspec, 125)
    run = ElasticRun(mesh, np.full(n, RHO, np.float32), np.full(n, ALPHA, np.float32),
                     np.full(n, BETA, np.float32), None, Receivers(mesh, c["rec"]), c["dt"], c["nstep"], 1, False)
    run.set_fk_injection(V, T)
    run.iterate_time()
    trace = run.seismograms()[:, 0, 2 if wave == "P" else 0].astype(np.float64)
    run.close()
    refl, err = reflection_measure(c, trace)
    print(f"{wave}: reflected/direct {refl:.2e}, direct arrival error {err:.2e}")
    assert refl < 1e-3 and err < 1e-3
