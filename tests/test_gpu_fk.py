"""FK injection of an incident wavefield through the Stacey faces (teleseismic
coupling, fk_coupling.f90:71-297): against the closed-form vertically incident P
plane wave (fk_plane_wave), and against the oracle's injection on a perturbed,
heterogeneous mesh with a generic incident field (seismograms and kernels)."""
import numpy as np
import pytest

from fk_plane_wave import ALPHA, BETA, RHO, incident_fields as _fields, plane_wave
from specfwat_amd.mesh import BoxMesh
from specfwat_amd.sources import Receivers, Sources
from specfwat_amd.solver import ElasticRun
from test_gpu_parity import KERN_TOL, SEIS_TOL, make_case, oracle_of, rel

pytestmark = pytest.mark.gpu


def incident_fields(mesh, nsamp, dt_s, L):
    """The plane wave's velocity and traction at every absorbing-face point of mesh."""
    e = mesh.abs_boundary_ispec - 1
    ijk = mesh.abs_boundary_ijk - 1
    loc = ijk[..., 0] + 5 * ijk[..., 1] + 25 * ijk[..., 2]
    z = mesh.xyz[mesh.ibool0[e[:, None], loc], 2].ravel()
    return _fields(z, mesh.abs_boundary_normal.reshape(-1, 3), nsamp, dt_s, L)


def setup(nstep=1125, dt=0.004):
    nex, ney, nez, h = 4, 4, 16, 500.0
    L = nez * h
    print(f"FK plane wave: box {nex}x{ney}x{nez}, h {h}, nstep {nstep}, dt {dt}", flush=True)
    mesh = BoxMesh(nex, ney, nez, nex * h, ney * h, L)
    n = mesh.nspec * 125
    rho = np.full(n, RHO, np.float32)
    vp = np.full(n, ALPHA, np.float32)
    vs = np.full(n, BETA, np.float32)
    depths = np.array([0.0, -1500.0, -3000.0, -5000.0, -6500.0])
    rxyz = np.stack([np.full(5, 1000.0), np.full(5, 1000.0), depths], axis=1)
    rec = Receivers(mesh, rxyz)
    return mesh, rho, vp, vs, rec, depths, L, nstep, dt


def run(mesh, rho, vp, vs, rec, nstep, dt, fk=None, resamp=1, fused=True):
    r = ElasticRun(mesh, rho, vp, vs, None, rec, dt, nstep, 1, False)
    if fk is not None:
        r.set_fk_injection(*fk, resamp=resamp)
    r.iterate_time(fused=fused)
    s = r.seismograms()
    r.close()
    return s


def test_zero_incident_field_is_plain_stacey_and_reconstruction():
    mesh, rho, vp, vs, rec, depths, L, _, dt = setup()
    nstep = 300
    src = Sources(mesh, [[1100.0, 900.0, -4100.0]], [[1e15, -5e14, 3e14, 2e14, -1e14, 4e14]], [2.0])
    z = np.zeros((nstep, mesh.num_abs_boundary_faces * 25, 3), np.float32)

    def fwd(fk, fused=True, save=False):
        r = ElasticRun(mesh, rho, vp, vs, src, rec, dt, nstep, 1, save)
        if fk is not None:
            r.set_fk_injection(*fk)
        r.iterate_time(fused=fused)
        return r

    a, b = fwd((z, z)), fwd(None)
    sa, sb = a.seismograms(), b.seismograms()
    a.close(); b.close()
    assert np.abs(sb).max() > 0 and np.array_equal(sa, sb)  # zero incident field == plain Stacey
    V, T = incident_fields(mesh, nstep, dt, L)
    f1, f2 = fwd((V, T)), fwd((V, T), fused=False)
    s1, s2 = f1.seismograms(), f2.seismograms()
    f1.close(); f2.close()
    rel = np.linalg.norm(s1 - s2) / np.linalg.norm(s1)
    print(f"SPECFEM stage order vs fused with injection: rel L2 {rel:.2e}")
    assert rel < 1e-5
    # the stored boundary terms include the injection: the reconstruction of an
    # injected forward run returns to the zero initial state
    f = fwd((V, T), save=True)
    u_scale = np.abs(f.seismograms()).max()
    adj = ElasticRun(mesh, rho, vp, vs, src, rec, dt, nstep, 3, False)
    adj.take_forward(f)
    f.close()
    adj.set_adjoint_sources(np.zeros((nstep, rec.nrec, 3), np.float32))
    adj.iterate_time()
    u_end = np.abs(adj.b_fields()[0]).max()
    adj.close()
    print(f"reconstruction end state {u_end:.2e} (forward max {u_scale:.2e})")
    assert u_end < 1e-3 * u_scale


def test_plane_wave_injection_matches_closed_form():
    mesh, rho, vp, vs, rec, depths, L, nstep, dt = setup()
    V, T = incident_fields(mesh, nstep, dt, L)
    seis = run(mesh, rho, vp, vs, rec, nstep, dt, fk=(V, T))
    t = np.arange(nstep) * dt
    want = np.stack([plane_wave(z, t, L)[0] for z in depths], axis=1)  # (nstep, nrec)
    err = np.linalg.norm(seis[:, :, 2] - want) / np.linalg.norm(want)
    side = np.abs(seis[:, :, :2]).max() / np.abs(want).max()
    shifted = np.linalg.norm(seis[1:, :, 2] - want[:-1]) / np.linalg.norm(want)
    print(f"u_z vs closed form: rel L2 {err:.2e} (one step later: {shifted:.2e}); "
          f"max |u_x|,|u_y| / max |u_z| {side:.1e}; max |u_z| {np.abs(want).max():.3f}")
    assert err < 2e-2 and side < 1e-3
    # the series every 4th step, interpolated on the device
    seis4 = run(mesh, rho, vp, vs, rec, nstep, dt, fk=(V[::4], T[::4]), resamp=4)
    err4 = np.linalg.norm(seis4[:, :, 2] - want) / np.linalg.norm(want)
    print(f"resampled every 4 steps: rel L2 {err4:.2e}")
    assert err4 < 3e-2
    # without injection the box stays at rest
    assert np.abs(run(mesh, rho, vp, vs, rec, 50, dt)).max() == 0.0


def test_injection_matches_oracle():
    """A perturbed heterogeneous mesh, a point source and a generic incident field
    (random per-point amplitudes and delays, stored every 3rd step): seismograms of
    the injected forward run and the kernels of its adjoint run vs the oracle."""
    c = make_case(seed=11, perturb=0.1, nstep=150)
    m = c["mesh"]
    o64 = oracle_of(c, 64)
    fo = o64.faces()
    assert np.array_equal(fo["ispec"] + 1, m.abs_boundary_ispec)
    assert np.array_equal(fo["ijk"] + 1, m.abs_boundary_ijk)
    rng = np.random.default_rng(5)
    npt, resamp = m.num_abs_boundary_faces * 25, 3
    nsamp = (c["nstep"] - 1) // resamp + 1
    tk = np.arange(nsamp) * resamp * c["dt"]
    amp = rng.standard_normal((npt, 3))
    dly = rng.uniform(0.5, 1.5, (npt, 1))
    pulse = np.exp(-(((tk[:, None] - dly.T) / 0.3) ** 2))  # (nsamp, npt)
    V = (1e-2 * pulse[..., None] * amp[None]).astype(np.float32)
    T = (3e5 * pulse[..., None] * amp[None, :, ::-1]).astype(np.float32)
    o64.set_fk(V, T, resamp=resamp)
    obs = o64.forward(save=True)
    src = Sources(m, c["sxyz"], c["mt"], c["f0"])
    rec = Receivers(m, c["rxyz"])
    f = ElasticRun(m, c["rho"], c["vp"], c["vs"], src, rec, c["dt"], c["nstep"], 1, True)
    f.set_fk_injection(V, T, resamp=resamp)
    f.iterate_time()
    syn = f.seismograms()
    e = rel(syn, obs)
    print(f"injected seismograms vs FP64 oracle: rel L2 {e:.2e}", flush=True)
    assert np.abs(obs).max() > 0 and e < SEIS_TOL
    adj = obs.astype(np.float32)
    a = ElasticRun(m, c["rho"], c["vp"], c["vs"], src, rec, c["dt"], c["nstep"], 3, False)
    a.take_forward(f)
    f.close()
    a.set_adjoint_sources(adj)
    a.iterate_time()
    k = a.kernels()
    a.close()
    ko = o64.adjoint(adj)
    for n in ("rho_kl", "mu_kl", "kappa_kl"):
        ek = rel(k[n], ko[n])
        print(f"{n} with injection: rel L2 vs FP64 oracle {ek:.2e}", flush=True)
        assert ek < KERN_TOL, n


@pytest.mark.parametrize("path", ["brick", "slabs"])
def test_injection_on_brick_and_slab_paths(path):
    """The brick-fused pass and the element-slab order inject the same incident
    field as the plain fused pass (plane wave on the regular lattice box)."""
    mesh, rho, vp, vs, rec, depths, L, _, dt = setup()
    nstep = 400
    V, T = incident_fields(mesh, nstep, dt, L)

    def fwd(tuning):
        r = ElasticRun(mesh, rho, vp, vs, None, rec, dt, nstep, 1, False)
        r.set_tuning(1, 0, 8, **tuning)
        r.set_fk_injection(V, T)
        r.iterate_time()
        s = r.seismograms()
        r.close()
        return s

    base = fwd({})
    other = fwd({"brick": (2, 2, 2)} if path == "brick" else {"slab_elems": 64})
    e = rel(other, base)
    print(f"{path} path with injection vs plain fused: rel L2 {e:.2e}", flush=True)
    assert np.abs(base).max() > 0 and e < 1e-6
