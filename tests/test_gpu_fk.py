"""FK injection of an incident wavefield through the Stacey faces (teleseismic
coupling, fk_coupling.f90:71-297), checked against the closed-form field of a
vertically incident P plane wave in a homogeneous half-space, whose free-surface
reflection is the same pulse travelling down:

    u_z(z, t) = g(t - tc - (z + L)/alpha) + g(t - tc - (L - z)/alpha),  u_x = u_y = 0
    sigma_zz = rho alpha^2 du_z/dz,  sigma_xx = sigma_yy = lambda du_z/dz

Injected on the bottom and the four sides (velocity and traction sigma.n), the
box must reproduce u_z in its interior.
"""
import numpy as np
import pytest

from specfwat_amd.mesh import BoxMesh
from specfwat_amd.sources import Receivers, Sources
from specfwat_amd.solver import ElasticRun

pytestmark = pytest.mark.gpu

RHO, ALPHA, BETA = 2700.0, 6000.0, 3464.0
SIG, TC = 0.4, 1.4  # Gaussian pulse g(t) = exp(-(t/SIG)^2), centred at the bottom at TC


def g(t):
    return np.exp(-(t / SIG) ** 2)


def dg(t):
    return -2.0 * t / SIG**2 * g(t)


def plane_wave(z, t, L):
    t1 = t - TC - (z + L) / ALPHA
    t2 = t - TC - (L - z) / ALPHA
    u = g(t1) + g(t2)
    v = dg(t1) + dg(t2)
    duz = (-dg(t1) + dg(t2)) / ALPHA
    return u, v, duz


def incident_fields(mesh, nsamp, dt_s, L):
    """Velocity and traction of the plane wave at every absorbing-face point,
    sample k at time k*dt_s: (nsamp, nfaces*25, 3)."""
    e = mesh.abs_boundary_ispec - 1
    ijk = mesh.abs_boundary_ijk - 1
    loc = ijk[..., 0] + 5 * ijk[..., 1] + 25 * ijk[..., 2]
    z = mesh.xyz[mesh.ibool0[e[:, None], loc], 2].ravel()
    n = mesh.abs_boundary_normal.reshape(-1, 3).astype(np.float64)
    lam = RHO * (ALPHA**2 - 2 * BETA**2)
    t = (np.arange(nsamp) * dt_s)[:, None]
    _, v, duz = plane_wave(z[None, :], t, L)
    V = np.zeros((nsamp, len(z), 3))
    V[..., 2] = v
    T = np.zeros_like(V)
    T[..., 0] = lam * duz * n[None, :, 0]
    T[..., 1] = lam * duz * n[None, :, 1]
    T[..., 2] = RHO * ALPHA**2 * duz * n[None, :, 2]
    return V.astype(np.float32), T.astype(np.float32)


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
