"""Closed-form field of a vertically incident P plane wave in a homogeneous
half-space (z up, free surface at z = 0, bottom of the box at z = -L), used to
check FK injection through the Stacey faces:

    u_z(z, t) = g(t - tc - (z + L)/alpha) + g(t - tc - (L - z)/alpha),  u_x = u_y = 0
    sigma_zz = rho alpha^2 du_z/dz,  sigma_xx = sigma_yy = lambda du_z/dz

(the free-surface reflection of the upgoing pulse is the same pulse going down).
"""
import numpy as np

RHO, ALPHA, BETA = 2700.0, 6000.0, 3464.0
SIG, TC = 0.4, 1.4  # Gaussian pulse g(t) = exp(-(t/SIG)^2), centred at the bottom at TC


def g(t):
    return np.exp(-(t / SIG) ** 2)


def dg(t):
    return -2.0 * t / SIG**2 * g(t)


def plane_wave(z, t, L):
    """u_z, v_z and du_z/dz at depth z (m, negative down) and time t."""
    t1 = t - TC - (z + L) / ALPHA
    t2 = t - TC - (L - z) / ALPHA
    return g(t1) + g(t2), dg(t1) + dg(t2), (-dg(t1) + dg(t2)) / ALPHA


def incident_fields(z, normal, nsamp, dt_s, L):
    """Velocity and traction sigma.n of the plane wave at face points of depths
    z (npt,) with outward normals (npt, 3), sample k at time k*dt_s:
    (nsamp, npt, 3) float32 each."""
    n = np.asarray(normal, np.float64).reshape(-1, 3)
    lam = RHO * (ALPHA**2 - 2 * BETA**2)
    t = (np.arange(nsamp) * dt_s)[:, None]
    _, v, duz = plane_wave(np.asarray(z, np.float64)[None, :], t, L)
    V = np.zeros((nsamp, len(n), 3))
    V[..., 2] = v
    T = np.zeros_like(V)
    T[..., 0] = lam * duz * n[None, :, 0]
    T[..., 1] = lam * duz * n[None, :, 1]
    T[..., 2] = RHO * ALPHA**2 * duz * n[None, :, 2]
    return V.astype(np.float32), T.astype(np.float32)
