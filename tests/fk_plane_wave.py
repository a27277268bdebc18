"""Closed-form fields of vertically incident plane waves in a homogeneous
half-space (z up, free surface at z = 0, bottom of the box at z = -L), used to
check FK injection and plain absorption through the Stacey faces.

P wave:  u_z(z, t) = g(t - tc - (z + L)/alpha) + g(t - tc - (L - z)/alpha),  u_x = u_y = 0
         sigma_zz = rho alpha^2 du_z/dz,  sigma_xx = sigma_yy = lambda du_z/dz
S wave (polarised along x):
         u_x(z, t) = g(t - tc - (z + L)/beta) + g(t - tc - (L - z)/beta),  u_y = u_z = 0
         sigma_xz = sigma_zx = mu du_x/dz

The first term is the upgoing pulse, the second its free-surface reflection
(the same pulse going down: at normal incidence the free surface reflects P
and S with coefficient +1 and no conversion).
"""
import numpy as np

RHO, ALPHA, BETA = 2700.0, 6000.0, 3464.0
SIG, TC = 0.4, 1.4  # Gaussian pulse g(t) = exp(-(t/SIG)^2), centred at the bottom at TC


def g(t):
    return np.exp(-(t / SIG) ** 2)


def dg(t):
    return -2.0 * t / SIG**2 * g(t)


def plane_wave(z, t, L, speed=ALPHA, parts="total"):
    """u, v and du/dz of the polarised component at depth z (m, negative down)
    and time t; parts = "total" (upgoing + reflection), "up" or "down"."""
    t1 = t - TC - (z + L) / speed
    t2 = t - TC - (L - z) / speed
    up = (g(t1), dg(t1), -dg(t1) / speed)
    down = (g(t2), dg(t2), dg(t2) / speed)
    if parts == "up":
        return up
    if parts == "down":
        return down
    return tuple(a + b for a, b in zip(up, down))


def incident_fields(z, normal, nsamp, dt_s, L, wave="P", parts=None):
    """Velocity and traction sigma.n of the plane wave at face points of depths
    z (npt,) with outward normals (npt, 3), sample k at time k*dt_s:
    (nsamp, npt, 3) float32 each.  parts: None = the total field everywhere;
    an array of "total"/"up"/"down" per face point selects the part injected
    at that point."""
    n = np.asarray(normal, np.float64).reshape(-1, 3)
    z = np.asarray(z, np.float64)
    t = (np.arange(nsamp) * dt_s)[:, None]
    speed = ALPHA if wave == "P" else BETA
    if parts is None:
        _, v, dudz = plane_wave(z[None, :], t, L, speed)
    else:
        parts = np.asarray(parts)
        v = np.zeros((nsamp, len(z)))
        dudz = np.zeros((nsamp, len(z)))
        for kind in ("total", "up", "down"):
            sel = parts == kind
            if sel.any():
                _, v[:, sel], dudz[:, sel] = plane_wave(z[None, sel], t, L, speed, kind)
    V = np.zeros((nsamp, len(n), 3))
    T = np.zeros_like(V)
    if wave == "P":
        lam = RHO * (ALPHA**2 - 2 * BETA**2)
        V[..., 2] = v
        T[..., 0] = lam * dudz * n[None, :, 0]
        T[..., 1] = lam * dudz * n[None, :, 1]
        T[..., 2] = RHO * ALPHA**2 * dudz * n[None, :, 2]
    else:
        mu = RHO * BETA**2
        V[..., 0] = v
        T[..., 0] = mu * dudz * n[None, :, 2]
        T[..., 2] = mu * dudz * n[None, :, 0]
    return V.astype(np.float32), T.astype(np.float32)
