"""Point moment-tensor sources, Ricker source time function, receivers.

Replaces what setup_sources_receivers_fwat (src/specfem_interface/
setup_sources_receivers_fwat.f90:29-88) and locate_source_fwat
(locate_source_fwat.f90:31-800) hand to the solver:
  * location: first element (ascending ispec) containing the point, reference
    coordinates by Newton on the element map;
  * sourcearray(c, i,j,k) = sum_d G[c][d] dphi_ijk/dref_d, G = M . grad(ref),
    i.e. the weak form of f = -M . grad(delta);
  * Ricker STF with dominant frequency f0 = the CMT "half duration" field
    (locate_source_fwat.f90:546-550), t0 = 1.2 / f0 (:502-503), evaluated at
    (it-1) dt - t0 - tshift for step it;
  * receivers: Lagrange interpolants h_i(xi) h_j(eta) h_k(gamma).
"""
import numpy as np

from .gll import lagrange_basis


def ricker(t, f0):
    a = (np.pi * f0) ** 2
    t = np.asarray(t, np.float64)
    return (1.0 - 2.0 * a * t * t) * np.exp(-a * t * t)


class Sources:
    def __init__(self, mesh, xyz, mt, f0, tshift=None):
        xyz = np.reshape(np.asarray(xyz, float), (-1, 3))
        self.xyz = xyz
        self.nsources = xyz.shape[0]
        self.mt = np.reshape(np.asarray(mt, float), (self.nsources, 6))  # Mxx Myy Mzz Mxy Mxz Myz
        self.f0 = np.reshape(np.asarray(f0, float), (self.nsources,))
        self.tshift = np.zeros(self.nsources) if tshift is None else np.reshape(np.asarray(tshift, float), (self.nsources,))
        self.ispec = np.zeros(self.nsources, np.int32)  # 1-based
        self.sourcearrays = np.zeros((self.nsources, 125, 3), np.float32)
        xi = mesh.gll.xi
        for s in range(self.nsources):
            e, r = mesh.locate(*xyz[s])
            self.ispec[s] = e + 1
            Ji = mesh.inverse_jacobian_at(e, r)  # Ji[d][c] = dref_d/dx_c
            hx, hxd = lagrange_basis(xi, r[0])
            hy, hyd = lagrange_basis(xi, r[1])
            hz, hzd = lagrange_basis(xi, r[2])
            m6 = self.mt[s]
            M = np.array([[m6[0], m6[3], m6[4]], [m6[3], m6[1], m6[5]], [m6[4], m6[5], m6[2]]])
            G = M @ Ji.T  # G[c][d] = sum_j M[c][j] dref_d/dx_j
            ii = np.arange(125)
            i, j, k = ii % 5, (ii // 5) % 5, ii // 25
            d = np.stack([hxd[i] * hy[j] * hz[k], hx[i] * hyd[j] * hz[k], hx[i] * hy[j] * hzd[k]], axis=1)
            self.sourcearrays[s] = (d @ G.T).astype(np.float32)

    def stf(self, nstep, dt):
        """(NSTEP, NSOURCES) float64 source time function samples."""
        it = np.arange(1, nstep + 1)
        t = (it - 1)[:, None] * dt - 1.2 / self.f0[None, :] - self.tshift[None, :]
        return np.ascontiguousarray(ricker(t, self.f0[None, :]))


class Receivers:
    def __init__(self, mesh, xyz):
        xyz = np.reshape(np.asarray(xyz, float), (-1, 3))
        self.nrec = xyz.shape[0]
        self.ispec = np.zeros(self.nrec, np.int32)
        self.hlagrange = np.zeros((self.nrec, 125), np.float32)
        xi = mesh.gll.xi
        ii = np.arange(125)
        i, j, k = ii % 5, (ii // 5) % 5, ii // 25
        for r in range(self.nrec):
            e, ref = mesh.locate(*xyz[r])
            self.ispec[r] = e + 1
            hx, _ = lagrange_basis(xi, ref[0])
            hy, _ = lagrange_basis(xi, ref[1])
            hz, _ = lagrange_basis(xi, ref[2])
            self.hlagrange[r] = (hx[i] * hy[j] * hz[k]).astype(np.float32)


def random_moment_tensor(rng, m0=1e15):
    return m0 * rng.standard_normal(6)


def surface_grid(mesh, n_x, n_y):
    """n_x x n_y receivers at the surface on a centred regular grid."""
    xs = (np.arange(n_x) + 0.5) * mesh.lx / n_x
    ys = (np.arange(n_y) + 0.5) * mesh.ly / n_y
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
