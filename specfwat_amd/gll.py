"""GLL points, weights and derivative tables for NGLL = 5.

Mirrors define_derivation_matrices as called by smooth_sem_pde
(src/postproc/smooth.f90:118-131) and SPECFEM's setup_GLL_points
(prepare_fwd.f90:90): tables are built in double precision and stored as
CUSTOM_REAL (float32), Fortran (i,j) order, with
    hprime_xx(i,j)     = l_j'(xi_i)
    hprimewgll_xx(i,j) = w_i * l_j'(xi_i)
    wgllwgll_xy(i,j)   = w_i * w_j
"""
import numpy as np

NGLL = 5


def gll_points_weights():
    s = np.sqrt(3.0 / 7.0)
    xi = np.array([-1.0, -s, 0.0, s, 1.0])
    # P4(x) = (35 x^4 - 30 x^2 + 3) / 8 ; w = 2 / (N (N+1) P_N(x)^2), N = 4
    p4 = (35 * xi**4 - 30 * xi**2 + 3) / 8.0
    w = 2.0 / (4 * 5 * p4 * p4)
    return xi, w


def lagrange_basis(xi_nodes, x):
    """Lagrange interpolants h_j(x) and derivatives h_j'(x) on the given nodes."""
    n = len(xi_nodes)
    h = np.ones(n)
    hd = np.zeros(n)
    for j in range(n):
        for k in range(n):
            if k != j:
                h[j] *= (x - xi_nodes[k]) / (xi_nodes[j] - xi_nodes[k])
        for m in range(n):
            if m == j:
                continue
            q = 1.0 / (xi_nodes[j] - xi_nodes[m])
            for k in range(n):
                if k != j and k != m:
                    q *= (x - xi_nodes[k]) / (xi_nodes[j] - xi_nodes[k])
            hd[j] += q
    return h, hd


def derivative_matrix(xi):
    """D[i, j] = l_j'(xi_i) via the barycentric form."""
    n = len(xi)
    c = np.array([np.prod([xi[i] - xi[k] for k in range(n) if k != i]) for i in range(n)])
    D = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                D[i, j] = c[i] / (c[j] * (xi[i] - xi[j]))
            else:
                D[i, i] = sum(1.0 / (xi[i] - xi[k]) for k in range(n) if k != i)
    return D


class GLL:
    """Double-precision nodes/weights and float32 Fortran-ordered tables."""

    def __init__(self):
        self.xi, self.w = gll_points_weights()
        D = derivative_matrix(self.xi)
        self.hprime64 = D
        self.hprimewgll64 = self.w[:, None] * D
        # Fortran column-major (i,j) -> flat index i + 5 j, i.e. C array [j][i]
        self.hprime_xx = np.ascontiguousarray(D.T, dtype=np.float32)
        self.hprimewgll_xx = np.ascontiguousarray(self.hprimewgll64.T, dtype=np.float32)
        ww = np.outer(self.w, self.w)  # ww[i, j] = w_i w_j (symmetric)
        self.wgllwgll_xy = np.ascontiguousarray(ww.T, dtype=np.float32)
        self.wgllwgll_xz = self.wgllwgll_xy.copy()
        self.wgllwgll_yz = self.wgllwgll_xy.copy()
        # transpose used by the smoother: hprime_xxT(j,i) = hprime_xx(i,j)
        self.hprime_xxT = np.ascontiguousarray(D, dtype=np.float32)
