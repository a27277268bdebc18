"""Structured hexahedral box mesher producing SPECFEM database arrays.

The arrays are those SpecFWAT reads in read_mesh_databases_fwat
(src/specfem_interface/read_mesh_databases.f90:158-189 geometry, :766-792
Stacey faces, :1129-1143 phases, :323-333 mass) after xfwat_mesh_databases
(create_regions_mesh_fwat.f90:113-131, 188, 257); the internal mesher itself
is SPECFEM code absent from the tree, so this is the project's own generator
for the Cartesian box (SUPPRESS_UTM_PROJECTION) of the benchmark configs.

Conventions (shared with the C ABI):
  * box x in [0, lx], y in [0, ly], z in [-lz, 0] (z up, free surface on top)
  * element ispec = ix + nex*(iy + ney*iz), iz = 0 at the bottom
  * local point L = i + 5 j + 25 k; ibool 1-based, numbered by first
    appearance sweeping elements then k, j, i (SPECFEM's get_global
    followed by its cache-friendly indirect renumbering)
  * optional deterministic perturbation of interior anchor nodes
    (general-hexahedron variant); boundary nodes stay on the box faces
  * an element is regular (no per-point Jacobian) iff it is an axis-aligned
    cube of the mesh's first cube size; force_irregular stores every element
  * Stacey faces: xmin, xmax, ymin, ymax, bottom, elements ascending, face
    points with the second tangential index outer
"""
import numpy as np

from .gll import GLL

_CORNER = np.array([[(n >> 0) & 1, (n >> 1) & 1, (n >> 2) & 1] for n in range(8)])


def element_colours(ibool0, nglob):
    """Mesh colouring (the role of setup_color_perm, create_regions_mesh_fwat.f90:283,
    whose algorithm is SPECFEM code absent from the tree): greedy first fit in element
    order -- element e takes the lowest colour not used by an earlier element that
    shares a point with it, so no two elements of a colour share a point.
    ibool0: [nspec, 125] 0-based.  Returns int32 colours [nspec]."""
    ib = np.asarray(ibool0).reshape(-1, 125)
    used = np.zeros(nglob, np.int64)  # bit c set: a colour-c element holds the point
    col = np.zeros(len(ib), np.int32)
    for e, pts in enumerate(ib):
        taken = int(np.bitwise_or.reduce(used[pts]))
        c = (~taken & (taken + 1)).bit_length() - 1  # lowest clear bit
        if c > 62:
            raise ValueError("more than 63 colours")
        col[e] = c
        used[pts] |= np.int64(1) << np.int64(c)
    return col


def _shape(xi, eta, gam):
    s = 2 * _CORNER - 1  # (8,3) signs
    fa = 1 + s[:, 0][None, :] * xi[:, None]
    fb = 1 + s[:, 1][None, :] * eta[:, None]
    fd = 1 + s[:, 2][None, :] * gam[:, None]
    N = fa * fb * fd / 8.0
    dNa = s[:, 0][None, :] * fb * fd / 8.0
    dNb = fa * s[:, 1][None, :] * fd / 8.0
    dNd = fa * fb * s[:, 2][None, :] / 8.0
    return N, np.stack([dNa, dNb, dNd], axis=-1)  # (P,8), (P,8,3)


class SemMesh:
    """Helpers over the SPECFEM database arrays shared by every mesh source
    (ibool0, xyz, jacobian_regular64, jacobian64_irr, irregular_element_number, ...)."""

    def distance_min_glob(self):
        """Minimum distance between neighbouring GLL points of an element
        (check_mesh_distances' distance_min_glob, used by smooth_sem_pde)."""
        i = np.arange(125)
        x = self.xyz[self.ibool0]  # (nspec,125,3)
        best = np.inf
        for stride, cond in ((1, i % 5 < 4), (5, (i // 5) % 5 < 4), (25, i // 25 < 4)):
            src = i[cond]
            d = np.linalg.norm(x[:, src + stride] - x[:, src], axis=-1)
            best = min(best, d.min())
        return float(best)

    def distance_max_glob(self):
        """Maximum distance between neighbouring GLL points of an element
        (check_mesh_distances' distance_max_glob, used by taper_kernel_grid)."""
        i = np.arange(125)
        x = self.xyz[self.ibool0]
        best = 0.0
        for stride, cond in ((1, i % 5 < 4), (5, (i // 5) % 5 < 4), (25, i // 25 < 4)):
            src = i[cond]
            d = np.linalg.norm(x[:, src + stride] - x[:, src], axis=-1)
            best = max(best, d.max())
        return float(best)

    def local_weights(self):
        """jacobian * w_i w_j w_k per local point (float64, [nspec,125])."""
        w = self.gll.w
        ii = np.arange(125)
        w3 = w[ii % 5] * w[(ii // 5) % 5] * w[ii // 25]
        jac = np.full((self.nspec, 125), self.jacobian_regular64)
        sel = self.irregular_element_number > 0
        jac[sel] = self.jacobian64_irr
        return jac * w3[None, :]

    def smooth_arrays(self):
        """The arrays smooth_sem_pde reads, flattened (1-based ibool)."""
        d = dict(nspec=self.nspec, nglob=self.nglob, ibool=self.ibool.ravel(),
                 xstore=self.xstore, ystore=self.ystore, zstore=self.zstore,
                 irregular_element_number=self.irregular_element_number,
                 xix_regular=float(self.xix_regular), jacobian_regular=float(self.jacobian_regular),
                 nspec_irregular=self.nspec_irregular)
        for nm in ("xix", "xiy", "xiz", "etax", "etay", "etaz", "gammax", "gammay", "gammaz", "jacobian"):
            d[nm] = getattr(self, nm).ravel()
        return d


class BoxMesh(SemMesh):
    def __init__(self, nex, ney, nez, lx, ly, lz, perturb=0.0, force_irregular=False):
        self.nex, self.ney, self.nez = nex, ney, nez
        self.lx, self.ly, self.lz = float(lx), float(ly), float(lz)
        self.nspec = nex * ney * nez
        self.perturb = float(perturb)
        self.gll = GLL()
        hx, hy, hz = self.lx / nex, self.ly / ney, self.lz / nez
        self.hx, self.hy, self.hz = hx, hy, hz
        # anchors
        ax, ay, az = np.meshgrid(np.arange(nex + 1), np.arange(ney + 1), np.arange(nez + 1), indexing="ij")
        ax, ay, az = ax.T.ravel(), ay.T.ravel(), az.T.ravel()  # x fastest
        P = np.stack([ax * hx, ay * hy, -self.lz + az * hz], axis=1).astype(np.float64)
        if perturb:
            inner = (ax > 0) & (ax < nex) & (ay > 0) & (ay < ney) & (az > 0) & (az < nez)
            d = np.stack([perturb * hx * np.sin(1.3 * ax + 2.1 * ay + 0.7 * az + 0.4),
                          perturb * hy * np.sin(0.9 * ax - 1.7 * ay + 2.3 * az + 1.1),
                          perturb * hz * np.sin(-1.1 * ax + 0.6 * ay + 1.9 * az + 2.2)], axis=1)
            P[inner] += d[inner]
        self.anchors = P
        e = np.arange(self.nspec)
        self.ex, self.ey, self.ez = e % nex, (e // nex) % ney, e // (nex * ney)
        cid = ((self.ex[:, None] + _CORNER[None, :, 0])
               + (nex + 1) * ((self.ey[:, None] + _CORNER[None, :, 1])
                              + (ney + 1) * (self.ez[:, None] + _CORNER[None, :, 2])))
        self.elem_anchors = P[cid]  # (nspec, 8, 3)

        # global numbering by first appearance (lattice key)
        ii = np.arange(125)
        i, j, k = ii % 5, (ii // 5) % 5, ii // 25
        NX, NY = 4 * nex + 1, 4 * ney + 1
        key = ((4 * self.ex[:, None] + i[None, :])
               + NX * ((4 * self.ey[:, None] + j[None, :]) + NY * (4 * self.ez[:, None] + k[None, :])))
        flat = key.ravel().astype(np.int64)
        uniq, first, inv = np.unique(flat, return_index=True, return_inverse=True)
        order = np.argsort(first, kind="stable")
        rank = np.empty(len(uniq), np.int64)
        rank[order] = np.arange(len(uniq))
        self.nglob = len(uniq)
        self.ibool0 = rank[inv].reshape(self.nspec, 125).astype(np.int32)  # 0-based
        self.ibool = self.ibool0 + 1

        # GLL point coordinates
        xi = self.gll.xi
        N, dN = _shape(xi[i], xi[j], xi[k])
        coords = np.einsum("pa,eac->epc", N, self.elem_anchors)  # (nspec,125,3)
        self.xyz = np.zeros((self.nglob, 3))
        self.xyz[self.ibool0.ravel()] = coords.reshape(-1, 3)
        self.xstore = self.xyz[:, 0].astype(np.float32)
        self.ystore = self.xyz[:, 1].astype(np.float32)
        self.zstore = self.xyz[:, 2].astype(np.float32)

        # regular / irregular classification
        X = self.elem_anchors
        ex_ = X[:, 1, 0] - X[:, 0, 0]
        ey_ = X[:, 2, 1] - X[:, 0, 1]
        ez_ = X[:, 4, 2] - X[:, 0, 2]
        want = X[:, 0:1, :] + _CORNER[None, :, :] * np.stack([ex_, ey_, ez_], axis=1)[:, None, :]
        tolh = 1e-9 * ex_
        cube = np.all(np.abs(X - want) <= tolh[:, None, None], axis=(1, 2))
        cube &= (np.abs(ey_ - ex_) <= tolh) & (np.abs(ez_ - ex_) <= tolh)
        if force_irregular:
            cube[:] = False
        href = -1.0
        if cube.any():
            href = ex_[np.argmax(cube)]
            cube &= np.abs(ex_ - href) <= 1e-9 * href
        irr = ~cube
        self.irregular_element_number = np.zeros(self.nspec, np.int32)
        self.irregular_element_number[irr] = np.arange(1, irr.sum() + 1, dtype=np.int32)
        self.nspec_irregular = int(irr.sum())
        self.xix_regular = np.float32(2.0 / href if href > 0 else 0.0)
        self.jacobian_regular = np.float32(href**3 / 8.0 if href > 0 else 0.0)
        self.xix_regular64 = 2.0 / href if href > 0 else 0.0
        self.jacobian_regular64 = href**3 / 8.0 if href > 0 else 0.0

        # Jacobian at GLL points (double) of the irregular elements, chunked
        self._dN = dN
        sel = np.nonzero(irr)[0]
        names = ["xix", "xiy", "xiz", "etax", "etay", "etaz", "gammax", "gammay", "gammaz"]
        nsel = max(len(sel), 1)
        for nm in names + ["jacobian"]:
            setattr(self, nm, np.zeros((nsel, 125), np.float32))
        self.jacobian64_irr = np.zeros((len(sel), 125))
        for c0 in range(0, len(sel), 32768):
            es = sel[c0:c0 + 32768]
            J = self._jac(es)
            det = np.linalg.det(J)
            Ji = np.linalg.inv(J).reshape(len(es), 125, 9)
            for q, nm in enumerate(names):
                getattr(self, nm)[c0:c0 + len(es)] = Ji[:, :, q]
            self.jacobian[c0:c0 + len(es)] = det
            self.jacobian64_irr[c0:c0 + len(es)] = det
        # per-element jacobian of every local point (smooth_sem_pde's jacobian_all)
        jall = np.full((self.nspec, 125), self.jacobian_regular, np.float32)
        if len(sel):
            jall[sel] = self.jacobian[: len(sel)]
        self.jacobian_all = jall
        self._faces()
        # one rank: every element is inner (phase_ispec_inner_elastic(:,2))
        self.nspec_outer = 0
        self.nspec_inner = self.nspec
        self.num_phase_ispec = self.nspec
        self.phase_ispec_inner = np.zeros((2, self.nspec), np.int32)
        self.phase_ispec_inner[1] = np.arange(1, self.nspec + 1, dtype=np.int32)

    def colour_sort_phases(self):
        """Order each phase list colour by colour (element_colours; elements of a
        colour in their list order), as SPECFEM's mesh colouring reorders the
        elements: the element loop, and so every point's assembly order, then
        runs colour by colour.  Returns the colours."""
        col = element_colours(self.ibool0, self.nglob)
        for iphase, n in ((0, self.nspec_outer), (1, self.nspec_inner)):
            lst = self.phase_ispec_inner[iphase, :n].copy()
            self.phase_ispec_inner[iphase, :n] = lst[np.argsort(col[lst - 1], kind="stable")]
        return col

    def _jac(self, es):
        """J[e, p, c, d] = dx_c/dref_d at the 125 GLL points of elements es."""
        return np.einsum("pad,eac->epcd", self._dN, self.elem_anchors[es])

    def _faces(self):
        nex, ney, nez = self.nex, self.ney, self.nez
        ispecs, ijks, normals, jws = [], [], [], []
        w = self.gll.w
        a = np.tile(np.arange(5), 5)
        b = np.repeat(np.arange(5), 5)
        for side in range(5):
            if side == 0:
                sel = self.ex == 0
            elif side == 1:
                sel = self.ex == nex - 1
            elif side == 2:
                sel = self.ey == 0
            elif side == 3:
                sel = self.ey == ney - 1
            else:
                sel = self.ez == 0
            es = np.nonzero(sel)[0]
            if side <= 1:
                i = np.full(25, 0 if side == 0 else 4); j = a; k = b
                ta, tb = 1, 2
                wab = w[j] * w[k]
            elif side <= 3:
                i = a; j = np.full(25, 0 if side == 2 else 4); k = b
                ta, tb = 0, 2
                wab = w[i] * w[k]
            else:
                i = a; j = b; k = np.zeros(25, int)
                ta, tb = 0, 1
                wab = w[i] * w[j]
            L = i + 5 * j + 25 * k
            Jf = self._jac(es)[:, L]  # (nf,25,3,3)
            t1, t2 = Jf[..., :, ta], Jf[..., :, tb]
            n = np.cross(t1, t2)
            nn = np.linalg.norm(n, axis=-1)
            out = np.zeros(3)
            out[[0, 0, 1, 1, 2][side]] = [-1, 1, -1, 1, -1][side]
            sgn = np.where((n @ out) >= 0, 1.0, -1.0)
            ispecs.append(es + 1)
            ijks.append(np.broadcast_to(np.stack([i, j, k], axis=1)[None] + 1, (len(es), 25, 3)))
            normals.append(sgn[..., None] * n / nn[..., None])
            jws.append(nn * wab[None, :])
        self.num_abs_boundary_faces = int(sum(len(x) for x in ispecs))
        self.abs_boundary_ispec = np.concatenate(ispecs).astype(np.int32)
        self.abs_boundary_ijk = np.ascontiguousarray(np.concatenate(ijks), dtype=np.int32)
        self.abs_boundary_normal = np.ascontiguousarray(np.concatenate(normals), dtype=np.float32)
        self.abs_boundary_jacobian2Dw = np.ascontiguousarray(np.concatenate(jws), dtype=np.float32)

    def locate(self, x, y, z):
        """First element (ascending ispec) containing the point and its
        reference coordinates (Newton on the trilinear map)."""
        if not self.perturb:
            # undeformed box: per-axis lowest element whose interval holds the
            # coordinate (the lowest ispec), reference coordinate in closed form
            out, ref = [], []
            for v, lo, h, n in ((x, 0.0, self.hx, self.nex), (y, 0.0, self.hy, self.ney),
                                (z, -self.lz, self.hz, self.nez)):
                t = (v - lo) / h
                if t < -1e-9 or t > n + 1e-9:
                    raise ValueError(f"point ({x}, {y}, {z}) is outside the mesh")
                e = min(max(int(np.ceil(t - 1e-9)) - 1, 0), n - 1)
                out.append(e)
                ref.append(min(max(2.0 * (t - e) - 1.0, -1.0), 1.0))
            return int(out[0] + self.nex * (out[1] + self.ney * out[2])), np.array(ref)
        X = self.elem_anchors
        if getattr(self, "_bbox", None) is None:
            lo, hi = X.min(axis=1), X.max(axis=1)
            self._bbox = (lo, hi, 1e-9 * (hi - lo))
        lo, hi, tol = self._bbox
        p = np.array([x, y, z], float)
        cand = np.nonzero(np.all((p >= lo - tol) & (p <= hi + tol), axis=1))[0]
        for e in cand:
            r = np.zeros(3)
            for _ in range(50):
                N, dN = _shape(r[0:1], r[1:2], r[2:3])
                xx = N[0] @ X[e]
                Jm = np.einsum("ad,ac->cd", dN[0], X[e])
                dr = np.linalg.solve(Jm, p - xx)
                r += dr
                if np.abs(dr).sum() < 1e-15:
                    break
            if np.all(np.abs(r) <= 1 + 1e-9):
                return int(e), np.clip(r, -1.0, 1.0)
        raise ValueError(f"point ({x}, {y}, {z}) is outside the mesh")

    def inverse_jacobian_at(self, e, r):
        N, dN = _shape(r[0:1], r[1:2], r[2:3])
        Jm = np.einsum("ad,ac->cd", dN[0], self.elem_anchors[e])
        return np.linalg.inv(Jm)
