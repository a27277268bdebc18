"""SPECFEM solver databases: proc%06d_external_mesh.bin read into the engine's mesh.

The file is Fortran sequential unformatted output: every record is a 4-byte
length marker, the payload, and the same marker again (gfortran splits a
record longer than 2**31-1 bytes into subrecords whose leading marker is
negative while the record continues).  The record order is the one
read_mesh_databases_fwat reads (src/specfem_interface/read_mesh_databases.f90):

  :158-193   NSPEC_AB, NGLOB_AB, NSPEC_IRREGULAR, ibool, x/y/zstore,
             irregular_element_number, xix_regular, jacobian_regular, the nine
             inverse-Jacobian arrays and jacobianstore (NSPEC_IRREGULAR elements,
             one value when there is none), kappa/mu/rhostore, the three
             ispec_is_* domain flags, sentinel 9999
  :287       rmass_acoustic                          (acoustic domain)
  :455-487   rmass, [rmass_ocean_load], rho_vp, rho_vs  (elastic domain)
  :597-609   eleven poroelastic arrays               (poroelastic domain)
  :652-760   C-PML records                           (PML_CONDITIONS)
  :766-815   absorbing faces, then rmassx/y/z and rmassz_acoustic (Stacey)
  :818-820   sentinel 9998
  :824-869   nspec2D_xmin..top and the ibelm_* lists
  :872-994   free surface and coupling faces
  :997-1000  sentinel 9997
  :1003-1047 MPI interfaces, sentinel 9996
  :1050-1073 21 c_ij                                 (elastic anisotropy)
  :1103-1167 ispec_is_inner, inner/outer phase lists per domain
  :1170-1215 mesh colouring                          (USE_MESH_COLORING_GPU)
  :1229-1261 surface flags, element adjacency, sentinel 9995

Which optional groups are present is decided by the Par_file flags passed to
the reader, as in the reference, and by counts read earlier in the file.
Logical arrays are 4-byte logicals; reals are CUSTOM_REAL (float32).
"""
import struct

import numpy as np

from .gll import GLL, lagrange_basis
from .mesh import SemMesh
from .model import CIJ_NAMES, moduli

F32 = np.float32
I32 = np.int32
NG3 = 125
NGLLSQUARE = 25
_JNAMES = ["xix", "xiy", "xiz", "etax", "etay", "etaz", "gammax", "gammay", "gammaz"]


class DatabaseError(ValueError):
    pass


class _Records:
    """Sequential reader of Fortran unformatted records (either byte order)."""

    def __init__(self, path):
        self.path = path
        self.f = open(path, "rb")
        head = self.f.read(4)
        if len(head) < 4:
            raise DatabaseError(f"{path}: empty file")
        little, big = struct.unpack("<i", head)[0], struct.unpack(">i", head)[0]
        # the first record is NSPEC_AB, one 4-byte integer
        if little == 4:
            self.bo = "<"
        elif big == 4:
            self.bo = ">"
        else:
            raise DatabaseError(f"{path}: first record is not a 4-byte integer (marker {little})")
        self.f.seek(0)
        self.n = 0

    def close(self):
        self.f.close()

    def _marker(self, name):
        b = self.f.read(4)
        if len(b) < 4:
            raise DatabaseError(f"{self.path}: file ends before record {self.n + 1} ({name})")
        return struct.unpack(self.bo + "i", b)[0]

    def read(self, dtype, name, count=None):
        dt = np.dtype(dtype).newbyteorder(self.bo)
        parts = []
        while True:
            h = self._marker(name)
            n = abs(h)
            buf = np.fromfile(self.f, np.uint8, n)
            if buf.size != n:
                raise DatabaseError(f"{self.path}: record {self.n + 1} ({name}) truncated")
            t = self._marker(name)
            if abs(t) != n:
                raise DatabaseError(f"{self.path}: record {self.n + 1} ({name}): markers {h} and {t} disagree")
            parts.append(buf)
            if h >= 0:
                break
        self.n += 1
        raw = parts[0] if len(parts) == 1 else np.concatenate(parts)
        if raw.size % dt.itemsize:
            raise DatabaseError(f"{self.path}: record {self.n} ({name}) holds {raw.size} bytes, "
                                f"not a whole number of {dt.itemsize}-byte values")
        a = raw.view(dt).astype(dt.newbyteorder("="), copy=False)
        if count is not None and a.size != count:
            raise DatabaseError(f"{self.path}: record {self.n} ({name}) holds {a.size} values, "
                                f"expected {count}")
        return a

    def scalar(self, dtype, name):
        return self.read(dtype, name, 1)[0]

    def ints(self, name, count):
        return self.read(I32, name, count)

    def reals(self, name, count):
        return self.read(F32, name, count)

    def logicals(self, name, count):
        return self.read(I32, name, count) != 0

    def sentinel(self, value, position):
        got = int(self.scalar(I32, f"itest {value}"))
        if got != value:
            raise DatabaseError(f"{self.path}: Error database read at position {position} "
                                f"(sentinel {got}, expected {value})")


def read_external_mesh(path, anisotropy=False, approximate_ocean_load=False, pml_conditions=False,
                       stacey_absorbing_conditions=True, use_mesh_coloring_gpu=False):
    """Every record of one proc%06d_external_mesh.bin, by the reference's array names.

    The flags are the Par_file switches the reference consults while reading
    (ANISOTROPY, APPROXIMATE_OCEAN_LOAD, PML_CONDITIONS,
    STACEY_ABSORBING_CONDITIONS, USE_MESH_COLORING_GPU).  Arrays keep the
    Fortran element order (first index fastest); multi-index arrays are
    returned with the last Fortran index first, e.g. ibool as (NSPEC, 125)
    and abs_boundary_normal as (faces, 25, 3).
    """
    r = _Records(path)
    d = {}
    try:
        # read_mesh_databases.f90:158-193
        nspec = d["NSPEC_AB"] = int(r.scalar(I32, "NSPEC_AB"))
        nglob = d["NGLOB_AB"] = int(r.scalar(I32, "NGLOB_AB"))
        nirr = d["NSPEC_IRREGULAR"] = int(r.scalar(I32, "NSPEC_IRREGULAR"))
        if nspec <= 0 or nglob <= 0 or not 0 <= nirr <= nspec:
            raise DatabaseError(f"{path}: NSPEC_AB {nspec}, NGLOB_AB {nglob}, NSPEC_IRREGULAR {nirr}")
        nloc = nspec * NG3
        d["ibool"] = r.ints("ibool", nloc).reshape(nspec, NG3)
        for c in "xyz":
            d[c + "store"] = r.reals(c + "store", nglob)
        d["irregular_element_number"] = r.ints("irregular_element_number", nspec)
        d["xix_regular"] = r.scalar(F32, "xix_regular")
        d["jacobian_regular"] = r.scalar(F32, "jacobian_regular")
        nj = max(nirr, 1) * NG3 if nirr else None
        for nm in _JNAMES + ["jacobian"]:
            a = r.reals(nm + "store", nj)
            if nirr == 0 and a.size not in (0, 1, NG3):
                raise DatabaseError(f"{path}: {nm}store holds {a.size} values with no irregular element")
            d[nm + "store"] = a.reshape(nirr, NG3) if nirr else a
        for nm in ("kappastore", "mustore", "rhostore"):
            d[nm] = r.reals(nm, nloc).reshape(nspec, NG3)
        for nm in ("ispec_is_acoustic", "ispec_is_elastic", "ispec_is_poroelastic"):
            d[nm] = r.logicals(nm, nspec)
        r.sentinel(9999, 1)
        acoustic = bool(d["ispec_is_acoustic"].any())
        elastic = bool(d["ispec_is_elastic"].any())
        poro = bool(d["ispec_is_poroelastic"].any())
        d["ACOUSTIC_SIMULATION"], d["ELASTIC_SIMULATION"], d["POROELASTIC_SIMULATION"] = acoustic, elastic, poro
        if not (acoustic or elastic or poro):
            raise DatabaseError(f"{path}: Error no simulation type defined")
        # :287, :455-487, :597-609
        if acoustic:
            d["rmass_acoustic"] = r.reals("rmass_acoustic", nglob)
        if elastic:
            d["rmass"] = r.reals("rmass", nglob)
            if approximate_ocean_load:
                d["rmass_ocean_load"] = r.reals("rmass_ocean_load", nglob)
            d["rho_vp"] = r.reals("rho_vp", nloc).reshape(nspec, NG3)
            d["rho_vs"] = r.reals("rho_vs", nloc).reshape(nspec, NG3)
        if poro:
            for nm in ("rmass_solid_poroelastic", "rmass_fluid_poroelastic", "rhoarraystore",
                       "kappaarraystore", "etastore", "tortstore", "permstore", "phistore",
                       "rho_vpI", "rho_vpII", "rho_vsI"):
                d[nm] = r.reals(nm, None)
        # :652-760
        d["NSPEC_CPML"] = 0
        if pml_conditions:
            d["NSPEC_CPML"] = ncpml = int(r.scalar(I32, "NSPEC_CPML"))
            for nm in ("CPML_width_x", "CPML_width_y", "CPML_width_z",
                       "min_distance_between_CPML_parameter"):
                d[nm] = r.scalar(F32, nm)
            if ncpml > 0:
                d["CPML_regions"] = r.ints("CPML_regions", ncpml)
                d["CPML_to_spec"] = r.ints("CPML_to_spec", ncpml)
                d["is_CPML"] = r.logicals("is_CPML", nspec)
                for nm in ("d_store_x", "d_store_y", "d_store_z", "k_store_x", "k_store_y",
                           "k_store_z", "alpha_store_x", "alpha_store_y", "alpha_store_z"):
                    d[nm] = r.reals(nm, ncpml * NG3)
                na = d["nglob_interface_PML_acoustic"] = int(r.scalar(I32, "nglob_interface_PML_acoustic"))
                ne = d["nglob_interface_PML_elastic"] = int(r.scalar(I32, "nglob_interface_PML_elastic"))
                if na > 0:
                    d["points_interface_PML_acoustic"] = r.ints("points_interface_PML_acoustic", na)
                if ne > 0:
                    d["points_interface_PML_elastic"] = r.ints("points_interface_PML_elastic", ne)
        # :766-815
        nf = d["num_abs_boundary_faces"] = int(r.scalar(I32, "num_abs_boundary_faces"))
        if nf < 0:
            raise DatabaseError(f"{path}: num_abs_boundary_faces {nf} < 0")
        if nf > 0:
            _faces(r, d, "abs_boundary", nf)
            if stacey_absorbing_conditions and not pml_conditions:
                if elastic:
                    for c in "xyz":
                        d["rmass" + c] = r.reals("rmass" + c, nglob)
                if acoustic:
                    d["rmassz_acoustic"] = r.reals("rmassz_acoustic", nglob)
        r.sentinel(9998, 2)
        # :824-869
        sides = ("xmin", "xmax", "ymin", "ymax", "bottom", "top")
        counts = [int(r.scalar(I32, "nspec2D_" + s)) for s in sides]
        for s, n in zip(sides, counts):
            d["nspec2D_" + s] = n
        for s, n in zip(sides, counts):
            if n > 0:
                d["ibelm_" + s] = r.ints("ibelm_" + s, n)
        # :872-994
        for key, prefix in (("num_free_surface_faces", "free_surface"),
                            ("num_coupling_ac_el_faces", "coupling_ac_el"),
                            ("num_coupling_ac_po_faces", "coupling_ac_po")):
            n = d[key] = int(r.scalar(I32, key))
            if n > 0:
                _faces(r, d, prefix, n)
        n = d["num_coupling_el_po_faces"] = int(r.scalar(I32, "num_coupling_el_po_faces"))
        if n > 0:
            d["coupling_el_po_ispec"] = r.ints("coupling_el_po_ispec", n)
            d["coupling_po_el_ispec"] = r.ints("coupling_po_el_ispec", n)
            d["coupling_el_po_ijk"] = r.ints("coupling_el_po_ijk", 3 * NGLLSQUARE * n).reshape(n, NGLLSQUARE, 3)
            d["coupling_po_el_ijk"] = r.ints("coupling_po_el_ijk", 3 * NGLLSQUARE * n).reshape(n, NGLLSQUARE, 3)
            d["coupling_el_po_jacobian2Dw"] = r.reals("coupling_el_po_jacobian2Dw", NGLLSQUARE * n).reshape(n, NGLLSQUARE)
            d["coupling_el_po_normal"] = r.reals("coupling_el_po_normal", 3 * NGLLSQUARE * n).reshape(n, NGLLSQUARE, 3)
        r.sentinel(9997, 3)
        # :1003-1047
        ni = d["num_interfaces_ext_mesh"] = int(r.scalar(I32, "num_interfaces_ext_mesh"))
        mx = d["max_nibool_interfaces_ext_mesh"] = int(r.scalar(I32, "max_nibool_interfaces_ext_mesh"))
        if ni > 0:
            d["my_neighbors_ext_mesh"] = r.ints("my_neighbors_ext_mesh", ni)
            d["nibool_interfaces_ext_mesh"] = r.ints("nibool_interfaces_ext_mesh", ni)
            d["ibool_interfaces_ext_mesh"] = r.ints("ibool_interfaces_ext_mesh", mx * ni).reshape(ni, mx)
        r.sentinel(9996, 3)
        # :1050-1073
        if elastic and anisotropy:
            for nm in CIJ_NAMES:
                d[nm + "store"] = r.reals(nm + "store", nloc).reshape(nspec, NG3)
        # :1103-1167
        d["ispec_is_inner"] = r.logicals("ispec_is_inner", nspec)
        for dom, present in (("acoustic", acoustic), ("elastic", elastic), ("poroelastic", poro)):
            if not present:
                continue
            io = r.ints(f"nspec_inner_{dom},nspec_outer_{dom}", 2)
            d[f"nspec_inner_{dom}"], d[f"nspec_outer_{dom}"] = int(io[0]), int(io[1])
            n = d[f"num_phase_ispec_{dom}"] = int(r.scalar(I32, f"num_phase_ispec_{dom}"))
            if n < 0:
                raise DatabaseError(f"{path}: num_phase_ispec_{dom} is < zero")
            if n > 0:
                d[f"phase_ispec_inner_{dom}"] = r.ints(f"phase_ispec_inner_{dom}", 2 * n).reshape(2, n)
        # :1170-1215
        if use_mesh_coloring_gpu:
            for dom, present in (("acoustic", acoustic), ("elastic", elastic)):
                if not present:
                    continue
                oi = r.ints(f"num_colors_outer_{dom},num_colors_inner_{dom}", 2)
                d[f"num_colors_outer_{dom}"], d[f"num_colors_inner_{dom}"] = int(oi[0]), int(oi[1])
                d[f"num_elem_colors_{dom}"] = r.ints(f"num_elem_colors_{dom}", int(oi[0] + oi[1]))
        # :1229-1261
        d["nfaces_surface"] = int(r.scalar(I32, "nfaces_surface"))
        d["ispec_is_surface_external_mesh"] = r.logicals("ispec_is_surface_external_mesh", nspec)
        d["iglob_is_surface_external_mesh"] = r.logicals("iglob_is_surface_external_mesh", nglob)
        nn = d["num_neighbors_all"] = int(r.scalar(I32, "num_neighbors_all"))
        d["neighbors_xadj"] = r.ints("neighbors_xadj", nspec + 1)
        d["neighbors_adjncy"] = r.ints("neighbors_adjncy", nn)
        r.sentinel(9995, "4 (end)")
        d["records"] = r.n
    finally:
        r.close()
    ib = d["ibool"]
    if ib.min() < 1 or ib.max() > nglob:
        raise DatabaseError(f"{path}: ibool outside 1..{nglob}")
    irr = d["irregular_element_number"]
    if irr.min() < 0 or irr.max() > nirr or int((irr > 0).sum()) != nirr:
        raise DatabaseError(f"{path}: irregular_element_number does not number {nirr} elements")
    return d


def _faces(r, d, prefix, n):
    d[prefix + "_ispec"] = r.ints(prefix + "_ispec", n)
    d[prefix + "_ijk"] = r.ints(prefix + "_ijk", 3 * NGLLSQUARE * n).reshape(n, NGLLSQUARE, 3)
    d[prefix + "_jacobian2Dw"] = r.reals(prefix + "_jacobian2Dw", NGLLSQUARE * n).reshape(n, NGLLSQUARE)
    d[prefix + "_normal"] = r.reals(prefix + "_normal", 3 * NGLLSQUARE * n).reshape(n, NGLLSQUARE, 3)


class DatabaseMesh(SemMesh):
    """A solver database as the engine's mesh (the attributes ElasticRun and the
    smoother read), plus its stored materials.

    Point location uses the element's 125 GLL points as its geometry (Newton on
    the degree-4 Lagrange map, exact for the 8- and 27-node elements SPECFEM
    meshes), first element in ascending ispec that holds the point.
    """

    def __init__(self, db, path=None):
        self.db, self.path = db, path
        if not db["ELASTIC_SIMULATION"] or db["ACOUSTIC_SIMULATION"] or db["POROELASTIC_SIMULATION"]:
            raise DatabaseError("the engine runs purely elastic meshes; this database has "
                                f"acoustic={db['ACOUSTIC_SIMULATION']} poroelastic={db['POROELASTIC_SIMULATION']}")
        if db["NSPEC_CPML"]:
            raise DatabaseError("C-PML databases are not supported (Stacey absorbing boundaries only)")
        if db["num_interfaces_ext_mesh"]:
            raise DatabaseError("the database is one slice of a partitioned mesh "
                                f"({db['num_interfaces_ext_mesh']} MPI interfaces); one GPU runs one whole mesh: "
                                "load every slice with load_partitioned_mesh")
        self.gll = GLL()
        self.nspec, self.nglob = db["NSPEC_AB"], db["NGLOB_AB"]
        self.nspec_irregular = db["NSPEC_IRREGULAR"]
        self.ibool = np.ascontiguousarray(db["ibool"], I32)
        self.ibool0 = self.ibool - 1
        self.xstore, self.ystore, self.zstore = db["xstore"], db["ystore"], db["zstore"]
        self.xyz = np.stack([self.xstore, self.ystore, self.zstore], axis=1).astype(np.float64)
        self.irregular_element_number = np.ascontiguousarray(db["irregular_element_number"], I32)
        self.xix_regular, self.jacobian_regular = F32(db["xix_regular"]), F32(db["jacobian_regular"])
        self.xix_regular64, self.jacobian_regular64 = float(self.xix_regular), float(self.jacobian_regular)
        nsel = max(self.nspec_irregular, 1)
        for nm in _JNAMES + ["jacobian"]:
            a = np.zeros((nsel, NG3), F32)
            if self.nspec_irregular:
                a[:] = db[nm + "store"]
            setattr(self, nm, a)
        self.jacobian64_irr = self.jacobian[: self.nspec_irregular].astype(np.float64)
        jall = np.full((self.nspec, NG3), self.jacobian_regular, F32)
        sel = self.irregular_element_number > 0
        jall[sel] = self.jacobian[self.irregular_element_number[sel] - 1]
        self.jacobian_all = jall
        nf = db["num_abs_boundary_faces"]
        self.stacey = bool(nf > 0 and "rmassx" in db)
        self.num_abs_boundary_faces = nf
        if nf:
            self.abs_boundary_ispec = np.ascontiguousarray(db["abs_boundary_ispec"], I32)
            self.abs_boundary_ijk = np.ascontiguousarray(db["abs_boundary_ijk"], I32)
            self.abs_boundary_normal = np.ascontiguousarray(db["abs_boundary_normal"], F32)
            self.abs_boundary_jacobian2Dw = np.ascontiguousarray(db["abs_boundary_jacobian2Dw"], F32)
        else:
            self.abs_boundary_ispec = np.zeros(0, I32)
            self.abs_boundary_ijk = np.zeros((0, NGLLSQUARE, 3), I32)
            self.abs_boundary_normal = np.zeros((0, NGLLSQUARE, 3), F32)
            self.abs_boundary_jacobian2Dw = np.zeros((0, NGLLSQUARE), F32)
        self.nspec_inner = db["nspec_inner_elastic"]
        self.nspec_outer = db["nspec_outer_elastic"]
        self.num_phase_ispec = db["num_phase_ispec_elastic"]
        self.phase_ispec_inner = np.ascontiguousarray(
            db.get("phase_ispec_inner_elastic", np.zeros((2, 1), I32)), I32)
        self._bbox = None

    def materials(self):
        """The stored material arrays in the form ElasticRun(materials=...) takes:
        rmassx/y/z are rmass where the database holds no Stacey masses."""
        d = self.db
        out = dict(rho=d["rhostore"], kappa=d["kappastore"], mu=d["mustore"], rho_vp=d["rho_vp"],
                   rho_vs=d["rho_vs"])
        for c in "xyz":
            out["rmass" + c] = d.get("rmass" + c, d["rmass"])
        return out

    def model(self):
        """rho, vp, vs (float32) recovered from the stored moduli (mu = rho vs^2,
        kappa = rho vp^2 - 4/3 mu)."""
        rho = self.db["rhostore"].astype(np.float64)
        mu, kappa = self.db["mustore"].astype(np.float64), self.db["kappastore"].astype(np.float64)
        vs = np.sqrt(mu / rho)
        vp = np.sqrt((kappa + 4.0 / 3.0 * mu) / rho)
        return rho.astype(F32), vp.astype(F32), vs.astype(F32)

    def cij(self):
        """(nspec, 125, 21) stored c_ij when the database holds them, else None."""
        if "c11store" not in self.db:
            return None
        return np.stack([self.db[nm + "store"] for nm in CIJ_NAMES], axis=-1)

    # ---------------------------------------------------------------- point location
    def _element_xyz(self, e):
        return self.xyz[self.ibool0[e]]  # (125, 3), L = i + 5 j + 25 k

    def _map(self, e, r):
        hx, dx = lagrange_basis(self.gll.xi, r[0])
        hy, dy = lagrange_basis(self.gll.xi, r[1])
        hz, dz = lagrange_basis(self.gll.xi, r[2])
        ii = np.arange(NG3)
        i, j, k = ii % 5, (ii // 5) % 5, ii // 25
        N = hx[i] * hy[j] * hz[k]
        dN = np.stack([dx[i] * hy[j] * hz[k], hx[i] * dy[j] * hz[k], hx[i] * hy[j] * dz[k]], axis=1)
        X = self._element_xyz(e)
        return N @ X, X.T @ dN  # x(r), J[c][d] = dx_c/dref_d

    def locate(self, x, y, z):
        if self._bbox is None:
            lo = np.empty((self.nspec, 3))
            hi = np.empty((self.nspec, 3))
            for c0 in range(0, self.nspec, 32768):
                X = self.xyz[self.ibool0[c0:c0 + 32768]]
                lo[c0:c0 + 32768], hi[c0:c0 + 32768] = X.min(axis=1), X.max(axis=1)
            self._bbox = (lo, hi, 1e-6 * (hi - lo).max(axis=1, keepdims=True))
        lo, hi, tol = self._bbox
        p = np.array([x, y, z], float)
        for e in np.nonzero(np.all((p >= lo - tol) & (p <= hi + tol), axis=1))[0]:
            r = np.zeros(3)
            for _ in range(50):
                xx, J = self._map(e, r)
                dr = np.linalg.solve(J, p - xx)
                r += dr
                if np.abs(dr).sum() < 1e-13 or np.abs(r).max() > 3:
                    break
            if np.all(np.abs(r) <= 1 + 1e-6):
                return int(e), np.clip(r, -1.0, 1.0)
        raise ValueError(f"point ({x}, {y}, {z}) is outside the mesh")

    def inverse_jacobian_at(self, e, r):
        return np.linalg.inv(self._map(e, r)[1])


def load_mesh(path, **flags):
    """read_external_mesh + DatabaseMesh."""
    return DatabaseMesh(read_external_mesh(path, **flags), path)


# ---------------------------------------------------------------------------- writer
class _Writer:
    def __init__(self, path):
        self.f = open(path, "wb")

    def rec(self, a, dtype):
        a = np.ascontiguousarray(a, dtype).ravel()
        n = a.nbytes
        if n > 2**31 - 1:
            raise DatabaseError("records above 2 GiB are not written by this writer")
        m = struct.pack("<i", n)
        self.f.write(m)
        a.tofile(self.f)
        self.f.write(m)

    def ints(self, *v):
        self.rec(np.array(v), I32)

    def close(self):
        self.f.close()


def _box_top_faces(mesh):
    """Free-surface faces of a BoxMesh (top side, k = 5): ispec, ijk, jacobian2Dw, normal."""
    es = np.nonzero(mesh.ez == mesh.nez - 1)[0]
    a, b = np.tile(np.arange(5), 5), np.repeat(np.arange(5), 5)
    i, j, k = a, b, np.full(25, 4)
    w = mesh.gll.w
    Jf = mesh._jac(es)[:, i + 5 * j + 25 * k]
    n = np.cross(Jf[..., :, 0], Jf[..., :, 1])
    nn = np.linalg.norm(n, axis=-1)
    n = np.where((n[..., 2:3] >= 0), 1.0, -1.0) * n / nn[..., None]
    ijk = np.broadcast_to(np.stack([i, j, k], axis=1)[None] + 1, (len(es), 25, 3))
    return es + 1, ijk, nn * (w[i] * w[j])[None, :], n


def _box_adjacency(mesh):
    """Elements sharing at least one corner with each element (ascending), CSR with
    0-based offsets into the 1-based element list."""
    nex, ney, nez = mesh.nex, mesh.ney, mesh.nez
    adj, xadj = [], [0]
    for e in range(mesh.nspec):
        ex, ey, ez = mesh.ex[e], mesh.ey[e], mesh.ez[e]
        nb = [int(x + nex * (y + ney * z)) + 1
              for z in range(max(ez - 1, 0), min(ez + 2, nez))
              for y in range(max(ey - 1, 0), min(ey + 2, ney))
              for x in range(max(ex - 1, 0), min(ex + 2, nex))
              if (x, y, z) != (ex, ey, ez)]
        adj.extend(nb)
        xadj.append(len(adj))
    return np.array(xadj, I32), np.array(adj, I32)


def write_external_mesh(path, mesh, rho, vp, vs, dt, cij=None, stacey=True, rmass_ocean_load=None,
                        elastic_colors=None):
    """Write a BoxMesh and its model as an elastic, single-slice
    proc%06d_external_mesh.bin in the record order read_external_mesh reads.

    Materials as get_model_fwat stores them (kappa, mu, rho_vp, rho_vs in
    float32 from rho, vp, vs); rmass and, with Stacey faces, rmassx/y/z =
    rmass + dt/2 C summed in float64 and stored in float32; free surface on
    top; every element inner; no PML, coupling or MPI interfaces.  Optional
    records: rmass_ocean_load (nglob, APPROXIMATE_OCEAN_LOAD) and
    elastic_colors = (num_colors_outer, num_colors_inner, num_elem_colors)
    (USE_MESH_COLORING_GPU).
    """
    from .solver import mass_sums  # the solver imports this module's users, not the reverse

    nspec, nglob = mesh.nspec, mesh.nglob
    loc = lambda a: np.ascontiguousarray(a, F32).reshape(nspec, NG3)
    rho = loc(rho)
    kappa, mu, rho_vp, rho_vs = moduli(rho, loc(vp), loc(vs))
    rm, rmx, rmy, rmz = mass_sums(mesh, rho, rho_vp, rho_vs, dt, stacey)
    w = _Writer(path)
    try:
        w.ints(nspec); w.ints(nglob); w.ints(mesh.nspec_irregular)
        w.rec(mesh.ibool, I32)
        for a in (mesh.xstore, mesh.ystore, mesh.zstore):
            w.rec(a, F32)
        w.rec(mesh.irregular_element_number, I32)
        w.rec([mesh.xix_regular], F32)
        w.rec([mesh.jacobian_regular], F32)
        for nm in ("xix", "xiy", "xiz", "etax", "etay", "etaz", "gammax", "gammay", "gammaz", "jacobian"):
            a = getattr(mesh, nm)
            w.rec(a[: mesh.nspec_irregular] if mesh.nspec_irregular else a.ravel()[:1], F32)
        for a in (kappa, mu, rho):
            w.rec(a, F32)
        w.rec(np.zeros(nspec), I32)      # ispec_is_acoustic
        w.rec(np.ones(nspec), I32)       # ispec_is_elastic
        w.rec(np.zeros(nspec), I32)      # ispec_is_poroelastic
        w.ints(9999)
        w.rec(rm, F32)
        if rmass_ocean_load is not None:
            w.rec(rmass_ocean_load, F32)
        w.rec(rho_vp, F32)
        w.rec(rho_vs, F32)
        nf = mesh.num_abs_boundary_faces
        w.ints(nf)
        if nf:
            w.rec(mesh.abs_boundary_ispec, I32)
            w.rec(mesh.abs_boundary_ijk, I32)
            w.rec(mesh.abs_boundary_jacobian2Dw, F32)
            w.rec(mesh.abs_boundary_normal, F32)
            if stacey:
                for a in (rmx, rmy, rmz):
                    w.rec(a, F32)
        w.ints(9998)
        sides = [mesh.ex == 0, mesh.ex == mesh.nex - 1, mesh.ey == 0, mesh.ey == mesh.ney - 1,
                 mesh.ez == 0, mesh.ez == mesh.nez - 1]
        ibelm = [np.nonzero(s)[0] + 1 for s in sides]
        for b in ibelm:
            w.ints(len(b))
        for b in ibelm:
            if len(b):
                w.rec(b, I32)
        fs = _box_top_faces(mesh)
        w.ints(len(fs[0]))
        if len(fs[0]):
            w.rec(fs[0], I32); w.rec(fs[1], I32); w.rec(fs[2], F32); w.rec(fs[3], F32)
        w.ints(0); w.ints(0); w.ints(0)  # no acoustic-elastic, acoustic-poro, elastic-poro coupling
        w.ints(9997)
        w.ints(0); w.ints(0)             # one slice: no MPI interface
        w.ints(9996)
        if cij is not None:
            c = np.asarray(cij, F32).reshape(nspec, NG3, 21)
            for q in range(21):
                w.rec(c[..., q], F32)
        w.rec(np.ones(nspec), I32)       # ispec_is_inner
        w.ints(nspec, 0)                 # nspec_inner_elastic, nspec_outer_elastic
        w.ints(nspec)                    # num_phase_ispec_elastic
        ph = np.zeros((2, nspec), I32)
        ph[1] = np.arange(1, nspec + 1)
        w.rec(ph, I32)
        if elastic_colors is not None:
            no, ni, ne = elastic_colors
            w.ints(no, ni)
            w.rec(ne, I32)
        surf = np.zeros(nspec, bool)
        for s in sides:
            surf |= s
        nfaces = 2 * (mesh.nex * mesh.ney + mesh.nex * mesh.nez + mesh.ney * mesh.nez)
        onb = np.zeros(nglob, bool)
        ii = np.arange(NG3)
        i, j, k = ii % 5, (ii // 5) % 5, ii // 25
        for s, m in zip(sides, (i == 0, i == 4, j == 0, j == 4, k == 0, k == 4)):
            onb[mesh.ibool0[np.nonzero(s)[0]][:, m].ravel()] = True
        w.ints(nfaces)
        w.rec(surf, I32)
        w.rec(onb, I32)
        xadj, adj = _box_adjacency(mesh)
        w.ints(len(adj))
        w.rec(xadj, I32)
        w.rec(adj, I32)
        w.ints(9995)
    finally:
        w.close()


# ---------------------------------------------------------------------------- partitioned meshes
def merge_slices(dbs):
    """One mesh from the proc%06d_external_mesh.bin records of every slice of a
    partitioned mesh (read_external_mesh of each rank, in rank order), for a
    single GPU running the whole mesh.

    Shared points are matched through the MPI interface lists: the k-th entry of
    rank r's list towards rank q and the k-th entry of rank q's list towards r
    are the same point (the pairing the reference's assembly of
    ibool_interfaces_ext_mesh relies on, read_mesh_databases.f90:1003-1041).
    Merged points are numbered by first appearance in rank order; elements,
    faces and per-element arrays are concatenated in rank order; the masses
    rmass and rmassx/y/z, which each slice stores unassembled, are summed over
    the slices that share a point (in rank order, float64, stored as float32).
    """
    nrank = len(dbs)
    for r, d in enumerate(dbs):
        if d["ACOUSTIC_SIMULATION"] or d["POROELASTIC_SIMULATION"] or not d["ELASTIC_SIMULATION"]:
            raise DatabaseError(f"slice {r}: only purely elastic slices are merged")
        if d["NSPEC_CPML"]:
            raise DatabaseError(f"slice {r}: C-PML slices are not merged")
    reg = [d for d in dbs if d["NSPEC_IRREGULAR"] < d["NSPEC_AB"]]  # slices holding regular elements
    if any(float(d["xix_regular"]) != float(reg[0]["xix_regular"]) for d in reg):
        raise DatabaseError("slices hold regular elements of different sizes")
    # union-find over (rank, 0-based point)
    base = np.cumsum([0] + [d["NGLOB_AB"] for d in dbs])
    parent = np.arange(base[-1])

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for r, d in enumerate(dbs):
        for i in range(d["num_interfaces_ext_mesh"]):
            q = int(d["my_neighbors_ext_mesh"][i])
            if not 0 <= q < nrank:
                raise DatabaseError(f"slice {r}: neighbour {q} outside 0..{nrank - 1}")
            dq = dbs[q]
            back = [j for j in range(dq["num_interfaces_ext_mesh"]) if int(dq["my_neighbors_ext_mesh"][j]) == r]
            if len(back) != 1:
                raise DatabaseError(f"slices {r} and {q}: interface lists do not pair up")
            j = back[0]
            n, nq = int(d["nibool_interfaces_ext_mesh"][i]), int(dq["nibool_interfaces_ext_mesh"][j])
            if n != nq:
                raise DatabaseError(f"slices {r} and {q}: interface of {n} and {nq} points")
            a = d["ibool_interfaces_ext_mesh"][i, :n].astype(np.int64) - 1 + base[r]
            b = dq["ibool_interfaces_ext_mesh"][j, :n].astype(np.int64) - 1 + base[q]
            for u, v in zip(a, b):
                ru, rv = find(u), find(v)
                if ru != rv:
                    parent[max(ru, rv)] = min(ru, rv)
    roots = np.array([find(a) for a in range(base[-1])])
    uniq, first = np.unique(roots, return_index=True)
    order = np.argsort(first)  # first appearance in rank order
    newid = np.empty(base[-1], np.int64)
    rank_of_root = np.empty(len(uniq), np.int64)
    rank_of_root[order] = np.arange(len(uniq))
    newid[:] = rank_of_root[np.searchsorted(uniq, roots)]
    nglob = len(uniq)
    out = {"NGLOB_AB": nglob, "ACOUSTIC_SIMULATION": False, "ELASTIC_SIMULATION": True,
           "POROELASTIC_SIMULATION": False, "NSPEC_CPML": 0, "num_interfaces_ext_mesh": 0,
           "xix_regular": (reg or dbs)[0]["xix_regular"], "jacobian_regular": (reg or dbs)[0]["jacobian_regular"]}
    out["NSPEC_AB"] = nspec = int(sum(d["NSPEC_AB"] for d in dbs))
    out["NSPEC_IRREGULAR"] = int(sum(d["NSPEC_IRREGULAR"] for d in dbs))
    out["ibool"] = np.concatenate([newid[d["ibool"].astype(np.int64) - 1 + base[r]] + 1
                                   for r, d in enumerate(dbs)]).astype(I32)
    for c in "xyz":
        a = np.zeros(nglob, F32)
        for r, d in enumerate(dbs):
            a[newid[base[r]:base[r + 1]]] = d[c + "store"]
        out[c + "store"] = a
    irr, off = [], 0
    for d in dbs:
        e = d["irregular_element_number"].astype(np.int64)
        irr.append(np.where(e > 0, e + off, 0))
        off += d["NSPEC_IRREGULAR"]
    out["irregular_element_number"] = np.concatenate(irr).astype(I32)
    for nm in _JNAMES + ["jacobian"]:
        parts = [d[nm + "store"].reshape(-1, NG3) for d in dbs if d["NSPEC_IRREGULAR"]]
        out[nm + "store"] = np.concatenate(parts) if parts else np.zeros(1, F32)
    for nm in ("kappastore", "mustore", "rhostore", "rho_vp", "rho_vs"):
        out[nm] = np.concatenate([d[nm] for d in dbs])
    if all(CIJ_NAMES[0] + "store" in d for d in dbs):
        for nm in CIJ_NAMES:
            out[nm + "store"] = np.concatenate([d[nm + "store"] for d in dbs])
    for nm in ("rmass", "rmassx", "rmassy", "rmassz"):
        if all(nm in d for d in dbs):
            acc = np.zeros(nglob)
            for r, d in enumerate(dbs):
                np.add.at(acc, newid[base[r]:base[r + 1]], d[nm].astype(np.float64))
            out[nm] = acc.astype(F32)
    eoff = np.cumsum([0] + [d["NSPEC_AB"] for d in dbs])
    for pre, key in (("abs_boundary", "num_abs_boundary_faces"), ("free_surface", "num_free_surface_faces")):
        have = [r for r, d in enumerate(dbs) if d[key] > 0]
        out[key] = int(sum(dbs[r][key] for r in have))
        if have:
            out[pre + "_ispec"] = np.concatenate([dbs[r][pre + "_ispec"] + eoff[r] for r in have]).astype(I32)
            for suf in ("_ijk", "_jacobian2Dw", "_normal"):
                out[pre + suf] = np.concatenate([dbs[r][pre + suf] for r in have])
    out["nspec_inner_elastic"], out["nspec_outer_elastic"] = nspec, 0
    out["num_phase_ispec_elastic"] = nspec
    ph = np.zeros((2, nspec), I32)
    ph[1] = np.arange(1, nspec + 1)
    out["phase_ispec_inner_elastic"] = ph
    return out


def load_partitioned_mesh(paths, **flags):
    """DatabaseMesh of the whole mesh from every slice's file (rank order)."""
    return DatabaseMesh(merge_slices([read_external_mesh(p, **flags) for p in paths]))
