"""Kernel post-processing on the GPU: SpecFWAT's xfwat_post_proc chain.

The reference (src/postproc/fwat_post_proc.f90:22-47) runs, per data type:
sum_kernel -> apply_precond or sum_precond -> pde_smooth -> taper_kernel_grid
-> write_gradient_grid.  Here each stage is a device pass behind the C ABI:

    PostFlow.add_event / add_event_files  event sums, post_processing.f90:108-149
                                          (HTI: kernel_cijkl2hti, kernel_io.f90:100-156)
    PostFlow.apply_precond                post_processing.f90:163-193, invert_hess :344-377
    PostFlow.sum_precond                  post_processing.f90:195-231
    PostFlow.pde_smooth                   post_processing.f90:256-274
    PostFlow.taper_kernel_grid            post_processing.f90:296-333
    GridProjection                        gll2grid, src/grid/model_grid_data.f90:192-211 ->
                                          projection_on_FD_grid_mod_fwat.f90:186-596
    write_kernel / read_kernel            kernel_io.f90:158-205 (Fortran sequential files)

Kernel sums are float64 on the device (the reference sums single-precision
files in single precision).  Event kernels enter either straight from a
finished adjoint run (ElasticRun) with no host round trip, or as the
single-precision arrays of EKERNEL files.
"""
import os

import numpy as np

from . import _lib
from .smooth import smooth_batch

KERNEL_ISO = ("alpha", "beta", "rhop")  # config.f90:35
KERNEL_AZI_ANI = ("alpha", "beta", "rhop", "gcp", "gsp")  # config.f90:36
HTI_INPUTS = ("rho", "c11", "c12", "c13", "c22", "c23", "c33", "c44", "c45", "c55", "c66")  # kernel_io.f90:121-131
THRESHOLD_HESS = 1.0e-3  # config.f90:42
RHO_SCALING_FAC = 0.33  # config.f90:42
DEFAULT_PRECOND = 1  # config.f90:46
NUM_ITER = 5  # Newton steps of locate_point_in_element (SPECFEM constant NUM_ITER)

F32 = np.float32


# ---------------------------------------------------------------- kernel files

def kernel_file(path, name, rank=0):
    """create_name_database + '<name>.bin': <path>/proc%06d_<name>.bin."""
    return os.path.join(path, "proc%06d_%s.bin" % (rank, name))


def write_kernel(path, name, data, rank=0):
    """write_kernel (kernel_io.f90:189-205): one Fortran unformatted sequential
    record holding the (NGLLX,NGLLY,NGLLZ,NSPEC) single-precision array."""
    a = np.ascontiguousarray(data, F32).ravel()
    n = a.nbytes
    if n > 2**31 - 9:
        raise ValueError("kernel record larger than one Fortran sequential record (2 GiB)")
    os.makedirs(path, exist_ok=True)
    marker = np.array([n], "<i4").tobytes()
    with open(kernel_file(path, name, rank), "wb") as f:
        f.write(marker)
        f.write(a.astype("<f4").tobytes())
        f.write(marker)


def read_kernel(path, name, nspec, rank=0):
    """read_kernel (kernel_io.f90:158-187): the record as float32 [nspec, 125]."""
    with open(kernel_file(path, name, rank), "rb") as f:
        raw = f.read()
    if len(raw) < 8:
        raise ValueError(f"{name}: truncated kernel file")
    n = int(np.frombuffer(raw[:4], "<i4")[0])
    if n != nspec * 125 * 4 or len(raw) != n + 8 or int(np.frombuffer(raw[-4:], "<i4")[0]) != n:
        raise ValueError(f"{name}: record of {n} bytes does not hold a (5,5,5,{nspec}) single-precision array")
    return np.frombuffer(raw[4:4 + n], "<f4").astype(F32).reshape(nspec, 125)


# ---------------------------------------------------------------- grid

class RegularGrid:
    """ext_grid of create_grid (model_grid_data.f90:18-53): origin, spacing, size."""

    def __init__(self, origin, spacing, size):
        self.ox, self.oy, self.oz = (float(F32(v)) for v in origin)
        self.hx, self.hy, self.hz = (float(F32(v)) for v in spacing)
        self.nx, self.ny, self.nz = (int(v) for v in size)
        if min(self.nx, self.ny, self.nz) <= 0 or min(self.hx, self.hy, self.hz) <= 0:
            raise ValueError("grid sizes and spacings must be positive")

    def axes(self):
        """x, y, z coordinates (single precision, ox + (i-1) hx)."""
        f = lambda o, h, n: (F32(o) + np.arange(n, dtype=F32) * F32(h)).astype(F32)
        return f(self.ox, self.hx, self.nx), f(self.oy, self.hy, self.ny), f(self.oz, self.hz, self.nz)

    @property
    def npts(self):
        return self.nx * self.ny * self.nz


class GridProjection:
    """gll2grid: compute_interpolation_coeff_FD_SEM once on the device, then
    Project_model_SEM2FD_grid for any number of fields.  Grid arrays are
    returned as float32 [nfield, nz, ny, nx] (Fortran (nx, ny, nz) order)."""

    def __init__(self, mesh, grid, num_iter=NUM_ITER):
        self.mesh, self.grid = mesh, grid
        self.h = _lib.HandleBox()
        out = _lib.call("fwat_gll2grid_prepare_", self.h, mesh.nspec, mesh.nglob,
                        np.ascontiguousarray(mesh.ibool, np.int32), np.ascontiguousarray(mesh.xstore, F32),
                        np.ascontiguousarray(mesh.ystore, F32), np.ascontiguousarray(mesh.zstore, F32),
                        grid.ox, grid.oy, grid.oz, grid.hx, grid.hy, grid.hz, grid.nx, grid.ny, grid.nz,
                        int(num_iter), 0)
        self.nb_fd_point = int(out[-1])

    def __call__(self, gll_fields):
        f = np.ascontiguousarray(gll_fields, F32).reshape(-1, self.mesh.nspec * 125)
        out = np.zeros((f.shape[0], self.grid.nz, self.grid.ny, self.grid.nx), F32)
        _lib.call("fwat_gll2grid_", self.h, f.shape[0], f, out)
        return out

    def owners(self):
        """(owner [npts] 1-based element or 0, ref [npts, 3])."""
        owner = np.zeros(self.grid.npts, np.int32)
        ref = np.zeros((self.grid.npts, 3))
        _lib.call("fwat_gll2grid_owner_", self.h, owner, ref)
        return owner, ref

    def last_time_ms(self):
        return float(_lib.call("fwat_gll2grid_last_time_", self.h, 0.0)[0])

    def close(self):
        if self.h:
            _lib.call("fwat_gll2grid_destroy_", self.h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def taper_grid(mesh_box, distance_min_glob, distance_max_glob, taper, grid, fields=None, with_mask=False):
    """taper_kernel_grid (post_processing.f90:296-333): multiply grid fields
    [nfield, nz, ny, nx] by the taper_cls mask built over the mesh box
    (xmin, xmax, ymin, ymax, zmin, zmax) with spacing (dmax + dmin) / 2;
    taper = (H_SUPPRESS, H_BUFFER, V_SUPPRESS, V_BUFFER).  Returns the tapered
    fields (and the mask when asked)."""
    if fields is None:
        data = np.zeros((0, grid.nz, grid.ny, grid.nx), F32)
    else:
        data = np.array(fields, F32).reshape(-1, grid.nz, grid.ny, grid.nx)
    mask = np.zeros((grid.nz, grid.ny, grid.nx), F32) if with_mask else None
    _lib.call("fwat_taper_grid_", *[float(v) for v in mesh_box], float(distance_min_glob),
              float(distance_max_glob), *[float(v) for v in taper], grid.ox, grid.oy, grid.oz, grid.hx, grid.hy,
              grid.hz, grid.nx, grid.ny, grid.nz, data.shape[0], data if data.size else np.zeros(1, F32), mask)
    return (data, mask) if with_mask else data


# ---------------------------------------------------------------- PostFlow

class PostFlow:
    """One data type's post-processing (PostFlow, post_processing.f90:20-32).

    model_type 1: isotropic kernels alpha, beta, rhop; 2: azimuthal (HTI)
    kernels alpha, beta, rhop, gcp, gsp from the c_ij kernels, which needs
    hti_model = dict(c44, c45, c55, rho_vp, rho_vs, rho) (float32 [nspec,125]).
    precond_type 0 / 1 sums the events' Hessians (1: absolute values).
    """

    def __init__(self, mesh, model_type=1, *, precond_type=DEFAULT_PRECOND, use_rho_scaling=False,
                 hti_model=None):
        if model_type not in (1, 2):
            raise ValueError("model_type must be 1 (isotropic) or 2 (azimuthal anisotropy)")
        self.mesh = mesh
        self.model_type = model_type
        self.precond_type = int(precond_type)
        self.use_rho_scaling = bool(use_rho_scaling)
        self.names = KERNEL_ISO if model_type == 1 else KERNEL_AZI_ANI
        self.h = _lib.HandleBox()
        _lib.call("fwat_post_create_", self.h, mesh.nspec, model_type, int(self.use_rho_scaling))
        if model_type == 2:
            if hti_model is None:
                raise ValueError("HTI post-processing needs the model (c44, c45, c55, rho_vp, rho_vs, rho)")
            m = [np.ascontiguousarray(hti_model[k], F32).reshape(mesh.nspec * 125)
                 for k in ("c44", "c45", "c55", "rho_vp", "rho_vs", "rho")]
            _lib.call("fwat_post_set_model_", self.h, *m)
        self.inverted = False
        self.last_smooth = (0, 0.0, 0.0)
        self._extent = None

    @property
    def nevents(self):
        return _lib.call("fwat_post_info_", self.h, 0, 0, 0)[1]

    def last_time_ms(self):
        """Device time of the last add / invert / apply call."""
        return float(_lib.call("fwat_post_last_time_", self.h, 0.0)[0])

    def add_event(self, run):
        """Add a finished adjoint run's kernels (and Hessian) straight from HBM."""
        _lib.call("fwat_post_add_event_", self.h, run.h, self.precond_type)

    def add_event_files(self, kernels, hess=None):
        """Add one event's single-precision kernels as read from its EKERNEL
        files: isotropic {alpha, beta, rhop}, HTI {rho, c11, ..., c66}."""
        names = KERNEL_ISO if self.model_type == 1 else HTI_INPUTS
        n = self.mesh.nspec * 125
        arr = np.zeros((len(names), n), F32)
        for q, nm in enumerate(names):
            if nm in kernels:
                arr[q] = np.asarray(kernels[nm], F32).reshape(n)
            elif not (nm == "rhop" and self.use_rho_scaling):
                raise KeyError(f"event kernel '{nm}' missing")
        h = None if hess is None else np.ascontiguousarray(hess, F32).reshape(n)
        if h is None and self.precond_type in (0, 1):
            raise ValueError("PRECOND_TYPE 0/1 needs each event's hess kernel")
        _lib.call("fwat_post_add_event_files_", self.h, arr, h, self.precond_type)

    def sum_kernel(self):
        """The summed kernels, float64 {name: [nspec, 125]}."""
        out = np.zeros((len(self.names), self.mesh.nspec, 125))
        _lib.call("fwat_post_get_kernels_", self.h, out)
        return dict(zip(self.names, out))

    def hessian(self):
        """The summed Hessian (after invert_hess: the inverse), float64 [nspec, 125]."""
        out = np.zeros((self.mesh.nspec, 125))
        _lib.call("fwat_post_get_hess_", self.h, out)
        return out

    def invert_hess(self):
        if self.precond_type not in (0, 1):
            raise ValueError("invert_hess applies to PRECOND_TYPE 0 and 1")
        if self.inverted:
            raise RuntimeError("the Hessian sum is already inverted")
        _lib.call("fwat_post_invert_hess_", self.h, self.precond_type)
        self.inverted = True
        return self.hessian()

    def apply_precond(self):
        """apply_precond: kernels *= invert_hess(sum of Hessians)."""
        if not self.inverted:
            self.invert_hess()
        _lib.call("fwat_post_apply_precond_", self.h)
        return self.sum_kernel()

    def sum_precond(self, sigma_h, sigma_v, projection, **smooth_kw):
        """sum_precond: the inverse Hessian, smoothed and projected onto the
        grid, normalised by its maximum |value|."""
        inv = self.invert_hess() if not self.inverted else self.hessian()
        smooth_kw.setdefault("distance_min_glob", self.mesh_extent()[1])
        dglob, _, _, _ = smooth_batch(self.mesh, inv[None].astype(F32), sigma_h, sigma_v, **smooth_kw)
        gm = projection(dglob[:, self.mesh.ibool0])[0]
        m = np.abs(gm).max()
        return gm / m if m > 0 else gm

    def pde_smooth(self, sigma_h, sigma_v, projection, kernels=None, **smooth_kw):
        """pde_smooth: smooth_sem_pde of every kernel (one batched device run),
        then gll2grid; with USE_RHO_SCALING rhop is 0.33 x the smoothed beta."""
        k = self.sum_kernel() if kernels is None else kernels
        names = [n for n in self.names if not (self.use_rho_scaling and n == "rhop")]
        fields = np.stack([np.asarray(k[n], F32) for n in names])
        smooth_kw.setdefault("distance_min_glob", self.mesh_extent()[1])
        dglob, nt, elem_ms, upd_ms = smooth_batch(self.mesh, fields, sigma_h, sigma_v, **smooth_kw)
        self.last_smooth = (nt, elem_ms, upd_ms)
        grid = projection(dglob[:, self.mesh.ibool0])
        out = {n: grid[q] for q, n in enumerate(names)}
        if self.use_rho_scaling:
            out["rhop"] = (out["beta"] * F32(RHO_SCALING_FAC)).astype(F32)
        return out

    def mesh_extent(self):
        """(box, distance_min_glob, distance_max_glob) of check_mesh_distances, computed once
        (read_mesh_databases_for_init, kernel_io.f90:15-74)."""
        if self._extent is None:
            m = self.mesh
            box = (m.xstore.min(), m.xstore.max(), m.ystore.min(), m.ystore.max(), m.zstore.min(), m.zstore.max())
            self._extent = (box, m.distance_min_glob(), m.distance_max_glob())
        return self._extent

    def taper_kernel_grid(self, grids, grid, taper):
        """taper = (TAPER_H_SUPPRESS, TAPER_H_BUFFER, TAPER_V_SUPPRESS, TAPER_V_BUFFER)."""
        box, dmin, dmax = self.mesh_extent()
        names = list(grids)
        data = taper_grid(box, dmin, dmax, taper, grid, np.stack([grids[n] for n in names]))
        return dict(zip(names, data))

    def write_gradient_grid(self, grids, grid, filename):
        """write_gradient_grid: '<name>_kernel_smooth' datasets plus x, y, z.
        The reference writes HDF5 (hdf5_interface); no HDF5 library is part
        of this package, so the same datasets go to a numpy .npz archive."""
        x, y, z = grid.axes()
        np.savez(filename, x=x, y=y, z=z, **{f"{n}_kernel_smooth": np.asarray(v, F32) for n, v in grids.items()})

    def close(self):
        if self.h:
            _lib.call("fwat_post_destroy_", self.h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
