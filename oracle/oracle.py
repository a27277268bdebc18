"""ctypes front end of the CPU oracle (liboracle_sem.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline.
The product package never imports this module.

Function-to-reference map (file:line in the SpecFWAT tree):
  Oracle.forward / Oracle.adjoint  SPECFEM3D elastic time loop driven by
                                   src/preproc/prepare_fwd.f90:328-368 (solver
                                   source absent: restated, pinned by invariants)
  smooth_pde                       src/postproc/smooth.f90:7-437, 603-689
  interp3                          src/shared/utils.f90:1162-1232
  cijkl2hti                        src/postproc/kernel_io.f90:100-156
  invert_hess                      src/postproc/post_processing.f90:344-377
  gll2grid_coeff / gll2grid_project
                                   src/shared/projection_on_FD_grid_mod_fwat.f90:186-596, 935-1125
  taper_grid                       src/postproc/post_processing.f90:296-333,
                                   src/postproc/taper_3D_mod.f90:16-141
  bandpass_design / _apply,        src/shared/signal.f90:7-36 (SAC xapiir, Butterworth bandpass),
  filter_adj                       src/preproc/adjoint_sources/adjoint_source.f90:92-114
  cc_adjoint                       src/preproc/adjoint_sources/cc_tt_misfit.f90:17-146, utils/cc.f90
  sd / lbfgs / cg_direction        src/optimization/optimize_grid.f90:229-389,
                                   src/shared/utils.f90:1527-1534
  model_update                     src/optimization/optimize_grid.f90:176-203, 391-404
  hti_cij                          src/patch/aniso.f90:55-162
  simpson                          src/shared/utils.f90:1630-1663
  window_taper                     src/shared/signal.f90:419-491
  waveform_adjoint                 src/preproc/adjoint_sources/waveform_misfit.f90:15-64
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_sem.so")
_lib = None

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


class OrcParams(C.Structure):
    _fields_ = [
        ("nex", C.c_int), ("ney", C.c_int), ("nez", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("perturb", C.c_double),
        ("force_irregular", C.c_int),
        ("numbering", C.c_int),
        ("aniso", C.c_int),
        ("rho", _fp), ("vp", _fp), ("vs", _fp), ("cij", _fp),
        ("dt", C.c_double),
        ("nstep", C.c_int),
        ("stacey", C.c_int),
        ("nsrc", C.c_int),
        ("src_xyz", _dp), ("src_mt", _dp), ("src_f0", _dp), ("src_tshift", _dp),
        ("nrec", C.c_int),
        ("rec_xyz", _dp),
        ("nthreads", C.c_int),
        ("elem_order", _ip),
    ]


def build(force=False):
    """Compile the oracle with g++ (OpenMP, portable x86-64-v3 code)."""
    srcs = [os.path.join(_HERE, n) for n in ("sem_oracle.cpp", "post_oracle.cpp", "opt_oracle.cpp",
                                              "measure_oracle.cpp")]
    if force or not os.path.exists(_LIB_PATH) or any(os.path.getmtime(_LIB_PATH) < os.path.getmtime(s)
                                                      for s in srcs):
        import subprocess
        cmd = ["g++", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-std=c++17",
               "-o", _LIB_PATH] + srcs
        subprocess.check_call(cmd)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcParams), C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_plant.argtypes = [C.c_int]
        L.orc_assembly_order.argtypes = [C.c_void_p, _ip, _ip]
        for n in ("orc_nglob", "orc_nfaces", "orc_nspec_irregular", "orc_ncolors"):
            getattr(L, n).restype = C.c_int
            getattr(L, n).argtypes = [C.c_void_p]
        L.orc_get_mesh.argtypes = [C.c_void_p, _ip, _dp, _ip, _dp, _dp, _dp, _ip]
        L.orc_get_faces.argtypes = [C.c_void_p, _ip, _ip, _dp, _dp]
        L.orc_get_gll.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_lagrange.argtypes = [C.c_double, _dp, _dp]
        L.orc_get_sources.argtypes = [C.c_void_p, _ip, _dp, _dp, _ip, _dp, _dp]
        L.orc_get_rmass.argtypes = [C.c_void_p, _fp, _fp, _fp]
        L.orc_forward.argtypes = [C.c_void_p, _fp, C.c_int]
        L.orc_get_final.argtypes = [C.c_void_p, _fp, _fp, _fp]
        L.orc_set_final.argtypes = [C.c_void_p, _fp, _fp, _fp]
        L.orc_get_stacey.argtypes = [C.c_void_p, _fp]
        L.orc_set_fk.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _fp, _fp]
        L.orc_adjoint.argtypes = [C.c_void_p, _fp, _dp, _dp, _dp, _dp, _fp]
        L.orc_energy_series.argtypes = [C.c_void_p, _fp, _dp]
        L.orc_apply_k.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_ricker.restype = C.c_double
        L.orc_ricker.argtypes = [C.c_double, C.c_double]
        L.orc_locate_bissection.restype = C.c_int
        L.orc_locate_bissection.argtypes = [_dp, C.c_int, C.c_double]
        L.orc_interp3.restype = C.c_float
        L.orc_interp3.argtypes = [_fp, C.c_int, _fp, C.c_int, _fp, C.c_int, _fp,
                                  C.c_float, C.c_float, C.c_float]
        L.orc_hti_cij.argtypes = [C.c_float] * 5 + [_fp]
        L.orc_simpson.restype = C.c_double
        L.orc_simpson.argtypes = [_dp, C.c_int, C.c_double]
        L.orc_window_taper.argtypes = [_dp, C.c_int, C.c_double, C.c_int]
        L.orc_waveform_adjoint.restype = C.c_double
        L.orc_waveform_adjoint.argtypes = [_dp, _dp, C.c_int, C.c_double, _dp, C.c_int,
                                           C.c_double, C.c_int, _dp, _dp, _dp]
        L.orc_smooth_pde.restype = C.c_int
        L.orc_smooth_pde.argtypes = ([C.c_int, C.c_int, _ip, _fp, _fp, _fp, C.c_int]
                                     + [_fp] * 10 + [_ip, C.c_float, C.c_float, C.c_int, C.c_int,
                                                     C.c_int, _ip, _ip, _fp, C.c_float, C.c_float,
                                                     C.c_float, C.c_int, C.c_int, C.c_int, _fp, _fp])
        L.orc_cijkl2hti.argtypes = [C.c_long] + [_fp] * 8
        L.orc_invert_hess.argtypes = [C.c_long, _fp]
        L.orc_gll2grid_coeff.restype = C.c_long
        L.orc_gll2grid_coeff.argtypes = ([C.c_int, _ip, _fp, _fp, _fp] + [C.c_float] * 6 + [C.c_int] * 4
                                         + [_ip, _dp])
        L.orc_gll2grid_project.argtypes = [C.c_long, _ip, _dp, _fp, _fp]
        L.orc_taper_grid.restype = C.c_int
        L.orc_taper_grid.argtypes = [C.c_float] * 11 + [_fp, C.c_int, _fp, C.c_int, _fp, C.c_int, _fp]
        L.orc_bandpass_design.restype = C.c_int
        L.orc_bandpass_design.argtypes = [C.c_double, C.c_float, C.c_float, _fp, _fp]
        L.orc_bandpass_design_d.restype = C.c_int
        L.orc_bandpass_design_d.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.orc_bandpass_apply.argtypes = [_fp, C.c_long, _fp, _fp, C.c_int, C.c_int]
        L.orc_filter_adj.argtypes = [_dp, C.c_int, C.c_double, C.c_float, C.c_float, _dp, C.c_int, C.c_double,
                                     C.c_int]
        L.orc_cc_adjoint.restype = C.c_double
        L.orc_cc_adjoint.argtypes = ([_dp, _dp, C.c_int, C.c_double, _dp, C.c_int, C.c_double, C.c_int, C.c_int,
                                      _dp] + [_dp] * 7 + [_ip])
        L.orc_fft.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int]
        L.orc_hilbert.argtypes = [_dp, C.c_int, _dp, _dp]
        L.orc_conv_adjoint.argtypes = [_dp] * 4 + [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_int, _dp, _dp, _dp]
        L.orc_win_sel.restype = C.c_int
        L.orc_win_sel.argtypes = [_dp, _dp, C.c_int] + [C.c_double] * 5 + [_dp, C.c_int, _dp, _dp, _dp]
        L.orc_ep_adjoint.restype = C.c_double
        L.orc_ep_adjoint.argtypes = [_dp, _dp, C.c_int, C.c_double, _dp, C.c_int, C.c_double, C.c_int, C.c_double,
                                     _dp, _dp]
        L.orc_dpss.argtypes = [C.c_int, C.c_double, C.c_int, _dp]
        L.orc_schedule_intervals.restype = C.c_int
        L.orc_schedule_intervals.argtypes = [C.c_int, _dp, _dp, _dp, _ip]
        L.orc_sta_lta.argtypes = [_dp, C.c_int, C.c_double, C.c_double, _dp]
        L.orc_flexwin.restype = C.c_int
        L.orc_flexwin.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.c_double, _dp, C.c_int, _dp, _dp]
        L.orc_mt_adjoint.restype = C.c_double
        L.orc_mt_adjoint.argtypes = [_dp, _dp, C.c_int, C.c_double, _dp, C.c_int, C.c_double, C.c_int, C.c_int,
                                     _dp, _dp, _dp, C.c_void_p]
        L.orc_sd_direction.argtypes = [C.c_long, _fp, _fp, _fp, C.c_int]
        L.orc_lbfgs_direction.argtypes = [C.c_long, C.c_long, C.c_int, _fp, _fp, _fp, C.c_float, C.c_float,
                                          C.c_float, _fp, _dp, C.c_int]
        L.orc_cg_direction.argtypes = [C.c_long, C.c_long, _fp, _fp, _fp, _fp, _dp, C.c_int]
        L.orc_model_update.argtypes = [C.c_long, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                       C.c_float, _fp, _fp]
        _lib = L
    return _lib


def _p(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(t))


def gll():
    xi, w, hp, hpw = (np.zeros(5), np.zeros(5), np.zeros(25), np.zeros(25))
    lib().orc_get_gll(_p(xi, C.c_double), _p(w, C.c_double), _p(hp, C.c_double), _p(hpw, C.c_double))
    return xi, w, hp.reshape(5, 5), hpw.reshape(5, 5)


def lagrange(x):
    h, hd = np.zeros(5), np.zeros(5)
    lib().orc_lagrange(float(x), _p(h, C.c_double), _p(hd, C.c_double))
    return h, hd


def ricker(t, f0):
    return lib().orc_ricker(float(t), float(f0))


class Oracle:
    """One simulation set-up (mesh, model, source, receivers) in FP32 or FP64.

    rho/vp/vs are float32 arrays [nspec*125] (local points, element-major,
    i fastest); cij is float32 [nspec*125, 21] for the anisotropic case.
    """

    def __init__(self, nex, ney, nez, lx, ly, lz, rho, vp, vs, *, cij=None, dt=0.005, nstep=10,
                 stacey=True, src_xyz=(), src_mt=(), src_f0=(), src_tshift=None, rec_xyz=(),
                 perturb=0.0, force_irregular=False, numbering=1, precision=32, nthreads=0,
                 elem_order=None):
        """elem_order: 0-based visiting order of the element loop that assembles the
        internal forces (SPECFEM's CPU loop over phase_ispec_inner_elastic); None
        assembles colour by colour."""
        self._keep = []

        def f32(a):
            a = np.ascontiguousarray(a, dtype=np.float32)
            self._keep.append(a)
            return a

        def f64(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            self._keep.append(a)
            return a

        src_xyz = f64(np.reshape(src_xyz, (-1, 3)))
        nsrc = src_xyz.shape[0]
        src_mt = f64(np.reshape(src_mt, (nsrc, 6)))
        src_f0 = f64(np.reshape(src_f0, (nsrc,)))
        src_tshift = f64(np.zeros(nsrc) if src_tshift is None else np.reshape(src_tshift, (nsrc,)))
        rec_xyz = f64(np.reshape(rec_xyz, (-1, 3)))
        self.nrec = rec_xyz.shape[0]
        self.nsrc = nsrc
        rho, vp, vs = f32(rho), f32(vp), f32(vs)
        cijp = f32(cij) if cij is not None else None
        self.params = OrcParams(
            nex, ney, nez, lx, ly, lz, perturb, int(force_irregular), numbering, int(cij is not None),
            _p(rho, C.c_float), _p(vp, C.c_float), _p(vs, C.c_float), _p(cijp, C.c_float),
            dt, nstep, int(stacey), nsrc, _p(src_xyz, C.c_double), _p(src_mt, C.c_double),
            _p(src_f0, C.c_double), _p(src_tshift, C.c_double), self.nrec, _p(rec_xyz, C.c_double),
            nthreads, None)
        if elem_order is not None:
            eo = np.ascontiguousarray(elem_order, np.int32)
            self._keep.append(eo)
            self.params.elem_order = _p(eo, C.c_int)
        self.nspec = nex * ney * nez
        self.nstep = nstep
        self.aniso = cij is not None
        self.h = lib().orc_create(C.byref(self.params), precision)
        if not self.h:
            raise ValueError("elem_order is not a permutation of the elements")
        self.nglob = lib().orc_nglob(self.h)
        self.nfaces = lib().orc_nfaces(self.h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_destroy(self.h)
            self.h = None

    def mesh(self):
        nloc = self.nspec * 125
        ibool = np.zeros(nloc, np.int32)
        xyz = np.zeros((self.nglob, 3))
        irreg = np.zeros(self.nspec, np.int32)
        reg = np.zeros(2)
        jinv = np.zeros((nloc, 9))
        jac = np.zeros(nloc)
        color = np.zeros(self.nspec, np.int32)
        lib().orc_get_mesh(self.h, _p(ibool, C.c_int), _p(xyz, C.c_double), _p(irreg, C.c_int),
                           _p(reg, C.c_double), _p(jinv, C.c_double), _p(jac, C.c_double),
                           _p(color, C.c_int))
        return dict(ibool=ibool, xyz=xyz, irregular=irreg, xix_regular=reg[0],
                    jacobian_regular=reg[1], jinv=jinv, jacobian=jac, color=color)

    def assembly_order(self):
        """(off [nglob+1], idx [nspec*125]): for every point, the element terms
        (e*125 + L) in the order the oracle adds them."""
        off = np.zeros(self.nglob + 1, np.int32)
        idx = np.zeros(self.nspec * 125, np.int32)
        lib().orc_assembly_order(self.h, _p(off, C.c_int), _p(idx, C.c_int))
        return off, idx

    def faces(self):
        nf = self.nfaces
        ispec = np.zeros(nf, np.int32)
        ijk = np.zeros((nf, 25, 3), np.int32)
        normal = np.zeros((nf, 25, 3))
        jw = np.zeros((nf, 25))
        lib().orc_get_faces(self.h, _p(ispec, C.c_int), _p(ijk, C.c_int), _p(normal, C.c_double),
                            _p(jw, C.c_double))
        return dict(ispec=ispec, ijk=ijk, normal=normal, jacobian2Dw=jw)

    def sources(self):
        si = np.zeros(self.nsrc, np.int32)
        sr = np.zeros((self.nsrc, 3))
        sa = np.zeros((self.nsrc, 125, 3))
        ri = np.zeros(self.nrec, np.int32)
        rr = np.zeros((self.nrec, 3))
        rh = np.zeros((self.nrec, 125))
        lib().orc_get_sources(self.h, _p(si, C.c_int), _p(sr, C.c_double), _p(sa, C.c_double),
                              _p(ri, C.c_int), _p(rr, C.c_double), _p(rh, C.c_double))
        return dict(src_ispec=si, src_ref=sr, sourcearrays=sa, rec_ispec=ri, rec_ref=rr, rec_h=rh)

    def rmass_inv(self):
        out = [np.zeros(self.nglob, np.float32) for _ in range(3)]
        lib().orc_get_rmass(self.h, *[_p(o, C.c_float) for o in out])
        return out

    def set_fk(self, veloc, tract, resamp=1):
        """Incident velocity and traction (nsamp, nfaces*25, 3) injected through the
        Stacey faces of later forward runs, sample k at step 1 + k*resamp."""
        v = np.ascontiguousarray(veloc, np.float32)
        t = np.ascontiguousarray(tract, np.float32)
        self._keep += [v, t]
        lib().orc_set_fk(self.h, v.shape[1], v.shape[0], int(resamp), _p(v, C.c_float), _p(t, C.c_float))

    def forward(self, save=True):
        seis = np.zeros((self.nstep, self.nrec, 3), np.float32)
        lib().orc_forward(self.h, _p(seis, C.c_float), int(save))
        return seis

    def final_state(self):
        out = [np.zeros((self.nglob, 3), np.float32) for _ in range(3)]
        lib().orc_get_final(self.h, *[_p(o, C.c_float) for o in out])
        return out

    def stacey_store(self):
        buf = np.zeros((self.nstep, self.nfaces, 25, 3), np.float32)
        lib().orc_get_stacey(self.h, _p(buf, C.c_float))
        return buf

    def adjoint(self, adj):
        """adj: float32 [nstep, nrec, 3] adjoint source traces (forward time order)."""
        adj = np.ascontiguousarray(adj, dtype=np.float32)
        nloc = self.nspec * 125
        rho = np.zeros(nloc)
        mu = np.zeros(nloc)
        kap = np.zeros(nloc)
        cij = np.zeros((nloc, 21)) if self.aniso else None
        bfin = np.zeros((3, self.nglob, 3), np.float32)
        lib().orc_adjoint(self.h, _p(adj, C.c_float), _p(rho, C.c_double), _p(mu, C.c_double),
                          _p(kap, C.c_double), _p(cij, C.c_double), _p(bfin, C.c_float))
        out = dict(rho_kl=rho, b_final=bfin)
        if self.aniso:
            out["cijkl_kl"] = cij
        else:
            out["mu_kl"] = mu
            out["kappa_kl"] = kap
        return out

    def energy_series(self, u0):
        u0 = np.ascontiguousarray(u0, dtype=np.float32)
        e = np.zeros(self.nstep)
        lib().orc_energy_series(self.h, _p(u0, C.c_float), _p(e, C.c_double))
        return e

    def apply_k(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        f = np.zeros_like(u)
        lib().orc_apply_k(self.h, _p(u, C.c_double), _p(f, C.c_double))
        return f


class planted:
    """Context manager: Oracles created inside run with a planted fault, so a
    test can show that its check fails on a plausible mistake.
    1 = Stacey impedances exchanged (rho*vs normal, rho*vp tangential),
    2 = Mxz and Myz exchanged when the moment tensor is unpacked."""

    def __init__(self, code):
        self.code = int(code)

    def __enter__(self):
        lib().orc_set_plant(self.code)
        return self

    def __exit__(self, *exc):
        lib().orc_set_plant(0)
        return False


def interp3(vx, vy, vz, valin_fortran, x, y, z):
    vx = np.ascontiguousarray(vx, np.float32)
    vy = np.ascontiguousarray(vy, np.float32)
    vz = np.ascontiguousarray(vz, np.float32)
    v = np.asfortranarray(valin_fortran, np.float32).ravel(order="F")
    return lib().orc_interp3(_p(vx, C.c_float), len(vx), _p(vy, C.c_float), len(vy),
                             _p(vz, C.c_float), len(vz), _p(v, C.c_float), x, y, z)


def locate_bissection(valx, x):
    valx = np.ascontiguousarray(valx, np.float64)
    return lib().orc_locate_bissection(_p(valx, C.c_double), len(valx), float(x))


def hti_cij(vp, vs, rho, gcp, gsp):
    out = np.zeros(21, np.float32)
    lib().orc_hti_cij(vp, vs, rho, gcp, gsp, _p(out, C.c_float))
    return out


def simpson(y, dx):
    y = np.ascontiguousarray(y, np.float64)
    return lib().orc_simpson(_p(y, C.c_double), len(y), float(dx))


def window_taper(sig, pct, itype):
    sig = np.array(sig, np.float64)
    lib().orc_window_taper(_p(sig, C.c_double), len(sig), float(pct), int(itype))
    return sig


def waveform_adjoint(dat, syn, dt, windows, taper_pct=0.05, itaper=1):
    dat = np.ascontiguousarray(dat, np.float64)
    syn = np.ascontiguousarray(syn, np.float64)
    win = np.ascontiguousarray(np.reshape(windows, (-1, 2)), np.float64)
    adj = np.zeros(len(dat))
    mf = np.zeros(win.shape[0])
    rs = np.zeros(win.shape[0])
    tot = lib().orc_waveform_adjoint(_p(dat, C.c_double), _p(syn, C.c_double), len(dat), dt,
                                     _p(win, C.c_double), win.shape[0], taper_pct, itaper,
                                     _p(adj, C.c_double), _p(mf, C.c_double), _p(rs, C.c_double))
    return tot, adj, mf, rs


def smooth_pde(mesh, dat_in, sigma_h, sigma_v, distance_min_glob, *, is_sph=False, faithful=True,
               max_steps=0, is_cpml=None, phase=None):
    """Loop-for-loop restatement of smooth_sem_pde's CPU branch.

    mesh: dict with nspec, nglob, ibool (1-based, [nspec*125]), xstore/ystore/zstore,
    xix..gammaz/jacobian ([nspec_irregular*125] each), irregular_element_number,
    xix_regular, jacobian_regular.  Returns (dat [nspec*125], dat_glob, ntstep).
    """
    nspec, nglob = mesh["nspec"], mesh["nglob"]
    keep = []

    def f(a):
        a = np.ascontiguousarray(a, np.float32)
        keep.append(a)
        return _p(a, C.c_float)

    def i(a):
        a = np.ascontiguousarray(a, np.int32)
        keep.append(a)
        return _p(a, C.c_int)

    if phase is None:
        nspec_outer, nspec_inner = 0, nspec
        num_phase = nspec
        ph = np.zeros((2, nspec), np.int32)
        ph[1, :] = np.arange(1, nspec + 1)
    else:
        nspec_outer, nspec_inner, ph = phase
        num_phase = ph.shape[1]
    cpml = np.zeros(nspec, np.int32) if is_cpml is None else np.asarray(is_cpml, np.int32)
    dat = np.zeros(nspec * 125, np.float32)
    dglob = np.zeros(nglob, np.float32)
    nsi = int(mesh.get("nspec_irregular", 0))
    n = lib().orc_smooth_pde(
        nspec, nglob, i(mesh["ibool"]), f(mesh["xstore"]), f(mesh["ystore"]), f(mesh["zstore"]), nsi,
        *[f(mesh[k]) for k in ("xix", "xiy", "xiz", "etax", "etay", "etaz", "gammax", "gammay",
                               "gammaz", "jacobian")],
        i(mesh["irregular_element_number"]), float(mesh["xix_regular"]),
        float(mesh["jacobian_regular"]), nspec_outer, nspec_inner, num_phase, i(ph), i(cpml),
        f(dat_in), float(sigma_h), float(sigma_v), float(distance_min_glob), int(is_sph),
        int(faithful), int(max_steps), _p(dat, C.c_float), _p(dglob, C.c_float))
    return dat, dglob, n


def cijkl2hti(c44, c45, c55, rho_vp, rho_vs, rho, k11):
    """One event's HTI kernels [5][n] (alpha, beta, rho, gcp, gsp) from the
    rho + 10 c_ij kernels k11 [11][n] (rho, c11, c12, c13, c22, c23, c33, c44,
    c45, c55, c66); single precision like the reference."""
    f = [np.ascontiguousarray(a, np.float32).ravel() for a in (c44, c45, c55, rho_vp, rho_vs, rho)]
    n = f[0].size
    k = np.ascontiguousarray(k11, np.float32).reshape(11, n)
    out = np.zeros((5, n), np.float32)
    lib().orc_cijkl2hti(n, *[_p(a, C.c_float) for a in f], _p(k, C.c_float), _p(out, C.c_float))
    return out


def invert_hess(h):
    h = np.array(h, np.float32).ravel()
    lib().orc_invert_hess(h.size, _p(h, C.c_float))
    return h


def gll2grid_coeff(ibool, xstore, ystore, zstore, grid, num_iter=5):
    """(owner [npt] 1-based or 0, ref [npt,3], nb_fd_point); grid = (ox, oy, oz,
    hx, hy, hz, nx, ny, nz); points in Fortran (nx, ny, nz) order."""
    ox, oy, oz, hx, hy, hz, nx, ny, nz = grid
    ib = np.ascontiguousarray(ibool, np.int32).ravel()
    xs, ys, zs = (np.ascontiguousarray(a, np.float32) for a in (xstore, ystore, zstore))
    npt = nx * ny * nz
    owner = np.zeros(npt, np.int32)
    ref = np.zeros((npt, 3), np.float64)
    nf = lib().orc_gll2grid_coeff(ib.size // 125, _p(ib, C.c_int), _p(xs, C.c_float), _p(ys, C.c_float),
                                  _p(zs, C.c_float), ox, oy, oz, hx, hy, hz, nx, ny, nz, num_iter,
                                  _p(owner, C.c_int), _p(ref, C.c_double))
    return owner, ref, int(nf)


def gll2grid_project(owner, ref, model):
    model = np.ascontiguousarray(model, np.float32).ravel()
    out = np.zeros(owner.size, np.float32)
    lib().orc_gll2grid_project(owner.size, _p(owner, C.c_int), _p(ref, C.c_double), _p(model, C.c_float),
                               _p(out, C.c_float))
    return out


def taper_grid(box, dh, h_suppress, h_buffer, v_suppress, v_buffer, gx, gy, gz):
    """Taper mask at the grid points, Fortran (nx, ny, nz) order; box =
    (xmin, xmax, ymin, ymax, zmin, zmax)."""
    gx, gy, gz = (np.ascontiguousarray(a, np.float32) for a in (gx, gy, gz))
    out = np.zeros(gx.size * gy.size * gz.size, np.float32)
    rc = lib().orc_taper_grid(*[float(v) for v in box], float(dh), float(h_suppress), float(h_buffer),
                              float(v_suppress), float(v_buffer), _p(gx, C.c_float), gx.size,
                              _p(gy, C.c_float), gy.size, _p(gz, C.c_float), gz.size, _p(out, C.c_float))
    if rc != 0:
        raise ValueError("taper ranges do not fit the taper grid")
    return out


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def sd_direction(grad, hess, dbl=False):
    g, h = _f32(grad).ravel(), _f32(hess).ravel()
    out = np.zeros(g.size, np.float32)
    lib().orc_sd_direction(g.size, _p(g, C.c_float), _p(h, C.c_float), _p(out, C.c_float), int(dbl))
    return out


def lbfgs_direction(grads, models, hess, npg, dx=1.0, dy=1.0, dz=1.0, dbl=False):
    """grads, models: [nstore+1, n] (iterations iter_store .. iter_current);
    returns (direction [n], angle in degrees)."""
    g, m, h = _f32(grads), _f32(models), _f32(hess).ravel()
    n = h.size
    out = np.zeros(n, np.float32)
    ang = np.zeros(1)
    lib().orc_lbfgs_direction(n, npg, g.shape[0] - 1, _p(g, C.c_float), _p(m, C.c_float), _p(h, C.c_float),
                              dx, dy, dz, _p(out, C.c_float), _p(ang, C.c_double), int(dbl))
    return out, float(ang[0])


def cg_direction(g1, g0, hess, npg, dbl=False):
    a, b, h = _f32(g1).ravel(), _f32(g0).ravel(), _f32(hess).ravel()
    out = np.zeros(a.size, np.float32)
    alpha = np.zeros(1)
    lib().orc_cg_direction(a.size, npg, _p(a, C.c_float), _p(b, C.c_float), _p(h, C.c_float),
                           _p(out, C.c_float), _p(alpha, C.c_double), int(dbl))
    return out, float(alpha[0])


def model_update(model, direction, npg, model_type, step, vpvs_range, update_iso=True, update_ani=True):
    m = np.array(model, np.float32).ravel()
    d = _f32(direction).ravel()
    lib().orc_model_update(npg, m.size // npg, model_type, int(update_iso), int(update_ani), step,
                           vpvs_range[0], vpvs_range[1], _p(m, C.c_float), _p(d, C.c_float))
    return m


def bandpass_design(dt, f1, f2):
    """(sn [2,3], sd [2,3]) float32 second-order sections."""
    sn = np.zeros(6, np.float32)
    sd = np.zeros(6, np.float32)
    ns = lib().orc_bandpass_design(dt, f1, f2, _p(sn, C.c_float), _p(sd, C.c_float))
    return sn.reshape(-1, 3)[:ns], sd.reshape(-1, 3)[:ns]


def bandpass_design_d(dt, f1, f2):
    sn = np.zeros(6)
    sd = np.zeros(6)
    ns = lib().orc_bandpass_design_d(dt, f1, f2, _p(sn, C.c_double), _p(sd, C.c_double))
    return sn.reshape(-1, 3)[:ns], sd.reshape(-1, 3)[:ns]


def bandpass_apply(x, sn, sd, zero_phase=True):
    x = np.array(x, np.float32)
    sn = np.ascontiguousarray(sn, np.float32)
    sd = np.ascontiguousarray(sd, np.float32)
    lib().orc_bandpass_apply(_p(x, C.c_float), x.size, _p(sn, C.c_float), _p(sd, C.c_float), sn.shape[0],
                             int(zero_phase))
    return x


def filter_adj(adj, dt, f1, f2, windows, taper_pct=0.05, itaper=1):
    a = np.array(adj, np.float64)
    win = np.ascontiguousarray(np.reshape(windows, (-1, 2)), np.float64)
    lib().orc_filter_adj(_p(a, C.c_double), a.size, dt, f1, f2, _p(win, C.c_double), win.shape[0], taper_pct,
                         itaper)
    return a


CC_DEFAULTS = dict(dt_sigma_min=1.0, dlna_sigma_min=0.5, tshift_min=-5.0, tshift_max=5.0, dlna_min=-1.5,
                   dlna_max=1.5, cc_min=0.7)  # adj_config.f90:18-23


def cc_adjoint(dat, syn, dt, windows, imeas=11, taper_pct=0.05, itaper=1, **cfg):
    """One trace's CC traveltime (11) or amplitude (12) measurement: (total,
    adj, per-window dict of misfit, residual, error, cc_max, tshift, dlna, selected)."""
    c = dict(CC_DEFAULTS, **cfg)
    cv = np.array([c["dt_sigma_min"], c["dlna_sigma_min"], c["tshift_min"], c["tshift_max"], c["dlna_min"],
                   c["dlna_max"], c["cc_min"]], np.float64)
    dat = np.ascontiguousarray(dat, np.float64)
    syn = np.ascontiguousarray(syn, np.float64)
    win = np.ascontiguousarray(np.reshape(windows, (-1, 2)), np.float64)
    nw = win.shape[0]
    adj = np.zeros(len(dat))
    o = {k: np.zeros(nw) for k in ("misfit", "residual", "error", "cc_max", "tshift", "dlna")}
    sel = np.zeros(nw, np.int32)
    tot = lib().orc_cc_adjoint(_p(dat, C.c_double), _p(syn, C.c_double), len(dat), dt, _p(win, C.c_double), nw,
                               taper_pct, itaper, imeas, _p(cv, C.c_double), _p(adj, C.c_double),
                               *[_p(o[k], C.c_double) for k in ("misfit", "residual", "error", "cc_max", "tshift",
                                                                 "dlna")], _p(sel, C.c_int))
    o["selected"] = sel.astype(bool)
    return tot, adj, o


def fft(x, inverse=False, single=False):
    """fft_raw_dp / fft_raw of the reference: unnormalised radix-2 transform,
    forward exp(-i...), inverse exp(+i...); len(x) a power of two."""
    x = np.asarray(x, np.complex128)
    re = np.ascontiguousarray(x.real)
    im = np.ascontiguousarray(x.imag)
    lib().orc_fft(_p(re, C.c_double), _p(im, C.c_double), len(x), 1 if inverse else -1, int(single))
    return re + 1j * im


def dpss(n, half_nbw, kmax):
    """dpss_windows: [kmax][n] unit-norm discrete prolate spheroidal sequences."""
    out = np.zeros((kmax, n))
    lib().orc_dpss(n, half_nbw, kmax, _p(out, C.c_double))
    return out


MT_DEFAULTS = dict(transfunc_waterlevel=1e-10, water_threshold=0.02, mt_nw=4.0, num_taper=5, dt_fac=2.0,
                   err_fac=2.5, dt_max_scale=3.5, min_cycle_in_window=3, use_cc_error=True,
                   use_mt_error=False)  # adj_config.f90:18-26
MT_FIELDS = ("misfit", "residual", "error", "cc_max", "tshift", "dlna", "selected", "ishift", "imeas", "misfit_other")


def mt_config(min_period, max_period, **cfg):
    c = dict(CC_DEFAULTS, **MT_DEFAULTS)
    c.update(cfg)
    return np.array([c["dt_sigma_min"], c["dlna_sigma_min"], c["tshift_min"], c["tshift_max"], c["dlna_min"],
                     c["dlna_max"], c["cc_min"], c["transfunc_waterlevel"], c["water_threshold"], c["mt_nw"],
                     c["num_taper"], c["dt_fac"], c["err_fac"], c["dt_max_scale"], c["min_cycle_in_window"],
                     1.0 if c["use_cc_error"] else 0.0, min_period, max_period,
                     1.0 if c["use_mt_error"] else 0.0], np.float64)


def mt_adjoint(dat, syn, dt, windows, min_period, max_period, imeas=13, taper_pct=0.05, itaper=1, errors=False,
               **cfg):
    """One trace's multitaper traveltime (13) or amplitude (14) measurement with
    the reference's cross-correlation fallback: (total, adj, per-window dict of
    MT_FIELDS; "imeas" is the measurement used, 11/12 after a fallback).  With
    use_mt_error=True the jackknife errors weight the misfit; errors=True also
    returns them, [nwin, 2, nf] (err_dtau, err_dlna per frequency)."""
    cv = mt_config(min_period, max_period, **cfg)
    dat = np.ascontiguousarray(dat, np.float64)
    syn = np.ascontiguousarray(syn, np.float64)
    win = np.ascontiguousarray(np.reshape(windows, (-1, 2)), np.float64)
    nw = win.shape[0]
    adj = np.zeros(len(dat))
    out = np.zeros((nw, 10))
    nf = 2 ** (int(np.frexp(np.float32(len(dat)))[1]))  # 2 ** exponent(real(nlen))
    err = np.zeros((nw, 2, nf))
    tot = lib().orc_mt_adjoint(_p(dat, C.c_double), _p(syn, C.c_double), len(dat), dt, _p(win, C.c_double), nw,
                               taper_pct, itaper, imeas, _p(cv, C.c_double), _p(adj, C.c_double),
                               _p(out, C.c_double), err.ctypes.data_as(C.c_void_p))
    o = {k: out[:, i].copy() for i, k in enumerate(MT_FIELDS)}
    o["selected"] = o["selected"] != 0
    if errors:
        return tot, adj, o, err
    return tot, adj, o


def hilbert(x):
    """The reference's analytic signal (fftpack hilbert): complex, len(x)."""
    x = np.ascontiguousarray(x, np.float64)
    re = np.zeros(len(x))
    im = np.zeros(len(x))
    lib().orc_hilbert(_p(x, C.c_double), len(x), _p(re, C.c_double), _p(im, C.c_double))
    return re + 1j * im


EP_FIELDS = ("misfit", "residual", "misfit_real", "misfit_imag", "residual_real", "residual_imag", "selected")


def ep_adjoint(dat, syn, dt, windows, taper_pct=0.05, itaper=1, wtr_env=0.2):
    """One trace's exponentiated-phase measurement: (total, adj, per-window dict of EP_FIELDS)."""
    dat = np.ascontiguousarray(dat, np.float64)
    syn = np.ascontiguousarray(syn, np.float64)
    win = np.ascontiguousarray(np.reshape(windows, (-1, 2)), np.float64)
    nw = win.shape[0]
    adj = np.zeros(len(dat))
    out = np.zeros((nw, 8))
    tot = lib().orc_ep_adjoint(_p(dat, C.c_double), _p(syn, C.c_double), len(dat), dt, _p(win, C.c_double), nw,
                               taper_pct, itaper, wtr_env, _p(adj, C.c_double), _p(out, C.c_double))
    return tot, adj, {k: out[:, i].copy() for i, k in enumerate(EP_FIELDS)}


WIN_DEFAULTS = dict(sliding_win_len_fac=2.0, min_velocity=2.4, threshold_corr=0.8, threshold_shift_fac=0.3,
                    jump_fac=0.1, min_win_len_fac=1.5, min_snr_window=5.0, min_energy_ratio=5.0,
                    min_peaks_troughs=3, is_split_phases=0.0)  # input_params.f90:389-398


def win_sel(dat, syn, dt, t0, tp, dis, min_period, maxwin=1000, **cfg):
    """win_sel_type%select_windows (is_split_phases: split at STA/LTA troughs):
    (windows [n][2] in s, sliding cc, sliding time shifts)."""
    c = dict(WIN_DEFAULTS, **cfg)
    cv = np.array([float(c[k]) for k in WIN_DEFAULTS], np.float64)
    dat = np.ascontiguousarray(dat, np.float64)
    syn = np.ascontiguousarray(syn, np.float64)
    n = len(dat)
    nlen = int(c["sliding_win_len_fac"] * min_period / dt) + 1
    cc = np.zeros(max(n - nlen, 1))
    ts = np.zeros(max(n - nlen, 1))
    tw = np.zeros((maxwin, 2))
    k = lib().orc_win_sel(_p(dat, C.c_double), _p(syn, C.c_double), n, dt, t0, tp, dis, min_period,
                          _p(cv, C.c_double), maxwin, _p(tw, C.c_double), _p(cc, C.c_double), _p(ts, C.c_double))
    return tw[:min(k, maxwin)].copy(), cc[:max(n - nlen, 0)], ts[:max(n - nlen, 0)]


FLEXWIN_DEFAULTS = dict(min_velocity=2.4, min_win_len_fac=1.5, threshold_corr=0.8, threshold_shift_fac=2.0,
                        min_peaks_troughs=3, threshold_dlna=1.0, min_snr_window=5.0, is_split_phases=0.0,
                        max_time_before_first_arrival=10.0, stalta_water_level=0.08, c_0=1.0, c_2=0.0,
                        c_3a=4.0, c_3b=2.5, c_4a=2.0, c_4b=6.0)  # input_params.f90:400-418 (WIN_TYPE_2)


def flexwin(dat, syn, dt, t0, tp, dis, min_period, max_period, maxwin=1000, **cfg):
    """flexwin_type%select_windows: (windows [n][2] in s, info [n][4] = weight,
    cc, time shift, dlnA)."""
    c = dict(FLEXWIN_DEFAULTS, **cfg)
    cv = np.array([float(c[k]) for k in FLEXWIN_DEFAULTS], np.float64)
    dat = np.ascontiguousarray(dat, np.float64)
    syn = np.ascontiguousarray(syn, np.float64)
    tw = np.zeros((maxwin, 2))
    info = np.zeros((maxwin, 4))
    k = lib().orc_flexwin(_p(dat, C.c_double), _p(syn, C.c_double), len(dat), dt, t0, tp, dis, min_period,
                          max_period, _p(cv, C.c_double), maxwin, _p(tw, C.c_double), _p(info, C.c_double))
    return tw[:min(k, maxwin)].copy(), info[:min(k, maxwin)].copy()


def schedule_intervals(left, right, weight):
    """schedule_weighted_intervals: input positions of the chosen intervals, in time order."""
    lf, rt, wt = (np.ascontiguousarray(v, np.float64) for v in (left, right, weight))
    sel = np.zeros(len(lf), np.int32)
    k = lib().orc_schedule_intervals(len(lf), _p(lf, C.c_double), _p(rt, C.c_double), _p(wt, C.c_double),
                                     _p(sel, C.c_int))
    return sel[:k].copy()


def sta_lta(env, dt, min_period):
    """STA_LTA (sta_lta.f90:6-104) of an envelope."""
    env = np.ascontiguousarray(env, np.float64)
    out = np.zeros(len(env))
    lib().orc_sta_lta(_p(env, C.c_double), len(env), dt, min_period, _p(out, C.c_double))
    return out


def conv_adjoint(dat_r, dat_z, syn_r, syn_z, dt, window, taper_pct=0.05, itaper=1):
    """Cross-convolution misfit of one receiver: (misfit, residual, adj_r, adj_z)."""
    a = [np.ascontiguousarray(v, np.float64) for v in (dat_r, dat_z, syn_r, syn_z)]
    n = len(a[0])
    ar = np.zeros(n)
    az = np.zeros(n)
    out = np.zeros(2)
    lib().orc_conv_adjoint(*[_p(v, C.c_double) for v in a], n, dt, float(window[0]), float(window[1]), taper_pct,
                           itaper, _p(ar, C.c_double), _p(az, C.c_double), _p(out, C.c_double))
    return out[0], out[1], ar, az
