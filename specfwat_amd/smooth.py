"""smooth_sem_pde host side: the reference's set-up, then smooth_pde_cuda_.

Restates src/postproc/smooth.f90:7-437 up to the GPU call: rotate_r
(:80-94), cv/ch/ntstep from distance_min_glob (:101-109), rvol assembly
(:142-164), projection of the element-local input onto global points
(:192-220), the call of smooth_pde_cuda_ (:392-402), and the final gather.
Single precision throughout, as CUSTOM_REAL.
"""
import numpy as np

from . import _lib

F32 = np.float32


def smoothing_constants(sigma_h, sigma_v, distance_min_glob):
    """(cv, ch, ntstep) as smooth.f90:101-109."""
    d = F32(distance_min_glob)
    sh, sv = F32(sigma_h), F32(sigma_v)
    cmax = d * d / F32(9.0)
    if sv >= sh:
        cv = cmax
        ch = cv * (sh * sh) / (sv * sv)
    else:
        ch = cmax
        cv = ch * (sv * sv) / (sh * sh)
    mx = max(sh, sv)
    ntstep = int(np.ceil((mx * mx) / (F32(2.0) * cmax)))
    return F32(cv), F32(ch), ntstep


def rvol_arrays(mesh):
    """rvol_local (float32 [nspec,125]) and inverted assembled rvol (float32 [nglob])."""
    w = mesh.gll.w
    ii = np.arange(125)
    weight = w[ii % 5] * w[(ii // 5) % 5] * w[ii // 25]
    rvol_local = (mesh.jacobian_all.astype(np.float64) * weight[None, :]).astype(F32)
    rvol = np.zeros(mesh.nglob, F32)
    np.add.at(rvol, mesh.ibool0.ravel(), rvol_local.ravel())
    return rvol_local, (F32(1.0) / rvol).astype(F32)


def project(mesh, dat_in, rvol_local, rvol_inv):
    dglob = np.zeros(mesh.nglob, F32)
    np.add.at(dglob, mesh.ibool0.ravel(), (np.asarray(dat_in, F32).reshape(mesh.nspec, 125) * rvol_local).ravel())
    return (dglob * rvol_inv).astype(F32)


def smooth_sem_pde(mesh, dat_in, sigma_h, sigma_v, *, is_sph=False, faithful=True,
                   distance_min_glob=None, nstep=None):
    """Returns (dat [nspec,125], dat_glob [nglob], ntstep)."""
    if distance_min_glob is None:
        distance_min_glob = mesh.distance_min_glob()
    cv, ch, ntstep = smoothing_constants(sigma_h, sigma_v, distance_min_glob)
    if nstep is not None:
        ntstep = int(nstep)
    rvol_local, rvol = rvol_arrays(mesh)
    dglob = project(mesh, dat_in, rvol_local, rvol)
    rot = np.zeros((mesh.nspec, 125, 3), F32)
    if is_sph:
        x = np.stack([mesh.xstore, mesh.ystore, mesh.zstore], axis=1)[mesh.ibool0]
        rot = (x / np.sqrt((x * x).sum(axis=-1, keepdims=True))).astype(F32)
    g = mesh.gll
    dat = np.zeros((mesh.nspec, 125), F32)
    ddat = np.zeros(mesh.nglob, F32)
    i32 = lambda a: np.ascontiguousarray(a, np.int32)
    _lib.call("fwat_smooth_set_faithful_", int(faithful))
    dummy_i = np.zeros(1, np.int32)
    _lib.call("smooth_pde_cuda_", int(is_sph), mesh.nspec, mesh.nglob, 1, ntstep, mesh.num_phase_ispec,
              i32(mesh.phase_ispec_inner), mesh.nspec_outer, mesh.nspec_inner,
              mesh.xix, mesh.xiy, mesh.xiz, mesh.etax, mesh.etay, mesh.etaz, mesh.gammax, mesh.gammay,
              mesh.gammaz, mesh.jacobian_all, rot, g.wgllwgll_xy, g.wgllwgll_xz, g.wgllwgll_yz,
              g.hprime_xxT, g.hprimewgll_xx, i32(mesh.ibool), np.zeros(mesh.nspec, np.int32),
              float(cv), float(ch), 0, 0, dummy_i, dummy_i, dummy_i,
              np.zeros((mesh.nspec, 125), F32), dat, rvol, dglob, ddat,
              i32(mesh.irregular_element_number), float(mesh.xix_regular), mesh.nspec_irregular)
    return dat, dglob, ntstep


def smooth_batch(mesh, fields, sigma_h, sigma_v, *, is_sph=False, faithful=True,
                 distance_min_glob=None, nstep=None, timed=False):
    """Smooth several element-local fields ([nfield, nspec, 125]) in one device
    run sharing the geometry reads.  Returns (dat_glob [nfield, nglob], ntstep, elem_ms,
    update_ms); the times are filled when timed."""
    if distance_min_glob is None:
        distance_min_glob = mesh.distance_min_glob()
    cv, ch, ntstep = smoothing_constants(sigma_h, sigma_v, distance_min_glob)
    if nstep is not None:
        ntstep = int(nstep)
    rvol_local, rvol = rvol_arrays(mesh)
    fields = np.asarray(fields, F32).reshape(-1, mesh.nspec, 125)
    dglob = np.ascontiguousarray(np.stack([project(mesh, f, rvol_local, rvol) for f in fields]))
    rot = np.zeros((mesh.nspec, 125, 3), F32)
    if is_sph:
        x = np.stack([mesh.xstore, mesh.ystore, mesh.zstore], axis=1)[mesh.ibool0]
        rot = (x / np.sqrt((x * x).sum(axis=-1, keepdims=True))).astype(F32)
    g = mesh.gll
    i32 = lambda a: np.ascontiguousarray(a, np.int32)
    _lib.call("fwat_smooth_set_faithful_", int(faithful))
    _lib.call("fwat_smooth_pde_batch_", int(is_sph), mesh.nspec, mesh.nglob, ntstep, len(fields),
              mesh.xix, mesh.xiy, mesh.xiz, mesh.etax, mesh.etay, mesh.etaz, mesh.gammax, mesh.gammay,
              mesh.gammaz, mesh.jacobian_all, rot, g.wgllwgll_xy, g.wgllwgll_xz, g.wgllwgll_yz,
              g.hprime_xxT, g.hprimewgll_xx, i32(mesh.ibool), float(cv), float(ch), rvol, dglob,
              i32(mesh.irregular_element_number), float(mesh.xix_regular), mesh.nspec_irregular,
              int(timed))
    ems, ums = np.zeros(1), np.zeros(1)
    _lib.call("fwat_smooth_last_time_", ems, ums)
    return dglob, ntstep, float(ems[0]), float(ums[0])
