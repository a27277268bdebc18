"""Grid optimiser on the GPU: SpecFWAT's OptGridFlow directions and model update
(src/optimization/optimize_grid.f90), with every inner product an fp64 sum.

    get_SD_direction     optimize_grid.f90:229-241
    get_lbfgs_direction  optimize_grid.f90:243-339 (inner_product, utils.f90:1527-1534)
    get_CG_direction     optimize_grid.f90:341-389
    model_update         optimize_grid.f90:176-203 with alpha_scaling :391-404

Grids are float32 arrays [nkernel, nz, ny, nx] (Fortran (nx,ny,nz,nkernel)).
The history of get_lbfgs_direction is passed as the arrays read_gradient_grid
and read_model_grid return for iterations iter_store .. iter_current (models
linear; the first three parameters enter the differences as logs).
"""
import numpy as np

from . import _lib

F32 = np.float32


def _grid(a):
    a = np.ascontiguousarray(a, F32)
    if a.ndim != 4:
        raise ValueError("grids are [nkernel, nz, ny, nx]")
    return a


def last_time_ms():
    return float(_lib.call("fwat_opt_last_time_", 0.0)[0])


def get_SD_direction(gradient, hess):
    g, h = _grid(gradient), _grid(hess)
    out = np.zeros_like(g)
    _lib.call("fwat_opt_sd_direction_", g.size, g, h, out)
    return out


def get_lbfgs_direction(grads, models, hess, spacing=(1.0, 1.0, 1.0)):
    """grads, models: sequences (or [nstore+1, nkernel, nz, ny, nx] arrays) of the
    iterations iter_store .. iter_current; returns (direction, angle in degrees).
    With a single iteration this is get_SD_direction (angle 0), as the reference
    does at ITER_START."""
    G = np.ascontiguousarray(np.stack([_grid(g) for g in grads]), F32)
    M = np.ascontiguousarray(np.stack([_grid(m) for m in models]), F32)
    h = _grid(hess)
    if G.shape != M.shape or G.shape[1:] != h.shape:
        raise ValueError("gradients, models and hess must share the grid shape")
    if G.shape[0] == 1:
        return get_SD_direction(G[0], h), 0.0
    npg = int(np.prod(h.shape[1:]))
    if np.any(M[:, :min(3, h.shape[0])] <= 0):
        raise ValueError("vp, vs and rho must be positive (they enter as logs)")
    out = np.zeros_like(h)
    res = _lib.call("fwat_opt_lbfgs_direction_", h.size, npg, G.shape[0] - 1, G, M, h, float(spacing[0]),
                    float(spacing[1]), float(spacing[2]), out, 0.0)
    return out, float(res[-1])


def get_CG_direction(g1, g0, hess):
    a, b, h = _grid(g1), _grid(g0), _grid(hess)
    out = np.zeros_like(a)
    res = _lib.call("fwat_opt_cg_direction_", a.size, int(np.prod(a.shape[1:])), a, b, h, out, 0.0)
    return out, float(res[-1])


def model_update(model, direction, model_type, step_len, vpvs_ratio_range, alt_inv=False, current_inv_type=1):
    """The updated model (a new array); model_type 1: vp, vs, rho (x exp(step d));
    2: plus gcp, gsp (+ step d); alternating inversion updates one half."""
    m = np.array(_grid(model), F32)
    d = _grid(direction)
    if m.shape != d.shape:
        raise ValueError("model and direction must share the grid shape")
    update_iso = (not alt_inv) or current_inv_type == 1
    update_ani = (not alt_inv) or current_inv_type == 2
    rng = np.asarray(vpvs_ratio_range, F32)
    _lib.call("fwat_opt_model_update_", int(np.prod(m.shape[1:])), m.shape[0], int(model_type), int(update_iso),
              int(update_ani), float(step_len), rng, m, d)
    return m
