"""Pins of the optimiser oracle (oracle/opt_oracle.cpp).

  lbfgs_direction  the explicit BFGS update of the inverse Hessian
                   H <- V^T H V + rho s s^T, V = I - rho y s^T, from
                   H0 = (y.s / y.y) diag(hess) of the latest pair, applied as
                   matrices (a different algorithm from the two-loop recursion);
                   the secant condition H y = s; invariance to the cell volume
  cg / sd          hand-computed cases
  model_update     exp update and the vp/vs clamp by hand
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def _history(nstore, nk=3, npg=12, seed=0):
    rng = np.random.default_rng(seed)
    models = [np.exp(rng.normal(0, 0.3, (nk, npg))) * 3000.0]
    grads = [rng.standard_normal((nk, npg))]
    for _ in range(nstore):
        models.append(models[-1] * np.exp(rng.normal(0, 0.05, (nk, npg))))
        # gradients of a convex quadratic in log-model space so that y.s > 0
        A = np.diag(rng.uniform(1.0, 4.0, nk * npg))
        s = (np.log(models[-1]) - np.log(models[-2])).ravel()
        grads.append((grads[-1].ravel() + A @ s).reshape(nk, npg))
    return np.array(grads, np.float32), np.array(models, np.float32)


def _explicit(grads, models, hess, npg):
    G = grads.reshape(grads.shape[0], -1).astype(np.float64)
    M = models.reshape(models.shape[0], -1).astype(np.float64)
    n = G.shape[1]
    logm = np.where((np.arange(n) // npg < 3)[None, :], np.log(M), M)
    S = np.diff(logm, axis=0)
    Y = np.diff(G, axis=0)
    gam = Y[-1] @ S[-1] / (Y[-1] @ Y[-1])
    H = gam * np.diag(hess.ravel().astype(np.float64))
    I = np.eye(n)
    for s, y in zip(S, Y):
        rho = 1.0 / (y @ s)
        V = I - rho * np.outer(y, s)
        H = V.T @ H @ V + rho * np.outer(s, s)
    r = H @ G[-1]
    d = -r / np.abs(r).max()
    g = G[-1]
    ang = np.degrees(np.arccos((-g @ d) / (np.linalg.norm(g) * np.linalg.norm(d))))
    return d, ang, H, S, Y


def test_lbfgs_is_the_explicit_bfgs_update():
    for nstore in (1, 2, 4):
        grads, models = _history(nstore, seed=nstore)
        npg = grads.shape[2]
        hess = np.random.default_rng(9).uniform(0.5, 2.0, grads.shape[1:]).astype(np.float32)
        d, ang = O.lbfgs_direction(grads, models, hess, npg, 250.0, 250.0, 125.0, dbl=True)
        de, ange, _, _, _ = _explicit(grads, models, hess, npg)
        np.testing.assert_allclose(d, de, atol=1e-6)
        assert abs(ang - ange) < 1e-6
        d32, ang32 = O.lbfgs_direction(grads, models, hess, npg, 250.0, 250.0, 125.0, dbl=False)
        np.testing.assert_allclose(d32, de, atol=2e-4)


def test_lbfgs_secant_condition_and_volume_invariance():
    grads, models = _history(3, seed=5)
    npg = grads.shape[2]
    hess = np.ones(grads.shape[1:], np.float32)
    _, _, H, S, Y = _explicit(grads, models, hess, npg)
    np.testing.assert_allclose(H @ Y[-1], S[-1], rtol=1e-9, atol=1e-12)
    a, _ = O.lbfgs_direction(grads, models, hess, npg, 1.0, 1.0, 1.0, dbl=True)
    b, _ = O.lbfgs_direction(grads, models, hess, npg, 500.0, 300.0, 70.0, dbl=True)
    np.testing.assert_allclose(a, b, atol=1e-12)


def test_sd_and_cg_hand_cases():
    g = np.array([1.0, -2.0, 4.0, 0.5], np.float32)
    h = np.array([1.0, 1.0, 0.5, 2.0], np.float32)
    np.testing.assert_allclose(O.sd_direction(g, h, dbl=True), [-0.5, 1.0, -1.0, -0.5])
    g0 = np.array([1.0, 0.0, 2.0, 0.0], np.float32)   # two parameters of two points
    g1 = np.array([2.0, 1.0, 1.0, 1.0], np.float32)
    # alpha = sum((g1 - g0) g1) / sum(g0 g0) = (2 + 1 - 1 + 1) / 5
    d, alpha = O.cg_direction(g1, g0, np.ones(4, np.float32), 2, dbl=True)
    assert abs(alpha - 0.6) < 1e-12
    r = -g1 - 0.6 * g0
    np.testing.assert_allclose(d, r / np.abs(r).max(), rtol=1e-6)


def test_model_update_and_alpha_scaling_by_hand():
    npg = 3
    vp = np.array([6000.0, 6000.0, 6000.0]); vs = np.array([3500.0, 2000.0, 4000.0]); rho = np.full(3, 2700.0)
    m = np.stack([vp, vs, rho]).astype(np.float32)
    d = np.zeros_like(m)
    d[2] = 1.0
    out = O.model_update(m, d, npg, 1, 0.1, (1.6, 2.0)).reshape(3, 3)
    np.testing.assert_allclose(out[2], 2700.0 * np.exp(0.1), rtol=1e-6)
    # vp/vs: 1.714 kept, 3.0 -> 2.0 x vs, 1.5 -> 1.6 x vs
    np.testing.assert_allclose(out[0], [6000.0, 4000.0, 6400.0], rtol=1e-6)
    # model type 2: gcp, gsp additive; alternating inversion of the anisotropic half only
    m5 = np.concatenate([m, np.full((2, npg), 0.01, np.float32)])
    d5 = np.ones_like(m5)
    out5 = O.model_update(m5, d5, npg, 2, 0.02, (1.6, 2.0), update_iso=False).reshape(5, 3)
    np.testing.assert_allclose(out5[:3], m5[:3])
    np.testing.assert_allclose(out5[3:], 0.03, rtol=1e-6)
