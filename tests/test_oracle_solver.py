"""Pins of the CPU oracle's solver path by closed forms and invariants.

The SPECFEM3D solver SpecFWAT drives is not in the reference tree, so these
checks (not a copy of the oracle's formulas) are what make the oracle
trustworthy: GLL closed forms and polynomial exactness, numbering by two
independent algorithms, Jacobians against finite differences, mass = rho V,
rigid-body and patch tests of the stiffness, K symmetry, energy conservation,
exact time reversal, iso/aniso consistency and a finite-difference gradient
check of the adjoint kernels.
"""
import numpy as np
import pytest

from oracle import oracle as orc


# ---------------------------------------------------------------- GLL basis
def test_gll_closed_form():
    xi, w, hp, hpw = orc.gll()
    s = np.sqrt(3.0 / 7.0)
    assert np.allclose(xi, [-1, -s, 0, s, 1], atol=1e-15)
    assert np.allclose(w, [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10], atol=1e-15)
    # quadrature exact to degree 2N-3 = 7
    for k in range(8):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, xi**k) - exact) < 1e-14
    # derivative matrix exact on polynomials of degree <= 4; rows sum to zero
    for k in range(5):
        d = k * xi ** max(k - 1, 0) if k else np.zeros(5)
        assert np.allclose(hp @ xi**k, d, atol=1e-12)
    assert np.allclose(hp.sum(axis=1), 0, atol=1e-13)
    assert np.allclose(hpw, w[:, None] * hp, atol=1e-15)


def test_lagrange_interpolants():
    xi, _, _, _ = orc.gll()
    for j, x in enumerate(xi):
        h, _ = orc.lagrange(x)
        assert np.allclose(h, np.eye(5)[j], atol=1e-14)
    x = 0.3172
    h, hd = orc.lagrange(x)
    for k in range(5):
        assert abs(h @ xi**k - x**k) < 1e-13
        assert abs(hd @ xi**k - (k * x ** (k - 1) if k else 0.0)) < 1e-12


def test_ricker_is_scaled_second_derivative_of_gaussian():
    f0 = 1.3
    a = (np.pi * f0) ** 2
    t = np.linspace(-1.0, 1.0, 41)
    eps = 1e-4
    g = lambda tt: np.exp(-a * tt * tt)
    d2 = (g(t + eps) - 2 * g(t) + g(t - eps)) / eps**2
    r = np.array([orc.ricker(tt, f0) for tt in t])
    assert np.allclose(r, -d2 / (2 * a), atol=1e-6)
    assert abs(orc.ricker(1.0 / (np.sqrt(2.0) * np.pi * f0), f0)) < 1e-14


# ---------------------------------------------------------------- mesh
def _oracle(nex=3, ney=2, nez=2, h=1000.0, perturb=0.0, firr=False, numbering=1, stacey=True,
            aniso_cij=None, rho=2700.0, **kw):
    nloc = nex * ney * nez * 125
    r = np.full(nloc, rho, np.float32)
    vp = np.full(nloc, 6000.0, np.float32)
    vs = np.full(nloc, 3500.0, np.float32)
    return orc.Oracle(nex, ney, nez, nex * h, ney * h, nez * h, r, vp, vs, cij=aniso_cij,
                      perturb=perturb, force_irregular=firr, numbering=numbering, stacey=stacey, **kw)


@pytest.mark.parametrize("perturb", [0.0, 0.15])
def test_numbering_two_algorithms_agree(perturb):
    a = _oracle(4, 3, 2, perturb=perturb, numbering=0).mesh()
    b = _oracle(4, 3, 2, perturb=perturb, numbering=1).mesh()
    assert np.array_equal(a["ibool"], b["ibool"])
    assert a["ibool"].max() + 1 == (4 * 4 + 1) * (4 * 3 + 1) * (4 * 2 + 1)


def test_numbering_is_first_appearance_and_consistent():
    o = _oracle(3, 3, 2, perturb=0.1, numbering=0)
    m = o.mesh()
    ib = m["ibool"]
    first = np.full(o.nglob, -1)
    for q, g in enumerate(ib):
        if first[g] < 0:
            first[g] = q
    assert np.all(np.diff(first) > 0)  # ids issued in order of first appearance
    # every local point of one id sits at the same place
    x = m["xyz"]
    assert np.allclose(x[ib], x[ib], atol=0)


def test_regular_jacobian_closed_form():
    h = 2500.0
    m = _oracle(2, 2, 2, h=h).mesh()
    assert np.all(m["irregular"] == 0)
    assert m["xix_regular"] == pytest.approx(2.0 / h, rel=1e-15)
    assert m["jacobian_regular"] == pytest.approx((h / 2) ** 3, rel=1e-15)
    mi = _oracle(2, 2, 2, h=h, firr=True).mesh()
    assert np.array_equal(mi["irregular"], np.arange(1, 9))
    J = mi["jinv"].reshape(-1, 3, 3)
    assert np.allclose(J, np.diag([2 / h] * 3)[None], atol=1e-18)
    assert np.allclose(mi["jacobian"], (h / 2) ** 3, rtol=1e-13)


def test_general_hex_jacobian_vs_finite_differences():
    o = _oracle(3, 3, 3, perturb=0.2)
    m = o.mesh()
    xi, _, _, _ = orc.gll()
    ib = m["ibool"].reshape(-1, 125)
    x = m["xyz"]
    e = 13  # interior element, all nodes perturbed
    corners = [0, 4, 20, 24, 100, 104, 120, 124]  # (i,j,k) in {0,4}^3, i fastest
    X = x[ib[e, corners]]

    def xmap(r):
        out = np.zeros(3)
        for n in range(8):
            s = [(n & 1) * 2 - 1, ((n >> 1) & 1) * 2 - 1, ((n >> 2) & 1) * 2 - 1]
            out += (1 + s[0] * r[0]) * (1 + s[1] * r[1]) * (1 + s[2] * r[2]) / 8 * X[n]
        return out

    eps = 1e-6
    for L in (0, 31, 62, 124):
        r = np.array([xi[L % 5], xi[(L // 5) % 5], xi[L // 25]])
        assert np.allclose(xmap(r), x[ib[e, L]], rtol=1e-12)
        Jfd = np.stack([(xmap(r + eps * np.eye(3)[d]) - xmap(r - eps * np.eye(3)[d])) / (2 * eps)
                        for d in range(3)], axis=1)
        Jinv = m["jinv"].reshape(-1, 125, 3, 3)[e, L]
        assert np.allclose(Jinv, np.linalg.inv(Jfd), rtol=1e-6)
        assert m["jacobian"].reshape(-1, 125)[e, L] == pytest.approx(np.linalg.det(Jfd), rel=1e-6)


@pytest.mark.parametrize("perturb", [0.0, 0.2])
def test_mass_integrates_density(perturb):
    o = _oracle(3, 2, 2, h=1000.0, perturb=perturb, stacey=False, rho=2500.0)
    rmx, rmy, rmz = o.rmass_inv()
    vol = 3000.0 * 2000.0 * 2000.0
    assert np.sum(1.0 / rmx.astype(np.float64)) == pytest.approx(2500.0 * vol, rel=1e-6)
    assert np.array_equal(rmx, rmz)


def test_stacey_faces_cover_the_absorbing_sides():
    nex, ney, nez, h = 3, 2, 2, 1000.0
    o = _oracle(nex, ney, nez, h=h)
    f = o.faces()
    assert o.nfaces == 2 * ney * nez + 2 * nex * nez + nex * ney
    n = f["normal"]
    assert np.allclose(np.linalg.norm(n, axis=-1), 1.0)
    areas = {}
    for side, (axis, sign) in enumerate([(0, -1), (0, 1), (1, -1), (1, 1), (2, -1)]):
        sel = np.all(np.isclose(n[:, :, axis], sign), axis=1)
        areas[side] = f["jacobian2Dw"][sel].sum()
    L = [nex * h, ney * h, nez * h]
    assert areas[0] == pytest.approx(L[1] * L[2]) and areas[1] == pytest.approx(L[1] * L[2])
    assert areas[2] == pytest.approx(L[0] * L[2]) and areas[4] == pytest.approx(L[0] * L[1])


def test_colouring_is_valid():
    o = _oracle(4, 3, 3, perturb=0.1)
    m = o.mesh()
    ib = m["ibool"].reshape(-1, 125)
    col = m["color"]
    assert col.max() + 1 == 8
    for c in range(8):
        pts = ib[col == c].ravel()
        assert len(np.unique(pts)) == len(pts)


# ---------------------------------------------------------------- stiffness
@pytest.mark.parametrize("perturb", [0.0, 0.2])
def test_rigid_motions_give_no_force(perturb):
    o = _oracle(3, 3, 2, perturb=perturb, precision=64)
    x = o.mesh()["xyz"]
    scale = np.abs(o.apply_k(np.random.default_rng(0).standard_normal(x.shape))).max()
    for u in (np.tile([1.0, -2.0, 0.5], (len(x), 1)), np.cross([0.3, -0.2, 0.7], x)):
        f = o.apply_k(u)
        assert np.abs(f).max() < 1e-9 * scale * np.abs(u).max()


@pytest.mark.parametrize("perturb", [0.0, 0.2])
def test_patch_test_linear_field(perturb):
    """A linear displacement has constant strain and stress: the assembled
    internal force vanishes at every interior point (trilinear elements map
    linear fields exactly)."""
    o = _oracle(3, 3, 3, perturb=perturb, precision=64)
    x = o.mesh()["xyz"]
    A = np.array([[1.0, 0.3, -0.2], [0.1, -0.5, 0.4], [0.2, 0.6, 0.8]]) * 1e-3
    f = o.apply_k(x @ A.T)
    L = np.array([3000.0, 3000.0, 3000.0])
    lo = np.array([0, 0, -3000.0])
    interior = np.all((x > lo + 1e-6) & (x < lo + L - 1e-6), axis=1)
    assert interior.sum() > 0
    assert np.abs(f[interior]).max() < 1e-9 * np.abs(f[~interior]).max()


def test_stiffness_symmetric_positive():
    o = _oracle(2, 2, 2, perturb=0.2, precision=64)
    rng = np.random.default_rng(1)
    u = rng.standard_normal((o.nglob, 3))
    v = rng.standard_normal((o.nglob, 3))
    Ku, Kv = -o.apply_k(u), -o.apply_k(v)
    assert np.vdot(u, Kv) == pytest.approx(np.vdot(v, Ku), rel=1e-11)
    assert np.vdot(u, Ku) > 0


def test_energy_conservation_closed_box():
    """Free surfaces everywhere (no Stacey, no source): the central-difference
    Newmark scheme keeps the energy to O(dt^2)."""
    drifts = []
    for dt in (0.01, 0.005):
        n = int(round(2.0 / dt))
        o = _oracle(2, 2, 2, h=1000.0, perturb=0.1, stacey=False, precision=64, dt=dt, nstep=n)
        x = o.mesh()["xyz"]
        u0 = (1e-3 * np.exp(-np.sum((x - [1000, 1000, -1000]) ** 2, axis=1) / (2 * 400.0**2)))[:, None] * [1, 0.5, -0.3]
        e = o.energy_series(u0.astype(np.float32))
        drifts.append((e.max() - e.min()) / e.mean())
    assert drifts[0] < 1e-2
    assert drifts[1] < drifts[0] / 3.0


# ---------------------------------------------------------------- time loop
def _event(precision=64, nstep=200, aniso=None, perturb=0.0, mu_scale=1.0, kappa_scale=1.0, rho_scale=1.0,
           bump=None, dt=0.02):
    nex, ney, nez, h = 4, 4, 3, 2500.0
    nloc = nex * ney * nez * 125
    rng = np.random.default_rng(5)
    f = 1 + 0.02 * rng.standard_normal(nloc)
    rho = 2700 * f
    vp = 6000 * f
    vs = 3500 * f
    if bump is not None:
        # perturb ln rho, ln kappa, ln mu by the given local field
        kap = rho * (vp**2 - 4 / 3 * vs**2) * np.exp(kappa_scale * bump)
        mu = rho * vs**2 * np.exp(mu_scale * bump)
        rho = rho * np.exp(rho_scale * bump)
        vs = np.sqrt(mu / rho)
        vp = np.sqrt((kap + 4 / 3 * mu) / rho)
    return orc.Oracle(nex, ney, nez, nex * h, ney * h, nez * h, rho.astype(np.float32),
                      vp.astype(np.float32), vs.astype(np.float32), cij=aniso, dt=dt, nstep=nstep,
                      src_xyz=[[4300.0, 5700.0, -3400.0]], src_mt=[[1e15, -0.7e15, 0.2e15, 0.4e15, -0.3e15, 0.5e15]],
                      src_f0=[0.8], rec_xyz=[[2000, 2000, 0], [8000, 3000, 0], [5000, 8000, 0], [3000, 6000, 0]],
                      perturb=perturb, precision=precision)


def test_reconstruction_is_exact_time_reversal():
    o = _event(precision=64)
    s = o.forward(save=True)
    u_end = o.final_state()[0]
    k = o.adjoint(np.zeros_like(s))
    b_u0 = k["b_final"][0]
    assert np.abs(b_u0).max() < 1e-9 * np.abs(u_end).max()
    assert np.abs(s).max() > 0


def test_fp32_oracle_close_to_fp64():
    s32 = _event(precision=32).forward(save=False)
    s64 = _event(precision=64).forward(save=False)
    assert np.linalg.norm(s32 - s64) / np.linalg.norm(s64) < 1e-5


def test_isotropic_cij_reproduces_isotropic_run():
    o_iso = _event(precision=64, nstep=120)
    nloc = 4 * 4 * 3 * 125
    rng = np.random.default_rng(5)
    f = 1 + 0.02 * rng.standard_normal(nloc)
    rho = (2700 * f).astype(np.float32)
    vp = (6000 * f).astype(np.float32)
    vs = (3500 * f).astype(np.float32)
    mu = rho.astype(np.float64) * vs.astype(np.float64) ** 2
    lam = rho.astype(np.float64) * vp.astype(np.float64) ** 2 - 2 * mu
    c = np.zeros((nloc, 21))
    for q in (0, 6, 11):  # c11 c22 c33
        c[:, q] = lam + 2 * mu
    for q in (1, 2, 7):  # c12 c13 c23
        c[:, q] = lam
    for q in (15, 18, 20):  # c44 c55 c66
        c[:, q] = mu
    o_an = _event(precision=64, nstep=120, aniso=c.astype(np.float32))
    s_iso = o_iso.forward(save=True)
    s_an = o_an.forward(save=True)
    assert np.linalg.norm(s_an - s_iso) / np.linalg.norm(s_iso) < 1e-6
    adj = s_iso.astype(np.float32)
    ki = o_iso.adjoint(adj)
    ka = o_an.adjoint(adj)
    # kappa/mu kernels are the c_ij kernels contracted with d c / d ln kappa, d ln mu
    kap = (lam + 2 / 3 * mu)
    cij = ka["cijkl_kl"]
    dk = np.zeros(21); dk[[0, 6, 11, 1, 2, 7]] = 1.0           # d c_IJ / d kappa
    dm = np.zeros(21); dm[[0, 6, 11]] = 4 / 3; dm[[1, 2, 7]] = -2 / 3; dm[[15, 18, 20]] = 1.0
    k_kappa = kap * (cij @ dk)
    k_mu = mu * (cij @ dm)
    assert np.linalg.norm(k_kappa - ki["kappa_kl"]) / np.linalg.norm(ki["kappa_kl"]) < 1e-5
    assert np.linalg.norm(k_mu - ki["mu_kl"]) / np.linalg.norm(ki["mu_kl"]) < 1e-5


@pytest.mark.parametrize("param", ["rho", "kappa", "mu"])
def test_kernels_match_finite_difference_gradient(param):
    """chi(m) = 1/2 sum_rec int |s - d|^2 dt; d chi = int K dln(m) dV.
    The kernel integrated against a smooth perturbation must match the
    centred finite difference of chi.  This pins the sign, the scaling and
    the time pairing of adjoint and reconstructed fields: pairing adjoint
    step it with forward state N - it instead of N + 1 - it leaves an O(dt)
    error of 6-16 % at this dt."""
    nstep = 220
    nloc = 4 * 4 * 3 * 125
    base = _event(precision=64, nstep=nstep)
    x = base.mesh()["xyz"][base.mesh()["ibool"]]
    bump = np.exp(-np.sum((x - [5000, 5000, -3500]) ** 2, axis=1) / (2 * 2000.0**2))
    sc = dict(rho_scale=0.0, kappa_scale=0.0, mu_scale=0.0)
    sc[param + "_scale"] = 1.0
    # "observed" data from a model with a different perturbation
    obs = _event(precision=64, nstep=nstep, bump=0.05 * np.roll(bump, 3000), rho_scale=1, kappa_scale=1,
                 mu_scale=1).forward(save=False)
    dt = 0.02

    def chi(eps):
        o = _event(precision=64, nstep=nstep, bump=eps * bump, **sc)
        s = o.forward(save=False)
        return 0.5 * np.sum((s - obs) ** 2) * dt

    syn = base.forward(save=True)
    k = base.adjoint((syn - obs).astype(np.float32))
    key = {"rho": "rho_kl", "kappa": "kappa_kl", "mu": "mu_kl"}[param]
    m = base.mesh()
    _, w, _, _ = orc.gll()
    ii = np.arange(125)
    w3 = np.tile(w[ii % 5] * w[(ii // 5) % 5] * w[ii // 25], 48)
    jac = m["jacobian_regular"]
    pred = np.sum(k[key] * bump * w3 * jac)
    eps = 1e-3
    fd = (chi(eps) - chi(-eps)) / (2 * eps)
    assert np.sign(pred) == np.sign(fd)
    assert abs(pred - fd) < 0.02 * abs(fd)
