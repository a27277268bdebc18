"""Host side of the product (mesher, model, sources, mass, misfit, smoother
set-up) against the oracle's independent construction, plus the C ABI's
export table.  CPU only."""
import os
import re

import numpy as np
import pytest

from oracle import oracle as orc
from specfwat_amd import _lib
from specfwat_amd.gll import GLL
from specfwat_amd.mesh import BoxMesh
from specfwat_amd.misfit import simpson, waveform_adjoint_source, window_taper
from specfwat_amd.model import (gaussian_random_field, hti_cij, interp3, locate_bissection, moduli,
                                perturbed_model)
from specfwat_amd.smooth import rvol_arrays, smoothing_constants
from specfwat_amd.solver import mass_matrices
from specfwat_amd.sources import Receivers, Sources, ricker

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("perturb,firr", [(0.0, False), (0.12, False), (0.0, True)])
def test_mesh_arrays_match_oracle(perturb, firr):
    nex, ney, nez, h = 4, 3, 3, 2500.0
    m = BoxMesh(nex, ney, nez, nex * h, ney * h, nez * h, perturb=perturb, force_irregular=firr)
    nloc = m.nspec * 125
    o = orc.Oracle(nex, ney, nez, m.lx, m.ly, m.lz, np.full(nloc, 2700, np.float32),
                   np.full(nloc, 6000, np.float32), np.full(nloc, 3500, np.float32), perturb=perturb,
                   force_irregular=firr, numbering=0)
    om = o.mesh()
    assert np.array_equal(om["ibool"], m.ibool0.ravel())
    assert np.array_equal(om["irregular"], m.irregular_element_number)
    assert np.allclose(om["xyz"], m.xyz, rtol=0, atol=1e-9)
    sel = m.irregular_element_number > 0
    if sel.any():
        J = om["jinv"].reshape(m.nspec, 125, 9)[sel]
        for q, nm in enumerate(["xix", "xiy", "xiz", "etax", "etay", "etaz", "gammax", "gammay", "gammaz"]):
            assert np.allclose(getattr(m, nm), J[:, :, q], rtol=1e-6, atol=1e-12)
        assert np.allclose(m.jacobian, om["jacobian"].reshape(m.nspec, 125)[sel], rtol=1e-6)
    else:
        assert m.xix_regular == np.float32(om["xix_regular"])
        assert m.jacobian_regular == np.float32(om["jacobian_regular"])
    f = o.faces()
    assert np.array_equal(f["ispec"] + 1, m.abs_boundary_ispec)
    assert np.array_equal(f["ijk"] + 1, m.abs_boundary_ijk)
    assert np.allclose(f["normal"], m.abs_boundary_normal, atol=1e-7)
    assert np.allclose(f["jacobian2Dw"], m.abs_boundary_jacobian2Dw, rtol=1e-6)


def test_gll_tables_match_oracle():
    xi, w, hp, hpw = orc.gll()
    g = GLL()
    assert np.allclose(g.xi, xi, atol=1e-15) and np.allclose(g.w, w, atol=1e-15)
    # Fortran (i,j) order: flat index i + 5 j
    assert np.array_equal(g.hprime_xx.ravel(order="C"), hp.T.ravel().astype(np.float32))
    assert np.array_equal(g.hprimewgll_xx.ravel(), hpw.T.ravel().astype(np.float32))


def test_sources_receivers_and_mass_match_oracle():
    nex, ney, nez, h = 4, 4, 3, 2500.0
    m = BoxMesh(nex, ney, nez, nex * h, ney * h, nez * h, perturb=0.1)
    rng = np.random.default_rng(0)
    f = (1 + 0.03 * rng.standard_normal(m.nglob))[m.ibool0]
    rho, vp, vs = (2700 * f).astype(np.float32), (6000 * f).astype(np.float32), (3500 * f).astype(np.float32)
    sx = [[4300.0, 5700.0, -3375.0]]
    mt = [list(1e15 * rng.standard_normal(6))]
    rx = [[1000.0, 2000.0, 0.0], [7000.0, 3000.0, 0.0], [5000.0, 5000.0, 0.0]]
    o = orc.Oracle(nex, ney, nez, m.lx, m.ly, m.lz, rho.ravel(), vp.ravel(), vs.ravel(), dt=0.02,
                   nstep=10, src_xyz=sx, src_mt=mt, src_f0=[1.0], rec_xyz=rx, perturb=0.1)
    s = o.sources()
    src = Sources(m, sx, mt, [1.0])
    rec = Receivers(m, rx)
    assert np.array_equal(s["src_ispec"] + 1, src.ispec)
    assert np.abs(s["sourcearrays"] - src.sourcearrays).max() < 1e-6 * np.abs(s["sourcearrays"]).max()
    assert np.array_equal(s["rec_ispec"] + 1, rec.ispec)
    assert np.abs(s["rec_h"] - rec.hlagrange).max() < 1e-6
    _, _, rvp, rvs = moduli(rho, vp, vs)
    got = mass_matrices(m, rho, rvp, rvs, 0.02)
    want = o.rmass_inv()
    for a, b in zip(got, want):
        assert np.abs(a / b - 1).max() < 1e-6
    stf = src.stf(10, 0.02)
    assert np.allclose(stf[:, 0], [orc.ricker((it - 1) * 0.02 - 1.2, 1.0) for it in range(1, 11)], rtol=1e-12)
    assert ricker(0.0, 1.0) == 1.0


def test_interp3_bit_exact_with_oracle():
    rng = np.random.default_rng(3)
    vx = np.linspace(0, 4000, 17).astype(np.float32)
    vy = np.linspace(0, 3000, 13).astype(np.float32)
    vz = np.linspace(-2000, 0, 9).astype(np.float32)
    val = rng.standard_normal((17, 13, 9)).astype(np.float32)
    x = rng.uniform(0, 4000, 200).astype(np.float32)
    y = rng.uniform(0, 3000, 200).astype(np.float32)
    z = rng.uniform(-2000, 0, 200).astype(np.float32)
    x[:3] = [0.0, 4000.0, 250.0]  # grid ends and a node
    got = interp3(vx, vy, vz, val, x, y, z)
    want = np.array([orc.interp3(vx, vy, vz, val, x[i], y[i], z[i]) for i in range(200)], np.float32)
    assert np.array_equal(got, want)
    assert list(locate_bissection(np.array([0.0, 1.0, 2.0]), np.array([-1.0, 0.0, 0.5, 2.0, 3.0]))) == [-1, 1, 1, 2, -1]


def test_hti_bit_exact_with_oracle():
    rng = np.random.default_rng(5)
    vp = rng.uniform(5000, 7000, 50).astype(np.float32)
    vs = rng.uniform(3000, 4000, 50).astype(np.float32)
    rho = rng.uniform(2500, 3000, 50).astype(np.float32)
    gc = rng.uniform(-0.04, 0.04, 50).astype(np.float32)
    gs = rng.uniform(-0.04, 0.04, 50).astype(np.float32)
    got = hti_cij(vp, vs, rho, gc, gs)
    for i in range(50):
        assert np.array_equal(got[i], orc.hti_cij(vp[i], vs[i], rho[i], gc[i], gs[i]))


def test_misfit_matches_oracle():
    rng = np.random.default_rng(6)
    n, dt = 1500, 0.01
    dat = rng.standard_normal(n)
    syn = dat + 0.1 * rng.standard_normal(n)
    for win, pct, it in [([[1.0, 9.0]], 0.05, 1), ([[0.5, 4.0], [6.0, 14.0]], 0.1, 2), ([[0.0, 15.0]], 0.2, 4)]:
        t1, a1, m1, r1 = waveform_adjoint_source(dat, syn, dt, win, pct, it)
        t2, a2, m2, r2 = orc.waveform_adjoint(dat, syn, dt, win, pct, it)
        assert t1 == pytest.approx(t2, rel=1e-12)
        assert np.allclose(a1, a2, rtol=1e-12, atol=1e-15)
        assert np.allclose(m1, m2) and np.allclose(r1, r2)
    for k in (2, 7, 10):
        y = rng.standard_normal(k)
        assert simpson(y, 0.3) == pytest.approx(orc.simpson(y, 0.3), rel=1e-13)
        assert np.allclose(window_taper(y, 0.4, 3), orc.window_taper(y, 0.4, 3))


def test_smoothing_constants_and_rvol():
    m = BoxMesh(16, 16, 8, 40000.0, 40000.0, 20000.0)
    cv, ch, n = smoothing_constants(5000.0, 2000.0, m.distance_min_glob())
    assert n == 604
    assert ch > cv
    _, rvol = rvol_arrays(m)
    assert np.sum(1.0 / rvol.astype(np.float64)) == pytest.approx(40000.0 * 40000.0 * 20000.0, rel=1e-6)


def test_random_field_statistics():
    # covariance ~ exp(-r^2/l^2): correlation e^-1/4 at l/2 and e^-1 at l
    c10, c20 = [], []
    for seed in range(4):
        f = gaussian_random_field((128, 128, 64), 250.0, 5000.0, 0.02, seed=seed)
        assert f.std() == pytest.approx(0.02, rel=1e-9)
        assert abs(f.mean()) < 1e-12
        c10.append(np.mean(f[:-10] * f[10:]) / f.var())
        c20.append(np.mean(f[:-20] * f[20:]) / f.var())
    assert np.mean(c10) == pytest.approx(np.exp(-0.25), abs=0.06)
    assert np.mean(c20) == pytest.approx(np.exp(-1.0), abs=0.1)


def test_model_recipe_is_seeded_and_bounded():
    m = BoxMesh(4, 4, 2, 10000.0, 10000.0, 5000.0)
    a = perturbed_model(m, seed=1)
    b = perturbed_model(m, seed=1)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert abs(a[1].mean() / 6000.0 - 1) < 0.02
    s = perturbed_model(m, smooth=True)
    assert np.all(s[1] == 6000.0)


# ------------------------------------------------------------------ C ABI
def _header_functions():
    txt = open(os.path.join(ROOT, "include", "specfwat_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    out = {}
    for m in re.finditer(r"void\s+(\w+)\s*\(([^;]*?)\);", txt, flags=re.S):
        args = [a for a in m.group(2).split(",") if a.strip()]
        out[m.group(1)] = len(args)
    return out


def test_library_exports_every_header_symbol():
    decl = _header_functions()
    assert len(decl) >= 35
    L = _lib.lib()
    for name, nargs in decl.items():
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, name
        assert len(_lib.SIGNATURES[name]) == nargs, name
    assert set(_lib.SIGNATURES) == set(decl)


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libspecfwat_hip.so")
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.lib()


def test_event_batch_refuses_unsupported_shapes():
    """Batch-size and mesh checks happen on the host before any device call."""
    from specfwat_amd.batch import EventBatch
    h = 2500.0
    m = BoxMesh(2, 2, 2, 2 * h, 2 * h, 2 * h)
    nloc = (m.nspec, 125)
    rho, vp, vs = (np.full(nloc, v, np.float32) for v in (2700, 6000, 3500))
    src = Sources(m, [[h, h, -h]], [[1e15, 0, 0, 0, 0, 0]], [1.0])
    with pytest.raises(ValueError, match="batch size"):
        EventBatch(m, rho, vp, vs, [src] * 3, None, 0.01, [10, 10, 10])
    with pytest.raises(ValueError, match="one NSTEP"):
        EventBatch(m, rho, vp, vs, [src] * 2, None, 0.01, [10])
    mi = BoxMesh(2, 2, 2, 2 * h, 2 * h, 2 * h, force_irregular=True)
    with pytest.raises(ValueError, match="regular"):
        EventBatch(mi, rho, vp, vs, [src] * 2, None, 0.01, [10, 10])


def test_measurement_rejects_windows_outside_the_trace():
    from specfwat_amd.misfit import event_adjoint_sources
    obs = np.zeros((100, 2, 3), np.float32)
    with pytest.raises(ValueError):
        event_adjoint_sources(obs, obs, 0.01, [[0.5, 2.0]])
    with pytest.raises(ValueError):
        event_adjoint_sources(obs, obs, 0.01, None, min_period=1.0)


def test_adjoint_source_dispatch_errors():
    """calculate_adjoint_source stops for the types it cannot dispatch, as the
    reference (adjoint_source.f90:30-42, 74-86); checked before any device call."""
    from specfwat_amd import misfit
    obs = np.zeros((100, 2, 3), np.float32)
    for im, msg in ((2, "WAVEFORM_CONV"), (3, "RF"), (7, "Unknown measurement type")):
        with pytest.raises(ValueError, match=msg):
            misfit.calculate_adjoint_sources(obs, obs, 0.01, [[0.1, 0.9]], 0.5, 10.0, im)
    with pytest.raises(ValueError, match="16383"):
        misfit.event_mt_adjoint_sources(np.zeros((20000, 1, 3), np.float32), np.zeros((20000, 1, 3), np.float32),
                                        0.01, [[1.0, 9.0]], 0.5, 10.0)


def test_fixed_window_types():
    """leq_data.f90:121-135: group-velocity and arrival windows, clipped to the trace."""
    from specfwat_amd import misfit
    w = misfit.fixed_windows(3, 1000, 0.05, 2.0, dist=[30.0, 100.0], groupvel=(2.0, 4.0), long_p=10.0)
    np.testing.assert_allclose(w, [[30 / 4 - 5, 30 / 2 + 5], [100 / 4 - 5, min(100 / 2 + 5, 998 * 0.05 - 2.0)]])
    w = misfit.fixed_windows(4, 1000, 0.05, 2.0, tp=[1.0, 40.0], time_win=(-5.0, 20.0))
    np.testing.assert_allclose(w, [[-2.0, 21.0], [35.0, 47.9]])
    with pytest.raises(ValueError):
        misfit.fixed_windows(1, 1000, 0.05, 2.0)
