"""Solver databases (proc%06d_external_mesh.bin): record framing, the record
order of read_mesh_databases.f90, and every stored array checked by what it
means (geometry from the coordinates, masses from the FP64 oracle, moduli from
the velocities, boundary faces by their areas) rather than by the writer that
made the file."""
import os
import struct

import numpy as np
import pytest

from oracle import oracle as orc
from specfwat_amd import databases as D
from specfwat_amd.gll import GLL
from specfwat_amd.mesh import BoxMesh
from specfwat_amd.model import hti_model
from specfwat_amd.sources import Receivers, Sources

NEX, NEY, NEZ, H = 4, 4, 3, 2500.0


def _case(perturb=0.12, seed=0, aniso=False):
    m = BoxMesh(NEX, NEY, NEZ, NEX * H, NEY * H, NEZ * H, perturb=perturb)
    rng = np.random.default_rng(seed)
    f = (1 + 0.03 * rng.standard_normal(m.nglob))[m.ibool0]
    rho, vp, vs = (2700 * f).astype(np.float32), (6000 * f).astype(np.float32), (3500 * f).astype(np.float32)
    cij = hti_model(m, rho, vp, vs, seed=seed + 5, block=2 * H) if aniso else None
    return m, rho, vp, vs, cij


@pytest.fixture(scope="module")
def written(tmp_path_factory):
    m, rho, vp, vs, _ = _case()
    p = str(tmp_path_factory.mktemp("db") / "proc000000_external_mesh.bin")
    D.write_external_mesh(p, m, rho, vp, vs, 0.02)
    return p, m, rho, vp, vs


def test_record_framing_subrecords_and_byte_order(tmp_path):
    """Markers around every record; a record split into gfortran subrecords
    (negative leading marker while it continues) reads as one; a big-endian
    file (every 4-byte word swapped) reads the same."""
    a = np.arange(10, dtype=np.int32)
    p = tmp_path / "r.bin"
    with open(p, "wb") as f:
        f.write(struct.pack("<i", 4) + struct.pack("<i", 7) + struct.pack("<i", 4))
        f.write(struct.pack("<i", -16) + a[:4].tobytes() + struct.pack("<i", 16))
        f.write(struct.pack("<i", 24) + a[4:].tobytes() + struct.pack("<i", -24))
    r = D._Records(str(p))
    assert r.scalar(np.int32, "n") == 7
    assert np.array_equal(r.ints("a", 10), a)
    r.close()
    be = tmp_path / "be.bin"
    np.fromfile(p, "<i4").astype(">i4").tofile(be)
    r = D._Records(str(be))
    assert r.bo == ">" and r.scalar(np.int32, "n") == 7 and np.array_equal(r.ints("a", 10), a)
    r.close()


def test_round_trip_and_big_endian(written, tmp_path):
    p, m, rho, vp, vs = written
    d = D.read_external_mesh(p)
    dm = D.DatabaseMesh(d)
    assert (d["NSPEC_AB"], d["NGLOB_AB"], d["NSPEC_IRREGULAR"]) == (m.nspec, m.nglob, m.nspec_irregular)
    for nm in ("ibool", "irregular_element_number", "xix", "xiy", "xiz", "etax", "etay", "etaz",
               "gammax", "gammay", "gammaz", "jacobian", "abs_boundary_ispec", "abs_boundary_ijk",
               "abs_boundary_normal", "abs_boundary_jacobian2Dw", "phase_ispec_inner", "xstore",
               "ystore", "zstore", "jacobian_all"):
        assert np.array_equal(getattr(dm, nm), getattr(m, nm)), nm
    assert dm.xix_regular == m.xix_regular and dm.jacobian_regular == m.jacobian_regular
    assert (dm.nspec_inner, dm.nspec_outer, dm.num_phase_ispec) == (m.nspec_inner, m.nspec_outer, m.num_phase_ispec)
    assert dm.stacey and d["records"] == 74
    assert np.array_equal(np.sort(np.unique(dm.ibool)), np.arange(1, m.nglob + 1))
    # all 4-byte words: swapping the whole file gives the big-endian database
    be = tmp_path / "be.bin"
    np.fromfile(p, "<i4").astype(">i4").tofile(be)
    d2 = D.read_external_mesh(str(be))
    for k, v in d.items():
        assert np.array_equal(np.asarray(v), np.asarray(d2[k])), k


def test_geometry_records_are_the_inverse_jacobian_of_the_coordinates(written):
    """xix..gammaz and jacobianstore against J = sum_l l_l'(xi) x_l computed from
    the stored coordinates (any two of the nine records swapped fails), and
    against the FP64 oracle's own mesh."""
    p, m, *_ = written
    dm = D.load_mesh(p)
    g = GLL()
    X = dm.xyz[dm.ibool0].reshape(dm.nspec, 5, 5, 5, 3)  # [e, k, j, i, c]
    Dm = g.hprime64
    dxi = np.einsum("il,ekjlc->ekjic", Dm, X)
    deta = np.einsum("jl,eklic->ekjic", Dm, X)
    dgam = np.einsum("kl,eljic->ekjic", Dm, X)
    J = np.stack([dxi, deta, dgam], axis=-1).reshape(dm.nspec, 125, 3, 3)  # [c][d] = dx_c/dref_d
    Ji = np.linalg.inv(J)  # [d][c] = dref_d/dx_c
    irr = dm.irregular_element_number
    assert (irr > 0).all()
    sel = irr - 1
    names = ["xix", "xiy", "xiz", "etax", "etay", "etaz", "gammax", "gammay", "gammaz"]
    for q, nm in enumerate(names):
        want = Ji[:, :, q // 3, q % 3]
        got = getattr(dm, nm)[sel]
        assert np.abs(got - want).max() < 2e-5 * np.abs(Ji).max(), nm
    det = np.linalg.det(J)
    assert np.abs(dm.jacobian[sel] / det - 1).max() < 2e-5
    o = orc.Oracle(NEX, NEY, NEZ, NEX * H, NEY * H, NEZ * H, np.ones(m.nspec * 125, np.float32),
                   np.ones(m.nspec * 125, np.float32), np.ones(m.nspec * 125, np.float32), dt=0.02,
                   nstep=2, src_xyz=[[1.0, 1.0, -1.0]], src_mt=[[0] * 6], src_f0=[1.0],
                   rec_xyz=[[1.0, 1.0, 0.0]], perturb=0.12)
    om = o.mesh()
    assert np.array_equal(om["ibool"] + 1, dm.ibool.ravel())
    jinv = om["jinv"].reshape(dm.nspec, 125, 9)
    for q, nm in enumerate(names):
        assert np.abs(getattr(dm, nm)[sel] - jinv[..., q]).max() < 1e-6 * np.abs(jinv).max(), nm
    assert np.abs(dm.xyz - om["xyz"]).max() < 1e-3


def test_mass_and_material_records(written):
    """rmass = sum rho J w; rmassx/y/z = the FP64 oracle's Stacey-augmented masses;
    rho_vp^2 = rho (kappa + 4/3 mu), rho_vs^2 = rho mu; velocities recovered."""
    p, m, rho, vp, vs = written
    dm = D.load_mesh(p)
    d = dm.db
    g = GLL()
    ii = np.arange(125)
    w3 = g.w[ii % 5] * g.w[(ii // 5) % 5] * g.w[ii // 25]
    want = np.bincount(dm.ibool0.ravel(), (dm.jacobian_all.astype(np.float64) * w3 *
                                            d["rhostore"].astype(np.float64)).ravel(), dm.nglob)
    assert np.abs(d["rmass"] / want - 1).max() < 2e-6
    o = orc.Oracle(NEX, NEY, NEZ, NEX * H, NEY * H, NEZ * H, rho.ravel(), vp.ravel(), vs.ravel(),
                   dt=0.02, nstep=2, src_xyz=[[1.0, 1.0, -1.0]], src_mt=[[0] * 6], src_f0=[1.0],
                   rec_xyz=[[1.0, 1.0, 0.0]], perturb=0.12)
    for c, inv in zip("xyz", o.rmass_inv()):
        assert np.abs(d["rmass" + c] * inv.astype(np.float64) - 1).max() < 2e-6, c
    assert not np.allclose(d["rmassx"], d["rmassz"])  # the Stacey terms differ by component
    r64 = d["rhostore"].astype(np.float64)
    assert np.abs(d["rho_vp"].astype(np.float64) ** 2 / (r64 * (d["kappastore"] + 4.0 / 3.0 * d["mustore"])) - 1).max() < 1e-5
    assert np.abs(d["rho_vs"].astype(np.float64) ** 2 / (r64 * d["mustore"]) - 1).max() < 1e-5
    mat = dm.materials()
    assert np.abs(mat["rho_vp"].astype(np.float64) ** 2 / (r64 * (mat["kappa"] + 4.0 / 3.0 * mat["mu"])) - 1).max() < 1e-5
    assert np.abs(mat["rho_vs"].astype(np.float64) ** 2 / (r64 * mat["mu"]) - 1).max() < 1e-5
    assert np.array_equal(mat["rho"], rho)
    for c, inv in zip("xyz", o.rmass_inv()):
        assert np.abs(mat["rmass" + c] * inv.astype(np.float64) - 1).max() < 2e-6, c
    r2, vp2, vs2 = dm.model()
    assert np.array_equal(r2, rho) and np.abs(vp2 / vp - 1).max() < 1e-6 and np.abs(vs2 / vs - 1).max() < 1e-6


def test_boundary_face_records(written):
    """Absorbing and free-surface faces: unit outward normals; the weighted face
    Jacobians of each flat side sum to that side's area; ibelm lists name the
    elements of each side; surface flags cover the box's outer points."""
    p, m, *_ = written
    d = D.read_external_mesh(p)
    lx, ly, lz = NEX * H, NEY * H, NEZ * H
    for pre, areas in (("abs_boundary", [ly * lz, ly * lz, lx * lz, lx * lz, lx * ly]),
                       ("free_surface", [lx * ly])):
        n = d[pre + "_normal"].astype(np.float64)
        assert np.abs(np.linalg.norm(n, axis=-1) - 1).max() < 1e-6
        jw = d[pre + "_jacobian2Dw"].astype(np.float64)
        # group faces by their dominant outward normal
        axis = np.argmax(np.abs(n.mean(axis=1)), axis=1)
        sign = np.sign(n.mean(axis=1)[np.arange(len(axis)), axis])
        got = []
        for ax, sg in ((0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1)):
            s = (axis == ax) & (sign == sg)
            if s.any():
                got.append(jw[s].sum())
        assert np.allclose(got, areas, rtol=1e-5), (pre, got)
    assert np.array_equal(d["ibelm_top"], d["free_surface_ispec"])
    assert d["nspec2D_bottom"] == NEX * NEY and d["nspec2D_xmin"] == NEY * NEZ
    onb = d["iglob_is_surface_external_mesh"]
    x, y, z = d["xstore"], d["ystore"], d["zstore"]
    tol = 1e-3
    outer = ((np.abs(x) < tol) | (np.abs(x - lx) < tol) | (np.abs(y) < tol) | (np.abs(y - ly) < tol)
             | (np.abs(z + lz) < tol) | (np.abs(z) < tol))
    assert np.array_equal(onb, outer)
    xadj, adj = d["neighbors_xadj"], d["neighbors_adjncy"]
    assert xadj[0] == 0 and xadj[-1] == len(adj) and xadj[1] - xadj[0] == 7  # corner element


def test_sources_and_receivers_located_on_the_database_mesh(written):
    p, m, *_ = written
    dm = D.load_mesh(p)
    rng = np.random.default_rng(4)
    sx = [[4300.0, 5700.0, -3375.0]]
    mt = [list(1e15 * rng.standard_normal(6))]
    rx = [[1000.0, 2000.0, 0.0], [7000.0, 3000.0, 0.0], [5100.0, 4900.0, -10.0]]
    s1, s2 = Sources(m, sx, mt, [1.0]), Sources(dm, sx, mt, [1.0])
    r1, r2 = Receivers(m, rx), Receivers(dm, rx)
    assert np.array_equal(s1.ispec, s2.ispec) and np.array_equal(r1.ispec, r2.ispec)
    assert np.abs(s1.sourcearrays - s2.sourcearrays).max() < 1e-5 * np.abs(s1.sourcearrays).max()
    assert np.abs(r1.hlagrange - r2.hlagrange).max() < 1e-5
    with pytest.raises(ValueError):
        dm.locate(-100.0, 10.0, -10.0)


def test_anisotropic_records_and_flags(tmp_path):
    m, rho, vp, vs, cij = _case(aniso=True, seed=2)
    p = str(tmp_path / "a.bin")
    D.write_external_mesh(p, m, rho, vp, vs, 0.02, cij=cij)
    with pytest.raises(D.DatabaseError, match="ispec_is_inner"):
        D.read_external_mesh(p)  # ANISOTROPY off: the first c_ij record is taken for ispec_is_inner
    dm = D.load_mesh(p, anisotropy=True)
    assert np.array_equal(dm.cij(), np.asarray(cij, np.float32).reshape(m.nspec, 125, 21))


def test_corrupt_files_are_refused(written, tmp_path):
    p = written[0]
    raw = bytearray(open(p, "rb").read())
    cut = tmp_path / "cut.bin"
    cut.write_bytes(bytes(raw[: len(raw) // 2]))
    with pytest.raises(D.DatabaseError, match="truncated|ends before"):
        D.read_external_mesh(str(cut))
    # the 9999 sentinel: find it after the three domain-flag records
    words = np.frombuffer(bytes(raw), "<i4").copy()
    pos = int(np.nonzero(words == 9999)[0][0])
    words[pos] = 1234
    bad = tmp_path / "bad.bin"
    words.tofile(bad)
    with pytest.raises(D.DatabaseError, match="position 1"):
        D.read_external_mesh(str(bad))
    words[pos] = 9999
    words[0] = 8  # first marker
    words.tofile(bad)
    with pytest.raises(D.DatabaseError):
        D.read_external_mesh(str(bad))
    with pytest.raises(D.DatabaseError):
        D.read_external_mesh(p, stacey_absorbing_conditions=False)  # rmassx taken for the sentinel


def test_ocean_load_and_colouring_records(tmp_path):
    """The optional APPROXIMATE_OCEAN_LOAD and USE_MESH_COLORING_GPU records sit
    where read_mesh_databases.f90:461-475 and :1186-1198 read them; reading with
    the switches off takes the next record for the wrong array and fails."""
    m, rho, vp, vs, _ = _case()
    ocean = np.linspace(1.0, 2.0, m.nglob).astype(np.float32)
    colors = (1, 3, np.array([5, 20, 15, 8], np.int32))
    p = str(tmp_path / "o.bin")
    D.write_external_mesh(p, m, rho, vp, vs, 0.02, rmass_ocean_load=ocean, elastic_colors=colors)
    d = D.read_external_mesh(p, approximate_ocean_load=True, use_mesh_coloring_gpu=True)
    assert np.array_equal(d["rmass_ocean_load"], ocean)
    assert (d["num_colors_outer_elastic"], d["num_colors_inner_elastic"]) == (1, 3)
    assert np.array_equal(d["num_elem_colors_elastic"], colors[2])
    assert d["records"] == 74 + 3
    assert np.array_equal(d["rho_vp"], D.read_external_mesh(str(tmp_path / "o.bin"), approximate_ocean_load=True,
                                                           use_mesh_coloring_gpu=True)["rho_vp"])
    plain = str(tmp_path / "p.bin")
    D.write_external_mesh(plain, m, rho, vp, vs, 0.02)
    assert np.array_equal(D.read_external_mesh(plain)["rho_vp"], d["rho_vp"])
    with pytest.raises(D.DatabaseError):
        D.read_external_mesh(p, use_mesh_coloring_gpu=True)  # ocean record taken for rho_vp
    with pytest.raises(D.DatabaseError):
        D.read_external_mesh(p, approximate_ocean_load=True)  # colour counts taken for the surface count


def test_engine_refuses_what_it_cannot_run(written):
    """Acoustic or poroelastic elements, C-PML layers and one slice of a
    partitioned mesh are read but refused by the engine's mesh adaptor."""
    d = D.read_external_mesh(written[0])
    D.DatabaseMesh(dict(d))  # the elastic single-slice file itself is accepted
    for key, val, msg in (("ACOUSTIC_SIMULATION", True, "purely elastic"),
                          ("POROELASTIC_SIMULATION", True, "purely elastic"),
                          ("NSPEC_CPML", 12, "C-PML"),
                          ("num_interfaces_ext_mesh", 2, "partitioned")):
        bad = dict(d)
        bad[key] = val
        with pytest.raises(D.DatabaseError, match=msg):
            D.DatabaseMesh(bad)


def _write_slice(path, d):
    """Write a record dict (read_external_mesh's keys) of an elastic slice with
    Stacey and free-surface faces and MPI interfaces, in the reader's order."""
    w = D._Writer(path)
    n, g = d["NSPEC_AB"], d["NGLOB_AB"]
    w.ints(n); w.ints(g); w.ints(d["NSPEC_IRREGULAR"])
    w.rec(d["ibool"], np.int32)
    for c in "xyz":
        w.rec(d[c + "store"], np.float32)
    w.rec(d["irregular_element_number"], np.int32)
    w.rec([d["xix_regular"]], np.float32); w.rec([d["jacobian_regular"]], np.float32)
    for nm in D._JNAMES + ["jacobian"]:
        w.rec(d[nm + "store"], np.float32)
    for nm in ("kappastore", "mustore", "rhostore"):
        w.rec(d[nm], np.float32)
    w.rec(np.zeros(n), np.int32); w.rec(np.ones(n), np.int32); w.rec(np.zeros(n), np.int32)
    w.ints(9999)
    for nm in ("rmass", "rho_vp", "rho_vs"):
        w.rec(d[nm], np.float32)
    w.ints(d["num_abs_boundary_faces"])
    if d["num_abs_boundary_faces"]:
        for suf, dt in (("_ispec", np.int32), ("_ijk", np.int32), ("_jacobian2Dw", np.float32), ("_normal", np.float32)):
            w.rec(d["abs_boundary" + suf], dt)
        for c in "xyz":
            w.rec(d["rmass" + c], np.float32)
    w.ints(9998)
    for sname in ("xmin", "xmax", "ymin", "ymax", "bottom", "top"):
        w.ints(0)
    w.ints(d["num_free_surface_faces"])
    if d["num_free_surface_faces"]:
        for suf, dt in (("_ispec", np.int32), ("_ijk", np.int32), ("_jacobian2Dw", np.float32), ("_normal", np.float32)):
            w.rec(d["free_surface" + suf], dt)
    w.ints(0); w.ints(0); w.ints(0)
    w.ints(9997)
    ni = d["num_interfaces_ext_mesh"]
    w.ints(ni); w.ints(d["max_nibool_interfaces_ext_mesh"])
    if ni:
        w.rec(d["my_neighbors_ext_mesh"], np.int32)
        w.rec(d["nibool_interfaces_ext_mesh"], np.int32)
        w.rec(d["ibool_interfaces_ext_mesh"], np.int32)
    w.ints(9996)
    w.rec(np.ones(n), np.int32)
    w.ints(n, 0); w.ints(n)
    ph = np.zeros((2, n), np.int32); ph[1] = np.arange(1, n + 1)
    w.rec(ph, np.int32)
    w.ints(0); w.rec(np.zeros(n), np.int32); w.rec(np.zeros(g), np.int32)
    w.ints(0); w.rec(np.zeros(n + 1), np.int32); w.rec(np.zeros(0), np.int32)
    w.ints(9995)
    w.close()


def split_database(p, m, rho, cuts, tmp_path):
    """Write the slices of the database at p cut at element indices cuts: each
    slice numbers its own points, stores its own unassembled masses (rmass and
    the Stacey terms of its own faces) and lists the points it shares with each
    neighbour in one common order (ascending whole-mesh number).  Returns the
    slice paths in rank order."""
    full = D.read_external_mesh(p)
    g = GLL()
    ii = np.arange(125)
    w3 = g.w[ii % 5] * g.w[(ii // 5) % 5] * g.w[ii // 25]
    dm = D.DatabaseMesh(full)
    elem_mass = dm.jacobian_all.astype(np.float64) * w3 * full["rhostore"].astype(np.float64)
    ib0 = full["ibool"] - 1
    owner_pts = [np.unique(ib0[cuts[r]:cuts[r + 1]]) for r in range(len(cuts) - 1)]
    # the Stacey part of rmassx/y/z per face point: rmassx - rmass of the whole mesh, split by face element
    fe = full["abs_boundary_ispec"] - 1
    fl = (full["abs_boundary_ijk"][..., 0] - 1) + 5 * (full["abs_boundary_ijk"][..., 1] - 1) + 25 * (full["abs_boundary_ijk"][..., 2] - 1)
    from specfwat_amd.solver import mass_sums
    _, rmx, rmy, rmz = mass_sums(m, rho.reshape(m.nspec, 125), full["rho_vp"], full["rho_vs"], 0.02, True)
    paths = []
    for r in range(len(cuts) - 1):
        e0, e1 = cuts[r], cuts[r + 1]
        pts = owner_pts[r]  # global 0-based points of this slice, local number = position
        loc = np.full(full["NGLOB_AB"], -1)
        loc[pts] = np.arange(len(pts))
        d = {"NSPEC_AB": e1 - e0, "NGLOB_AB": len(pts), "xix_regular": full["xix_regular"],
             "jacobian_regular": full["jacobian_regular"]}
        d["ibool"] = (loc[ib0[e0:e1]] + 1).astype(np.int32)
        for c in "xyz":
            d[c + "store"] = full[c + "store"][pts]
        irr = full["irregular_element_number"][e0:e1]
        sel = irr > 0
        d["NSPEC_IRREGULAR"] = int(sel.sum())
        d["irregular_element_number"] = np.where(sel, np.cumsum(sel), 0).astype(np.int32)
        for nm in D._JNAMES + ["jacobian"]:
            d[nm + "store"] = full[nm + "store"][irr[sel] - 1]
        for nm in ("kappastore", "mustore", "rhostore", "rho_vp", "rho_vs"):
            d[nm] = full[nm][e0:e1]
        d["rmass"] = np.bincount(loc[ib0[e0:e1]].ravel(), elem_mass[e0:e1].ravel(), len(pts)).astype(np.float32)
        fs = (fe >= e0) & (fe < e1)
        d["num_abs_boundary_faces"] = int(fs.sum())
        d["abs_boundary_ispec"] = (fe[fs] - e0 + 1).astype(np.int32)
        for suf in ("_ijk", "_jacobian2Dw", "_normal"):
            d["abs_boundary" + suf] = full["abs_boundary" + suf][fs]
        gface = ib0[fe[fs][:, None], fl[fs]]
        for c, whole in zip("xyz", (rmx, rmy, rmz)):
            # this slice's rmass plus the Stacey terms of its own faces (from the whole mesh's split)
            stacey = np.zeros(full["NGLOB_AB"])
            d["rmass" + c] = d["rmass"].astype(np.float64)
            term = (whole - np.bincount(ib0.ravel(), elem_mass.ravel(), full["NGLOB_AB"]))
            # attribute each point's Stacey term to the slice of the first face that holds it
            firstface = np.full(full["NGLOB_AB"], 10**9)
            np.minimum.at(firstface, ib0[fe[:, None], fl].ravel(), np.repeat(fe, 25))
            mine = np.zeros(full["NGLOB_AB"], bool)
            mine[gface.ravel()] = True
            mine &= (firstface >= e0) & (firstface < e1)
            stacey[mine] = term[mine]
            d["rmass" + c] = (d["rmass" + c] + stacey[pts]).astype(np.float32)
        ft = (full["free_surface_ispec"] - 1 >= e0) & (full["free_surface_ispec"] - 1 < e1)
        d["num_free_surface_faces"] = int(ft.sum())
        d["free_surface_ispec"] = (full["free_surface_ispec"][ft] - e0).astype(np.int32)
        for suf in ("_ijk", "_jacobian2Dw", "_normal"):
            d["free_surface" + suf] = full["free_surface" + suf][ft]
        nbrs, lists = [], []
        for q in range(len(cuts) - 1):
            if q == r:
                continue
            shared = np.intersect1d(pts, owner_pts[q])  # ascending global number: the same order on both sides
            if len(shared):
                nbrs.append(q)
                lists.append(loc[shared] + 1)
        mx = max(len(a) for a in lists)
        d["num_interfaces_ext_mesh"] = len(nbrs)
        d["max_nibool_interfaces_ext_mesh"] = mx
        d["my_neighbors_ext_mesh"] = np.array(nbrs, np.int32)
        d["nibool_interfaces_ext_mesh"] = np.array([len(a) for a in lists], np.int32)
        ibi = np.zeros((len(nbrs), mx), np.int32)
        for k, a in enumerate(lists):
            ibi[k, :len(a)] = a
        d["ibool_interfaces_ext_mesh"] = ibi
        path = str(tmp_path / f"proc{r:06d}_external_mesh.bin")
        _write_slice(path, d)
        paths.append(path)
    return paths


def test_partitioned_mesh_merges_back_to_the_whole(written, tmp_path):
    """Three slices cut by element ranges, written, read and merged: the same
    elements, points, geometry, faces and assembled masses as the whole mesh."""
    p, m, rho, vp, vs = written
    full = D.read_external_mesh(p)
    dm = D.DatabaseMesh(full)
    ib0 = full["ibool"] - 1
    paths = split_database(p, m, rho, [0, 17, 31, m.nspec], tmp_path)
    slices = [D.read_external_mesh(q) for q in paths]
    assert all(s["num_interfaces_ext_mesh"] == 2 for s in slices)
    with pytest.raises(D.DatabaseError, match="partitioned"):
        D.DatabaseMesh(slices[1])
    mm = D.load_partitioned_mesh(paths)
    assert (mm.nspec, mm.nglob, mm.nspec_irregular) == (m.nspec, m.nglob, m.nspec_irregular)
    # a bijection between merged and whole-mesh points, consistent on every local point
    pi = np.full(m.nglob, -1)
    pi[mm.ibool0.ravel()] = ib0.ravel()
    assert np.array_equal(pi[mm.ibool0], ib0) and len(np.unique(pi)) == m.nglob
    assert np.array_equal(mm.xyz, dm.xyz[pi])
    for nm in ("xix", "etay", "gammaz", "jacobian", "jacobian_all"):
        assert np.array_equal(getattr(mm, nm), getattr(dm, nm)), nm
    # absorbing faces come slice by slice instead of side by side: the same set
    def faces(x):
        key = np.lexsort((x.abs_boundary_ijk[:, 12, 2], x.abs_boundary_ijk[:, 12, 1], x.abs_boundary_ijk[:, 12, 0],
                          x.abs_boundary_ispec))
        return [x.abs_boundary_ispec[key], x.abs_boundary_ijk[key], x.abs_boundary_normal[key],
                x.abs_boundary_jacobian2Dw[key]]
    for a, b in zip(faces(mm), faces(dm)):
        assert np.array_equal(a, b)
    mat, whole = mm.materials(), dm.materials()
    for nm in ("rho", "kappa", "mu", "rho_vp", "rho_vs"):
        assert np.array_equal(mat[nm], whole[nm]), nm
    for c in "xyz":
        assert np.abs(mat["rmass" + c] / whole["rmass" + c][pi] - 1).max() < 2e-6, c
