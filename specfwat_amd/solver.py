"""Host driver of the GPU solver: the run sequence of SpecFWAT's PrepareFWD.

Mirrors src/preproc/prepare_fwd.f90:
  run_simulation(run_opt)   :328-368  InitSpecfem -> setup_sources_receivers ->
                                      restore_rmass -> prepare_timerun (prepare_GPU) ->
                                      iterate_time -> FinalizeSpecfem
  simulation()              :297-326  forward (SIMULATION_TYPE 1, SAVE_FORWARD) ->
                                      measure_adj -> adjoint (SIMULATION_TYPE 3) ->
                                      postproc_adjoint (save_adjoint_kernels)
All device work goes through the C ABI of libspecfwat_hip.so; the forward
state and the Stacey store are handed to the adjoint run inside HBM instead
of through save_forward_arrays.bin / absorb_field.bin.
"""
import numpy as np

from . import _lib
from .model import moduli, CIJ_NAMES

F32 = np.float32


def mass_matrices(mesh, rho, rho_vp, rho_vs, dt, stacey=True):
    """Inverted rmassx/y/z (float32): lumped mass plus the dt/2 Stacey term."""
    _, rmx, rmy, rmz = mass_sums(mesh, rho, rho_vp, rho_vs, dt, stacey)
    return (1.0 / rmx).astype(F32), (1.0 / rmy).astype(F32), (1.0 / rmz).astype(F32)


def mass_sums(mesh, rho, rho_vp, rho_vs, dt, stacey=True):
    """float64 rmass and rmassx/y/z (rmass plus the dt/2 Stacey term on absorbing faces)."""
    w = mesh.local_weights() * rho.astype(np.float64)
    rm = np.bincount(mesh.ibool0.ravel(), weights=w.ravel(), minlength=mesh.nglob)
    rmx, rmy, rmz = rm.copy(), rm.copy(), rm.copy()
    if stacey and mesh.num_abs_boundary_faces:
        e = mesh.abs_boundary_ispec - 1
        ijk = mesh.abs_boundary_ijk - 1
        L = ijk[..., 0] + 5 * ijk[..., 1] + 25 * ijk[..., 2]
        g = mesh.ibool0[e[:, None], L]
        rvp = rho_vp[e[:, None], L].astype(np.float64)
        rvs = rho_vs[e[:, None], L].astype(np.float64)
        n = mesh.abs_boundary_normal.astype(np.float64)
        jw = mesh.abs_boundary_jacobian2Dw.astype(np.float64)
        dto2 = 0.5 * dt
        vn = dto2 * n.sum(axis=-1)
        for comp, arr in enumerate((rmx, rmy, rmz)):
            t = rvp * vn * n[..., comp] + rvs * (dto2 - vn * n[..., comp])
            np.add.at(arr, g.ravel(), (t * jw).ravel())
    return rm, rmx, rmy, rmz


class ElasticRun:
    """One solver run (SIMULATION_TYPE 1 or 3) held on the GPU."""

    def __init__(self, mesh, rho, vp, vs, sources, receivers, dt, nstep, simulation_type=1,
                 save_forward=False, cij=None, stacey=True, compute_and_store_strain=False,
                 approximate_hess=False, static_precision=0, materials=None):
        """materials: the stored arrays of a solver database (rho, kappa, mu, rho_vp, rho_vs
        per local point, rmassx/y/z per point, not inverted); when given they replace the
        ones derived from rho, vp and vs, as restore_rmass (specfem_api.f90:393-411) and
        read_mesh_databases (read_mesh_databases.f90:183-185, 455-487, 802-804) do."""
        self.mesh, self.dt, self.nstep = mesh, float(dt), int(nstep)
        self.stacey = bool(stacey)
        self.simulation_type = simulation_type
        self.aniso = cij is not None
        self.nrec = receivers.nrec if receivers is not None else 0
        self.nsrc = sources.nsources if sources is not None else 0
        loc = lambda a: np.ascontiguousarray(a, F32).reshape(mesh.nspec, 125)
        if materials is not None:
            rho, kappa, mu = loc(materials["rho"]), loc(materials["kappa"]), loc(materials["mu"])
            rho_vp, rho_vs = loc(materials["rho_vp"]), loc(materials["rho_vs"])
            rmx, rmy, rmz = (F32(1.0) / np.ascontiguousarray(materials[k], F32)
                             for k in ("rmassx", "rmassy", "rmassz"))
        else:
            rho = loc(rho)
            kappa, mu, rho_vp, rho_vs = moduli(rho, loc(vp), loc(vs))
            rmx, rmy, rmz = mass_matrices(mesh, rho, rho_vp, rho_vs, self.dt, stacey)
        self.rho, self.kappa, self.mu = rho, kappa, mu
        self.h = _lib.HandleBox()
        g = mesh.gll
        i32 = lambda a: np.ascontiguousarray(a, np.int32)
        f32 = lambda a: np.ascontiguousarray(a, F32)
        src_arr = f32(sources.sourcearrays) if self.nsrc else np.zeros((1, 125, 3), F32)
        src_isp = i32(sources.ispec) if self.nsrc else np.ones(1, np.int32)
        rec_isp = i32(receivers.ispec) if self.nrec else np.ones(1, np.int32)
        rec_h = f32(receivers.hlagrange) if self.nrec else np.zeros((1, 125), F32)
        nf = mesh.num_abs_boundary_faces if stacey else 0
        _lib.call("prepare_constants_device_", self.h, 5, mesh.nspec, mesh.nglob, mesh.nspec_irregular,
                  mesh.xix, mesh.xiy, mesh.xiz, mesh.etax, mesh.etay, mesh.etaz, mesh.gammax,
                  mesh.gammay, mesh.gammaz, mesh.jacobian, float(mesh.xix_regular),
                  float(mesh.jacobian_regular), i32(mesh.ibool), i32(mesh.irregular_element_number),
                  g.hprime_xx, g.hprimewgll_xx, g.wgllwgll_xy, g.wgllwgll_xz, g.wgllwgll_yz,
                  int(stacey), nf, i32(mesh.abs_boundary_ispec), i32(mesh.abs_boundary_ijk),
                  f32(mesh.abs_boundary_normal), f32(mesh.abs_boundary_jacobian2Dw), self.nsrc,
                  src_arr, src_isp, self.nrec, rec_isp, rec_h, self.nstep, simulation_type,
                  int(save_forward))
        if self.aniso:
            cij = np.ascontiguousarray(cij, F32).reshape(mesh.nspec, 125, 21)
            cs = [np.ascontiguousarray(cij[..., q]) for q in range(21)]
        else:
            cs = [None] * 21
        _lib.call("prepare_fields_elastic_device_", self.h, rmx, rmy, rmz, rho_vp, rho_vs, kappa, mu,
                  int(self.aniso), *cs, mesh.num_phase_ispec, i32(mesh.phase_ispec_inner),
                  mesh.nspec_outer, mesh.nspec_inner, int(compute_and_store_strain))
        if simulation_type == 3:
            _lib.call("prepare_fields_elastic_adj_dev_", self.h, 3 * mesh.nglob,
                      int(compute_and_store_strain), rho, int(self.aniso), int(approximate_hess))
            _lib.call("prepare_sim2_or_3_const_device_", self.h, self.nrec, self.nstep)
        if static_precision:
            _lib.call("fwat_set_static_precision_", self.h, int(static_precision))
        # a box mesh of regular elements is an element lattice: declare it (checked on the
        # device side) so that the brick-fused pass can be switched on
        self.lattice = (mesh.nspec_irregular == 0 and hasattr(mesh, "nex")
                        and mesh.nex * mesh.ney * mesh.nez == mesh.nspec)
        if self.lattice:
            _lib.call("fwat_set_lattice_", self.h, mesh.nex, mesh.ney, mesh.nez)
        if self.nsrc:
            self.stf = sources.stf(self.nstep, self.dt)
            _lib.call("fwat_set_stf_", self.h, self.stf, self.nstep)

    def set_tuning(self, regular_path=1, blocks_per_cu_fwd=0, blocks_per_cu_adj=8, brick=(0, 0, 0),
                   brick_blocks_per_cu=0, slab_elems=0, interior=True):
        """Element-pass launch shape; brick=(bx, by, bz) switches the brick-fused
        element + assembly pass on (lattice meshes of regular elements only);
        slab_elems > 0 runs element pass + assembly slab after slab; interior: the
        pipelined passes finish the element-interior points (set_interior)."""
        _lib.call("fwat_set_tuning_", self.h, int(regular_path), int(blocks_per_cu_fwd), int(blocks_per_cu_adj))
        _lib.call("fwat_set_interior_", self.h, int(interior))
        _lib.call("fwat_set_slabs_", self.h, int(slab_elems))
        if brick[0] and not self.lattice:
            raise ValueError("the brick pass needs a lattice mesh of regular elements")
        if brick[0] or self.lattice:
            _lib.call("fwat_set_brick_", self.h, int(brick[0]), int(brick[1]), int(brick[2]),
                      int(brick_blocks_per_cu))

    def set_assembly_lanes(self, points_per_lane):
        """Points per lane of the gather assembly (1 or 2; the same sums), 0 for
        the default (2 for the one-field pass, 1 for the two-field pass)."""
        _lib.call("fwat_set_assembly_lanes_", self.h, int(points_per_lane))

    def set_assembly_block(self, lanes):
        """Lanes per workgroup of the gather assembly (64..1024, default 256)."""
        _lib.call("fwat_set_assembly_block_", self.h, int(lanes))

    def set_floc_layout(self, grouped=True):
        """Element-local force rows grouped by reader (default) or in local-point order."""
        _lib.call("fwat_set_floc_layout_", self.h, int(bool(grouped)))

    def set_interior(self, on=True):
        """Element-interior points finished by the element pass of the fused loops
        (pipelined passes 3 and 5; the same results bit for bit) or assembled by the gather."""
        _lib.call("fwat_set_interior_", self.h, int(on))

    def interior_info(self):
        """(points the gather assembles, element-interior points, forward pass finishes them,
        two-field pass finishes them) with the current launch shape."""
        out = [np.zeros(1, np.int32) for _ in range(4)]
        _lib.call("fwat_interior_info_", self.h, *out)
        return tuple(int(o[0]) for o in out)

    def set_graph(self, on=True):
        """Capture every whole-run loop of this handle into one HIP graph and launch
        it as a unit (fused gather path; captured runs are not timed)."""
        _lib.call("fwat_set_graph_", self.h, int(bool(on)))

    def set_kernel_block(self, nsteps):
        """Steps per fp32 partial-sum block of the mu/kappa (c_ij) kernels; 1 = fp64 every step."""
        _lib.call("fwat_set_kernel_block_", self.h, int(nsteps))

    def brick_info(self):
        """(halo points, partial-sum slots per field and step, halo CSR length) of the brick pass."""
        nh, ns, nc = _lib.call("fwat_brick_info_", self.h, 0, 0, 0)
        return int(nh), int(ns), int(nc)

    # ------------------------------------------------------------ runs
    def set_fk_injection(self, veloc, tract, resamp=1):
        """Incident wavefield injected through the Stacey faces of a forward run
        (teleseismic FK coupling, fk_coupling.f90:71-297): veloc, tract
        (nsamp, nfaces*25, 3) velocity and traction at the absorbing-face points
        in face order, sample k at step 1 + k*resamp (time k*resamp*dt, the
        source-time convention), linear in between.  veloc=None switches it off."""
        if veloc is None:
            _lib.call("fwat_set_fk_injection_", self.h, 0, 0, 1, None, None)
            return
        v = np.ascontiguousarray(veloc, F32)
        t = np.ascontiguousarray(tract, F32)
        npt = self.mesh.num_abs_boundary_faces * 25
        if v.shape != t.shape or v.ndim != 3 or v.shape[1:] != (npt, 3):
            raise ValueError(f"FK fields must be (nsamp, {npt}, 3), got {v.shape} and {t.shape}")
        _lib.call("fwat_set_fk_injection_", self.h, npt, v.shape[0], int(resamp), v, t)

    def set_adjoint_sources(self, adj):
        adj = np.ascontiguousarray(adj, F32).reshape(self.nstep, self.nrec, 3)
        _lib.call("fwat_set_adjoint_sources_", self.h, adj, self.nstep)

    def iterate_time(self, fused=True, timed=False):
        _lib.call("fwat_iterate_time_", self.h, self.simulation_type, self.dt, int(fused), int(timed))

    def force_time(self):
        ms = np.zeros(1)
        n = np.zeros(1, np.int32)
        _lib.call("fwat_last_force_time_", self.h, ms, n)
        return float(ms[0]), int(n[0])

    def assembly_time(self):
        ms = np.zeros(1)
        n = np.zeros(1, np.int32)
        _lib.call("fwat_last_assembly_time_", self.h, ms, n)
        return float(ms[0]), int(n[0])

    def seismograms(self):
        out = np.zeros((self.nstep, self.nrec, 3), F32)
        if self.nrec:
            _lib.call("transfer_seismograms_from_device_", self.h, out)
        return out

    def fields(self):
        n3 = 3 * self.mesh.nglob
        u, v, a = (np.zeros((self.mesh.nglob, 3), F32) for _ in range(3))
        _lib.call("transfer_fields_el_from_device_", n3, u, v, a, self.h)
        return u, v, a

    def set_fields(self, u, v, a):
        """Initial state of a forward run (displ, veloc, accel [nglob, 3]), e.g. the state
        of the same event at an earlier step: the run then continues from it."""
        n3 = 3 * self.mesh.nglob
        f = lambda x: np.ascontiguousarray(x, F32).reshape(self.mesh.nglob, 3)
        _lib.call("transfer_fields_el_to_device_", n3, f(u), f(v), f(a), self.h)

    def b_fields(self):
        n3 = 3 * self.mesh.nglob
        u, v, a = (np.zeros((self.mesh.nglob, 3), F32) for _ in range(3))
        _lib.call("transfer_b_fields_from_device_", n3, u, v, a, self.h)
        return u, v, a

    def take_forward(self, fwd):
        """In-HBM hand-off: forward final state -> b_ fields, Stacey store moved."""
        _lib.call("fwat_copy_forward_to_backward_", fwd.h, self.h)
        if self.stacey and self.mesh.num_abs_boundary_faces:
            _lib.call("fwat_move_stacey_store_", fwd.h, self.h)

    def kernels(self):
        """save_adjoint_kernels-scaled kernels, float64 [nspec,125] (or [nspec,125,21])."""
        nloc = (self.mesh.nspec, 125)
        rho_kl = np.zeros(nloc)
        if self.aniso:
            c = np.zeros(nloc + (21,))
            mu = np.zeros(nloc)
            kap = np.zeros(nloc)
            _lib.call("transfer_kernels_el_to_host_dp_", self.h, rho_kl, mu, kap, self.mesh.nspec)
            _lib.call("transfer_kernels_ani_to_host_dp_", self.h, c, self.mesh.nspec)
            return {"rho_kl": rho_kl, "cijkl_kl": c}
        mu = np.zeros(nloc)
        kap = np.zeros(nloc)
        _lib.call("transfer_kernels_el_to_host_dp_", self.h, rho_kl, mu, kap, self.mesh.nspec)
        return {"rho_kl": rho_kl, "mu_kl": mu, "kappa_kl": kap}

    def close(self):
        if self.h:
            _lib.call("prepare_cleanup_device_", self.h, 0, 1, 0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def derived_kernels(k, rho, kappa, mu):
    """alpha/beta/rho' kernels from the scaled rho/mu/kappa kernels
    (names of kernel_io.f90:123-133; Tromp, Tape & Liu 2005)."""
    kappa = kappa.astype(np.float64)
    mu = mu.astype(np.float64)
    r = 4.0 * mu / (3.0 * kappa)
    return {"alpha_kl": 2.0 * (1.0 + r) * k["kappa_kl"],
            "beta_kl": 2.0 * (k["mu_kl"] - r * k["kappa_kl"]),
            "rhop_kl": k["rho_kl"] + k["kappa_kl"] + k["mu_kl"]}


def event_gradient(mesh, model, obs, sources, receivers, dt, nstep, *, cij=None, fused=True,
                   timed=False, windows=None, static_precision=0, approximate_hess=False):
    """PrepareFWD%simulation for one event: forward (SAVE_FORWARD), waveform
    misfit + adjoint source, adjoint run with reconstruction, kernels."""
    from .misfit import event_adjoint_sources
    rho, vp, vs = model
    fwd = ElasticRun(mesh, rho, vp, vs, sources, receivers, dt, nstep, 1, True, cij=cij,
                     static_precision=static_precision)
    fwd.iterate_time(fused=fused, timed=timed)
    syn = fwd.seismograms()
    misfit, adj = event_adjoint_sources(obs, syn, dt, windows)
    adjr = ElasticRun(mesh, rho, vp, vs, sources, receivers, dt, nstep, 3, False, cij=cij,
                      compute_and_store_strain=not fused, static_precision=static_precision,
                      approximate_hess=approximate_hess)
    adjr.take_forward(fwd)
    fwd.close()
    adjr.set_adjoint_sources(adj)
    adjr.iterate_time(fused=fused, timed=timed)
    k = adjr.kernels()
    adjr.close()
    return misfit, syn, k
