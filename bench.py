"""Benchmark: one event's forward + adjoint + kernel path on one MI355X.

Workload (production-size plain FP32 configuration): isotropic elastic box
40 x 40 x 20 km, 96 x 96 x 48 hexahedra (442,368 elements, 28,607,425 unique
GLL points, NGLL = 5), vp/vs/rho = 6.0 km/s / 3.5 km/s / 2.7 g/cc times
(1 + eps), eps a seeded Gaussian random field (2 % std, 5 km correlation);
one moment-tensor source at 10 km depth (seeded MT, Ricker 1 Hz), 400 surface
receivers on a 20 x 20 grid, free surface on top, Stacey elsewhere,
dt = 0.004 s.

A "step" is one time step of the whole event pipeline: a forward step of the
starting model (with the Stacey store for reconstruction) and, after the
waveform misfit, an adjoint step plus a reconstruction step plus kernel
accumulation.  --steps K sets NSTEP = K for the timed event (default 10000,
the configuration's NSTEP); --warmup W runs a W-step event first.  The
observed data come from an untimed forward run in the true model.

value = unique GLL points * K / wall time of the timed event
        (GLL-point-steps per second per event, fwd + adj).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
# steps per fp32 partial-sum block of the time-integrated kernels (fp64 added at each block's end)
KERNEL_BLOCK = 16


def element_kernel_name(tuning, nirr):
    """Name of the element-pass kernel the chosen launch shape runs: the general
    pass on meshes with irregular elements or regular_path 0, else the all-regular
    pass (1), its pipelined form (2), the pipelined form with ping-pong LDS
    buffers (3) or that form with non-temporal streams (4)."""
    if nirr > 0 or tuning.get("regular_path", 1) == 0:
        return "k_forces"
    return {1: "k_forces_reg", 2: "k_forces_reg_pipe", 3: "k_forces_reg_pp",
            4: "k_forces_reg_pp_nt", 5: "k_forces_reg_pp"}[tuning["regular_path"]]


def kernel_acc_bytes(kb):
    """HBM bytes per point-step of one time-integrated kernel value: fp32 partial sum read +
    written every step, fp64 accumulator read + written once per kb-step block."""
    return 8.0 + 16.0 / kb if kb > 1 else 16.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=None,
                   help="NSTEP of the timed event (default: the configuration's, 10000 for c2, 2000 for c1)")
    p.add_argument("--warmup", type=int, default=100)
    p.add_argument("--config", choices=["c2", "c1", "c3", "c4", "c5", "post", "measure"], default="c2",
                   help="c2: one production-size event (default); c1: the small event; "
                        "c3: c1 with bf16 / fp16 statics against the FP64 oracle; "
                        "c4: 32 uneven events batched vs one at a time on the 64x64x32 mesh; "
                        "c5: HTI anisotropic event with 21 c_ij kernels + smoothing of all 21; "
                        "post: kernel post-processing chain (event sums, preconditioner, smoothing, "
                        "grid projection, taper) on the 64x64x32 mesh; measure: waveform and cross-"
                        "correlation measurements of a C2-sized event (400 receivers x 3 x 10000 samples)")
    p.add_argument("--post-events", type=int, default=32, help="events summed by the post leg")
    p.add_argument("--c4-events", type=int, default=32)
    p.add_argument("--c5-events", type=int, default=16)
    p.add_argument("--c5-plain-events", type=int, default=2)
    p.add_argument("--c4-nstep-scale", type=float, default=1.0,
                   help="NSTEP of the c4 events is U[4000,12000] times this factor")
    p.add_argument("--c4-mesh", default="c4", choices=["c4", "c1"],
                   help="mesh of the c4 event set: the 64x64x32 box (default) or the 16x16x8 box, "
                        "where every launch is short and batching amortises it")
    p.add_argument("--c4-store", type=int, default=1, choices=[0, 1],
                   help="Stacey store of the batches: 0 fp32, 1 fp16 with a scale per (step, event) record")
    p.add_argument("--c4-plain-events", type=int, default=4,
                   help="events also run one at a time on the plain single-event path (its rate per "
                        "event-step is taken from these)")
    p.add_argument("--events-per-batch", type=int, default=0, help="0: autotuned")
    p.add_argument("--batch-interior", type=int, default=0, choices=[0, 1],
                   help="isotropic batches: 1 element-interior points finished in the element pass, "
                        "0 every point in the assembly")
    p.add_argument("--hbm-query", type=int, default=1, help="c4: size batches from the free HBM (torch)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-steps", type=int, default=None,
                   help="steps of the CPU baseline sample (default: 4 for c2/c4, the whole run for c1)")
    p.add_argument("--unfused", action="store_true", help="time the SPECFEM-order stage sequence")
    p.add_argument("--graph", action="store_true",
                   help="c1/c2: capture each run's step loop into one HIP graph (no per-kernel timing)")
    p.add_argument("--asm-lanes", type=int, default=0,
                   help="points per lane of the gather assembly (1 or 2; 0: 2 for the forward pass, "
                        "1 for the adjoint pass)")
    p.add_argument("--asm-block", type=int, default=0,
                   help="lanes per workgroup of the gather assembly (64..1024; 0: the library's 256)")
    p.add_argument("--floc-order", action="store_true",
                   help="element-local force rows in local-point order instead of grouped by reader")
    p.add_argument("--static-precision", type=int, default=0, help="0 fp32, 1 bf16, 2 fp16")
    p.add_argument("--force-irregular", action="store_true",
                   help="c1/c2: every element stores its inverse Jacobian per point (the general-hexahedron "
                        "layout of meshes with topography), 40 B per local point and pass")
    p.add_argument("--no-tune", action="store_true", help="skip the element-pass launch-shape autotuner")
    p.add_argument("--fk", type=int, default=0, metavar="R",
                   help="c2: inject a synthetic incident field (FK coupling) through the Stacey faces of the "
                        "forward run, one sample every R steps, the series resident in HBM")
    p.add_argument("--tuning", default="", metavar="R,F,A[,I]",
                   help="fixed launch shape instead of the autotuner: regular_path, forward and adjoint "
                        "workgroups per CU (e.g. 1,16,8), as the autotuner picks them, and whether the "
                        "pipelined passes finish the element-interior points (1, the default, or 0)")
    p.add_argument("--window-start", type=int, default=-1, metavar="S0",
                   help="the timed event covers steps S0+1..S0+K, the state at S0 from an untimed forward run "
                        "(default: 0 for K >= 2000, else the first P arrival at the nearest receiver + 50 steps)")
    p.add_argument("--copy-sweep", action="store_true",
                   help="with --copy-probe: also the copy with 1-8 loads in flight per lane and several grid sizes")
    p.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                   help="only run a plain device copy of BYTES (counter calibration) and exit")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(cfg):
    if cfg == "c2":
        return dict(nex=96, ney=96, nez=48, lx=40000.0, ly=40000.0, lz=20000.0, dt=0.004, nrx=20, nstep=10000)
    if cfg == "c4":
        return dict(nex=64, ney=64, nez=32, lx=40000.0, ly=40000.0, lz=20000.0, dt=0.005, nrx=20, nstep=8000)
    return dict(nex=16, ney=16, nez=8, lx=40000.0, ly=40000.0, lz=20000.0, dt=0.005, nrx=4, nstep=2000)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(events would shard across GPUs, one process each; only the one-GPU share is measured)",
              file=sys.stderr)
        sys.exit(2)
    from specfwat_amd import _lib
    if a.copy_probe:
        ms = np.zeros(1)
        _lib.call("fwat_copy_probe_", a.copy_probe, 5, ms)
        out = {"copy_probe_bytes": a.copy_probe, "avg_ms": float(ms[0]),
               "GBps_read_plus_write": 2 * a.copy_probe / (ms[0] * 1e-3) / 1e9}
        if a.copy_sweep:
            sweep = []
            for u in (1, 2, 4, 8):
                for bpc in (0, 4, 8, 16, 32):
                    _lib.call("fwat_copy_probe_unrolled_", a.copy_probe, 5, u, bpc, ms)
                    sweep.append({"unroll": u, "blocks_per_cu": bpc, "avg_ms": float(ms[0]),
                                  "GBps": 2 * a.copy_probe / (ms[0] * 1e-3) / 1e9})
            out["sweep"] = sweep
            out["best_GBps"] = max(r["GBps"] for r in sweep)
        print(json.dumps(out), flush=True)
        return
    if a.config == "c4":
        return batched(a)
    if a.config == "c3":
        return precision(a)
    if a.config == "post":
        return post_chain(a)
    if a.config == "measure":
        return measure(a)
    if a.config == "c5":
        return aniso_smooth(a)
    from specfwat_amd.mesh import BoxMesh
    from specfwat_amd.model import perturbed_model
    from specfwat_amd.sources import Receivers, Sources, surface_grid
    from specfwat_amd.solver import ElasticRun
    from specfwat_amd import misfit as misfit_mod
    from specfwat_amd.misfit import event_adjoint_sources

    w = workload(a.config)
    K, W = a.steps or w["nstep"], a.warmup
    cpu_steps = a.cpu_steps or (K if a.config == "c1" else 6)
    t0 = time.time()
    mesh = BoxMesh(w["nex"], w["ney"], w["nez"], w["lx"], w["ly"], w["lz"], force_irregular=a.force_irregular)
    log(f"mesh {mesh.nspec} elements, {mesh.nglob} points, {mesh.nspec_irregular} with stored Jacobians "
        f"({time.time() - t0:.1f}s)")
    true_model = perturbed_model(mesh, seed=2024)
    smooth_model = perturbed_model(mesh, smooth=True)
    rng = np.random.Generator(np.random.Philox(7))
    sxyz = [[rng.uniform(0.3, 0.7) * w["lx"], rng.uniform(0.3, 0.7) * w["ly"], -10000.0]]
    mt = [list(1e15 * rng.standard_normal(6))]
    src = Sources(mesh, sxyz, mt, [1.0])
    rxyz = surface_grid(mesh, w["nrx"], w["nrx"])
    rec = Receivers(mesh, rxyz)
    dt = w["dt"]
    # A short timed event (the driver's --steps 20) would end before any wave reaches a
    # receiver, leaving a zero misfit and adjoint field.  It then covers the event's steps
    # S0+1..S0+K, S0 being the first P arrival at the nearest receiver plus 50 steps: the
    # forward state at step S0 comes from an untimed run of the same event (starting and true
    # model alike) and the reconstruction returns to it.
    S0 = a.window_start
    if S0 < 0:
        dmin = float(np.min(np.linalg.norm(rxyz - np.asarray(sxyz[0]), axis=1)))
        S0 = 0 if K >= 2000 else int(np.ceil((1.2 / 1.0 + dmin / 6000.0) / dt)) + 50
    src_w = Sources(mesh, sxyz, mt, [1.0], tshift=[-S0 * dt]) if S0 else src
    log(f"model/source/receivers ready ({time.time() - t0:.1f}s); timed window: steps {S0 + 1}..{S0 + K}")
    fused = not a.unfused
    tuning = {"regular_path": 1, "blocks_per_cu_fwd": 0, "blocks_per_cu_adj": 8}
    if a.tuning:
        vals = [int(v) for v in a.tuning.split(",")]
        tuning = {"regular_path": vals[0], "blocks_per_cu_fwd": vals[1], "blocks_per_cu_adj": vals[2],
                  "interior": vals[3] if len(vals) > 3 else 1}
    elif fused and not a.no_tune:
        from specfwat_amd.tune import tune_element_passes
        tuning, table = tune_element_passes(mesh, smooth_model, src, rec, dt,
                                            static_precision=a.static_precision)
        log("autotuner (regular_path, blocks/CU, fwd ms, adj ms): "
            + ", ".join(f"({r},{b},{x:.3f},{y:.3f})" for r, b, x, y in table) + f" -> {tuning}")

    def primed_state(model):
        """Forward state of the event at step S0 (None when the window starts at rest)."""
        if not S0:
            return None
        r = ElasticRun(mesh, *model, src, rec, dt, S0, 1, False)
        r.set_tuning(**tuning)
        r.iterate_time()
        st = r.fields()
        r.close()
        return st

    state_true = primed_state(true_model)
    state_smooth = primed_state(smooth_model)

    def observed(nstep, window=False):
        r = ElasticRun(mesh, *true_model, src_w if window else src, rec, dt, nstep, 1, False)
        r.set_tuning(**tuning)
        if window and state_true is not None:
            r.set_fields(*state_true)
        r.iterate_time()
        obs = r.seismograms()
        r.close()
        return obs

    def prepare(nstep, window=False):
        s_ = src_w if window else src
        f = ElasticRun(mesh, *smooth_model, s_, rec, dt, nstep, 1, True,
                       static_precision=a.static_precision)
        ad = ElasticRun(mesh, *smooth_model, s_, rec, dt, nstep, 3, False,
                        compute_and_store_strain=not fused, static_precision=a.static_precision)
        f.set_tuning(**tuning)
        ad.set_tuning(**tuning)
        ad.set_kernel_block(KERNEL_BLOCK)
        if a.graph:
            f.set_graph(True)
            ad.set_graph(True)
        if a.asm_lanes:
            f.set_assembly_lanes(a.asm_lanes)
            ad.set_assembly_lanes(a.asm_lanes)
        if a.asm_block:
            f.set_assembly_block(a.asm_block)
            ad.set_assembly_block(a.asm_block)
        if a.floc_order:
            f.set_floc_layout(False)
            ad.set_floc_layout(False)
        if window and state_smooth is not None:
            f.set_fields(*state_smooth)
        if a.fk:
            # synthetic incident field: a separable pulse times a fixed pattern over the face points
            nsamp = (nstep - 1) // a.fk + 1
            nfp3 = mesh.num_abs_boundary_faces * 75
            tt = np.arange(nsamp) * a.fk * dt
            pulse = (np.exp(-((tt - 2.0) / 0.5) ** 2) * 1e-3).astype(np.float32)
            pat = np.cos(np.arange(nfp3) * 0.001).astype(np.float32)
            vfk = np.multiply.outer(pulse, pat).reshape(nsamp, nfp3 // 3, 3)
            f.set_fk_injection(vfk, vfk * np.float32(1e7), resamp=a.fk)
            log(f"FK injection: {nsamp} samples x {nfp3 // 3} face points "
                f"({2 * vfk.nbytes / 1e9:.2f} GB in HBM)")
        return f, ad

    def event(f, ad, obs, timed):
        f.iterate_time(fused=fused, timed=timed)
        tf = time.time()
        syn = f.seismograms()
        # waveform misfit, adjoint sources band-passed 0.4-10 s (filter_adj) around the 1 Hz source
        misfit, adj = event_adjoint_sources(obs, syn, dt, min_period=0.4, max_period=10.0)
        ad.set_adjoint_sources(adj)
        ad.take_forward(f)
        ad.iterate_time(fused=fused, timed=timed)
        return misfit, tf

    if W > 0:
        obs_w = observed(W)
        f, ad = prepare(W)
        event(f, ad, obs_w, False)
        f.close(); ad.close()
        log(f"warmup event of {W} steps done ({time.time() - t0:.1f}s)")
    obs = observed(K, window=True)
    log(f"observed data ({K} steps) generated ({time.time() - t0:.1f}s)")
    f, ad = prepare(K, window=True)
    _lib.call("fwat_synchronize_", f.h)
    _lib.call("fwat_synchronize_", ad.h)
    ts = time.time()
    misfit, tf = event(f, ad, obs, not a.graph)
    _lib.call("fwat_synchronize_", ad.h)
    te = time.time()
    wall = te - ts
    fwd_ms, fwd_n = f.force_time()
    adj_ms, adj_n = ad.force_time()
    f_asm, a_asm = f.assembly_time(), ad.assembly_time()
    brick = tuning.get("brick") if fused else None
    binfo = ad.brick_info() if brick else None
    # points the gather assembles and element-interior points the element passes finish
    n_sh_f, n_in_f, act_f, _ = f.interior_info()
    n_sh_a, n_in_a, _, act_a = ad.interior_info()
    n_in_f = n_in_f if (fused and act_f) else 0
    n_in_a = n_in_a if (fused and act_a) else 0
    # of those, the ones at an element's 27 inner slots (the rest: inner points of a
    # free-surface element's top face), whose element-local forces the pass does not store
    n_special = len(set(src.ispec.tolist()) | set(rec.ispec.tolist()))
    n_slot_f = min(n_in_f, 27 * (mesh.nspec - n_special))
    n_slot_a = min(n_in_a, 27 * (mesh.nspec - n_special))
    k = ad.kernels()
    f.close(); ad.close()
    if not all(np.isfinite(v).all() for v in k.values()):
        raise RuntimeError("non-finite kernels")
    log(f"timed event: {wall:.2f}s (forward {tf - ts:.2f}s), misfit {misfit:.4e}")

    nglob, nloc = mesh.nglob, mesh.nspec * 125
    nirr = mesh.nspec_irregular * 125
    nfp = mesh.num_abs_boundary_faces * 25
    ncontrib = nloc + nfp + 125 * (src.nsources + rec.nrec)
    value = nglob * K / wall
    fwd_asm_ms, fwd_asm_n = f_asm
    adj_asm_ms, adj_asm_n = a_asm
    # Algorithmic bytes per launch (each array touched once, in bytes):
    kmu = 8.0 if a.static_precision == 0 else 4.0  # kappa,mu bytes per local point
    kacc = kernel_acc_bytes(KERNEL_BLOCK)  # one time-integrated kernel value per point-step
    if brick:
        nh, nslot, nch = binfo
        ninner = nglob - nh
        bname = "x".join(str(v) for v in brick)
        kinds = {
            # brick pass, two fields: ibool 4 + kappa,mu per local point, mu/kappa accumulators
            # 2 x kacc per local point, displacement gathers 12 + 12 per unique point; brick-inner
            # points: rmass_inv 12, u,v of both fields read + written 96, rho kernel kacc;
            # partial sums of both fields written 24 per slot
            f"k_brick<{bname},two-field>": ((4.0 + kmu + 2 * kacc) * nloc + 24.0 * nglob
                                            + (108.0 + kacc) * ninner + 24.0 * nslot, adj_ms, adj_n),
            # halo pass, two fields: point list + CSR offsets 8 per halo point, 4 per contribution,
            # partial sums 24 per slot, Stacey terms 24 per face point, rmass_inv 12, u,v of both
            # fields 96 and the rho kernel kacc per halo point
            "k_halo<two-field>": (8.0 * nh + 4.0 * nch + 24.0 * nslot + 24.0 * nfp + (108.0 + kacc) * nh,
                                  adj_asm_ms, adj_asm_n),
        }
        fwd_bytes = (4.0 + kmu) * nloc + 12.0 * nglob + 60.0 * ninner + 12.0 * nslot
        fwd_asm_bytes = 8.0 * nh + 4.0 * nch + 12.0 * nslot + 12.0 * nfp + 60.0 * nh
    elif fused:
        # two-field element pass: ibool 4 + kappa,mu 8 (+ 40 irregular Jacobian) per local
        # point, mu/kappa accumulators 2 x kacc, element-local forces written for both fields
        # 24; displacement gather of both fields 12 + 12 per unique point
        # With element-interior points finished by the element pass (n_in of them, one
        # element-local slot each): the pass reads their rmass_inv 12 and velocities (12
        # per field), writes u and v (24 per field) and, two fields, the rho kernel kacc,
        # and stores no element-local forces for the n_slot of them at an element's inner
        # slots; the assembly covers the other nglob - n_in points, their contributions
        # and element-local slots only.
        ename = element_kernel_name(tuning, nirr)
        ngs = nglob - n_in_a
        kinds = {
            f"{ename}<2,iso,fused-kernels>": ((28.0 + kmu + 2 * kacc) * nloc + 40.0 * nirr + 24.0 * nglob
                                              + (12.0 + 72.0 + kacc) * n_in_a - 24.0 * n_slot_a, adj_ms, adj_n),
            # two-field assembly: CSR offsets 4 per point + indices 4 per contribution,
            # element-local forces of both fields 24 per local point, Stacey terms of both
            # fields 24 per face point, rmass_inv 12, u,v of both fields read 48 + write 48,
            # rho kernel kacc per unique point
            "k_assemble<two-field>": (4.0 * ngs + 4.0 * (ncontrib - n_in_a) + 24.0 * (nloc - n_in_a) + 24.0 * nfp
                                      + 12.0 * ngs + 96.0 * ngs + kacc * ngs, adj_asm_ms, adj_asm_n),
        }
        ngf = nglob - n_in_f
        fwd_bytes = (16.0 + kmu) * nloc + 40.0 * nirr + 12.0 * nglob + (12.0 + 36.0) * n_in_f - 12.0 * n_slot_f
        fwd_asm_bytes = (4.0 * ngf + 4.0 * (ncontrib - n_in_f) + 12.0 * (nloc - n_in_f) + 12.0 * nfp
                         + 12.0 * ngf + 48.0 * ngf)
    else:
        # SPECFEM stage order: one field per launch, atomic scatter, strain store
        kinds = {"k_forces<1,iso,strain-store,atomic>": (12.0 * nloc + 24.0 * nloc + 40.0 * nirr + 36.0 * nglob,
                                                         adj_ms, adj_n)}
        fwd_bytes = 12.0 * nloc + 40.0 * nirr + 36.0 * nglob
        fwd_asm_bytes = 0.0
    # element slabs: every pass is split into nsl launches over equal element ranges
    nsl = -(-mesh.nspec // tuning["slab_elems"]) if (fused and tuning.get("slab_elems")) else 1
    if nsl > 1:
        kinds = {k: (v[0] / nsl, v[1], v[2]) for k, v in kinds.items()}
        fwd_bytes /= nsl
        fwd_asm_bytes /= nsl
    kname = max(kinds, key=lambda k: kinds[k][1])
    kbytes, kms, kn = kinds[kname]
    avg_ms = kms / max(kn, 1)
    # a graph-captured run carries no per-kernel times
    achieved = kbytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else None
    traffic = None
    tf_path = os.path.join(HERE, "profiles", "c2_fwd_adj_hbm_traffic.md")
    if os.path.exists(tf_path) and fused and a.config == "c2" and a.static_precision == 0 and nsl == 1:
        for line in open(tf_path):
            if line.startswith("hbm_bytes_per_launch " + kname):
                traffic = float(line.split("=")[1])
    fwd_avg_ms = fwd_ms / max(fwd_n, 1)
    step_bytes = nsl * (sum(v[0] for v in kinds.values()) + fwd_bytes + fwd_asm_bytes)
    result = {
        "metric": "GLL-pt-steps/s/event fwd+adj",
        "value": value,
        "unit": "GLL-pt-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if a.static_precision == 0 else ("f32+bf16-statics" if a.static_precision == 1 else "f32+fp16-statics"),
        "data": "synthetic",
        "config": {"workload": (f"{'C2' if a.config == 'c2' else 'C1'} isotropic box "
                                + ("(every element with stored per-point Jacobians) " if a.force_irregular else "")
                                + 
                                f"{w['nex']}x{w['ney']}x{w['nez']} el, {nglob} GLL pts, NGLL=5, "
                                f"{rec.nrec} receivers, dt={dt}, NSTEP={K}, fwd + waveform misfit (adjoint sources "
                                "band-passed 0.4-10 s) + adjoint/reconstruction + rho/mu/kappa kernels "
                                f"(fp64 accum, fp32 partial sums over {KERNEL_BLOCK} steps)"),
                   "window": {"steps": [S0 + 1, S0 + K], "misfit": misfit,
                              "note": ("whole event from rest" if S0 == 0 else
                                       f"event steps {S0 + 1}..{S0 + K}: the forward state at step {S0} (after "
                                       "the first P arrival at the nearest receiver) from an untimed run; "
                                       "the reconstruction returns to it")},
                   "path": ((f"brick-fused {'x'.join(map(str, brick))}" if brick else
                             (f"fused, {nsl} element slabs" if nsl > 1 else "fused"))
                            if fused else "SPECFEM stage order (strain stores)")
                           + (", step loops captured as HIP graphs" if a.graph else ""),
                   "element_pass_tuning": tuning if fused else None,
                   **({"fk_injection": {"every_steps": a.fk}} if a.fk else {})},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                     "algorithmic_bytes_per_launch": kbytes, "avg_launch_ms": avg_ms, "launches": kn,
                     # the whole step: the four passes' algorithmic bytes over the wall time per step
                     "step": {"algorithmic_bytes": step_bytes, "achieved": step_bytes / (wall / K) / 1e9,
                              "frac": step_bytes / (wall / K) / 1e9 / HBM_PEAK_GBS}},
        "breakdown": {"wall_s": wall, "forward_s": tf - ts, "adjoint_s": te - tf,
                      "kernels": {k: {"avg_ms": v[1] / max(v[2], 1),
                                      "GBps": v[0] / (v[1] / max(v[2], 1) * 1e-3) / 1e9 if v[1] > 0 else None}
                                  for k, v in kinds.items()},
                      "fwd_force_avg_ms": fwd_avg_ms,
                      "fwd_force_GBps": fwd_bytes / (fwd_avg_ms * 1e-3) / 1e9 if fwd_avg_ms > 0 else None,
                      "measurement_device_ms": misfit_mod.last_time_ms(),
                      "points_assembled": {"forward": nglob - n_in_f, "adjoint": nglob - n_in_a,
                                           "element_interior_finished_in_element_pass": n_in_a},
                      "fwd_assembly_avg_ms": fwd_asm_ms / max(fwd_asm_n, 1),
                      "fwd_assembly_GBps": (fwd_asm_bytes / (fwd_asm_ms / fwd_asm_n * 1e-3) / 1e9
                                            if fwd_asm_n and fwd_asm_ms > 0 else None)},
    }
    if not a.no_cpu:
        result["cpu_baseline"] = cpu_baseline(w, mesh, smooth_model, sxyz, mt, rec, cpu_steps,
                                              label=("C1 event" if a.config == "c1" else "C2 event"),
                                              force_irregular=a.force_irregular,
                                              nstep_event=w["nstep"],
                                              one_thread_steps=min(cpu_steps, 20 if a.config == "c1" else 1))
    print(json.dumps(result), flush=True)


def batched(a):
    """Many uneven events on one mesh: the batched path (events grouped by
    NSTEP, events per launch and workgroups per CU autotuned) against the
    plain single-event path on the same events, with the per-event
    differences between the two stated."""
    from specfwat_amd import _lib
    from specfwat_amd.batch import EventBatch, plan_batches, tune_batch
    from specfwat_amd.mesh import BoxMesh
    from specfwat_amd.misfit import event_adjoint_sources
    from specfwat_amd.model import perturbed_model
    from specfwat_amd.solver import ElasticRun
    from specfwat_amd.sources import Receivers, Sources, surface_grid
    from specfwat_amd.tune import tune_element_passes

    w = workload(a.c4_mesh)
    t0 = time.time()
    mesh = BoxMesh(w["nex"], w["ney"], w["nez"], w["lx"], w["ly"], w["lz"])
    true_model = perturbed_model(mesh, seed=2024)
    smooth_model = perturbed_model(mesh, smooth=True)
    rec = Receivers(mesh, surface_grid(mesh, w["nrx"], w["nrx"]))
    dt = w["dt"]
    rng = np.random.Generator(np.random.Philox(31))
    srcs, nsteps = [], []
    for _ in range(a.c4_events):
        sxyz = [[rng.uniform(0.0, 1.0) * w["lx"], rng.uniform(0.0, 1.0) * w["ly"], -rng.uniform(2000.0, 18000.0)]]
        srcs.append(Sources(mesh, sxyz, [list(1e15 * rng.standard_normal(6))], [rng.uniform(0.5, 1.5)]))
        nsteps.append(max(2, int(round(rng.uniform(4000.0, 12000.0) * a.c4_nstep_scale))))
    log(f"c4: mesh {mesh.nspec} el, {mesh.nglob} pts; {len(srcs)} events, NSTEP {min(nsteps)}..{max(nsteps)} "
        f"({time.time() - t0:.1f}s)")
    sp = a.static_precision
    # events per batch: what the tuner finds fastest among the sizes whose batch (fields,
    # element-local forces, kernels, Stacey store of the longest event) fits the HBM left
    # beside the mesh handle
    from specfwat_amd.batch import batch_device_bytes
    free_b, _ = __import__("torch").cuda.mem_get_info() if a.hbm_query else (288e9, None)
    budget = 0.85 * free_b
    fits = [e for e in (2, 4, 8) if batch_device_bytes(mesh, e, max(nsteps), a.c4_store) <= budget]
    if not fits:
        raise RuntimeError("no batch size fits the HBM budget")
    if a.events_per_batch:
        E, bpc, ttable = a.events_per_batch, 8, []
    else:
        (E, bpc), ttable = tune_batch(mesh, smooth_model, srcs, rec, dt, static_precision=sp, sizes=tuple(fits),
                                      interior=bool(a.batch_interior))
        log("batch tuner (events, blocks/CU, fwd ms/event-step, adj ms/event-step): "
            + ", ".join(f"({e},{b},{x:.3f},{y:.3f})" for e, b, x, y in ttable) + f" -> E={E} bpc={bpc}")
    log(f"batch sizes fitting {budget / 1e9:.0f} GB with the {'fp16' if a.c4_store else 'fp32'} store: {fits}")
    groups = plan_batches(nsteps, E)

    # observed data: batched forward runs in the true model (untimed)
    obs = [None] * len(srcs)
    for g in groups:
        b = EventBatch(mesh, *true_model, [srcs[i] for i in g], rec, dt, [nsteps[i] for i in g])
        b.set_store_precision(a.c4_store)
        b.forward()
        s = b.seismograms()
        for k, i in enumerate(g):
            obs[i] = s[k, :nsteps[i]].copy()
        b.close()
    log(f"observed data ready ({time.time() - t0:.1f}s)")

    def run_group(g, keep):
        b = EventBatch(mesh, *smooth_model, [srcs[i] for i in g], rec, dt, [nsteps[i] for i in g],
                       static_precision=sp)
        b.set_tuning(bpc)
        b.set_store_precision(a.c4_store)
        b.set_interior(a.batch_interior)
        # points the element pass finishes (the source elements' own, at most 27 per
        # event, stay in the assembly and are counted with the others here)
        n_fin[0] = b.run.interior_info()[1] if a.batch_interior else 0
        _lib.call("fwat_synchronize_", b.run.h)
        ts = time.time()
        b.forward(timed=True)
        fwd_t = b.pass_times()
        syn = b.seismograms()
        adj = np.zeros_like(syn)
        for k, i in enumerate(g):
            _, adj[k, :nsteps[i]] = event_adjoint_sources(obs[i], syn[k, :nsteps[i]], dt)
        b.set_adjoint_sources(adj)
        b.adjoint(timed=True)
        _lib.call("fwat_synchronize_", b.run.h)
        wall = time.time() - ts
        adj_t = b.pass_times()
        kern = b.kernels() if keep else None
        b.close()
        return wall, syn, kern, fwd_t, adj_t

    n_fin = [0]
    bwall, bsyn, bkern, times = 0.0, {}, {}, {}
    check = set(groups[0][:2])  # the two longest events: kernels compared with the plain path
    for gi, g in enumerate(groups):
        wall, syn, kern, ft, at = run_group(g, any(i in check for i in g))
        log(f"batch {gi + 1}/{len(groups)} ({len(g)} events, NSTEP up to {max(nsteps[i] for i in g)}): {wall:.1f}s")
        bwall += wall
        for k, i in enumerate(g):
            bsyn[i] = syn[k, :nsteps[i]]
            if kern is not None and i in check:
                bkern[i] = {q: v[k] for q, v in kern.items()}
        times.setdefault(len(g), (ft, at))
    total = mesh.nglob * float(sum(nsteps))
    bvalue = total / bwall
    log(f"batched: {bwall:.2f}s over {len(groups)} batches -> {bvalue:.4e} GLL-pt-steps/s/event")

    # plain path: the same events one at a time (fused single-event path, tuned shapes)
    tuning, _ = tune_element_passes(mesh, smooth_model, srcs[0], rec, dt, static_precision=sp)
    pwall, perr, kerr = 0.0, 0.0, {}
    # the plain path on the first c4_plain_events events of the first batch (those whose kernels
    # are compared), its rate per event-step applied to the whole set
    plain = [i for g in groups for i in g][:max(1, a.c4_plain_events)]
    for i in plain:
        src = srcs[i]
        f = ElasticRun(mesh, *smooth_model, src, rec, dt, nsteps[i], 1, True, static_precision=sp)
        ad = ElasticRun(mesh, *smooth_model, src, rec, dt, nsteps[i], 3, False, static_precision=sp)
        f.set_tuning(**tuning)
        ad.set_tuning(**tuning)
        _lib.call("fwat_synchronize_", f.h)
        ts = time.time()
        f.iterate_time()
        syn = f.seismograms()
        _, adj = event_adjoint_sources(obs[i], syn, dt)
        ad.set_adjoint_sources(adj)
        ad.take_forward(f)
        ad.iterate_time()
        _lib.call("fwat_synchronize_", ad.h)
        pwall += time.time() - ts
        log(f"plain event {i} (NSTEP {nsteps[i]}): {time.time() - ts:.1f}s")
        f.close()
        if np.linalg.norm(syn) > 0:
            perr = max(perr, float(np.linalg.norm(bsyn[i] - syn) / np.linalg.norm(syn)))
        if i in bkern:
            k = ad.kernels()
            for q in k:
                kerr[q] = max(kerr.get(q, 0.0),
                              float(np.linalg.norm(bkern[i][q] - k[q]) / np.linalg.norm(k[q])))
        ad.close()
    pvalue = mesh.nglob * float(sum(nsteps[i] for i in plain)) / pwall
    log(f"plain: {pwall:.2f}s for {len(plain)} events -> {pvalue:.4e} GLL-pt-steps/s/event")

    # roofline of the largest batch's dominant pass (algorithmic bytes per launch)
    ftimes, atimes = times[max(times)]
    e = max(times)
    nglob, nloc = mesh.nglob, mesh.nspec * 125
    nfp = mesh.num_abs_boundary_faces * 25
    ncontrib = nloc + nfp + 125 * (e + rec.nrec)
    kmu = 8.0 if sp == 0 else 4.0
    kacc = kernel_acc_bytes(KERNEL_BLOCK)  # the batch handle's partial-sum blocks (library default)
    n_in = n_fin[0]
    n_sh = nglob - n_in
    kinds = {
        # statics once (ibool 4 + kappa,mu) per local point; per event: both fields gathered
        # 24 per unique point, forces of both fields written 24 per local point that the
        # assembly reads and the mu/kappa accumulators 2 x kacc per local point; for the
        # points finished here rmass_inv 12 once and per event v of both fields read 24,
        # u, v of both fields written 48 and the rho kernel kacc
        f"k_batch_elem<{e},2>": ((4.0 + kmu) * nloc + 12.0 * n_in
                                 + e * (24.0 * nglob + 2 * kacc * nloc + 24.0 * (nloc - n_in)
                                        + (72.0 + kacc) * n_in),
                                 atimes[0], atimes[1]),
        # over the points it assembles: CSR 4 per point + 4 per contribution, rmass_inv 12 per
        # point once; per event: forces of both fields 24 per element term, Stacey terms 24 per
        # face point, u,v of both fields read + written 96 and the rho kernel kacc per point
        f"k_batch_asm<{e},two-field>": (16.0 * n_sh + 4.0 * (ncontrib - n_in)
                                        + e * (24.0 * (nloc - n_in) + 24.0 * nfp + (96.0 + kacc) * n_sh),
                                        atimes[2], atimes[3]),
    }
    kname = max(kinds, key=lambda k: kinds[k][1])
    kbytes, kms, kn = kinds[kname]
    avg_ms = kms / max(kn, 1)
    achieved = kbytes / (avg_ms * 1e-3) / 1e9
    result = {
        "metric": "GLL-pt-steps/s/event fwd+adj",
        "value": bvalue,
        "unit": "GLL-pt-steps/s",
        "n_gpus": 1,
        "steps": int(sum(nsteps)),
        "warmup": 0,
        "ms_per_step": bwall / sum(nsteps) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if sp == 0 else ("f32+bf16-statics" if sp == 1 else "f32+fp16-statics"),
        "data": "synthetic",
        "config": {"workload": (f"{'C4' if a.c4_mesh == 'c4' else 'C4 event set on the C1 mesh'} batched: "
                                f"isotropic box {w['nex']}x{w['ney']}x{w['nez']} el, {nglob} GLL pts, "
                                f"{len(srcs)} events (depth U[2,18] km, random MT, f0 U[0.5,1.5] Hz), "
                                f"NSTEP U[4000,12000] x {a.c4_nstep_scale} = {min(nsteps)}..{max(nsteps)}, "
                                f"{rec.nrec} receivers, dt={dt}; fwd + misfit + adjoint/reconstruction + "
                                "rho/mu/kappa kernels per event"),
                   "events_per_batch": E, "blocks_per_cu": bpc, "batches": [len(g) for g in groups],
                   "element_interior_points_finished_in_element_pass": n_in,
                   "tuner": [list(r) for r in ttable]},
        "batched_vs_plain": {"plain_value": pvalue, "speedup": bvalue / pvalue,
                             "plain_events_timed": len(plain),
                             "max_seismogram_rel_l2": perr, "kernel_rel_l2": kerr,
                             "kernel_events_checked": sorted(check),
                             "stacey_store": "fp16 + scale per (step, event)" if a.c4_store else "fp32"},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": kbytes, "avg_launch_ms": avg_ms, "launches": kn},
        "breakdown": {"batched_s": bwall, "plain_s": pwall,
                      "kernels": {k: {"avg_ms": v[1] / max(v[2], 1),
                                      "GBps": v[0] / (v[1] / max(v[2], 1) * 1e-3) / 1e9 if v[1] > 0 else None}
                                  for k, v in kinds.items()},
                      "fwd_elem_avg_ms": ftimes[0] / max(ftimes[1], 1),
                      "fwd_asm_avg_ms": ftimes[2] / max(ftimes[3], 1)},
    }
    if not a.no_cpu:
        s0 = srcs[0]
        result["cpu_baseline"] = cpu_baseline(w, mesh, smooth_model, s0.xyz.tolist(), s0.mt.tolist(), rec,
                                              a.cpu_steps or 4, f0=float(s0.f0[0]), label="C4 event",
                                              nstep_event=int(np.mean(nsteps)))
    print(json.dumps(result), flush=True)


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else None


def precision(a):
    """The small event with the static arrays (kappa/mu, and the inverse
    Jacobian of irregular elements) in fp32, bf16 and fp16-with-scaling, each
    against an FP64 oracle run of the same pipeline (forward, waveform misfit
    and adjoint source from its own synthetics, adjoint + kernels), on the
    regular mesh and on the forced-irregular (general hexahedron) variant."""
    from oracle import oracle as orc
    from specfwat_amd import _lib
    from specfwat_amd.mesh import BoxMesh
    from specfwat_amd.misfit import event_adjoint_sources
    from specfwat_amd.model import perturbed_model
    from specfwat_amd.solver import ElasticRun
    from specfwat_amd.sources import Receivers, Sources, surface_grid

    w = workload("c1")
    K = a.steps or 2000
    dt = w["dt"]
    rows, timing = {}, {}
    for firr, kind in ((False, "starting"), (False, "perturbed"), (True, "starting"), (True, "perturbed")):
        mesh = BoxMesh(w["nex"], w["ney"], w["nez"], w["lx"], w["ly"], w["lz"], force_irregular=firr)
        true_model = perturbed_model(mesh, seed=2024)
        smooth_model = perturbed_model(mesh, smooth=True)
        if kind == "perturbed":
            # simulate in the 2 % random model (varying statics), data from the homogeneous one
            true_model, smooth_model = smooth_model, true_model
        rng = np.random.Generator(np.random.Philox(7))
        sxyz = [[rng.uniform(0.3, 0.7) * w["lx"], rng.uniform(0.3, 0.7) * w["ly"], -10000.0]]
        mt = [list(1e15 * rng.standard_normal(6))]
        src = Sources(mesh, sxyz, mt, [1.0])
        rxyz = surface_grid(mesh, w["nrx"], w["nrx"])
        rec = Receivers(mesh, rxyz)
        r = ElasticRun(mesh, *true_model, src, rec, dt, K, 1, False)
        r.iterate_time()
        obs = r.seismograms()
        r.close()
        t0 = time.time()
        o = orc.Oracle(w["nex"], w["ney"], w["nez"], w["lx"], w["ly"], w["lz"], *[m.ravel() for m in smooth_model],
                       dt=dt, nstep=K, src_xyz=sxyz, src_mt=mt, src_f0=[1.0], rec_xyz=rxyz,
                       force_irregular=firr, precision=64)
        s64 = o.forward(save=True)
        _, adj64 = event_adjoint_sources(obs, s64, dt)
        k64 = o.adjoint(adj64)
        log(f"c3 {'forced-irregular' if firr else 'regular'} {kind} model: FP64 oracle {time.time() - t0:.1f}s")
        for sp, name in ((0, "fp32"), (1, "bf16"), (2, "fp16-scaled")):
            f = ElasticRun(mesh, *smooth_model, src, rec, dt, K, 1, True, static_precision=sp)
            ad = ElasticRun(mesh, *smooth_model, src, rec, dt, K, 3, False, static_precision=sp)
            _lib.call("fwat_synchronize_", f.h)
            ts = time.time()
            f.iterate_time()
            syn = f.seismograms()
            _, adj = event_adjoint_sources(obs, syn, dt)
            ad.set_adjoint_sources(adj)
            ad.take_forward(f)
            ad.iterate_time()
            _lib.call("fwat_synchronize_", ad.h)
            wall = time.time() - ts
            k = ad.kernels()
            f.close(); ad.close()
            key = f"{'irregular' if firr else 'regular'}/{kind}-model/{name}"
            rows[key] = {"seismograms": rel_l2(syn, s64),
                         **{q: rel_l2(k[q].ravel(), k64[q]) for q in ("rho_kl", "mu_kl", "kappa_kl")}}
            timing[key] = mesh.nglob * K / wall
            log(f"c3 {key}: {rows[key]}  {timing[key]:.3e} GLL-pt-steps/s")
    best = "regular/starting-model/bf16"
    result = {
        "metric": "GLL-pt-steps/s/event fwd+adj",
        "value": timing[best],
        "unit": "GLL-pt-steps/s",
        "n_gpus": 1, "steps": K, "warmup": 0,
        "ms_per_step": 1e3 * mesh.nglob / timing[best],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+bf16-statics", "data": "synthetic",
        "config": {"workload": (f"C3: C1 event (16x16x8 el, 139,425 GLL pts, NSTEP={K}, dt={dt}, 16 receivers) "
                                "with static arrays in bf16 / fp16-with-scaling (per-element fp32 mean + "
                                "reduced-precision deviation), regular and forced-irregular meshes, simulated in "
                                "the homogeneous starting model and in the 2 % random model; errors = relative L2 "
                                "against an FP64 oracle run of the same pipeline")},
        "precision_errors_vs_fp64_oracle": rows,
        "rates": timing,
    }
    print(json.dumps(result), flush=True)


def aniso_smooth(a):
    """C5: HTI anisotropic events on the 64x64x32 mesh, batched (E events per
    launch, the c_ij read once per element for all of them, fp16 Stacey store),
    each with its rho and 21 c_ij kernels in fp64, against the plain single-event
    path on a sample of the events; then Gaussian PDE smoothing (sigma_h 5 km,
    sigma_v 2 km) of all 21 kernels of one event in one device run."""
    from specfwat_amd import _lib
    from specfwat_amd.batch import EventBatch, batch_device_bytes, plan_batches, tune_batch
    from specfwat_amd.mesh import BoxMesh
    from specfwat_amd.misfit import event_adjoint_sources
    from specfwat_amd.model import hti_model, perturbed_model
    from specfwat_amd.smooth import smooth_batch
    from specfwat_amd.solver import ElasticRun
    from specfwat_amd.sources import Receivers, Sources, surface_grid
    from specfwat_amd.tune import tune_element_passes

    w = workload("c4")
    K = a.steps or 8000
    dt = w["dt"]
    t0 = time.time()
    mesh = BoxMesh(w["nex"], w["ney"], w["nez"], w["lx"], w["ly"], w["lz"])
    true_model = perturbed_model(mesh, seed=2024)
    smooth_model = perturbed_model(mesh, smooth=True)
    cij_true = hti_model(mesh, *true_model, seed=77)
    cij_smooth = hti_model(mesh, *smooth_model, seed=78, max_strength=0.0)
    rng = np.random.Generator(np.random.Philox(5))
    srcs = []
    for _ in range(a.c5_events):
        sxyz = [[rng.uniform(0.2, 0.8) * w["lx"], rng.uniform(0.2, 0.8) * w["ly"], -rng.uniform(2000.0, 18000.0)]]
        srcs.append(Sources(mesh, sxyz, [list(1e15 * rng.standard_normal(6))], [rng.uniform(0.5, 1.5)]))
    rec = Receivers(mesh, surface_grid(mesh, w["nrx"], w["nrx"]))
    nsteps = [K] * len(srcs)
    log(f"c5: mesh {mesh.nspec} el, {mesh.nglob} pts, HTI c_ij ready, {len(srcs)} events of {K} steps "
        f"({time.time() - t0:.1f}s)")
    free_b = __import__("torch").cuda.mem_get_info()[0] if a.hbm_query else 288e9
    # the anisotropic batch adds the 21 fp64 kernels and their fp32 partial sums per event
    per_aniso = lambda e: batch_device_bytes(mesh, e, K, a.c4_store) + e * mesh.nspec * 125 * 21 * 12
    fits = [e for e in (1, 2, 4, 8) if per_aniso(e) <= 0.85 * free_b and e <= len(srcs)]
    bpc, ttable = 8, []
    if a.events_per_batch:
        E = a.events_per_batch
    else:
        (E, bpc), ttable = tune_batch(mesh, smooth_model, srcs, rec, dt, sizes=tuple(fits), cij=cij_smooth)
        log("batch tuner (events, blocks/CU, fwd ms/event-step, adj ms/event-step): "
            + ", ".join(f"({e},{b},{x:.3f},{y:.3f})" for e, b, x, y in ttable) + f" -> E={E} bpc={bpc}")
    groups = plan_batches(nsteps, E)
    log(f"c5: {E} events per batch ({len(groups)} batches; sizes fitting the HBM: {fits})")

    obs = [None] * len(srcs)
    for gi, g in enumerate(groups):
        b = EventBatch(mesh, *true_model, [srcs[i] for i in g], rec, dt, [nsteps[i] for i in g], cij=cij_true)
        b.set_store_precision(a.c4_store)
        b.forward()
        sy = b.seismograms()
        for k, i in enumerate(g):
            obs[i] = sy[k].copy()
        b.close()
        log(f"observed data batch {gi + 1}/{len(groups)} ({time.time() - t0:.1f}s)")
    log(f"observed data ({len(srcs)} events, HTI true model) ({time.time() - t0:.1f}s)")

    bwall, times, bkern, misfits = 0.0, None, {}, []
    check = set(groups[0][:a.c5_plain_events])
    for gi, g in enumerate(groups):
        b = EventBatch(mesh, *smooth_model, [srcs[i] for i in g], rec, dt, [nsteps[i] for i in g], cij=cij_smooth)
        b.set_store_precision(a.c4_store)
        b.set_tuning(bpc)
        _lib.call("fwat_synchronize_", b.run.h)
        ts = time.time()
        b.forward(timed=True)
        ft = b.pass_times()
        log(f"batch {gi + 1}/{len(groups)} forward: {time.time() - ts:.1f}s")
        syn = b.seismograms()
        adj = np.zeros_like(syn)
        for k, i in enumerate(g):
            m_, adj[k] = event_adjoint_sources(obs[i], syn[k], dt)
            misfits.append(m_)
        b.set_adjoint_sources(adj)
        b.adjoint(timed=True)
        _lib.call("fwat_synchronize_", b.run.h)
        bwall += time.time() - ts
        log(f"batch {gi + 1}/{len(groups)} forward + adjoint: {time.time() - ts:.1f}s")
        at = b.pass_times()
        times = times or (ft, at)
        if any(i in check for i in g):
            k_ = b.kernels()
            for k, i in enumerate(g):
                if i in check:
                    bkern[i] = {q: v[k] for q, v in k_.items()}
        b.close()
    total = mesh.nglob * float(sum(nsteps))
    bvalue = total / bwall
    log(f"batched: {bwall:.2f}s -> {bvalue:.4e} GLL-pt-steps/s/event")

    # plain single-event path on the first events (tuned launch shapes)
    tuning, _ = tune_element_passes(mesh, smooth_model, srcs[0], rec, dt)
    pwall, kerr, k_plain = 0.0, {}, None
    for i in sorted(check):
        f = ElasticRun(mesh, *smooth_model, srcs[i], rec, dt, K, 1, True, cij=cij_smooth)
        ad = ElasticRun(mesh, *smooth_model, srcs[i], rec, dt, K, 3, False, cij=cij_smooth)
        f.set_tuning(**tuning)
        ad.set_tuning(**tuning)
        _lib.call("fwat_synchronize_", f.h)
        ts = time.time()
        f.iterate_time(timed=True)
        syn = f.seismograms()
        _, adj = event_adjoint_sources(obs[i], syn, dt)
        ad.set_adjoint_sources(adj)
        ad.take_forward(f)
        ad.iterate_time(timed=True)
        _lib.call("fwat_synchronize_", ad.h)
        pwall += time.time() - ts
        log(f"plain event {i}: {time.time() - ts:.1f}s")
        fwd_ms, fwd_n = f.force_time()
        adj_ms, adj_n = ad.force_time()
        a_asm = ad.assembly_time()
        k_plain = ad.kernels()
        f.close(); ad.close()
        for q in ("rho_kl", "cijkl_kl"):
            kerr[q] = max(kerr.get(q, 0.0), rel_l2(bkern[i][q], k_plain[q]))
    pvalue = mesh.nglob * float(K * len(check)) / pwall
    log(f"plain: {pwall:.2f}s for {len(check)} events -> {pvalue:.4e}; batched/plain {bvalue / pvalue:.2f}")

    # smoothing of the 21 c_ij kernels of the last plain event (diagonal map on the all-regular mesh)
    fields = np.ascontiguousarray(np.moveaxis(k_plain["cijkl_kl"], -1, 0), np.float32)  # [21][nspec][125]
    ts = time.time()
    sm, ntstep, sm_elem_ms, sm_upd_ms = smooth_batch(mesh, fields, 5000.0, 2000.0, faithful=False, timed=True)
    sm_wall = time.time() - ts
    if not np.isfinite(sm).all():
        raise RuntimeError("non-finite smoothed kernels")
    log(f"smoothing 21 kernels: ntstep {ntstep}, {sm_wall:.2f}s (element {sm_elem_ms / 1e3:.2f}s, "
        f"gather/update {sm_upd_ms / 1e3:.2f}s)")

    nglob, nloc = mesh.nglob, mesh.nspec * 125
    nfp = mesh.num_abs_boundary_faces * 25
    ncontrib = nloc + nfp + 125 * (1 + rec.nrec)
    kacc = kernel_acc_bytes(KERNEL_BLOCK)
    e = len(groups[0])
    ftimes, atimes = times
    # batched adjoint element pass: ibool 4 + 21 c_ij 84 per local point once; per event: both
    # fields gathered 24 per unique point, forces written 24 and 21 c_ij accumulators 21 x kacc
    # per local point
    kinds = {
        f"k_batch_elem_aniso<{e},2>": (88.0 * nloc + e * (24.0 * nglob + (24.0 + 21 * kacc) * nloc),
                                       atimes[0], atimes[1]),
        f"k_batch_asm<{e},two-field>": (16.0 * nglob + 4.0 * (ncontrib + 125 * (e - 1))
                                        + e * (24.0 * nloc + 24.0 * nfp + (96.0 + kacc) * nglob),
                                        atimes[2], atimes[3]),
    }
    kname = max(kinds, key=lambda q: kinds[q][1])
    kbytes, kms, kn = kinds[kname]
    avg_ms = kms / max(kn, 1)
    achieved = kbytes / (avg_ms * 1e-3) / 1e9
    sm_elem_b = 8.0 * nloc + 96.0 * nglob + 96.0 * nloc
    sm_upd_b = 4.0 * nglob + 4.0 * nloc + 96.0 * nloc + 4.0 * nglob + 168.0 * nglob
    result = {
        "metric": "GLL-pt-steps/s/event fwd+adj",
        "value": bvalue,
        "unit": "GLL-pt-steps/s",
        "n_gpus": 1, "steps": int(sum(nsteps)), "warmup": 0,
        "ms_per_step": bwall / sum(nsteps) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": (f"C5: HTI anisotropic box {w['nex']}x{w['ney']}x{w['nez']} el, {nglob} GLL pts, "
                                f"21 c_ij (strength U[0,4%], azimuth per 10 km block), {len(srcs)} events batched "
                                f"{E} per launch, NSTEP={K}, dt={dt}, {rec.nrec} receivers, rho + 21 c_ij kernels in "
                                "fp64; then PDE smoothing of all 21 kernels of one event, sigma_h 5 km, sigma_v 2 km "
                                "(diagonal regular-element map)"),
                   "events_per_batch": E, "blocks_per_cu": bpc, "tuner": [list(r) for r in ttable],
                   "stacey_store": "fp16 + scale per (step, event)" if a.c4_store else "fp32"},
        "batched_vs_plain": {"plain_value": pvalue, "speedup": bvalue / pvalue, "plain_events_timed": len(check),
                             "kernel_rel_l2": kerr, "plain_element_pass_tuning": tuning},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": kbytes, "avg_launch_ms": avg_ms, "launches": kn},
        "smoothing": {"fields": 21, "ntstep": ntstep, "wall_s": sm_wall,
                      "GLL-pt-steps/s_per_field": nglob * ntstep * 21 / sm_wall,
                      "elem_avg_ms": sm_elem_ms / ntstep, "update_avg_ms": sm_upd_ms / ntstep,
                      "elem_GBps": sm_elem_b / (sm_elem_ms / ntstep * 1e-3) / 1e9,
                      "update_GBps": sm_upd_b / (sm_upd_ms / ntstep * 1e-3) / 1e9},
        "breakdown": {"batched_s": bwall, "plain_s": pwall,
                      "plain_fwd_force_avg_ms": fwd_ms / max(fwd_n, 1),
                      "plain_adj_force_avg_ms": adj_ms / max(adj_n, 1),
                      "plain_adj_assembly_avg_ms": a_asm[0] / max(a_asm[1], 1),
                      "kernels": {q: {"avg_ms": v[1] / max(v[2], 1),
                                      "GBps": v[0] / (v[1] / max(v[2], 1) * 1e-3) / 1e9 if v[1] > 0 else None}
                                  for q, v in kinds.items()}},
    }
    if not a.no_cpu:
        s0 = srcs[0]
        result["cpu_baseline"] = cpu_baseline(w, mesh, smooth_model, s0.xyz.tolist(), s0.mt.tolist(), rec,
                                              a.cpu_steps or 3, f0=float(s0.f0[0]), label="C5 HTI event",
                                              nstep_event=K, cij=cij_smooth)
        result["cpu_baseline"]["smoothing"] = smoothing_cpu_baseline(mesh, fields[0], 5000.0, 2000.0, ntstep)
    print(json.dumps(result), flush=True)


def smoothing_cpu_baseline(mesh, field, sigma_h, sigma_v, ntstep, steps=3):
    """The smoother's loop restatement (one thread) on one field for a few
    diffusion steps, extrapolated to ntstep steps and 21 fields."""
    from oracle import oracle as orc
    arrays, dmin = mesh.smooth_arrays(), mesh.distance_min_glob()
    ts = time.time()
    orc.smooth_pde(arrays, np.ascontiguousarray(field).ravel(), sigma_h, sigma_v, dmin, faithful=False,
                   max_steps=steps)
    wall = time.time() - ts
    return {"threads": 1, "sample": f"one field, {steps} diffusion steps (set-up included)",
            "s_per_step_per_field": wall / steps, "extrapolated_21_fields_s": wall / steps * ntstep * 21}


def measure(a):
    """Adjoint-source measurements of one C2-sized event on the device: every
    trace (400 receivers x 3 components, 10000 samples at dt 0.004 s, a wavelet
    plus 1 % noise) through the
    waveform misfit, the CC, coda-correlation, exponentiated-phase and multitaper
    measurements over two windows, adjoint sources band-passed 0.4-10 s, and the
    window selection (sliding correlation at every sample).  Synthetic traces: a random
    wavelet per trace, the synthetic shifted and scaled."""
    from specfwat_amd import misfit
    nrec, nstep, dt = 400, a.steps or 10000, 0.004
    rng = np.random.Generator(np.random.Philox(3))
    t = np.arange(nstep) * dt
    obs = np.zeros((nstep, nrec, 3), np.float32)
    syn = np.zeros_like(obs)
    f = rng.uniform(0.6, 1.4, (nrec, 3))
    t0 = rng.uniform(8, 30, (nrec, 3))
    for r in range(nrec):
        for c in range(3):
            w = np.exp(-((t - t0[r, c]) * f[r, c]) ** 2) * np.sin(2 * np.pi * f[r, c] * (t - t0[r, c]))
            obs[:, r, c] = w
            syn[:, r, c] = np.roll(w, int(rng.integers(-60, 60))) * rng.uniform(0.7, 1.3)
    # background noise: no window is empty (an all-zero window makes the reference's
    # zero-lag correlation coefficient 0/0)
    obs += (0.01 * rng.standard_normal(obs.shape)).astype(np.float32)
    syn += (0.01 * rng.standard_normal(syn.shape)).astype(np.float32)
    win = [[2.0, 22.0], [15.0, 39.0]]
    stages = {}
    for label, fn in (("waveform", lambda: misfit.event_adjoint_sources(obs, syn, dt, win, min_period=0.4,
                                                                          max_period=10.0)),
                      ("cc_tt", lambda: misfit.event_cc_adjoint_sources(obs, syn, dt, win, imeas=11,
                                                                          min_period=0.4, max_period=10.0)),
                      ("cc_dlna", lambda: misfit.event_cc_adjoint_sources(obs, syn, dt, win, imeas=12,
                                                                            min_period=0.4, max_period=10.0)),
                      ("ccc", lambda: misfit.event_cc_adjoint_sources(obs, syn, dt, win, imeas=5, min_period=0.4,
                                                                      max_period=10.0)),
                      ("exp_phase", lambda: misfit.event_ep_adjoint_sources(obs, syn, dt, win, min_period=0.4,
                                                                            max_period=10.0)),
                      ("mt_tt", lambda: misfit.event_mt_adjoint_sources(obs, syn, dt, win, 0.4, 10.0, imeas=13)),
                      ("mt_dlna", lambda: misfit.event_mt_adjoint_sources(obs, syn, dt, win, 0.4, 10.0,
                                                                          imeas=14))):
        fn()  # warm-up
        ts = time.time()
        out = fn()
        stages[label] = {"wall_s": time.time() - ts, "device_ms": misfit.last_time_ms(), "misfit": out[0]}
        if label.startswith("mt"):
            used = out[2][:, :, 8]
            stages[label]["windows_by_measurement"] = {str(int(k)): int((used == k).sum()) for k in np.unique(used)}
        if not np.isfinite(out[1]).all():
            raise RuntimeError(f"non-finite {label} adjoint sources")
    # window selection (WIN_SELECTOR_TYPE) of the same traces: first arrival 3 s before each wavelet
    tp = t0[:, 0] - 3.0
    dist = 2.4 * (t0[:, 0] + 8.0)
    misfit.select_windows(obs, syn, dt, 0.0, tp, dist, 0.4)  # warm-up
    ts = time.time()
    wins = misfit.select_windows(obs, syn, dt, 0.0, tp, dist, 0.4)
    stages["win_sel"] = {"wall_s": time.time() - ts, "device_ms": misfit.last_time_ms(),
                         "windows": int(sum(len(w) for w in wins))}
    # FLEXWIN (WIN_TYPE 2) on traces of its usual band: 1200 traces of 4000 samples at dt 0.05 s
    # (three arrivals each, 1 % noise), periods 2-20 s; merging and interval scheduling
    fw_dt, fw_n = 0.05, 4000
    tf = np.arange(fw_n) * fw_dt
    fobs = np.zeros((fw_n, nrec, 3), np.float32)
    fsyn = np.zeros_like(fobs)
    ftp = rng.uniform(20, 30, nrec)
    fdist = rng.uniform(200, 260, nrec)
    for r in range(nrec):
        for c in range(3):
            arr = ftp[r] + np.array([5.0, 28.0, 62.0]) + rng.uniform(-2, 2, 3)
            fq, am = rng.uniform(0.18, 0.32, 3), rng.uniform(0.6, 1.6, 3)
            for k in range(3):
                fsyn[:, r, c] += am[k] * np.exp(-((tf - arr[k]) * fq[k] / 2) ** 2) * np.sin(2 * np.pi * fq[k] * (tf - arr[k]))
                sh = arr[k] + rng.uniform(-0.6, 0.6)
                fobs[:, r, c] += am[k] * rng.uniform(0.8, 1.25) * np.exp(-((tf - sh) * fq[k] / 2) ** 2) * \
                    np.sin(2 * np.pi * fq[k] * (tf - sh))
    fobs += (0.01 * rng.standard_normal(fobs.shape)).astype(np.float32)
    for label, split in (("flexwin_merge", False), ("flexwin_split", True)):
        misfit.flexwin_windows(fobs, fsyn, fw_dt, 0.0, ftp, fdist, 2.0, 20.0, is_split_phases=split)  # warm-up
        ts = time.time()
        fw, _ = misfit.flexwin_windows(fobs, fsyn, fw_dt, 0.0, ftp, fdist, 2.0, 20.0, is_split_phases=split)
        stages[label] = {"wall_s": time.time() - ts, "device_ms": misfit.last_time_ms(),
                         "windows": int(sum(len(w) for w in fw)),
                         "traces": f"{nrec * 3} x {fw_n} samples, dt {fw_dt}, periods 2-20 s"}
    log("measure: " + ", ".join(f"{k} {v['device_ms']:.1f} ms" for k, v in stages.items()))
    ntr = nrec * 3
    dev = stages["cc_tt"]["device_ms"]
    result = {
        "metric": "traces/s (CC traveltime measurement with adjoint sources)",
        "value": ntr / (dev * 1e-3), "unit": "traces/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": dev, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": (f"{nrec} receivers x 3 components x {nstep} samples (dt {dt}), windows {win}, "
                                "adjoint sources band-passed 0.4-10 s")},
        "roofline": None,
        "stages": stages,
    }
    if not a.no_cpu:
        from oracle import oracle as orc
        sample = 300
        ts = time.time()
        for q in range(sample):
            r, c = divmod(q, 3)
            _, adj_o, _ = orc.cc_adjoint(obs[:, r, c], syn[:, r, c], dt, win, imeas=11)
            orc.filter_adj(adj_o, dt, np.float32(1 / 10.0), np.float32(1 / 0.4), win)
        wall = time.time() - ts
        result["cpu_baseline"] = {"value": sample / wall, "unit": "traces/s", "cores": 1, "kind": "port",
                                  "sample": f"{sample} traces of the same event through the oracle's CC-TT "
                                            f"measurement + filter_adj (C++, one thread), {wall:.1f}s"}
        sample = 36
        ts = time.time()
        for q in range(sample):
            r, c = divmod(q, 3)
            _, adj_o, _ = orc.mt_adjoint(obs[:, r, c], syn[:, r, c], dt, win, 0.4, 10.0, imeas=13)
            orc.filter_adj(adj_o, dt, np.float32(1 / 10.0), np.float32(1 / 0.4), win)
        wall = time.time() - ts
        stages["mt_tt"]["cpu_traces_per_s"] = sample / wall
        stages["mt_tt"]["cpu_sample"] = f"{sample} traces through the oracle's MT-TT + filter_adj, one thread, " \
                                        f"{wall:.1f}s"
        sample = 3
        ts = time.time()
        for q in range(sample):
            r, c = divmod(q, 3)
            orc.win_sel(obs[:, r, c], syn[:, r, c], dt, 0.0, tp[r], dist[r], 0.4)
        wall = time.time() - ts
        stages["win_sel"]["cpu_traces_per_s"] = sample / wall
        stages["win_sel"]["cpu_sample"] = f"{sample} traces through the oracle's window selection, one thread, " \
                                          f"{wall:.1f}s"
        for label, split in (("flexwin_merge", 0.0), ("flexwin_split", 1.0)):
            sample = 120
            ts = time.time()
            for q in range(sample):
                r, c = divmod(q, 3)
                orc.flexwin(fobs[:, r, c], fsyn[:, r, c], fw_dt, 0.0, ftp[r], fdist[r], 2.0, 20.0,
                            is_split_phases=split)
            wall = time.time() - ts
            stages[label]["cpu_traces_per_s"] = sample / wall
            stages[label]["cpu_sample"] = f"{sample} traces through the oracle's FLEXWIN, one thread, {wall:.1f}s"
    print(json.dumps(result), flush=True)


def post_chain(a):
    """xfwat_post_proc on the 64x64x32 mesh: E events' isotropic kernels summed
    in fp64 straight from a finished adjoint run in HBM (the run's kernels are
    added E times: the pass costs the same whatever the values), then four events
    through the EKERNEL-file path, the Hessian preconditioner, PDE smoothing of
    the three kernels, projection onto a 250 m inversion grid and the taper."""
    from specfwat_amd import optimize, postproc
    from specfwat_amd.mesh import BoxMesh
    from specfwat_amd.model import perturbed_model
    from specfwat_amd.solver import ElasticRun
    from specfwat_amd.sources import Receivers, Sources, surface_grid

    w = workload("c4")
    E = a.post_events
    t0 = time.time()
    mesh = BoxMesh(w["nex"], w["ney"], w["nez"], w["lx"], w["ly"], w["lz"])
    model = perturbed_model(mesh, smooth=True)
    rng = np.random.Generator(np.random.Philox(9))
    src = Sources(mesh, [[0.4 * w["lx"], 0.55 * w["ly"], -8000.0]], [list(1e15 * rng.standard_normal(6))], [1.0])
    rec = Receivers(mesh, surface_grid(mesh, w["nrx"], w["nrx"]))
    nstep = 200
    f = ElasticRun(mesh, *model, src, rec, w["dt"], nstep, 1, True)
    f.iterate_time()
    syn = f.seismograms()
    ad = ElasticRun(mesh, *model, src, rec, w["dt"], nstep, 3, False, approximate_hess=True)
    ad.take_forward(f)
    f.close()
    ad.set_adjoint_sources(np.ascontiguousarray(syn[::-1], np.float32))
    ad.iterate_time()
    log(f"post: mesh {mesh.nspec} el, adjoint run with Hessian ready ({time.time() - t0:.1f}s)")
    nloc, nglob = mesh.nspec * 125, mesh.nglob
    stages = {}
    pf = postproc.PostFlow(mesh, 1, precond_type=1)
    pf.add_event(ad)  # warm-up
    pf.close()
    pf = postproc.PostFlow(mesh, 1, precond_type=1)
    dev_ms = 0.0
    ts = time.time()
    for _ in range(E):
        pf.add_event(ad)
        dev_ms += pf.last_time_ms()
    stages["sum_device_runs"] = {"events": E, "wall_s": time.time() - ts, "device_ms_per_event": dev_ms / E}
    ad.close()
    # EKERNEL-file path: single-precision arrays in host memory, four events
    files = [{n: rng.standard_normal((mesh.nspec, 125), dtype=np.float32) for n in postproc.KERNEL_ISO}
             for _ in range(4)]
    hfiles = [np.abs(rng.standard_normal((mesh.nspec, 125), dtype=np.float32)) for _ in range(4)]
    pf2 = postproc.PostFlow(mesh, 1, precond_type=1)
    ts = time.time()
    fms = 0.0
    for e, h in zip(files, hfiles):
        pf2.add_event_files(e, h)
        fms += pf2.last_time_ms()
    stages["sum_event_files"] = {"events": 4, "wall_s": time.time() - ts, "device_ms_per_event": fms / 4,
                                 "note": "host-to-device copy of 4 x 66 MB per event dominates the wall time"}
    pf2.close()
    ts = time.time()
    pf.invert_hess()
    inv_ms = pf.last_time_ms()
    kp = pf.apply_precond()
    stages["precond"] = {"wall_s": time.time() - ts, "invert_device_ms": inv_ms, "apply_device_ms": pf.last_time_ms()}
    grid = postproc.RegularGrid((0.0, 0.0, -w["lz"]), (250.0, 250.0, 250.0),
                                (int(w["lx"] / 250) + 1, int(w["ly"] / 250) + 1, int(w["lz"] / 250) + 1))
    ts = time.time()
    gp = postproc.GridProjection(mesh, grid)
    stages["gll2grid_prepare"] = {"wall_s": time.time() - ts, "device_ms": gp.last_time_ms(),
                                  "grid": [grid.nx, grid.ny, grid.nz], "nb_fd_point": gp.nb_fd_point}
    pf.mesh_extent()  # check_mesh_distances: set-up, not part of the timed stages
    ts = time.time()
    smoothed = pf.pde_smooth(5000.0, 2000.0, gp, kernels=kp, faithful=False, timed=True)
    nt, sm_e, sm_u = pf.last_smooth
    stages["pde_smooth_and_project"] = {"wall_s": time.time() - ts, "ntstep": nt,
                                        "smooth_device_ms": sm_e + sm_u, "project_device_ms": gp.last_time_ms()}
    ts = time.time()
    tapered = pf.taper_kernel_grid(smoothed, grid, (2000.0, 5000.0, 2000.0, 4000.0))
    stages["taper"] = {"wall_s": time.time() - ts}
    if not all(np.isfinite(v).all() for v in tapered.values()):
        raise RuntimeError("non-finite gradient grid")
    gp.close()
    pf.close()
    # optimiser: L-BFGS direction from 5 stored iterations of the gradient grid, model update
    grads, models = _opt_history(tapered, grid, 5)
    hgrid = np.ones_like(grads[0])
    ts = time.time()
    direction, angle = optimize.get_lbfgs_direction(grads, models, hgrid, spacing=(grid.hx, grid.hy, grid.hz))
    lb_ms = optimize.last_time_ms()
    new_model = optimize.model_update(models[-1], direction, 1, 0.02, (1.5, 2.0))
    stages["lbfgs_direction_and_update"] = {"wall_s": time.time() - ts, "lbfgs_device_ms": lb_ms,
                                            "update_device_ms": optimize.last_time_ms(), "nstore": 5,
                                            "angle_deg": angle}
    if not np.isfinite(new_model).all():
        raise RuntimeError("non-finite updated model")
    log("post stages: " + ", ".join(f"{k} {v['wall_s']:.3f}s" for k, v in stages.items()))
    per = stages["sum_device_runs"]["device_ms_per_event"]
    # per local point: ibool 4, mu/kappa kernels 16, rho/kappa/mu statics 12, three fp64 sums
    # read+write 48, Hessian sum read+write 16; per unique point the rho and Hessian
    # accumulators 16
    sum_bytes = 96.0 * nloc + 16.0 * nglob
    achieved = sum_bytes / (per * 1e-3) / 1e9
    value = nloc * E / (stages["sum_device_runs"]["wall_s"])
    result = {
        "metric": "kernel-point-events/s (post-processing event sums)",
        "value": value, "unit": "kernel-point-events/s",
        "n_gpus": 1, "steps": E, "warmup": 1, "ms_per_step": stages["sum_device_runs"]["wall_s"] / E * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": (f"post-processing on the {w['nex']}x{w['ney']}x{w['nez']} el mesh ({nloc} kernel "
                                f"points, {nglob} GLL pts): {E} isotropic events summed from HBM, precondition, "
                                "smooth 3 kernels (sigma 5/2 km), project onto a 250 m grid, taper")},
        "roofline": {"bound": "hbm", "kernel": "k_post_add_iso_run", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": sum_bytes, "avg_launch_ms": per, "launches": E},
        "stages": stages,
    }
    if not a.no_cpu:
        result["cpu_baseline"] = post_cpu_baseline(mesh, grid, files, hfiles, (grads, models, hgrid))
    print(json.dumps(result), flush=True)


def _opt_history(gradient, grid, nstore):
    """A synthetic optimisation history on the inversion grid: nstore + 1
    models (vp, vs, rho) and gradients, the last gradient the tapered one."""
    rng = np.random.Generator(np.random.Philox(11))
    g = np.stack([gradient[n] for n in ("alpha", "beta", "rhop")]).astype(np.float32)
    scale = np.abs(g).max() or 1.0
    base = np.stack([np.full(g.shape[1:], v, np.float32) for v in (6000.0, 3500.0, 2700.0)])
    models, grads = [base], [g / scale]
    for _ in range(nstore):
        models.append((models[-1] * np.exp(rng.normal(0, 0.01, g.shape))).astype(np.float32))
        grads.append((grads[-1] + rng.uniform(0.5, 2.0, g.shape) * np.log(models[-1] / models[-2])).astype(np.float32))
    return grads, models


def post_cpu_baseline(mesh, grid, files, hfiles, opt_hist):
    """The oracle's post-processing stages on the host (single thread, single
    precision like the reference): four events' sums, invert_hess and the
    preconditioner, gll2grid location + projection of the three kernels, taper,
    and the L-BFGS direction.  Smoothing is not part of the sample."""
    from oracle import oracle as orc
    from specfwat_amd import postproc
    # mesh distances and box: read_mesh_databases_for_init's set-up, not a timed stage
    dh = np.float32(np.float32(mesh.distance_max_glob()) + np.float32(mesh.distance_min_glob())) / np.float32(2)
    box = (mesh.xstore.min(), mesh.xstore.max(), mesh.ystore.min(), mesh.ystore.max(), mesh.zstore.min(),
           mesh.zstore.max())
    gx, gy, gz = grid.axes()
    ts = time.time()
    tot = {n: np.zeros((mesh.nspec, 125), np.float32) for n in postproc.KERNEL_ISO}
    hs = np.zeros((mesh.nspec, 125), np.float32)
    for e, h in zip(files, hfiles):
        for n in postproc.KERNEL_ISO:
            tot[n] += e[n]
        hs += np.abs(h)
    inv = orc.invert_hess(hs).reshape(mesh.nspec, 125)
    for n in tot:
        tot[n] *= inv
    t_sum = time.time() - ts
    t1 = time.time()
    owner, ref, _ = orc.gll2grid_coeff(mesh.ibool, mesh.xstore, mesh.ystore, mesh.zstore,
                                       (grid.ox, grid.oy, grid.oz, grid.hx, grid.hy, grid.hz, grid.nx, grid.ny,
                                        grid.nz))
    t_loc = time.time() - t1
    t1 = time.time()
    g = [orc.gll2grid_project(owner, ref, tot[n]) for n in tot]
    t_proj = time.time() - t1
    t1 = time.time()
    mask = orc.taper_grid(box, dh, 2000.0, 5000.0, 2000.0, 4000.0, gx, gy, gz)
    for q in range(len(g)):
        g[q] = g[q] * mask
    t_tap = time.time() - t1
    t1 = time.time()
    G, M, H = opt_hist
    orc.lbfgs_direction(np.stack(G).reshape(len(G), -1), np.stack(M).reshape(len(M), -1), H, grid.npts,
                        grid.hx, grid.hy, grid.hz, dbl=False)
    t_opt = time.time() - t1
    wall = time.time() - ts
    log(f"post cpu baseline: sums+precond {t_sum:.2f}s, locate {t_loc:.2f}s, project {t_proj:.2f}s, "
        f"taper {t_tap:.2f}s, lbfgs {t_opt:.2f}s ({wall:.1f}s)")
    return {"value": mesh.nspec * 125 * len(files) / t_sum, "unit": "kernel-point-events/s", "cores": 1,
            "kind": "port",
            "sample": (f"4 events' file kernels summed + preconditioned in single precision ({t_sum:.2f}s), "
                       f"gll2grid location {t_loc:.2f}s and projection of 3 kernels {t_proj:.2f}s, taper "
                       f"{t_tap:.2f}s, L-BFGS direction from 5 stored iterations {t_opt:.2f}s; oracle (C++) "
                       "and numpy, one thread; smoothing not sampled")}


def host_info():
    """Host threads available to this process and the CPU model name."""
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count()
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return threads, model


def cpu_baseline(w, mesh, model, sxyz, mt, rec, nsteps, f0=1.0, label="C2 event", nstep_event=None, cij=None,
                 one_thread_steps=1, repeats=3, force_irregular=False):
    """The oracle (C++/OpenMP FP32 restatement of the SPECFEM CPU time loop) on the
    host cores, on a bounded sample of the same event: nsteps forward steps with the
    Stacey store, then nsteps adjoint + reconstruction + kernel steps, each timed
    apart; the median of `repeats` runs on all host threads, plus one run of
    one_thread_steps on one thread.  The per-step times extrapolate linearly to
    the event's NSTEP (set-up excluded)."""
    from oracle import oracle as orc
    threads, cpu_model = host_info()
    rho, vp, vs = model
    from specfwat_amd.sources import surface_grid
    rxyz = surface_grid(mesh, w["nrx"], w["nrx"])

    def make(n, nthreads):
        return orc.Oracle(w["nex"], w["ney"], w["nez"], w["lx"], w["ly"], w["lz"], rho.ravel(), vp.ravel(),
                          vs.ravel(), cij=None if cij is None else cij.reshape(-1, 21), dt=w["dt"], nstep=n,
                          src_xyz=sxyz, src_mt=mt, src_f0=[f0], rec_xyz=rxyz, precision=32, nthreads=nthreads,
                          force_irregular=force_irregular)

    def timed_run(o):
        ts = time.time()
        s = o.forward(save=True)
        tf = time.time()
        o.adjoint(s.astype(np.float32))
        return tf - ts, time.time() - tf

    t0 = time.time()
    o = make(nsteps, threads)
    log(f"cpu baseline set-up {time.time() - t0:.1f}s ({threads} threads, {cpu_model})")
    runs = [timed_run(o) for _ in range(repeats)]
    del o
    fwd = float(np.median([r[0] for r in runs]))
    adj = float(np.median([r[1] for r in runs]))
    log(f"cpu baseline {nsteps} fwd + {nsteps} adj steps, median of {repeats}: forward {fwd:.2f}s, "
        f"adjoint+kernels {adj:.2f}s")
    o1 = make(one_thread_steps, 1)
    f1, a1 = timed_run(o1)
    del o1
    log(f"cpu baseline one thread, {one_thread_steps} fwd + adj steps: forward {f1:.2f}s, adjoint {a1:.2f}s")
    ne = nstep_event or nsteps
    return {"value": mesh.nglob * nsteps / (fwd + adj), "unit": "GLL-pt-steps/s", "cores": threads,
            "kind": "port",
            "sample": (f"same {label} from rest, {nsteps} forward + {nsteps} adjoint/reconstruction/kernel "
                       f"steps, FP32 oracle (C++/OpenMP, colour-ordered assembly), median of {repeats}"),
            "host": {"threads": threads, "cpu_model": cpu_model},
            "forward_s_per_step": fwd / nsteps, "adjoint_kernels_s_per_step": adj / nsteps,
            "one_thread": {"forward_s_per_step": f1 / one_thread_steps,
                           "adjoint_kernels_s_per_step": a1 / one_thread_steps,
                           "value": mesh.nglob * one_thread_steps / (f1 + a1)},
            "event_extrapolated_s": {"nstep": ne, "forward": fwd / nsteps * ne, "adjoint_kernels": adj / nsteps * ne,
                                     "note": "linear in NSTEP from the sample, set-up excluded"}}


if __name__ == "__main__":
    main()
