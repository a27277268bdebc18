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


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10000)
    p.add_argument("--warmup", type=int, default=100)
    p.add_argument("--config", choices=["c2", "c1"], default="c2")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-steps", type=int, default=4)
    p.add_argument("--unfused", action="store_true", help="time the SPECFEM-order stage sequence")
    p.add_argument("--static-precision", type=int, default=0, help="0 fp32, 1 bf16, 2 fp16")
    p.add_argument("--no-tune", action="store_true", help="skip the element-pass launch-shape autotuner")
    p.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                   help="only run a plain device copy of BYTES (counter calibration) and exit")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(cfg):
    if cfg == "c2":
        return dict(nex=96, ney=96, nez=48, lx=40000.0, ly=40000.0, lz=20000.0, dt=0.004, nrx=20)
    return dict(nex=16, ney=16, nez=8, lx=40000.0, ly=40000.0, lz=20000.0, dt=0.005, nrx=4)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the event-parallel multi-GPU design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    from specfwat_amd import _lib
    if a.copy_probe:
        ms = np.zeros(1)
        _lib.call("fwat_copy_probe_", a.copy_probe, 5, ms)
        print(json.dumps({"copy_probe_bytes": a.copy_probe, "avg_ms": float(ms[0]),
                          "GBps_read_plus_write": 2 * a.copy_probe / (ms[0] * 1e-3) / 1e9}), flush=True)
        return
    from specfwat_amd.mesh import BoxMesh
    from specfwat_amd.model import perturbed_model
    from specfwat_amd.sources import Receivers, Sources, surface_grid
    from specfwat_amd.solver import ElasticRun
    from specfwat_amd.misfit import event_adjoint_sources

    w = workload(a.config)
    K, W = a.steps, a.warmup
    t0 = time.time()
    mesh = BoxMesh(w["nex"], w["ney"], w["nez"], w["lx"], w["ly"], w["lz"])
    log(f"mesh {mesh.nspec} elements, {mesh.nglob} points ({time.time() - t0:.1f}s)")
    true_model = perturbed_model(mesh, seed=2024)
    smooth_model = perturbed_model(mesh, smooth=True)
    rng = np.random.Generator(np.random.Philox(7))
    sxyz = [[rng.uniform(0.3, 0.7) * w["lx"], rng.uniform(0.3, 0.7) * w["ly"], -10000.0]]
    mt = [list(1e15 * rng.standard_normal(6))]
    src = Sources(mesh, sxyz, mt, [1.0])
    rec = Receivers(mesh, surface_grid(mesh, w["nrx"], w["nrx"]))
    dt = w["dt"]
    log(f"model/source/receivers ready ({time.time() - t0:.1f}s)")
    fused = not a.unfused
    tuning = {"regular_path": 1, "blocks_per_cu_fwd": 0, "blocks_per_cu_adj": 8}
    if fused and not a.no_tune:
        from specfwat_amd.tune import tune_element_passes
        tuning, table = tune_element_passes(mesh, smooth_model, src, rec, dt,
                                            static_precision=a.static_precision)
        log("autotuner (regular_path, blocks/CU, fwd ms, adj ms): "
            + ", ".join(f"({r},{b},{x:.3f},{y:.3f})" for r, b, x, y in table) + f" -> {tuning}")

    def observed(nstep):
        r = ElasticRun(mesh, *true_model, src, rec, dt, nstep, 1, False)
        r.iterate_time()
        obs = r.seismograms()
        r.close()
        return obs

    def prepare(nstep):
        f = ElasticRun(mesh, *smooth_model, src, rec, dt, nstep, 1, True,
                       static_precision=a.static_precision)
        ad = ElasticRun(mesh, *smooth_model, src, rec, dt, nstep, 3, False,
                        compute_and_store_strain=not fused, static_precision=a.static_precision)
        f.set_tuning(**tuning)
        ad.set_tuning(**tuning)
        return f, ad

    def event(f, ad, obs, timed):
        f.iterate_time(fused=fused, timed=timed)
        tf = time.time()
        syn = f.seismograms()
        misfit, adj = event_adjoint_sources(obs, syn, dt)
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
    obs = observed(K)
    log(f"observed data ({K} steps) generated ({time.time() - t0:.1f}s)")
    f, ad = prepare(K)
    _lib.call("fwat_synchronize_", f.h)
    _lib.call("fwat_synchronize_", ad.h)
    ts = time.time()
    misfit, tf = event(f, ad, obs, True)
    _lib.call("fwat_synchronize_", ad.h)
    te = time.time()
    wall = te - ts
    fwd_ms, fwd_n = f.force_time()
    adj_ms, adj_n = ad.force_time()
    f_asm, a_asm = f.assembly_time(), ad.assembly_time()
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
    if fused:
        # two-field element pass: ibool 4 + kappa,mu 8 (+ 40 irregular Jacobian) per local
        # point, fp64 mu/kappa accumulators read+write 32, element-local forces written
        # for both fields 24; displacement gather of both fields 12 + 12 per unique point
        kinds = {
            "k_forces<2,iso,fused-kernels>": ((60.0 + kmu) * nloc + 40.0 * nirr + 24.0 * nglob, adj_ms, adj_n),
            # two-field assembly: CSR offsets 4 per point + indices 4 per contribution,
            # element-local forces of both fields 24 per local point, Stacey terms of both
            # fields 24 per face point, rmass_inv 12, u,v of both fields read 48 + write 48,
            # fp64 rho kernel read+write 16 per unique point
            "k_assemble<two-field>": (4.0 * nglob + 4.0 * ncontrib + 24.0 * nloc + 24.0 * nfp
                                      + 12.0 * nglob + 96.0 * nglob + 16.0 * nglob, adj_asm_ms, adj_asm_n),
        }
        fwd_bytes = (16.0 + kmu) * nloc + 40.0 * nirr + 12.0 * nglob
        fwd_asm_bytes = 4.0 * nglob + 4.0 * ncontrib + 12.0 * nloc + 12.0 * nfp + 12.0 * nglob + 48.0 * nglob
    else:
        # SPECFEM stage order: one field per launch, atomic scatter, strain store
        kinds = {"k_forces<1,iso,strain-store,atomic>": (12.0 * nloc + 24.0 * nloc + 40.0 * nirr + 36.0 * nglob,
                                                         adj_ms, adj_n)}
        fwd_bytes = 12.0 * nloc + 40.0 * nirr + 36.0 * nglob
        fwd_asm_bytes = 0.0
    kname = max(kinds, key=lambda k: kinds[k][1])
    kbytes, kms, kn = kinds[kname]
    avg_ms = kms / max(kn, 1)
    achieved = kbytes / (avg_ms * 1e-3) / 1e9
    traffic = None
    tf_path = os.path.join(HERE, "profiles", "c2_fwd_adj_hbm_traffic.md")
    if os.path.exists(tf_path) and fused and a.config == "c2" and a.static_precision == 0:
        for line in open(tf_path):
            if line.startswith("hbm_bytes_per_launch " + kname):
                traffic = float(line.split("=")[1])
    fwd_avg_ms = fwd_ms / max(fwd_n, 1)
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
                                f"{w['nex']}x{w['ney']}x{w['nez']} el, {nglob} GLL pts, NGLL=5, "
                                f"{rec.nrec} receivers, dt={dt}, NSTEP={K}, fwd + waveform misfit + "
                                "adjoint/reconstruction + rho/mu/kappa kernels (fp64 accum)"),
                   "path": "fused" if fused else "SPECFEM stage order (strain stores)",
                   "element_pass_tuning": tuning if fused else None},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": kbytes, "avg_launch_ms": avg_ms, "launches": kn},
        "breakdown": {"wall_s": wall, "forward_s": tf - ts, "adjoint_s": te - tf,
                      "kernels": {k: {"avg_ms": v[1] / max(v[2], 1),
                                      "GBps": v[0] / (v[1] / max(v[2], 1) * 1e-3) / 1e9 if v[1] > 0 else None}
                                  for k, v in kinds.items()},
                      "fwd_force_avg_ms": fwd_avg_ms,
                      "fwd_force_GBps": fwd_bytes / (fwd_avg_ms * 1e-3) / 1e9 if fwd_avg_ms > 0 else None,
                      "fwd_assembly_avg_ms": fwd_asm_ms / max(fwd_asm_n, 1),
                      "fwd_assembly_GBps": (fwd_asm_bytes / (fwd_asm_ms / fwd_asm_n * 1e-3) / 1e9
                                            if fwd_asm_n and fwd_asm_ms > 0 else None)},
    }
    if not a.no_cpu:
        result["cpu_baseline"] = cpu_baseline(w, mesh, smooth_model, sxyz, mt, rec, a.cpu_steps)
    print(json.dumps(result), flush=True)


def cpu_baseline(w, mesh, model, sxyz, mt, rec, nsteps):
    """The oracle (C++/OpenMP FP32 restatement) on the host cores, bounded
    sample: the same mesh/model/source/receivers, nsteps forward steps with
    the Stacey store then nsteps adjoint+reconstruction+kernel steps."""
    from oracle import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    rho, vp, vs = model
    rxyz = None
    from specfwat_amd.sources import surface_grid
    rxyz = surface_grid(mesh, w["nrx"], w["nrx"])
    t0 = time.time()
    o = orc.Oracle(w["nex"], w["ney"], w["nez"], w["lx"], w["ly"], w["lz"], rho.ravel(), vp.ravel(),
                   vs.ravel(), dt=w["dt"], nstep=nsteps, src_xyz=sxyz, src_mt=mt, src_f0=[1.0],
                   rec_xyz=rxyz, precision=32, nthreads=threads)
    log(f"cpu baseline set-up {time.time() - t0:.1f}s ({threads} threads)")
    ts = time.time()
    s = o.forward(save=True)
    o.adjoint(s.astype(np.float32))
    wall = time.time() - ts
    log(f"cpu baseline {nsteps} fwd+adj steps: {wall:.2f}s")
    return {"value": mesh.nglob * nsteps / wall, "unit": "GLL-pt-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"same C2 event, {nsteps} forward + {nsteps} adjoint/reconstruction/kernel steps, "
                      f"FP32 oracle (C++/OpenMP, colour-ordered assembly), {wall:.1f}s"}


if __name__ == "__main__":
    main()
