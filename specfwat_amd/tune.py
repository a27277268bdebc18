"""Autotuner of the element-pass launch shape.

Times short forward and adjoint runs on the real mesh for a set of
candidate shapes (general kernel, folded-table kernel with one workgroup
per element, persistent workgroups per CU) and returns the fastest; the
caller applies it with ElasticRun.set_tuning.  The measurement is the
live HIP-event time of the element passes on the simulation stream.
"""
import numpy as np

from .solver import ElasticRun

CANDIDATES = [(0, 0), (1, 0), (1, 8), (1, 10), (1, 16), (1, 20), (1, 32)]


def tune_element_passes(mesh, model, sources, receivers, dt, nstep=20, candidates=CANDIDATES,
                        static_precision=0):
    """Returns ({'regular_path', 'blocks_per_cu_fwd', 'blocks_per_cu_adj'}, table)."""
    rho, vp, vs = model
    f = ElasticRun(mesh, rho, vp, vs, sources, receivers, dt, nstep, 1, True,
                   static_precision=static_precision)
    a = ElasticRun(mesh, rho, vp, vs, sources, receivers, dt, nstep, 3, False,
                   static_precision=static_precision)
    a.set_adjoint_sources(np.zeros((nstep, receivers.nrec, 3), np.float32))
    table = []
    f.iterate_time()  # warm-up, also fills the Stacey store for the adjoint runs
    a.take_forward(f)
    for reg, bpc in candidates:
        f.set_tuning(reg, bpc, bpc)
        a.set_tuning(reg, bpc, bpc)
        f.iterate_time(timed=True)
        fms, fn = f.force_time()
        a.iterate_time(timed=True)
        ams, an = a.force_time()
        table.append((reg, bpc, fms / max(fn, 1), ams / max(an, 1)))
    f.close()
    a.close()
    best_f = min(table, key=lambda r: r[2])
    best_a = min(table, key=lambda r: r[3])
    # one regular_path switch for both passes: keep the better total
    reg = 1 if (best_f[0] + best_a[0]) >= 1 else 0
    bf = min((r for r in table if r[0] == reg), key=lambda r: r[2])[1]
    ba = min((r for r in table if r[0] == reg), key=lambda r: r[3])[1]
    return {"regular_path": reg, "blocks_per_cu_fwd": bf, "blocks_per_cu_adj": ba}, table
