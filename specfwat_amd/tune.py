"""Autotuner of the element-pass launch shape.

Times short forward and adjoint runs on the real mesh for a set of
candidate shapes (general kernel, folded-table kernel with one workgroup
per element, persistent workgroups per CU, its software-pipelined forms,
element order or column blocks) and returns the fastest; for the pipelined
passes it then compares element pass + assembly with and without the
element-interior points finished in the element pass (at several workgroup
counts); on lattice meshes it compares the brick-fused element + assembly
pass (brick shapes 2x2x2, 4x2x2, 4x4x2) against the best gather shapes, and
on any mesh element slabs.  The caller applies the result with
ElasticRun.set_tuning.  The measurement is the live HIP-event time of the
passes on the simulation stream.
"""
import numpy as np

from .solver import ElasticRun

# (regular_path, workgroups per CU): 0 general kernel, 1 folded-table kernel, 2 its
# software-pipelined form, 3 the pipelined form with two barriers per element (isotropic),
# 5 form 3 visiting a lattice mesh's elements in column blocks (form 3 elsewhere);
# 0 workgroups per CU = one workgroup per element
CANDIDATES = [(0, 0), (1, 0), (1, 8), (1, 10), (1, 16), (1, 20), (1, 32), (2, 6), (2, 8), (2, 12), (2, 16),
              (3, 6), (3, 8), (3, 12), (3, 16), (5, 6), (5, 8), (5, 12), (5, 16)]
BRICKS = [(2, 2, 2), (4, 2, 2), (4, 4, 2)]
SLABS = [6, 12, 24]  # slabs per mesh
INTERIOR_BPC = [6, 7, 8, 9, 10, 12]  # workgroups per CU tried with the element-interior points in the pipelined pass


def tune_element_passes(mesh, model, sources, receivers, dt, nstep=20, candidates=CANDIDATES,
                        static_precision=0, bricks=BRICKS, slabs=SLABS, interior_bpc=INTERIOR_BPC):
    """Returns ({'regular_path', 'blocks_per_cu_fwd', 'blocks_per_cu_adj', 'interior'[, 'brick'
    | 'slab_elems']}, table); the dict is ElasticRun.set_tuning's keyword arguments."""
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
        # element passes alone, every point through the gather (interior points below)
        f.set_tuning(reg, bpc, bpc, interior=False)
        a.set_tuning(reg, bpc, bpc, interior=False)
        f.iterate_time(timed=True)
        fms, fn = f.force_time()
        a.iterate_time(timed=True)
        ams, an = a.force_time()
        table.append((reg, bpc, fms / max(fn, 1), ams / max(an, 1)))
    # one regular_path switch for both passes: the value with the best forward + adjoint total
    regs = sorted({r[0] for r in table})
    reg = min(regs, key=lambda g: min(r[2] for r in table if r[0] == g) + min(r[3] for r in table if r[0] == g))
    bf = min((r for r in table if r[0] == reg), key=lambda r: r[2])[1]
    ba = min((r for r in table if r[0] == reg), key=lambda r: r[3])[1]
    best = {"regular_path": reg, "blocks_per_cu_fwd": bf, "blocks_per_cu_adj": ba, "interior": False}

    def step_ms(**kw):
        # element + assembly time per step of the forward and the adjoint pass
        f.set_tuning(**kw)
        a.set_tuning(**kw)
        f.iterate_time(timed=True)
        fe, fa = f.force_time(), f.assembly_time()
        a.iterate_time(timed=True)
        ae, aa = a.force_time(), a.assembly_time()
        return (fe[0] + fa[0]) / max(fe[1], 1), (ae[0] + aa[0]) / max(ae[1], 1)

    # the pipelined passes (3, 5) can finish the element-interior points themselves: the
    # element pass grows, the gather shrinks, so whole element + assembly time decides,
    # against the best shape above whatever its regular_path
    pipe = [g for g in (3, 5) if g in regs] if mesh.nspec_irregular == 0 else []
    if pipe:
        ireg = min(pipe, key=lambda g: min(r[2] for r in table if r[0] == g) + min(r[3] for r in table if r[0] == g))
        off = step_ms(**best)
        table.append(("interior-off", best["blocks_per_cu_fwd"], off[0], off[1]))
        # the pass with interior points holds more state: its own workgroup counts
        for bpc in interior_bpc:
            on = step_ms(regular_path=ireg, blocks_per_cu_fwd=bpc, blocks_per_cu_adj=bpc, interior=True)
            table.append(("interior-on", bpc, on[0], on[1]))
        # one interior switch for both passes: the total of the best forward and adjoint shapes
        ons = [r for r in table if r[0] == "interior-on"]
        if min(r[2] for r in ons) + min(r[3] for r in ons) < off[0] + off[1]:
            best = {"regular_path": ireg, "blocks_per_cu_fwd": min(ons, key=lambda r: r[2])[1],
                    "blocks_per_cu_adj": min(ons, key=lambda r: r[3])[1], "interior": True}
    # brick-fused pass (lattice meshes): element + assembly time per step against the
    # element + assembly time of the best gather shapes
    if bricks and f.lattice:
        base_f, base_a = step_ms(**best)
        table.append(("gather", best["blocks_per_cu_fwd"], base_f, base_a))
        top = (base_f + base_a, None)
        for br in bricks:
            m = mesh
            if m.nex % br[0] or m.ney % br[1] or m.nez % br[2]:
                continue
            tf, ta = step_ms(**best, brick=br)
            table.append(("brick%dx%dx%d" % br, 0, tf, ta))
            if tf + ta < top[0]:
                top = (tf + ta, br)
        if top[1] is not None:
            best = dict(best, brick=top[1])
    # element slabs (any mesh): element pass + assembly slab after slab (every point
    # through the gather)
    if slabs and "brick" not in best and not best["interior"]:
        def step_ms_s(nsl, **kw):
            # element + assembly time per step; with slabs every step has nsl launches of each
            f.set_tuning(**kw)
            a.set_tuning(**kw)
            f.iterate_time(timed=True)
            fe, fa = f.force_time(), f.assembly_time()
            a.iterate_time(timed=True)
            ae, aa = a.force_time(), a.assembly_time()
            return ((fe[0] + fa[0]) * nsl / max(fe[1], 1), (ae[0] + aa[0]) * nsl / max(ae[1], 1))
        base = step_ms_s(1, **best)
        table.append(("whole-mesh", 0, base[0], base[1]))
        top = (base[0] + base[1], 0)
        for nsl in slabs:
            c = -(-mesh.nspec // nsl)
            tf, ta = step_ms_s(-(-mesh.nspec // c), **best, slab_elems=c)
            table.append((f"slabs{nsl}", c, tf, ta))
            if tf + ta < top[0]:
                top = (tf + ta, c)
        if top[1]:
            best = dict(best, slab_elems=top[1])
    f.close()
    a.close()
    return best, table
