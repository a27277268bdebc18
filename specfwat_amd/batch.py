"""Multi-event batches on the GPU: E events advanced together on one mesh.

SpecFWAT runs the events of a set one after the other
(src/preproc/fwat_fwd_measure_adj.f90:43-51, each a PrepareFWD%simulation,
src/preproc/prepare_fwd.f90:297-326).  `EventBatch` runs the same sequence
(forward with SAVE_FORWARD -> waveform misfit and adjoint sources -> adjoint
run with reconstruction -> rho/mu/kappa kernels) for E = 1, 2, 4 or 8 events
at once through the fwat_batch_* entry points, so that one element pass
reads the mesh and model once for all events.  Events may have different
NSTEP; every event keeps its own source and source time function and shares
the receivers.  Regular (all-cube) meshes, isotropic or with 21 c_ij.
"""
import numpy as np

from . import _lib
from .solver import ElasticRun

F32 = np.float32
BATCH_SIZES = (1, 2, 4, 8)


class EventBatch:
    def __init__(self, mesh, rho, vp, vs, sources, receivers, dt, nsteps, static_precision=0, cij=None):
        """sources: one Sources object per event (its first source is used);
        nsteps: NSTEP per event; cij: [nspec, 125, 21] anisotropic model (rho/vp/vs
        then give the density and the Stacey impedances), kernels rho and 21 c_ij."""
        e = len(sources)
        if e not in BATCH_SIZES:
            raise ValueError(f"batch size {e} not supported {BATCH_SIZES}")
        if len(nsteps) != e:
            raise ValueError("one NSTEP per event")
        if mesh.nspec_irregular:
            raise ValueError("batches need a regular (all-cube) mesh")
        self.mesh, self.dt, self.nevents = mesh, float(dt), e
        self.nsteps = np.asarray(nsteps, np.int32)
        self.nmax = int(self.nsteps.max())
        self.nrec = receivers.nrec if receivers is not None else 0
        self.rho = np.ascontiguousarray(rho, F32).reshape(mesh.nspec, 125)
        # the mesh handle holds model, receivers and Stacey faces; no sources of its own
        self.aniso = cij is not None
        self.run = ElasticRun(mesh, rho, vp, vs, None, receivers, dt, self.nmax, 1, False,
                              static_precision=static_precision, cij=cij)
        sarr = np.ascontiguousarray(np.stack([s.sourcearrays[0] for s in sources]), F32)
        isp = np.array([s.ispec[0] for s in sources], np.int32)
        stf = np.zeros((e, self.nmax))
        for q, s in enumerate(sources):
            stf[q, :self.nsteps[q]] = s.stf(int(self.nsteps[q]), self.dt)[:, 0]
        self.stf = stf
        self.h = _lib.HandleBox()
        self.store_precision = 0
        _lib.call("fwat_batch_create_", self.run.h, self.h, e, self.nsteps, self.dt, sarr, isp, stf)

    def set_tuning(self, blocks_per_cu=8):
        _lib.call("fwat_batch_set_tuning_", self.h, int(blocks_per_cu))

    def set_interior(self, on=True):
        """Element-interior points finished in the element pass (isotropic batches;
        default off; the results are the same bits either way)."""
        _lib.call("fwat_batch_set_interior_", self.h, int(bool(on)))

    def set_store_precision(self, precision):
        """Stacey store of the next forward run: 0 fp32, 1 fp16 with one scale per
        step and event (half the HBM; the reconstruction replays the rounded terms)."""
        _lib.call("fwat_batch_set_store_precision_", self.h, int(precision))
        self.store_precision = int(precision)

    def forward(self, timed=False):
        _lib.call("fwat_batch_forward_", self.h, int(timed))

    def seismograms(self):
        """[E][NSTEP_max][nrec][3]; rows past an event's NSTEP are zero."""
        out = np.zeros((self.nevents, self.nmax, self.nrec, 3), F32)
        if self.nrec:
            _lib.call("fwat_batch_seismograms_", self.h, out)
        return out

    def set_adjoint_sources(self, adj):
        """adj [E][NSTEP_max][nrec][3] in forward time order."""
        adj = np.ascontiguousarray(adj, F32).reshape(self.nevents, self.nmax, self.nrec, 3)
        _lib.call("fwat_batch_set_adjoint_sources_", self.h, adj)

    def adjoint(self, timed=False):
        _lib.call("fwat_batch_adjoint_", self.h, int(timed))

    def kernels(self):
        """save_adjoint_kernels-scaled kernels per event, float64 [E][nspec][125]
        (cijkl_kl [E][nspec][125][21] for an anisotropic batch)."""
        shape = (self.nevents, self.mesh.nspec, 125)
        if self.aniso:
            rho_kl, cijkl = np.zeros(shape), np.zeros(shape + (21,))
            _lib.call("fwat_batch_kernels_ani_", self.h, self.rho, rho_kl, cijkl)
            return {"rho_kl": rho_kl, "cijkl_kl": cijkl}
        rho_kl, mu_kl, kappa_kl = np.zeros(shape), np.zeros(shape), np.zeros(shape)
        _lib.call("fwat_batch_kernels_", self.h, self.rho, rho_kl, mu_kl, kappa_kl)
        return {"rho_kl": rho_kl, "mu_kl": mu_kl, "kappa_kl": kappa_kl}

    def pass_times(self):
        """(element ms, element passes, assembly ms, assembly passes) of the last timed run."""
        em, am = np.zeros(1), np.zeros(1)
        en, an = np.zeros(1, np.int32), np.zeros(1, np.int32)
        _lib.call("fwat_batch_last_time_", self.h, em, en, am, an)
        return float(em[0]), int(en[0]), float(am[0]), int(an[0])

    def close(self):
        if self.h:
            _lib.call("fwat_batch_destroy_", self.h)
        if self.run is not None:
            self.run.close()
            self.run = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_gradient(mesh, model, obs, sources, receivers, dt, nsteps, windows=None, timed=False,
                   static_precision=0):
    """fwat_fwd_measure_adj's event loop for a batch: per-event misfit,
    synthetics and kernels.  obs [E][NSTEP_max][nrec][3]."""
    from .misfit import event_adjoint_sources
    b = EventBatch(mesh, *model, sources, receivers, dt, nsteps, static_precision=static_precision)
    try:
        b.forward(timed=timed)
        syn = b.seismograms()
        adj = np.zeros_like(syn)
        misfit = np.zeros(b.nevents)
        for e in range(b.nevents):
            n = int(b.nsteps[e])
            misfit[e], adj[e, :n] = event_adjoint_sources(obs[e, :n], syn[e, :n], dt, windows)
        b.set_adjoint_sources(adj)
        b.adjoint(timed=timed)
        k = b.kernels()
    finally:
        b.close()
    return misfit, syn, k


def plan_batches(nsteps, events_per_batch=8):
    """Event groups for batched runs: events sorted by NSTEP (longest first) so
    that a batch's events have similar NSTEP (idle slots cost a full step),
    cut into batches of `events_per_batch`; a remainder is split into the
    supported sizes (1, 2, 4, 8)."""
    if events_per_batch not in BATCH_SIZES:
        raise ValueError(f"events per batch {events_per_batch} not in {BATCH_SIZES}")
    order = [int(k) for k in np.argsort(-np.asarray(nsteps), kind="stable")]
    groups, i = [], 0
    while i < len(order):
        e = max(s for s in BATCH_SIZES if s <= min(len(order) - i, events_per_batch))
        groups.append(order[i:i + e])
        i += e
    return groups


def batch_device_bytes(mesh, nevents, nmax, store_precision=0):
    """HBM held by one EventBatch (fields, element-local forces, kernels and their
    fp32 partial sums, Stacey rows, CSR), besides the mesh handle."""
    nloc, ng = mesh.nspec * 125, mesh.nglob
    nfp = mesh.num_abs_boundary_faces * 25
    e = nevents
    store = nmax * nfp * e * (6 if store_precision else 12) + (nmax * e * 4 + nfp * e * 12 if store_precision else 0)
    return (6 * ng * e * 12 + 2 * nloc * e * 12 + (2 * nloc + ng) * e * 12
            + store + nfp * e * 12 + 4 * (ng + nloc + nfp) + 8 * e * nmax)


def tune_batch(mesh, model, sources, receivers, dt, sizes=(2, 4, 8), blocks_per_cu=(4, 6, 8, 12, 16), nstep=12,
               static_precision=0, hbm_budget=None, cij=None, interior=False):
    """Events per launch and persistent workgroups per CU of the batched
    element pass, from short timed runs on the real mesh (anisotropic with
    cij; interior: EventBatch.set_interior).  The figure of merit is the
    element + assembly time per event-step of a forward and an adjoint step.
    Returns ((events, blocks_per_cu), table)."""
    table = []
    for e in sizes:
        if e > len(sources):
            continue
        if hbm_budget is not None and batch_device_bytes(mesh, e, nstep) > hbm_budget:
            continue
        b = EventBatch(mesh, *model, sources[:e], receivers, dt, [nstep] * e,
                       static_precision=static_precision, cij=cij)
        b.set_interior(interior)
        b.forward()  # warm-up; fills the Stacey rows the adjoint replays
        b.set_adjoint_sources(np.zeros((e, nstep, b.nrec, 3), F32))
        for bpc in blocks_per_cu:
            b.set_tuning(bpc)
            b.forward(timed=True)
            fe, fn, fa, _ = b.pass_times()
            b.adjoint(timed=True)
            ae, an, aa, _ = b.pass_times()
            table.append((e, bpc, (fe + fa) / (fn * e), (ae + aa) / (an * e)))
        b.close()
    best = min(table, key=lambda r: r[2] + r[3])
    return (best[0], best[1]), table
