// Time stepping on the device: Newmark predictor/corrector, Stacey absorbing
// boundary, point and adjoint sources, seismograms, sensitivity kernels, and
// the whole-loop driver.
//
// Replaces SPECFEM3D's per-step GPU entry points that SpecFWAT's
// iterate_time calls (prepare_fwd.f90:364; sources absent from the tree), and
// adds fwat_iterate_time_, which runs a whole forward or adjoint run on the
// device with no host round trip per step.
//
// Scheme (state n at time t_n = n dt - t0, u_0 = v_0 = a_0 = 0):
//   predictor  u += dt v + dt^2/2 a ;  v += dt/2 a ;  a = 0
//   forces     a -= K u ;  a -= S(v) (Stacey) ;  a += stf((it-1) dt - t0) s_src
//   corrector  a *= 1/(M + dt/2 C) ;  v += dt/2 a
// The reconstructed field runs the same scheme with -dt from the forward
// run's last state, replaying the stored Stacey terms S and the source, which
// inverts the forward steps exactly in exact arithmetic.  Adjoint step it
// pairs with reconstructed state m = NSTEP + 1 - it (the pairing of the
// discrete adjoint of the central-difference scheme, lambda_n with u_n):
// step 1 uses the saved last state, each later step goes one step back and
// replays Stacey record m and the source of forward step m.  The adjoint
// source sample NSTEP - it (0-based) is injected at adjoint step it.
//
// Kernels (Tromp, Tape & Liu 2005), accumulated in fp64:
//   rho_kl   += dt  a_adj . u_rec              (per global point, expanded on transfer)
//   mu_kl    += dt  D_adj : D_rec              (per local point)
//   kappa_kl += dt  tr(e_adj) tr(e_rec)
//   cijkl_kl += dt  (e_I e_J + e_J e_I)/(1 + delta_IJ)  engineering Voigt strains
// and save_adjoint_kernels' scaling on transfer: rho_kl *= -rho,
// mu_kl *= -2 mu, kappa_kl *= -kappa, cijkl_kl *= -1.
#include "sem_device.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

constexpr int TPB = 256;

inline int nblk(size_t n, int tpb = TPB) { return (int)((n + tpb - 1) / tpb); }

__global__ void k_fill_f64(double* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ void k_fill_f32(realw* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

// predictor on one field, n3 = 3*nglob scalars
__global__ void k_predictor(size_t n3, realw* __restrict__ u, realw* __restrict__ v,
                            realw* __restrict__ a, realw dt, realw dt2o2, realw dto2) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  realw ai = a[i], vi = v[i];
  u[i] += dt * vi + dt2o2 * ai;
  v[i] = vi + dto2 * ai;
  a[i] = 0.0f;
}

// corrector on one field (SPECFEM kernel_3)
__global__ void k_corrector(int nglob, realw* __restrict__ v, realw* __restrict__ a,
                            const realw* __restrict__ rm, realw dto2) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nglob) return;
  const size_t b = 3 * (size_t)g;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    realw ac = a[b + c] * rm[b + c];
    a[b + c] = ac;
    v[b + c] += dto2 * ac;
  }
}

// Gather assembly fused with the Newmark update.  For every global point:
//   f = sum of its contributions (CSR, fixed order): element-local forces,
//       minus the Stacey face traction, plus sourcearray * stf, plus
//       h_rec * adjoint trace
//   a = rmass_inv * f ; v += dt/2 a            (corrector of step it)
//   unless LAST: u += dt v + dt^2/2 a ; v += dt/2 a   (predictor of it+1)
// TWO: the adjoint field (dt > 0) and the reconstructed field (dt < 0) are
// updated together and the density kernel a_adj . u_rec is accumulated in fp64.

// glist: the points of this launch (one slab's), or null for points 0..npts-1.
template <bool LAST, bool TWO>
__global__ void k_assemble(int npts, const int* __restrict__ glist, const int* __restrict__ off,
                           const int* __restrict__ idx, int nloc, int nfp, int nsp, const realw* __restrict__ sarr,
                           const realw* __restrict__ rech, const realw* __restrict__ rm,
                           AsmField F0, AsmField F1, double* __restrict__ rho_kl,
                           float* __restrict__ rho_part, int flush, double* __restrict__ hess_kl, double dtk) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= npts) return;
  const int g = glist ? glist[h] : h;
  realw f0x = 0, f0y = 0, f0z = 0, f1x = 0, f1y = 0, f1z = 0;
  const PointState P = load_point<TWO>((size_t)g, rm, F0, F1, rho_part);
  const int q1 = off[g + 1];
  int q = off[g];
  // pairs of element-local contributions (most of them): both rows are loaded
  // before either is added, two loads in flight per lane; same summation order
  for (; q + 1 < q1; q += 2) {
    const int c0 = idx[q], c1 = idx[q + 1];
    if (c1 >= nloc) break;  // element terms lead each row: c0 is one too
    const realw* P0 = F0.floc + 3 * (size_t)c0;
    const realw* P1 = F0.floc + 3 * (size_t)c1;
    const realw a0 = P0[0], a1 = P0[1], a2 = P0[2], b0 = P1[0], b1 = P1[1], b2 = P1[2];
    if (TWO) {
      const realw* B0 = F1.floc + 3 * (size_t)c0;
      const realw* B1 = F1.floc + 3 * (size_t)c1;
      const realw d0 = B0[0], d1 = B0[1], d2 = B0[2], e0 = B1[0], e1 = B1[1], e2 = B1[2];
      f1x += d0; f1y += d1; f1z += d2;
      f1x += e0; f1y += e1; f1z += e2;
    }
    f0x += a0; f0y += a1; f0z += a2;
    f0x += b0; f0y += b1; f0z += b2;
  }
  for (; q < q1; ++q) {
    const int c = idx[q];
    if (c < nloc) {
      const realw* P = F0.floc + 3 * (size_t)c;
      f0x += P[0]; f0y += P[1]; f0z += P[2];
      if (TWO) {
        const realw* B = F1.floc + 3 * (size_t)c;
        f1x += B[0]; f1y += B[1]; f1z += B[2];
      }
    } else if (c < nloc + nfp) {
      const size_t p = 3 * (size_t)(c - nloc);
      if (F0.face) { f0x -= F0.face[p]; f0y -= F0.face[p + 1]; f0z -= F0.face[p + 2]; }
      if (TWO && F1.face) { f1x -= F1.face[p]; f1y -= F1.face[p + 1]; f1z -= F1.face[p + 2]; }
    } else if (c < nloc + nfp + nsp) {
      const int p = c - nloc - nfp;
      const realw* S = sarr + 3 * (size_t)p;
      if (F0.stf) { const realw w = (realw)F0.stf[p / 125]; f0x = fmaf(S[0], w, f0x); f0y = fmaf(S[1], w, f0y); f0z = fmaf(S[2], w, f0z); }
      if (TWO && F1.stf) { const realw w = (realw)F1.stf[p / 125]; f1x = fmaf(S[0], w, f1x); f1y = fmaf(S[1], w, f1y); f1z = fmaf(S[2], w, f1z); }
    } else {
      const int p = c - nloc - nfp - nsp;
      const realw h = rech[p];
      const size_t r = 3 * (size_t)(p / 125);
      if (F0.adj) { f0x = fmaf(h, F0.adj[r], f0x); f0y = fmaf(h, F0.adj[r + 1], f0y); f0z = fmaf(h, F0.adj[r + 2], f0z); }
      if (TWO && F1.adj) { f1x = fmaf(h, F1.adj[r], f1x); f1y = fmaf(h, F1.adj[r + 1], f1y); f1z = fmaf(h, F1.adj[r + 2], f1z); }
    }
  }
  newmark_point<TWO>((size_t)g, P, f0x, f0y, f0z, f1x, f1y, f1z, F0, F1, LAST, rho_kl, rho_part, flush != 0,
                     hess_kl, dtk);
}

// Boundary and single element terms of point g's row from q on (the tail of
// k_assemble's loop), added in row order.
template <bool TWO>
__device__ __forceinline__ void assemble_tail(int q, int q1, const int* __restrict__ idx, int nloc, int nfp, int nsp,
                                              const realw* __restrict__ sarr, const realw* __restrict__ rech,
                                              const AsmField& F0, const AsmField& F1, realw* f) {
  for (; q < q1; ++q) {
    const int c = idx[q];
    if (c < nloc) {
      const realw* P = F0.floc + 3 * (size_t)c;
      f[0] += P[0]; f[1] += P[1]; f[2] += P[2];
      if (TWO) {
        const realw* B = F1.floc + 3 * (size_t)c;
        f[3] += B[0]; f[4] += B[1]; f[5] += B[2];
      }
    } else if (c < nloc + nfp) {
      const size_t p = 3 * (size_t)(c - nloc);
      if (F0.face) { f[0] -= F0.face[p]; f[1] -= F0.face[p + 1]; f[2] -= F0.face[p + 2]; }
      if (TWO && F1.face) { f[3] -= F1.face[p]; f[4] -= F1.face[p + 1]; f[5] -= F1.face[p + 2]; }
    } else if (c < nloc + nfp + nsp) {
      const int p = c - nloc - nfp;
      const realw* S = sarr + 3 * (size_t)p;
      if (F0.stf) { const realw w = (realw)F0.stf[p / 125]; f[0] = fmaf(S[0], w, f[0]); f[1] = fmaf(S[1], w, f[1]); f[2] = fmaf(S[2], w, f[2]); }
      if (TWO && F1.stf) { const realw w = (realw)F1.stf[p / 125]; f[3] = fmaf(S[0], w, f[3]); f[4] = fmaf(S[1], w, f[4]); f[5] = fmaf(S[2], w, f[5]); }
    } else {
      const int p = c - nloc - nfp - nsp;
      const realw h = rech[p];
      const size_t r = 3 * (size_t)(p / 125);
      if (F0.adj) { f[0] = fmaf(h, F0.adj[r], f[0]); f[1] = fmaf(h, F0.adj[r + 1], f[1]); f[2] = fmaf(h, F0.adj[r + 2], f[2]); }
      if (TWO && F1.adj) { f[3] = fmaf(h, F1.adj[r], f[3]); f[4] = fmaf(h, F1.adj[r + 1], f[4]); f[5] = fmaf(h, F1.adj[r + 2], f[5]); }
    }
  }
}

// k_assemble with two points per lane (points h and h + blockDim.x of the
// block's 2 * blockDim.x): both points' state, offsets, indices and
// element-local rows are requested before either is summed, so a lane keeps
// twice the loads in flight.  Each point's row is summed in the same order as
// k_assemble (pairs of element terms, then the tail), so results are identical.
template <bool LAST, bool TWO>
__global__ void k_assemble2(int npts, const int* __restrict__ glist, const int* __restrict__ off,
                            const int* __restrict__ idx, int nloc, int nfp, int nsp, const realw* __restrict__ sarr,
                            const realw* __restrict__ rech, const realw* __restrict__ rm,
                            AsmField F0, AsmField F1, double* __restrict__ rho_kl,
                            float* __restrict__ rho_part, int flush, double* __restrict__ hess_kl, double dtk) {
  const int ha = blockIdx.x * 2 * blockDim.x + threadIdx.x, hb = ha + blockDim.x;
  if (ha >= npts) return;
  const bool vb = hb < npts;
  const int ga = glist ? glist[ha] : ha;
  const int gb = vb ? (glist ? glist[hb] : hb) : ga;
  const PointState Pa = load_point<TWO>((size_t)ga, rm, F0, F1, rho_part);
  const PointState Pb = load_point<TWO>((size_t)gb, rm, F0, F1, rho_part);
  int qa = off[ga], qb = off[gb];
  const int qa1 = off[ga + 1], qb1 = vb ? off[gb + 1] : qb;
  realw fa[6] = {0, 0, 0, 0, 0, 0}, fb[6] = {0, 0, 0, 0, 0, 0};
  bool ra = true, rb = vb;  // still in the paired part of the row
  while (ra || rb) {
    ra = ra && qa + 1 < qa1;
    rb = rb && qb + 1 < qb1;
    int a0 = 0, a1 = 0, b0 = 0, b1 = 0;
    if (ra) { a0 = idx[qa]; a1 = idx[qa + 1]; }
    if (rb) { b0 = idx[qb]; b1 = idx[qb + 1]; }
    ra = ra && a1 < nloc;  // element terms lead each row: a0 is one too
    rb = rb && b1 < nloc;
    realw x[12], y[12];
    if (ra) {
      const realw* P0 = F0.floc + 3 * (size_t)a0;
      const realw* P1 = F0.floc + 3 * (size_t)a1;
#pragma unroll
      for (int c = 0; c < 3; ++c) { x[c] = P0[c]; x[3 + c] = P1[c]; }
      if (TWO) {
        const realw* B0 = F1.floc + 3 * (size_t)a0;
        const realw* B1 = F1.floc + 3 * (size_t)a1;
#pragma unroll
        for (int c = 0; c < 3; ++c) { x[6 + c] = B0[c]; x[9 + c] = B1[c]; }
      }
    }
    if (rb) {
      const realw* P0 = F0.floc + 3 * (size_t)b0;
      const realw* P1 = F0.floc + 3 * (size_t)b1;
#pragma unroll
      for (int c = 0; c < 3; ++c) { y[c] = P0[c]; y[3 + c] = P1[c]; }
      if (TWO) {
        const realw* B0 = F1.floc + 3 * (size_t)b0;
        const realw* B1 = F1.floc + 3 * (size_t)b1;
#pragma unroll
        for (int c = 0; c < 3; ++c) { y[6 + c] = B0[c]; y[9 + c] = B1[c]; }
      }
    }
    if (ra) {
      if (TWO) {
#pragma unroll
        for (int c = 0; c < 3; ++c) fa[3 + c] += x[6 + c];
#pragma unroll
        for (int c = 0; c < 3; ++c) fa[3 + c] += x[9 + c];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) fa[c] += x[c];
#pragma unroll
      for (int c = 0; c < 3; ++c) fa[c] += x[3 + c];
      qa += 2;
    }
    if (rb) {
      if (TWO) {
#pragma unroll
        for (int c = 0; c < 3; ++c) fb[3 + c] += y[6 + c];
#pragma unroll
        for (int c = 0; c < 3; ++c) fb[3 + c] += y[9 + c];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) fb[c] += y[c];
#pragma unroll
      for (int c = 0; c < 3; ++c) fb[c] += y[3 + c];
      qb += 2;
    }
  }
  assemble_tail<TWO>(qa, qa1, idx, nloc, nfp, nsp, sarr, rech, F0, F1, fa);
  if (vb) assemble_tail<TWO>(qb, qb1, idx, nloc, nfp, nsp, sarr, rech, F0, F1, fb);
  newmark_point<TWO>((size_t)ga, Pa, fa[0], fa[1], fa[2], fa[3], fa[4], fa[5], F0, F1, LAST, rho_kl, rho_part,
                     flush != 0, hess_kl, dtk);
  if (vb)
    newmark_point<TWO>((size_t)gb, Pb, fb[0], fb[1], fb[2], fb[3], fb[4], fb[5], F0, F1, LAST, rho_kl, rho_part,
                       flush != 0, hess_kl, dtk);
}

// Incident wavefield of one step for FK injection: rows k and k + 1 of the
// stored velocity and traction and the weight w of row k + 1 (no injection
// when v0 is null)
struct FkStep {
  const realw *v0 = nullptr, *v1 = nullptr, *t0 = nullptr, *t1 = nullptr;
  realw w = 0;
};

// Stacey term of face point q, SPECFEM's compute_stacey_viscoelastic with
// injection: the absorbing traction of the velocity relative to the
// incident one, minus the incident traction, times jacobian2Dw
__device__ __forceinline__ void stacey_point(size_t q, const realw* __restrict__ C, realw vx, realw vy, realw vz,
                                             const FkStep& fk, realw& tx, realw& ty, realw& tz) {
  const realw nx = C[0], ny = C[1], nz = C[2], jw = C[3], rvp = C[4], rvs = C[5];
  realw ix = 0, iy = 0, iz = 0;
  if (fk.v0) {
    const realw w0 = (realw)1 - fk.w, w1 = fk.w;
    vx -= w0 * fk.v0[3 * q] + w1 * fk.v1[3 * q];
    vy -= w0 * fk.v0[3 * q + 1] + w1 * fk.v1[3 * q + 1];
    vz -= w0 * fk.v0[3 * q + 2] + w1 * fk.v1[3 * q + 2];
    ix = w0 * fk.t0[3 * q] + w1 * fk.t1[3 * q];
    iy = w0 * fk.t0[3 * q + 1] + w1 * fk.t1[3 * q + 1];
    iz = w0 * fk.t0[3 * q + 2] + w1 * fk.t1[3 * q + 2];
  }
  const realw vn = vx * nx + vy * ny + vz * nz;
  tx = (rvp * vn * nx + rvs * (vx - vn * nx) - ix) * jw;
  ty = (rvp * vn * ny + rvs * (vy - vn * ny) - iy) * jw;
  tz = (rvp * vn * nz + rvs * (vz - vn * nz) - iz) * jw;
}

// Stacey traction of one field written per face point (no scatter):
// out[q] = (rho_vp (v.n) n + rho_vs (v - (v.n) n)) * jacobian2Dw, with v
// relative to the incident field and its traction subtracted under FK injection
__global__ void k_stacey_face(int nfp, const int* __restrict__ figlob, const realw* __restrict__ fcoef,
                              const realw* __restrict__ v, realw* __restrict__ out, FkStep fk) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nfp) return;
  const size_t g = (size_t)figlob[q];
  realw tx, ty, tz;
  stacey_point((size_t)q, fcoef + 8 * (size_t)q, v[3 * g], v[3 * g + 1], v[3 * g + 2], fk, tx, ty, tz);
  out[3 * (size_t)q] = tx;
  out[3 * (size_t)q + 1] = ty;
  out[3 * (size_t)q + 2] = tz;
}

// Start of a reconstruction: the saved last state (u_N, v_N, a_N) becomes
// the state of adjoint step 1; its velocity is set to v_N + dt/2 a_N so that
// the corrector of step 1 returns v_N.
__global__ void k_b_init(size_t n3, realw* __restrict__ v, const realw* __restrict__ a, realw dto2) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) v[i] += dto2 * a[i];
}

// rho (and Hessian) kernel accumulation of the per-step API path
__global__ void k_rho_kernel_glob(int nglob, const realw* __restrict__ a, const realw* __restrict__ bu,
                                  const realw* __restrict__ ba, double* __restrict__ rho_kl,
                                  double* __restrict__ hess_kl, double dtk) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nglob) return;
  const size_t b = 3 * (size_t)g;
  realw d = a[b] * bu[b] + a[b + 1] * bu[b + 1] + a[b + 2] * bu[b + 2];
  rho_kl[g] += dtk * (double)d;
  if (hess_kl) {
    realw h = a[b] * ba[b] + a[b + 1] * ba[b + 1] + a[b + 2] * ba[b + 2];
    hess_kl[g] += dtk * (double)h;
  }
}

// mu/kappa or c_ij kernels from the two stored strain fields (SPECFEM-style
// compute_kernels path; the fused path does this inside the force kernel).
template <bool ANISO>
__global__ void k_kernels_from_strain(int nspec, const realw* __restrict__ ea_all,
                                      const realw* __restrict__ eb_all, double* __restrict__ mu_kl,
                                      double* __restrict__ kappa_kl, double* __restrict__ cijkl_kl,
                                      double dtk) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)nspec * 125) return;
  const size_t e = idx / 125, L = idx % 125;
  const realw* EA = ea_all + e * 750 + L;
  const realw* EB = eb_all + e * 750 + L;
  realw ea[6], eb[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) { ea[c] = EA[c * 125]; eb[c] = EB[c * 125]; }
  if (!ANISO) {
    const realw tra = ea[0] + ea[1] + ea[2], trb = eb[0] + eb[1] + eb[2];
    const realw t3a = tra / 3.0f, t3b = trb / 3.0f;
    const realw prod = (ea[0] - t3a) * (eb[0] - t3b) + (ea[1] - t3a) * (eb[1] - t3b) +
                       (ea[2] - t3a) * (eb[2] - t3b) + 2.0f * (ea[3] * eb[3] + ea[4] * eb[4] + ea[5] * eb[5]);
    mu_kl[idx] += dtk * (double)prod;
    kappa_kl[idx] += dtk * (double)(tra * trb);
  } else {
    const realw va[6] = {ea[0], ea[1], ea[2], 2.0f * ea[5], 2.0f * ea[4], 2.0f * ea[3]};
    const realw vb[6] = {eb[0], eb[1], eb[2], 2.0f * eb[5], 2.0f * eb[4], 2.0f * eb[3]};
    double* K = cijkl_kl + e * 21 * 125 + L;
    int q = 0;
#pragma unroll
    for (int P = 0; P < 6; ++P)
#pragma unroll
      for (int Q = P; Q < 6; ++Q) {
        realw pr = va[P] * vb[Q];
        if (Q > P) pr += va[Q] * vb[P];
        K[q * 125] += dtk * (double)pr;
        ++q;
      }
  }
}

// Stacey: mode 0 absorb, 1 absorb + store, 2 replay a stored record
__global__ void k_stacey(int nfp, const int* __restrict__ figlob, const realw* __restrict__ fcoef,
                         const realw* __restrict__ v, realw* __restrict__ a, realw* __restrict__ rec,
                         int mode, FkStep fk) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nfp) return;
  const size_t g = (size_t)figlob[q];
  realw tx, ty, tz;
  if (mode == 2) {
    tx = rec[3 * (size_t)q]; ty = rec[3 * (size_t)q + 1]; tz = rec[3 * (size_t)q + 2];
  } else {
    stacey_point((size_t)q, fcoef + 8 * (size_t)q, v[3 * g], v[3 * g + 1], v[3 * g + 2], fk, tx, ty, tz);
    if (mode == 1) { rec[3 * (size_t)q] = tx; rec[3 * (size_t)q + 1] = ty; rec[3 * (size_t)q + 2] = tz; }
  }
  unsafeAtomicAdd(a + 3 * g, -tx);
  unsafeAtomicAdd(a + 3 * g + 1, -ty);
  unsafeAtomicAdd(a + 3 * g + 2, -tz);
}

// point sources: a += stf(source) * sourcearray
__global__ void k_add_sources(int nsrc, const int* __restrict__ siglob, const realw* __restrict__ sarr,
                              const double* __restrict__ stf, realw* __restrict__ a) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nsrc * 125) return;
  const int s = q / 125;
  const realw f = (realw)stf[s];
  const size_t g = (size_t)siglob[q];
  unsafeAtomicAdd(a + 3 * g, sarr[3 * (size_t)q] * f);
  unsafeAtomicAdd(a + 3 * g + 1, sarr[3 * (size_t)q + 1] * f);
  unsafeAtomicAdd(a + 3 * g + 2, sarr[3 * (size_t)q + 2] * f);
}

// adjoint sources: a += h_rec * adj(sample)
__global__ void k_add_adjoint(int nrec, const int* __restrict__ riglob, const realw* __restrict__ rh,
                              const realw* __restrict__ adj_row, realw* __restrict__ a) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nrec * 125) return;
  const int r = q / 125;
  const realw h = rh[q];
  const size_t g = (size_t)riglob[q];
  unsafeAtomicAdd(a + 3 * g, h * adj_row[3 * r]);
  unsafeAtomicAdd(a + 3 * g + 1, h * adj_row[3 * r + 1]);
  unsafeAtomicAdd(a + 3 * g + 2, h * adj_row[3 * r + 2]);
}

// seismograms: one 128-lane block per receiver, Lagrange interpolation of u
__global__ __launch_bounds__(128) void k_seismo(int nrec, const int* __restrict__ riglob,
                                                const realw* __restrict__ rh, const realw* __restrict__ u,
                                                realw* __restrict__ row) {
  __shared__ realw s[3][128];
  const int r = blockIdx.x, t = threadIdx.x;
  realw x = 0, y = 0, z = 0;
  if (t < 125) {
    const realw h = rh[(size_t)r * 125 + t];
    const size_t g = (size_t)riglob[(size_t)r * 125 + t];
    x = h * u[3 * g]; y = h * u[3 * g + 1]; z = h * u[3 * g + 2];
  }
  s[0][t] = x; s[1][t] = y; s[2][t] = z;
  __syncthreads();
  for (int w = 64; w > 0; w >>= 1) {
    if (t < w) { s[0][t] += s[0][t + w]; s[1][t] += s[1][t + w]; s[2][t] += s[2][t + w]; }
    __syncthreads();
  }
  if (t == 0) { row[3 * r] = s[0][0]; row[3 * r + 1] = s[1][0]; row[3 * r + 2] = s[2][0]; }
}

__global__ void k_copy4(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// The same copy with U float4 loads issued before the stores (U in flight per
// lane), lanes of a workgroup on contiguous 16-byte slots of each of U chunks.
template <int U>
__global__ void k_copy4u(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n4; i += U * stride) {
    float4 v[U];
#pragma unroll
    for (int k = 0; k < U; ++k) v[k] = a[i + k * stride];
#pragma unroll
    for (int k = 0; k < U; ++k) b[i + k * stride] = v[k];
  }
  for (; i < n4; i += stride) b[i] = a[i];
}

__global__ void k_expand_glob(int nspec, const int* __restrict__ ibool, const double* __restrict__ g,
                              double scale_sign, const realw* __restrict__ rho, double* __restrict__ out) {
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)nspec * 125) return;
  double v = g[ibool[idx]];
  out[idx] = rho ? scale_sign * (double)rho[idx] * v : v;
}

}  // namespace

// ------------------------------------------------------------ internal stages

void sem_zero_kernels(SemMesh* m) {
  hipStream_t s = m->stream;
  const size_t nloc = (size_t)m->nspec * 125;
  if (m->d_rho_kl_glob) hipLaunchKernelGGL(k_fill_f64, dim3(nblk(m->nglob)), dim3(TPB), 0, s, m->d_rho_kl_glob, (size_t)m->nglob);
  if (m->d_hess_kl_glob) hipLaunchKernelGGL(k_fill_f64, dim3(nblk(m->nglob)), dim3(TPB), 0, s, m->d_hess_kl_glob, (size_t)m->nglob);
  if (m->d_mu_kl) hipLaunchKernelGGL(k_fill_f64, dim3(nblk(nloc)), dim3(TPB), 0, s, m->d_mu_kl, nloc);
  if (m->d_kappa_kl) hipLaunchKernelGGL(k_fill_f64, dim3(nblk(nloc)), dim3(TPB), 0, s, m->d_kappa_kl, nloc);
  if (m->d_cijkl_kl) hipLaunchKernelGGL(k_fill_f64, dim3(nblk(nloc * 21)), dim3(TPB), 0, s, m->d_cijkl_kl, nloc * 21);
  HIP_CHECK(hipGetLastError());
}

namespace {

void stage_predictor(SemMesh* m, bool backward, realw dt) {
  realw* u = backward ? m->d_b_displ : m->d_displ;
  realw* v = backward ? m->d_b_veloc : m->d_veloc;
  realw* a = backward ? m->d_b_accel : m->d_accel;
  const realw sdt = backward ? -dt : dt;
  const size_t n3 = (size_t)m->nglob * 3;
  hipLaunchKernelGGL(k_predictor, dim3(nblk(n3)), dim3(TPB), 0, m->stream, n3, u, v, a, sdt,
                     0.5f * dt * dt, 0.5f * sdt);
}

void stage_corrector(SemMesh* m, bool backward, realw dt) {
  realw* v = backward ? m->d_b_veloc : m->d_veloc;
  realw* a = backward ? m->d_b_accel : m->d_accel;
  const realw sdt = backward ? -dt : dt;
  hipLaunchKernelGGL(k_corrector, dim3(nblk(m->nglob)), dim3(TPB), 0, m->stream, m->nglob, v, a,
                     m->d_rmass_inv, 0.5f * sdt);
}

// the incident field of forward step it (1-based) under FK injection
FkStep fk_step(const SemMesh* m, int it) {
  FkStep f;
  if (m->simulation_type != 1 || m->fk_nsamp == 0 || it < 1) return f;
  const int nfp3 = m->nfaces * 75;
  const int k = std::min((it - 1) / m->fk_resamp, m->fk_nsamp - 1);
  const int k1 = std::min(k + 1, m->fk_nsamp - 1);
  f.v0 = m->d_fk_veloc + (size_t)k * nfp3;
  f.v1 = m->d_fk_veloc + (size_t)k1 * nfp3;
  f.t0 = m->d_fk_tract + (size_t)k * nfp3;
  f.t1 = m->d_fk_tract + (size_t)k1 * nfp3;
  f.w = k1 == k ? (realw)0 : (realw)((it - 1) - k * m->fk_resamp) / (realw)m->fk_resamp;
  return f;
}

void stage_stacey(SemMesh* m, bool backward, int mode, size_t record) {
  if (!m->absorbing || m->nfaces == 0) return;
  const int nfp = m->nfaces * 25;
  realw* v = backward ? m->d_b_veloc : m->d_veloc;
  realw* a = backward ? m->d_b_accel : m->d_accel;
  realw* rec = nullptr;
  if (mode != 0) {
    if (!m->d_stacey_store) SEM_FAIL("Stacey store requested but not allocated");
    if (record >= m->stacey_store_steps) SEM_FAIL("Stacey record %zu outside store of %zu steps", record, m->stacey_store_steps);
    rec = m->d_stacey_store + record * (size_t)nfp * 3;
  }
  const FkStep fk = backward ? FkStep{} : fk_step(m, (int)record + 1);
  hipLaunchKernelGGL(k_stacey, dim3(nblk(nfp)), dim3(TPB), 0, m->stream, nfp, m->d_face_iglob,
                     m->d_face_coef, v, a, rec, mode, fk);
}

void stage_sources(SemMesh* m, bool backward, const double* d_stf_row) {
  if (m->nsources == 0) return;
  realw* a = backward ? m->d_b_accel : m->d_accel;
  hipLaunchKernelGGL(k_add_sources, dim3(nblk((size_t)m->nsources * 125)), dim3(TPB), 0, m->stream,
                     m->nsources, m->d_src_iglob, m->d_sourcearrays, d_stf_row, a);
}

void stage_adjoint_sources(SemMesh* m, const realw* d_adj_row) {
  if (m->nrec == 0) return;
  hipLaunchKernelGGL(k_add_adjoint, dim3(nblk((size_t)m->nrec * 125)), dim3(TPB), 0, m->stream,
                     m->nrec, m->d_rec_iglob, m->d_rec_h, d_adj_row, m->d_accel);
}

// it: the forward step (1-based) whose incident field is injected, 0 for none
void stage_stacey_face(SemMesh* m, const realw* v, realw* out, int it = 0) {
  if (!m->absorbing || m->nfaces == 0) return;
  const int nfp = m->nfaces * 25;
  hipLaunchKernelGGL(k_stacey_face, dim3(nblk(nfp)), dim3(TPB), 0, m->stream, nfp, m->d_face_iglob,
                     m->d_face_coef, v, out, fk_step(m, it));
}

// Assembly + Newmark of points 0..nglob-1 (0..n_shared-1 when the element pass
// finished the others: interior), or of slab `slab`'s point list.
void stage_assemble(SemMesh* m, bool last, bool two, const AsmField& F0, const AsmField& F1, double dtk,
                    int slab = -1, bool interior = false) {
  const int nfp = m->absorbing ? m->nfaces * 25 : 0;
  float* rho_part = m->kernel_block > 1 ? m->d_rho_part : nullptr;
  const int nloc = m->nspec * 125, nsp = m->nsources * 125;
  const int* glist = slab >= 0 ? m->d_slab_g + m->h_slab_goff[slab] : nullptr;
  const int npts = slab >= 0 ? m->h_slab_goff[slab + 1] - m->h_slab_goff[slab] : (interior ? m->n_shared : m->nglob);
  if (npts == 0) return;
  const int tpb = m->asm_block;
  const dim3 bl(tpb);
  // 0 (default): two points per lane for the one-field pass, one for the two-field pass
  // (measured on the C2 mesh: 0.656 vs 0.668 ms one-field, 1.32 vs 1.21 ms two-field)
  const int ppl = m->asm_points_per_lane ? m->asm_points_per_lane : (two ? 1 : 2);
  if (ppl == 2) {
    const dim3 g2((unsigned)((npts + 2 * tpb - 1) / (2 * tpb)));
#define ASM2(L, T, RK, RP, FL, HK)                                                                              \
  SEM_LAUNCH((k_assemble2<L, T>), g2, bl, 0, m->stream, npts, glist, m->d_asm_off, m->d_asm_idx, nloc, nfp, nsp, \
             m->d_sourcearrays, m->d_rec_h, m->d_rmass_inv, F0, F1, RK, RP, FL, HK, dtk)
    if (two) {
      if (last) ASM2(true, true, m->d_rho_kl_glob, rho_part, m->kernel_flush, m->d_hess_kl_glob);
      else ASM2(false, true, m->d_rho_kl_glob, rho_part, m->kernel_flush, m->d_hess_kl_glob);
    } else {
      if (last) ASM2(true, false, (double*)nullptr, (float*)nullptr, 1, (double*)nullptr);
      else ASM2(false, false, (double*)nullptr, (float*)nullptr, 1, (double*)nullptr);
    }
#undef ASM2
    return;
  }
  const dim3 gr(nblk(npts, tpb));
  if (two) {
    if (last)
      SEM_LAUNCH((k_assemble<true, true>), gr, bl, 0, m->stream, npts, glist, m->d_asm_off, m->d_asm_idx, nloc, nfp,
                         nsp, m->d_sourcearrays, m->d_rec_h, m->d_rmass_inv, F0, F1, m->d_rho_kl_glob, rho_part, m->kernel_flush,
                         m->d_hess_kl_glob, dtk);
    else
      SEM_LAUNCH((k_assemble<false, true>), gr, bl, 0, m->stream, npts, glist, m->d_asm_off, m->d_asm_idx, nloc, nfp,
                         nsp, m->d_sourcearrays, m->d_rec_h, m->d_rmass_inv, F0, F1, m->d_rho_kl_glob, rho_part, m->kernel_flush,
                         m->d_hess_kl_glob, dtk);
  } else {
    if (last)
      SEM_LAUNCH((k_assemble<true, false>), gr, bl, 0, m->stream, npts, glist, m->d_asm_off, m->d_asm_idx, nloc, nfp,
                         nsp, m->d_sourcearrays, m->d_rec_h, m->d_rmass_inv, F0, F1, (double*)nullptr, (float*)nullptr, 1,
                         (double*)nullptr, dtk);
    else
      SEM_LAUNCH((k_assemble<false, false>), gr, bl, 0, m->stream, npts, glist, m->d_asm_off, m->d_asm_idx, nloc, nfp,
                         nsp, m->d_sourcearrays, m->d_rec_h, m->d_rmass_inv, F0, F1, (double*)nullptr, (float*)nullptr, 1,
                         (double*)nullptr, dtk);
  }
}

void stage_seismo(SemMesh* m, int it0) {
  if (m->nrec == 0) return;
  if (it0 < 0 || it0 >= m->nstep) SEM_FAIL("seismogram index %d outside NSTEP=%d", it0 + 1, m->nstep);
  hipLaunchKernelGGL(k_seismo, dim3(m->nrec), dim3(128), 0, m->stream, m->nrec, m->d_rec_iglob,
                     m->d_rec_h, m->d_displ, m->d_seismograms + (size_t)it0 * m->nrec * 3);
}

// Timing samples every timer_stride(N)-th step (at most about 256 steps of a
// run), so a timed run records a bounded number of events whatever NSTEP is.
inline int timer_stride(int nstep) { return std::max(1, nstep / 256); }

// Steps between host synchronisations of a whole-run loop: bounds the work
// queued ahead of the device (tens of thousands of launches at NSTEP 10,000).
constexpr int kSyncEvery = 1024;

// Force or assembly launches of the sampled steps timed by their dispatch
// timestamps (an event pair armed for the launch, SEM_LAUNCH); exactly two
// events per sampled launch are created, and the elapsed times are summed
// after the run, so timing adds no host sync and no packet between kernels.
struct ForceTimer {
  std::vector<hipEvent_t> ev;
  size_t used = 0;
  int stride = 1;
  bool active = false;
  int launches = 0;
  // room for per_step bracketed launches on every sampled step of an nstep run
  void begin(int nstep, size_t per_step) {
    for (auto e : ev) HIP_CHECK(hipEventDestroy(e));
    stride = timer_stride(nstep);
    const size_t samples = (size_t)(nstep + stride - 1) / stride;
    ev.assign(2 * samples * per_step, nullptr);
    for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
    used = 0;
    launches = 0;
    active = false;
  }
  // it: 1-based step about to be enqueued
  void at_step(int it) { active = (it - 1) % stride == 0; }
  // arm the next launch's start/end events (one kernel launch per bracket)
  void open(hipStream_t) {
    if (!active) return;
    if (used + 2 > ev.size()) SEM_FAIL("force timer overflow");
    sem_arm_launch_events(ev[used], ev[used + 1]);
  }
  void close(hipStream_t, int nlaunch = 1) {
    if (!active) return;
    if (!sem_launch_events_consumed()) return;  // the bracket launched nothing (an empty slab)
    used += 2;
    launches += nlaunch;
  }
  double total_ms() {
    double t = 0;
    for (size_t q = 0; q + 1 < used; q += 2) {
      float ms = 0;
      HIP_CHECK(hipEventElapsedTime(&ms, ev[q], ev[q + 1]));
      t += ms;
    }
    return t;
  }
  ~ForceTimer() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};

// HIP graph of a whole-run step loop (fwat_set_graph_): the loop's launches
// are captured from the simulation stream into one graph, which is
// instantiated and launched as one unit, so the GPU runs the steps' kernels
// back to back without waiting on the host's launch rate.  The captured loops
// issue kernels only (no memset, copy or host synchronisation).
struct LoopGraph {
  SemMesh* m;
  bool on;
  LoopGraph(SemMesh* m_, bool on_) : m(m_), on(on_) {
    if (on) HIP_CHECK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
  }
  void launch() {
    if (!on) return;
    on = false;
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    HIP_CHECK(hipStreamEndCapture(m->stream, &g));
    HIP_CHECK(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
    HIP_CHECK(hipGraphLaunch(x, m->stream));
    HIP_CHECK(hipStreamSynchronize(m->stream));
    HIP_CHECK(hipGraphExecDestroy(x));
    HIP_CHECK(hipGraphDestroy(g));
  }
};

void stage_forces_timed(SemMesh* m, int nfields, int kmode, realw dt, ForceTimer* tm, bool gather,
                        const SemInterior* in = nullptr) {
  if (tm) tm->open(m->stream);
  sem_launch_forces(m, nfields, kmode, 0, m->nspec, dt, nullptr, gather, in);
  if (tm) tm->close(m->stream);
}

}  // namespace

// ------------------------------------------------------------------ C-ABI

extern "C" {

// FORWARD_OR_ADJOINT: 1 = forward/adjoint field, 3 = reconstructed (b_) field
void update_displacement_cuda_(long* Mesh_pointer, realw* deltat, realw* deltatsqover2,
                               realw* deltatover2, int* FORWARD_OR_ADJOINT) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  (void)deltatsqover2; (void)deltatover2;
  stage_predictor(m, *FORWARD_OR_ADJOINT == 3, *deltat);
  HIP_CHECK(hipGetLastError());
}

// Internal forces of one inner/outer phase (iphase 1 = outer, 2 = inner).
// COMPUTE_AND_STORE_STRAIN stores the strain for compute_kernels_elastic_cuda_.
void compute_forces_viscoelastic_cuda_(long* Mesh_pointer, int* iphase, realw* deltat,
                                       int* nspec_outer_elastic, int* nspec_inner_elastic,
                                       int* COMPUTE_AND_STORE_STRAIN, int* ATTENUATION,
                                       int* FORWARD_OR_ADJOINT) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (*ATTENUATION) SEM_FAIL("ATTENUATION is not supported on this path (SpecFWAT disables it for adjoint runs)");
  const int nfields = *FORWARD_OR_ADJOINT == 3 ? 3 : 1;
  const int kmode = *COMPUTE_AND_STORE_STRAIN ? 1 : 0;
  if (kmode && (nfields == 1 ? !m->d_epsilon : !m->d_b_epsilon)) SEM_FAIL("strain store not allocated");
  // iphase 1 = outer (halo) elements, 2 = inner elements, from
  // phase_ispec_inner_elastic; without phase lists every element is inner.
  if (m->d_phase_ispec && m->num_phase_ispec > 0) {
    if (*nspec_outer_elastic != m->nspec_outer || *nspec_inner_elastic != m->nspec_inner)
      SEM_FAIL("phase sizes differ from prepare_fields_elastic_device_");
    const int* list = m->d_phase_ispec + (*iphase == 1 ? 0 : m->num_phase_ispec);
    const int count = *iphase == 1 ? m->nspec_outer : m->nspec_inner;
    sem_launch_forces(m, nfields, kmode, 0, count, *deltat, list, false);
    return;
  }
  if (*iphase != 2) return;
  sem_launch_forces(m, nfields, kmode, 0, m->nspec, *deltat, nullptr, false);
}

// Stacey: the forward run of a SAVE_FORWARD simulation stores record it
// (1-based); the reconstructed field replays it.
void compute_stacey_viscoelastic_cuda_(long* Mesh_pointer, int* iphase, int* it_record,
                                       int* FORWARD_OR_ADJOINT) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (*iphase != 1) return;
  const bool back = *FORWARD_OR_ADJOINT == 3;
  int mode = back ? 2 : ((m->save_forward && m->simulation_type == 1) ? 1 : 0);
  stage_stacey(m, back, mode, (size_t)std::max(*it_record - 1, 0));
  HIP_CHECK(hipGetLastError());
}

// h_stf_pre_compute (NSOURCES): source time function values of this step
void compute_add_sources_viscoelastic_cuda_(long* Mesh_pointer, int* NSOURCES, double* h_stf_pre_compute,
                                            int* FORWARD_OR_ADJOINT) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (*NSOURCES != m->nsources) SEM_FAIL("NSOURCES mismatch");
  if (m->nsources == 0) return;
  HIP_CHECK(hipMemcpyAsync(m->d_stf_step, h_stf_pre_compute, sizeof(double) * m->nsources, hipMemcpyHostToDevice, m->stream));
  stage_sources(m, *FORWARD_OR_ADJOINT == 3, m->d_stf_step);
  HIP_CHECK(hipGetLastError());
}

// h_source_adjoint (NDIM, nadj_rec_local): adjoint source samples of this step
void add_sources_el_sim_type_2_or_3_(long* Mesh_pointer, int* nrec, realw* h_source_adjoint, int* it) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  (void)it;
  if (*nrec != m->nrec) SEM_FAIL("nrec mismatch");
  if (m->nrec == 0) return;
  HIP_CHECK(hipMemcpyAsync(m->d_adj_step, h_source_adjoint, sizeof(realw) * 3 * m->nrec, hipMemcpyHostToDevice, m->stream));
  stage_adjoint_sources(m, m->d_adj_step);
  HIP_CHECK(hipGetLastError());
}

// kernel_3: accel *= rmass_inv, veloc += deltatover2 * accel
void kernel_3_a_cuda_(long* Mesh_pointer, realw* deltatover2, int* FORWARD_OR_ADJOINT) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  const bool back = *FORWARD_OR_ADJOINT == 3;
  stage_corrector(m, back, 2.0f * std::fabs(*deltatover2));
  HIP_CHECK(hipGetLastError());
}

// record seismograms of step it (1-based) into the device buffer
void compute_seismograms_cuda_(long* Mesh_pointer, int* it) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  stage_seismo(m, *it - 1);
  HIP_CHECK(hipGetLastError());
}

// seismograms (NDIM, nrec, NSTEP)
void transfer_seismograms_from_device_(long* Mesh_pointer, realw* h_seismograms) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (m->nrec == 0) return;
  HIP_CHECK(hipMemcpyAsync(h_seismograms, m->d_seismograms, sizeof(realw) * (size_t)m->nstep * m->nrec * 3,
                           hipMemcpyDeviceToHost, m->stream));
  HIP_CHECK(hipStreamSynchronize(m->stream));
}

// Kernel accumulation of one adjoint step from the stored strains
// (COMPUTE_AND_STORE_STRAIN path) and the current accel / b_displ.
void compute_kernels_elastic_cuda_(long* Mesh_pointer, realw* deltat) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (!m->d_rho_kl_glob) SEM_FAIL("kernels not prepared (prepare_fields_elastic_adj_dev_)");
  if (!m->d_epsilon || !m->d_b_epsilon) SEM_FAIL("compute_kernels_elastic_cuda_ needs COMPUTE_AND_STORE_STRAIN");
  const double dtk = (double)*deltat;
  hipLaunchKernelGGL(k_rho_kernel_glob, dim3(nblk(m->nglob)), dim3(TPB), 0, m->stream, m->nglob,
                     m->d_accel, m->d_b_displ, m->d_b_accel, m->d_rho_kl_glob, m->d_hess_kl_glob, dtk);
  const size_t nloc = (size_t)m->nspec * 125;
  if (m->anisotropic_kl)
    hipLaunchKernelGGL(k_kernels_from_strain<true>, dim3(nblk(nloc)), dim3(TPB), 0, m->stream, m->nspec,
                       m->d_epsilon, m->d_b_epsilon, nullptr, nullptr, m->d_cijkl_kl, dtk);
  else
    hipLaunchKernelGGL(k_kernels_from_strain<false>, dim3(nblk(nloc)), dim3(TPB), 0, m->stream, m->nspec,
                       m->d_epsilon, m->d_b_epsilon, m->d_mu_kl, m->d_kappa_kl, nullptr, dtk);
  HIP_CHECK(hipGetLastError());
}

// Scaled kernels (save_adjoint_kernels): rho_kl = -rho K_rho, mu_kl = -2 mu K_mu,
// kappa_kl = -kappa K_kappa, each (NGLLX,NGLLY,NGLLZ,NSPEC_AB) fp64.
void transfer_kernels_el_to_host_dp_(long* Mesh_pointer, double* h_rho_kl, double* h_mu_kl,
                                     double* h_kappa_kl, int* NSPEC_AB) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (*NSPEC_AB != m->nspec) SEM_FAIL("NSPEC_AB mismatch");
  const size_t nloc = (size_t)m->nspec * 125;
  hipStream_t s = m->stream;
  double* tmp = dev_alloc<double>(nloc);
  hipLaunchKernelGGL(k_expand_glob, dim3(nblk(nloc)), dim3(TPB), 0, s, m->nspec, m->d_ibool,
                     m->d_rho_kl_glob, -1.0, m->d_rho, tmp);
  HIP_CHECK(hipMemcpyAsync(h_rho_kl, tmp, nloc * 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  dev_free(tmp);
  if (!m->anisotropic_kl) {
    std::vector<float> kap(nloc), mu(nloc);
    HIP_CHECK(hipMemcpyAsync(h_mu_kl, m->d_mu_kl, nloc * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(h_kappa_kl, m->d_kappa_kl, nloc * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(kap.data(), m->d_kappa, nloc * 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(mu.data(), m->d_mu, nloc * 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    for (size_t L = 0; L < nloc; ++L) {
      h_mu_kl[L] *= -2.0 * (double)mu[L];
      h_kappa_kl[L] *= -(double)kap[L];
    }
  }
}

// fp32 variant with SPECFEM's realw output type
void transfer_kernels_el_to_host_(long* Mesh_pointer, realw* h_rho_kl, realw* h_mu_kl,
                                  realw* h_kappa_kl, int* NSPEC_AB) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  const size_t nloc = (size_t)m->nspec * 125;
  std::vector<double> r(nloc), mu(nloc), k(nloc);
  transfer_kernels_el_to_host_dp_(Mesh_pointer, r.data(), mu.data(), k.data(), NSPEC_AB);
  for (size_t L = 0; L < nloc; ++L) {
    h_rho_kl[L] = (realw)r[L];
    if (!m->anisotropic_kl) { h_mu_kl[L] = (realw)mu[L]; h_kappa_kl[L] = (realw)k[L]; }
  }
}

// cijkl_kl (21,NGLLX,NGLLY,NGLLZ,NSPEC_AB) fp64, scaled by -1
void transfer_kernels_ani_to_host_dp_(long* Mesh_pointer, double* h_cijkl_kl, int* NSPEC_AB) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (*NSPEC_AB != m->nspec) SEM_FAIL("NSPEC_AB mismatch");
  if (!m->d_cijkl_kl) SEM_FAIL("anisotropic kernels not prepared");
  const size_t nloc = (size_t)m->nspec * 125;
  std::vector<double> tmp(nloc * 21);
  HIP_CHECK(hipMemcpyAsync(tmp.data(), m->d_cijkl_kl, tmp.size() * 8, hipMemcpyDeviceToHost, m->stream));
  HIP_CHECK(hipStreamSynchronize(m->stream));
  for (size_t e = 0; e < (size_t)m->nspec; ++e)
    for (int q = 0; q < 21; ++q)
      for (int L = 0; L < 125; ++L) h_cijkl_kl[((e * 125) + L) * 21 + q] = -tmp[(e * 21 + q) * 125 + L];
}

// approximate Hessian (sum of dt * a_adj . a_rec), (NGLLX,NGLLY,NGLLZ,NSPEC_AB)
void transfer_kernels_hess_el_tohost_dp_(long* Mesh_pointer, double* h_hess_kl, int* NSPEC_AB) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (*NSPEC_AB != m->nspec) SEM_FAIL("NSPEC_AB mismatch");
  if (!m->d_hess_kl_glob) SEM_FAIL("Hessian not prepared (APPROXIMATE_HESS_KL)");
  const size_t nloc = (size_t)m->nspec * 125;
  double* tmp = dev_alloc<double>(nloc);
  hipLaunchKernelGGL(k_expand_glob, dim3(nblk(nloc)), dim3(TPB), 0, m->stream, m->nspec, m->d_ibool,
                     m->d_hess_kl_glob, 1.0, (const realw*)nullptr, tmp);
  HIP_CHECK(hipMemcpyAsync(h_hess_kl, tmp, nloc * 8, hipMemcpyDeviceToHost, m->stream));
  HIP_CHECK(hipStreamSynchronize(m->stream));
  dev_free(tmp);
}

// Upload the whole-run source time functions (NSOURCES, NSTEP) and adjoint
// source traces (NDIM, nrec, NSTEP, forward time order) for fwat_iterate_time_.
void fwat_set_stf_(long* Mesh_pointer, double* h_stf, int* NSTEP) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (*NSTEP != m->nstep) SEM_FAIL("NSTEP mismatch");
  if (m->nsources == 0) return;
  dev_free(m->d_stf);
  m->d_stf = dev_upload(h_stf, (size_t)m->nstep * m->nsources, m->stream);
  HIP_CHECK(hipStreamSynchronize(m->stream));
}

// FK injection for teleseismic events: the incident velocity and traction
// (Veloc_FK, Tract_FK of couple_with_injection_prepare_boundary_fwat,
// fk_coupling.f90:71-297, layout [nsamp][npt][NDIM], npt = nfaces*25 face
// points in absorbing-face order) that the forward run's Stacey term injects;
// sample k belongs to step 1 + k*resamp, steps between samples interpolate
// linearly.  nsamp = 0 switches injection off.  The series stays in HBM.
void fwat_set_fk_injection_(long* Mesh_pointer, int* npt, int* nsamp, int* resamp, const realw* veloc_fk,
                            const realw* tract_fk) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  dev_free(m->d_fk_veloc);
  dev_free(m->d_fk_tract);
  m->d_fk_veloc = m->d_fk_tract = nullptr;
  m->fk_nsamp = 0;
  if (*nsamp == 0) return;
  if (!m->absorbing || m->nfaces == 0) SEM_FAIL("FK injection needs the absorbing faces");
  if (*npt != m->nfaces * 25) SEM_FAIL("FK injection: npt %d, the mesh has %d face points", *npt, m->nfaces * 25);
  if (*nsamp < 0 || *resamp < 1) SEM_FAIL("FK injection: nsamp %d, resamp %d", *nsamp, *resamp);
  if ((size_t)(m->nstep - 1) / (size_t)*resamp >= (size_t)*nsamp)
    SEM_FAIL("FK injection: %d samples every %d steps do not cover NSTEP %d", *nsamp, *resamp, m->nstep);
  const size_t n = (size_t)*nsamp * *npt * 3;
  m->d_fk_veloc = dev_upload(veloc_fk, n, m->stream);
  m->d_fk_tract = dev_upload(tract_fk, n, m->stream);
  m->fk_nsamp = *nsamp;
  m->fk_resamp = *resamp;
  HIP_CHECK(hipStreamSynchronize(m->stream));
}

void fwat_set_adjoint_sources_(long* Mesh_pointer, realw* h_adj, int* NSTEP) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (*NSTEP != m->nstep) SEM_FAIL("NSTEP mismatch");
  if (m->nrec == 0) return;
  dev_free(m->d_adj);
  m->d_adj = dev_upload(h_adj, (size_t)m->nstep * m->nrec * 3, m->stream);
  HIP_CHECK(hipStreamSynchronize(m->stream));
}

// Copy the forward run's final state (displ, veloc, accel) of handle `from`
// into the reconstructed field of handle `to` (both on this device): the
// in-HBM replacement of save_forward_arrays.bin (specfem_api.f90:312-358).
void fwat_copy_forward_to_backward_(long* Mesh_pointer_from, long* Mesh_pointer_to) {
  SemMesh* a = sem_mesh(Mesh_pointer_from);
  SemMesh* b = sem_mesh(Mesh_pointer_to);
  if (a->nglob != b->nglob || !b->d_b_displ) SEM_FAIL("fwat_copy_forward_to_backward_: incompatible handles");
  const size_t n = (size_t)a->nglob * 3 * sizeof(realw);
  HIP_CHECK(hipStreamSynchronize(a->stream));
  if (a->n_shared != b->n_shared || a->perm_hash != b->perm_hash || a->h_perm.size() != b->h_perm.size()) {
    // the handles number their points differently (other source or receiver elements):
    // through the host, in the caller's numbering
    std::vector<realw> u(3 * (size_t)a->nglob), v(u.size()), w(u.size());
    realw* const h[3] = {u.data(), v.data(), w.data()};
    realw* const da[3] = {a->d_displ, a->d_veloc, a->d_accel};
    realw* const db[3] = {b->d_b_displ, b->d_b_veloc, b->d_b_accel};
    sem_fields_from_device(a, 3 * a->nglob, h, da);
    sem_fields_to_device(b, 3 * b->nglob, h, db);
    return;
  }
  HIP_CHECK(hipMemcpyAsync(b->d_b_displ, a->d_displ, n, hipMemcpyDeviceToDevice, b->stream));
  HIP_CHECK(hipMemcpyAsync(b->d_b_veloc, a->d_veloc, n, hipMemcpyDeviceToDevice, b->stream));
  HIP_CHECK(hipMemcpyAsync(b->d_b_accel, a->d_accel, n, hipMemcpyDeviceToDevice, b->stream));
  HIP_CHECK(hipStreamSynchronize(b->stream));
}

// Whole run on the device.  SIMULATION_TYPE 1: forward steps 1..NSTEP with
// seismograms (and the Stacey store when SAVE_FORWARD).  SIMULATION_TYPE 3:
// adjoint + reconstruction + kernels.  fused = 1 uses the fused Newmark and
// the two-field force pass with in-pass kernel accumulation; fused = 0 runs
// the SPECFEM-style stage sequence (separate predictor/corrector, strain
// stores, compute_kernels).  timed = 1 brackets the force and assembly
// launches of every timer_stride(NSTEP)-th step with HIP events on the
// simulation stream (fwat_last_force_time_ / fwat_last_assembly_time_ report
// the sums over the sampled launches and their count).
void fwat_iterate_time_(long* Mesh_pointer, int* SIMULATION_TYPE, realw* DT, int* fused, int* timed) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  const realw dt = *DT;
  const int N = m->nstep;
  const bool fz = *fused != 0;
  m->last_force_ms = 0.0;
  m->last_force_launches = 0;
  m->last_asm_ms = 0.0;
  m->last_asm_launches = 0;
  ForceTimer timer, atimer;
  ForceTimer* tm = nullptr;
  ForceTimer* ta = nullptr;
  // graph capture covers the plain fused loops (not the brick or slab passes);
  // a captured run is not timed
  const bool graph = m->graph && fz && m->nslabs == 0 && !sem_brick_active(m, *SIMULATION_TYPE == 3 ? 2 : 1);
  if (*timed && !graph) {
    const size_t per_step = m->nslabs > 0 ? (size_t)m->nslabs : 1;  // bracketed launches per step
    timer.begin(N, per_step);
    atimer.begin(N, per_step);
    tm = &timer;
    ta = &atimer;
  }
  // called at the top of every step of every path below
  const auto step_begin = [&](int it) {
    if (tm) tm->at_step(it);
    if (ta) ta->at_step(it);
    if (!graph && it > 1 && (it - 1) % kSyncEvery == 0) HIP_CHECK(hipStreamSynchronize(m->stream));
  };
  if (m->nsources > 0 && !m->d_stf) SEM_FAIL("fwat_set_stf_ not called");
  hipStream_t s = m->stream;
  const realw dt2o2 = 0.5f * dt * dt, dto2 = 0.5f * dt;
  if (*SIMULATION_TYPE == 1) {
    const int smode = (m->save_forward && m->d_stacey_store) ? 1 : 0;
    if (fz && sem_brick_active(m, 1)) {
      // brick pass: elements + Newmark of brick-inner points; halo pass for the rest
      const int nfp = m->nfaces * 25;
      stage_predictor(m, false, dt);  // u_1, v_1/2 from the initial state
      for (int it = 1; it <= N; ++it) {
        step_begin(it);
        realw* face = nullptr;
        if (m->absorbing && m->nfaces)
          face = smode ? m->d_stacey_store + (size_t)(it - 1) * nfp * 3 : m->d_face_scratch;
        AsmField F0{nullptr, face, m->nsources ? m->d_stf + (size_t)(it - 1) * m->nsources : nullptr,
                    nullptr, m->d_displ, m->d_veloc, m->d_accel, dt};
        if (tm) tm->open(s);
        sem_launch_brick(m, 1, F0, F0, it == N, 0.0);
        if (tm) tm->close(s);
        if (face) stage_stacey_face(m, m->d_veloc, face, it);
        stage_seismo(m, it - 1);
        if (ta) ta->open(s);
        sem_launch_halo(m, 1, F0, F0, it == N, 0.0);
        if (ta) ta->close(s);
      }
    } else if (fz && m->nslabs > 0) {
      // slab order: Stacey terms and seismograms read the predicted state only,
      // so they come first; then element pass + assembly slab after slab
      const int nfp = m->nfaces * 25;
      stage_predictor(m, false, dt);  // u_1, v_1/2 from the initial state
      for (int it = 1; it <= N; ++it) {
        step_begin(it);
        realw* face = nullptr;
        if (m->absorbing && m->nfaces) {
          face = smode ? m->d_stacey_store + (size_t)(it - 1) * nfp * 3 : m->d_face_scratch;
          stage_stacey_face(m, m->d_veloc, face, it);
        }
        stage_seismo(m, it - 1);
        AsmField F0{m->d_floc, face, m->nsources ? m->d_stf + (size_t)(it - 1) * m->nsources : nullptr,
                    nullptr, m->d_displ, m->d_veloc, m->d_accel, dt};
        for (int sl = 0; sl < m->nslabs; ++sl) {
          const int e0 = sl * m->slab_elems, ne = std::min(m->slab_elems, m->nspec - e0);
          if (tm) tm->open(s);
          sem_launch_forces(m, 1, 0, e0, ne, dt, nullptr, true);
          if (tm) tm->close(s);
          if (ta) ta->open(s);
          stage_assemble(m, it == N, false, F0, F0, 0.0, sl);
          if (ta) ta->close(s);
        }
      }
    } else if (fz) {
      const int nfp = m->nfaces * 25;
      LoopGraph lg(m, graph);
      stage_predictor(m, false, dt);  // u_1, v_1/2 from the initial state
      // the element pass finishes the element-interior points (their Newmark update
      // needs nothing but their own element's term); the assembly covers the others
      const bool inner = sem_interior_active(m, 1);
      for (int it = 1; it <= N; ++it) {
        step_begin(it);
        const SemInterior in{AsmField{nullptr, nullptr, nullptr, nullptr, m->d_displ, m->d_veloc, m->d_accel, dt},
                             AsmField{}, it == N, 0.0};
        stage_forces_timed(m, 1, 0, dt, tm, true, inner ? &in : nullptr);
        realw* face = nullptr;
        if (m->absorbing && m->nfaces) {
          face = smode ? m->d_stacey_store + (size_t)(it - 1) * nfp * 3 : m->d_face_scratch;
          stage_stacey_face(m, m->d_veloc, face, it);
        }
        stage_seismo(m, it - 1);
        AsmField F0{m->d_floc, face, m->nsources ? m->d_stf + (size_t)(it - 1) * m->nsources : nullptr,
                    nullptr, m->d_displ, m->d_veloc, m->d_accel, dt};
        if (ta) ta->open(s);
        stage_assemble(m, it == N, false, F0, F0, 0.0, -1, inner);
        if (ta) ta->close(s);
      }
      lg.launch();
    } else {
      for (int it = 1; it <= N; ++it) {
        step_begin(it);
        stage_predictor(m, false, dt);
        stage_forces_timed(m, 1, 0, dt, tm, false);
        stage_stacey(m, false, smode, (size_t)(it - 1));
        if (m->nsources) stage_sources(m, false, m->d_stf + (size_t)(it - 1) * m->nsources);
        stage_corrector(m, false, dt);
        stage_seismo(m, it - 1);
      }
    }
  } else if (*SIMULATION_TYPE == 3) {
    if (!m->d_b_displ) SEM_FAIL("adjoint fields not prepared");
    if (m->absorbing && m->nfaces && !m->d_stacey_store) SEM_FAIL("adjoint run needs the forward Stacey store");
    if (m->absorbing && m->nfaces && m->stacey_store_steps < (size_t)N)
      SEM_FAIL("Stacey store holds %zu steps, adjoint run needs %d", m->stacey_store_steps, N);
    if (m->nrec && !m->d_adj) SEM_FAIL("fwat_set_adjoint_sources_ not called");
    if (!fz && (!m->d_epsilon || !m->d_b_epsilon)) SEM_FAIL("unfused adjoint run needs COMPUTE_AND_STORE_STRAIN");
    const double dtk = (double)dt;
    const size_t n3 = (size_t)m->nglob * 3;
    hipLaunchKernelGGL(k_b_init, dim3(nblk(n3)), dim3(TPB), 0, s, n3, m->d_b_veloc, m->d_b_accel, 0.5f * dt);
    const int kb = m->kernel_block;
    if (fz && kb > 1) {
      // fp32 partial sums of the element-pass kernels, zero at the start of the run
      const size_t nloc = (size_t)m->nspec * 125;
      if (m->anisotropic_kl) {
        if (!m->d_cij_part) m->d_cij_part = dev_alloc<float>(nloc * 21);
        hipLaunchKernelGGL(k_fill_f32, dim3(nblk(nloc * 21)), dim3(TPB), 0, s, m->d_cij_part, nloc * 21);
      }
      if (!m->d_rho_part) m->d_rho_part = dev_alloc<float>((size_t)m->nglob);
      hipLaunchKernelGGL(k_fill_f32, dim3(nblk((size_t)m->nglob)), dim3(TPB), 0, s, m->d_rho_part, (size_t)m->nglob);
      if (!m->anisotropic_kl) {
        if (!m->d_mu_part) m->d_mu_part = dev_alloc<float>(nloc);
        if (!m->d_kappa_part) m->d_kappa_part = dev_alloc<float>(nloc);
        hipLaunchKernelGGL(k_fill_f32, dim3(nblk(nloc)), dim3(TPB), 0, s, m->d_mu_part, nloc);
        hipLaunchKernelGGL(k_fill_f32, dim3(nblk(nloc)), dim3(TPB), 0, s, m->d_kappa_part, nloc);
      }
    }
    // a block closes every kb steps and at the last step
    const auto closes_block = [kb, N](int it) { return kb <= 1 || it % kb == 0 || it == N; };
    if (fz && sem_brick_active(m, 2) && !m->anisotropic_kl) {
      const int nfp = m->nfaces * 25;
      stage_predictor(m, false, dt);  // adjoint field (from rest)
      for (int it = 1; it <= N; ++it) {
        step_begin(it);
        const int rec0 = N - it;  // 0-based record / source row of forward step N + 1 - it
        realw* face_a = (m->absorbing && m->nfaces) ? m->d_face_scratch : nullptr;
        const realw* face_b = (m->absorbing && m->nfaces) ? m->d_stacey_store + (size_t)rec0 * nfp * 3 : nullptr;
        AsmField F0{nullptr, face_a, nullptr, m->nrec ? m->d_adj + (size_t)rec0 * m->nrec * 3 : nullptr,
                    m->d_displ, m->d_veloc, m->d_accel, dt};
        AsmField F1{nullptr, face_b, m->nsources ? m->d_stf + (size_t)rec0 * m->nsources : nullptr,
                    nullptr, m->d_b_displ, m->d_b_veloc, m->d_b_accel, -dt};
        m->kernel_flush = closes_block(it);
        if (tm) tm->open(s);
        sem_launch_brick(m, 2, F0, F1, it == N, dtk);
        if (tm) tm->close(s);
        if (face_a) stage_stacey_face(m, m->d_veloc, face_a);
        if (ta) ta->open(s);
        sem_launch_halo(m, 2, F0, F1, it == N, dtk);
        if (ta) ta->close(s);
      }
    } else if (fz && m->nslabs > 0) {
      const int nfp = m->nfaces * 25;
      stage_predictor(m, false, dt);  // adjoint field (from rest)
      for (int it = 1; it <= N; ++it) {
        step_begin(it);
        const int rec0 = N - it;  // 0-based record / source row of forward step N + 1 - it
        m->kernel_flush = closes_block(it);
        realw* face_a = nullptr;
        if (m->absorbing && m->nfaces) {
          face_a = m->d_face_scratch;
          stage_stacey_face(m, m->d_veloc, face_a);
        }
        const realw* face_b = (m->absorbing && m->nfaces) ? m->d_stacey_store + (size_t)rec0 * nfp * 3 : nullptr;
        AsmField F0{m->d_floc, face_a, nullptr, m->nrec ? m->d_adj + (size_t)rec0 * m->nrec * 3 : nullptr,
                    m->d_displ, m->d_veloc, m->d_accel, dt};
        AsmField F1{m->d_floc_b, face_b, m->nsources ? m->d_stf + (size_t)rec0 * m->nsources : nullptr,
                    nullptr, m->d_b_displ, m->d_b_veloc, m->d_b_accel, -dt};
        for (int sl = 0; sl < m->nslabs; ++sl) {
          const int e0 = sl * m->slab_elems, ne = std::min(m->slab_elems, m->nspec - e0);
          if (tm) tm->open(s);
          sem_launch_forces(m, 2, 2, e0, ne, dt, nullptr, true);
          if (tm) tm->close(s);
          if (ta) ta->open(s);
          stage_assemble(m, it == N, true, F0, F1, dtk, sl);
          if (ta) ta->close(s);
        }
      }
    } else if (fz) {
      const int nfp = m->nfaces * 25;
      LoopGraph lg(m, graph);
      stage_predictor(m, false, dt);  // adjoint field (from rest)
      const bool inner = sem_interior_active(m, 2);
      for (int it = 1; it <= N; ++it) {
        step_begin(it);
        const int rec0 = N - it;  // 0-based record / source row of forward step N + 1 - it
        m->kernel_flush = closes_block(it);
        const SemInterior in{AsmField{nullptr, nullptr, nullptr, nullptr, m->d_displ, m->d_veloc, m->d_accel, dt},
                             AsmField{nullptr, nullptr, nullptr, nullptr, m->d_b_displ, m->d_b_veloc, m->d_b_accel, -dt},
                             it == N, dtk};
        stage_forces_timed(m, 2, 2, dt, tm, true, inner ? &in : nullptr);
        realw* face_a = nullptr;
        if (m->absorbing && m->nfaces) {
          face_a = m->d_face_scratch;
          stage_stacey_face(m, m->d_veloc, face_a);
        }
        const realw* face_b = (m->absorbing && m->nfaces) ? m->d_stacey_store + (size_t)rec0 * nfp * 3 : nullptr;
        AsmField F0{m->d_floc, face_a, nullptr, m->nrec ? m->d_adj + (size_t)rec0 * m->nrec * 3 : nullptr,
                    m->d_displ, m->d_veloc, m->d_accel, dt};
        AsmField F1{m->d_floc_b, face_b, m->nsources ? m->d_stf + (size_t)rec0 * m->nsources : nullptr,
                    nullptr, m->d_b_displ, m->d_b_veloc, m->d_b_accel, -dt};
        if (ta) ta->open(s);
        stage_assemble(m, it == N, true, F0, F1, dtk, -1, inner);
        if (ta) ta->close(s);
      }
      lg.launch();
    } else {
      for (int it = 1; it <= N; ++it) {
        step_begin(it);
        const int rec0 = N - it;
        if (it >= 2) stage_predictor(m, true, dt);
        else hipLaunchKernelGGL(k_fill_f32, dim3(nblk(n3)), dim3(TPB), 0, s, m->d_b_accel, n3);
        stage_predictor(m, false, dt);
        if (tm) tm->open(s);
        sem_launch_forces(m, 3, 1, 0, m->nspec, dt, nullptr, false);
        sem_launch_forces(m, 1, 1, 0, m->nspec, dt, nullptr, false);
        if (tm) tm->close(s, 2);
        stage_stacey(m, true, 2, (size_t)rec0);
        if (m->nsources) stage_sources(m, true, m->d_stf + (size_t)rec0 * m->nsources);
        stage_stacey(m, false, 0, 0);
        if (m->nrec) stage_adjoint_sources(m, m->d_adj + (size_t)rec0 * m->nrec * 3);
        stage_corrector(m, true, dt);
        stage_corrector(m, false, dt);
        compute_kernels_elastic_cuda_(Mesh_pointer, DT);
      }
    }
  } else {
    SEM_FAIL("SIMULATION_TYPE %d not supported", *SIMULATION_TYPE);
  }
  m->kernel_flush = 1;
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(s));
  if (tm) {
    m->last_force_ms = tm->total_ms();
    m->last_force_launches = tm->launches;
  }
  if (ta) {
    m->last_asm_ms = ta->total_ms();
    m->last_asm_launches = ta->launches;
  }
}

// Sum of the force-kernel durations of the last fwat_iterate_time_ call (ms)
// and the number of timed launches.
void fwat_last_force_time_(long* Mesh_pointer, double* ms, int* launches) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  *ms = m->last_force_ms;
  *launches = m->last_force_launches;
}

// Plain device copy of nbytes (float4 loads/stores), repeated reps times;
// avg_ms is the mean duration.  A reference kernel of known traffic for
// scaling hardware byte counters.
void fwat_copy_probe_(long* nbytes, int* reps, double* avg_ms) {
  const size_t n4 = (size_t)*nbytes / 16;
  if (n4 == 0 || *reps <= 0) SEM_FAIL("fwat_copy_probe_: empty probe");
  float4 *a = nullptr, *b = nullptr;
  HIP_CHECK(hipMalloc((void**)&a, n4 * 16));
  HIP_CHECK(hipMalloc((void**)&b, n4 * 16));
  HIP_CHECK(hipMemset(a, 0, n4 * 16));
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  const int blocks = (int)std::min<size_t>((n4 + 255) / 256, 1u << 20);
  hipLaunchKernelGGL(k_copy4, dim3(blocks), dim3(256), 0, 0, a, b, n4);
  HIP_CHECK(hipEventRecord(e0, 0));
  for (int r = 0; r < *reps; ++r) hipLaunchKernelGGL(k_copy4, dim3(blocks), dim3(256), 0, 0, a, b, n4);
  HIP_CHECK(hipEventRecord(e1, 0));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  *avg_ms = ms / *reps;
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  HIP_CHECK(hipFree(a));
  HIP_CHECK(hipFree(b));
}

// Copy-rate probe: the plain copy of nbytes with unroll float4 loads in flight
// per lane (1, 2, 4 or 8) on a grid of ncu * blocks_per_cu workgroups of 256
// (0: one float4 per lane over the whole array), repeated reps times; avg_ms
// is the mean duration.  Measures what a simple streaming kernel reaches on
// this device, beside the nominal HBM peak.
void fwat_copy_probe_unrolled_(long* nbytes, int* reps, int* unroll, int* blocks_per_cu, double* avg_ms) {
  const size_t n4 = (size_t)*nbytes / 16;
  if (n4 == 0 || *reps <= 0) SEM_FAIL("fwat_copy_probe_unrolled_: empty probe");
  if (*unroll != 1 && *unroll != 2 && *unroll != 4 && *unroll != 8) SEM_FAIL("unroll must be 1, 2, 4 or 8");
  int dev = 0, ncu = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  float4 *a = nullptr, *b = nullptr;
  HIP_CHECK(hipMalloc((void**)&a, n4 * 16));
  HIP_CHECK(hipMalloc((void**)&b, n4 * 16));
  hipLaunchKernelGGL(k_copy4, dim3((unsigned)std::min<size_t>((n4 + 255) / 256, 1u << 20)), dim3(256), 0, 0,
                     b, a, n4);  // touch both arrays
  const int blocks = *blocks_per_cu > 0 ? ncu * *blocks_per_cu
                                        : (int)std::min<size_t>((n4 / *unroll + 255) / 256, 1u << 20);
  const auto launch = [&]() {
    switch (*unroll) {
      case 1: hipLaunchKernelGGL(k_copy4u<1>, dim3(blocks), dim3(256), 0, 0, a, b, n4); break;
      case 2: hipLaunchKernelGGL(k_copy4u<2>, dim3(blocks), dim3(256), 0, 0, a, b, n4); break;
      case 4: hipLaunchKernelGGL(k_copy4u<4>, dim3(blocks), dim3(256), 0, 0, a, b, n4); break;
      default: hipLaunchKernelGGL(k_copy4u<8>, dim3(blocks), dim3(256), 0, 0, a, b, n4); break;
    }
  };
  launch();
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  HIP_CHECK(hipEventRecord(e0, 0));
  for (int r = 0; r < *reps; ++r) launch();
  HIP_CHECK(hipEventRecord(e1, 0));
  HIP_CHECK(hipEventSynchronize(e1));
  HIP_CHECK(hipGetLastError());
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  *avg_ms = ms / *reps;
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  HIP_CHECK(hipFree(a));
  HIP_CHECK(hipFree(b));
}

// Points per lane of the gather assembly: 1 (k_assemble), 2 (k_assemble2, twice
// the loads in flight per lane; the same sums) or 0 (default: 2 for the one-field
// pass, 1 for the two-field pass)
// Lanes per workgroup of the gather assembly launches (a multiple of 64 in 64..1024;
// default 256).  Consecutive workgroups go to the XCDs in turn, so larger workgroups keep
// longer runs of consecutive points (and the element rows they share) on one XCD.
void fwat_set_assembly_block_(long* Mesh_pointer, int* lanes) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (*lanes < 64 || *lanes > 1024 || *lanes % 64) SEM_FAIL("fwat_set_assembly_block_: 64..1024, a multiple of 64");
  m->asm_block = *lanes;
}

void fwat_set_assembly_lanes_(long* Mesh_pointer, int* points_per_lane) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (*points_per_lane < 0 || *points_per_lane > 2) SEM_FAIL("fwat_set_assembly_lanes_: 0, 1 or 2 points per lane");
  m->asm_points_per_lane = *points_per_lane;
}

// Element-interior points (one element-local term, no face, source or receiver
// term; numbered after the shared points by prepare_constants_device_) finished by
// the element pass of the fused loops: 1 = on (default; pipelined passes 3 and 5 on
// isotropic all-regular meshes, without slabs or bricks), 0 = every point through
// the gather assembly.  Results are the same bit for bit.
void fwat_set_interior_(long* Mesh_pointer, int* on) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  m->interior_newmark = *on != 0;
}

// Points the gather assembly of a fused loop covers, and the element-interior points
// the element pass finishes when fwat_set_interior_ is on (for a forward and a
// two-field pass with the current launch shape).
void fwat_interior_info_(long* Mesh_pointer, int* n_shared, int* n_interior, int* active_fwd, int* active_adj) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  *n_shared = m->n_shared;
  *n_interior = m->nglob - m->n_shared;
  *active_fwd = sem_interior_active(m, 1) ? 1 : 0;
  *active_adj = sem_interior_active(m, 2) ? 1 : 0;
}

// HIP graph capture of the whole-run loop (fused gather path): 1 = every
// fwat_iterate_time_ run captures its step loop into one graph and launches
// it (no per-kernel timing), 0 = launches on the stream (default).
void fwat_set_graph_(long* Mesh_pointer, int* on) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  m->graph = *on != 0;
}

// Launch-shape knobs of the element passes (set by the autotuner in
// specfwat_amd/tune.py): regular_path 1 = folded-table kernel on all-regular
// meshes, 2 = its software-pipelined form, 3 = the pipelined form with two
// barriers per element, 4 = form 3 with non-temporal hints on the streamed
// arrays, 5 = form 3 visiting the elements of a lattice mesh in column blocks
// (isotropic; anisotropic passes keep 1), 0 = general
// kernel; blocks_per_cu_* = persistent workgroups per CU
// for the one-field and two-field passes, 0 = one workgroup per element.
void fwat_set_tuning_(long* Mesh_pointer, int* regular_path, int* blocks_per_cu_fwd, int* blocks_per_cu_adj) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (*blocks_per_cu_fwd < 0 || *blocks_per_cu_adj < 0 || *blocks_per_cu_fwd > 64 || *blocks_per_cu_adj > 64)
    SEM_FAIL("fwat_set_tuning_: blocks per CU must be in 0..64");
  if (*regular_path < 0 || *regular_path > 5) SEM_FAIL("fwat_set_tuning_: regular_path must be 0 to 5");
  m->regular_fast_path = *regular_path;
  m->blocks_per_cu_fwd = *blocks_per_cu_fwd;
  m->blocks_per_cu_adj = *blocks_per_cu_adj;
}

// Element slabs of the fused loops: contiguous ranges of elems_per_slab
// elements (0: off).  Each point is assembled right after the slab holding
// its last contributing element, so the slab's element-local forces are read
// back while still in the memory-side cache.  Valid for any mesh; worth it
// when a point's elements are close in element order (lattice-ordered
// meshes).  Irregular meshes, anisotropy and the per-step entry points are
// unaffected by the ordering.
void fwat_set_slabs_(long* Mesh_pointer, int* elems_per_slab) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  dev_free(m->d_slab_g);
  m->h_slab_goff.clear();
  m->nslabs = 0;
  m->slab_elems = 0;
  const int C = *elems_per_slab;
  if (C <= 0 || C >= m->nspec) return;
  const size_t nloc = (size_t)m->nspec * 125;
  std::vector<int> ib(nloc);
  HIP_CHECK(hipMemcpyAsync(ib.data(), m->d_ibool, nloc * sizeof(int), hipMemcpyDeviceToHost, m->stream));
  HIP_CHECK(hipStreamSynchronize(m->stream));
  std::vector<int> last(m->nglob, -1);
  for (size_t L = 0; L < nloc; ++L) last[ib[L]] = std::max(last[ib[L]], (int)(L / 125));
  const int ns = (m->nspec + C - 1) / C;
  std::vector<int> off(ns + 1, 0), pts(m->nglob);
  for (int g = 0; g < m->nglob; ++g) {
    if (last[g] < 0) SEM_FAIL("fwat_set_slabs_: point %d belongs to no element", g + 1);
    off[last[g] / C + 1]++;
  }
  for (int q = 0; q < ns; ++q) off[q + 1] += off[q];
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (int g = 0; g < m->nglob; ++g) pts[fill[last[g] / C]++] = g;  // ascending within a slab
  m->d_slab_g = dev_upload(pts.data(), pts.size(), m->stream);
  m->h_slab_goff = off;
  m->nslabs = ns;
  m->slab_elems = C;
  HIP_CHECK(hipStreamSynchronize(m->stream));
}

// Steps per fp32 partial-sum block of the element-pass kernels (mu/kappa or
// c_ij) in the fused adjoint loop: 1 accumulates every step straight into the
// fp64 arrays; n > 1 sums n steps in fp32 before adding them into fp64.
void fwat_set_kernel_block_(long* Mesh_pointer, int* nsteps) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (*nsteps < 1 || *nsteps > 256) SEM_FAIL("fwat_set_kernel_block_: steps per block must be in 1..256");
  m->kernel_block = *nsteps;
}

// Summed durations (ms) and launch counts of the assembly passes of the
// last timed fwat_iterate_time_ call.
void fwat_last_assembly_time_(long* Mesh_pointer, double* ms, int* launches) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  *ms = m->last_asm_ms;
  *launches = m->last_asm_launches;
}

void fwat_synchronize_(long* Mesh_pointer) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  HIP_CHECK(hipStreamSynchronize(m->stream));
}

}  // extern "C"
