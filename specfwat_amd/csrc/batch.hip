// Multi-event batches: E events (own source, source time function, NSTEP
// and adjoint sources) on one mesh and model, advanced together.
//
// No counterpart in the reference: SpecFWAT loops over events one at a time
// (src/preproc/fwat_fwd_measure_adj.f90:43-51) and SPECFEM's batched-run
// option is acoustic forward only (read_mesh_databases.f90:261-268).  The
// per-event arithmetic is the single-event fused path's; batching changes the
// data layout and what is read once for all events:
//   fields u, v, a              [nglob][E][3]      (events innermost)
//   element-local forces        [nloc][E][3] (element rows grouped by reader, sem_floc_slot)
//   Stacey store                [NSTEP_max][nfp][E][3]  fp32, or fp16 with one scale per
//                                                        (record, event) (fwat_batch_set_store_precision_)
//   rho kernel                  [nglob][E] fp64;  mu, kappa kernels [E][nloc] fp64
//   statics (ibool, kappa/mu), CSR, rmass_inv, face geometry: once per element
//   or point for all E events.
// Element pass: one workgroup per element, every lane loads the 3E values of
// its point in one contiguous read and writes its 3E forces in one contiguous
// write.  Assembly, Stacey and Newmark: one lane per (point, event), so a
// wavefront touches contiguous memory.
// NSTEP may differ per event: the forward run advances event e for its first
// NSTEP_e steps (its state is frozen afterwards); the adjoint run is aligned
// at the end (event e idles for the first NSTEP_max - NSTEP_e steps), so at
// global adjoint step it every active event replays Stacey row NSTEP_max - it,
// which is its own forward step NSTEP_e + 1 - it_e.
// Regular (all-cube) meshes, isotropic (kappa, mu; rho/mu/kappa kernels) or
// anisotropic (21 c_ij; rho and 21 c_ij kernels [E][nspec][21][125]); meshes
// with irregular elements are refused.
#include "sem_device.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

constexpr int EMAX = 8;
constexpr int TPB = 256;

inline int nblk(size_t n) { return (int)((n + TPB - 1) / TPB); }

struct BatchElemArgs {
  const int* ibool;
  const realw* kappa;
  const realw* mu;
  const uint16_t* kmu_lp;
  const float* kmu_base;  // [nspec][2] per-element means (reduced-precision statics)
  const realw* tables;
  int nspec;
  realw xix_regular, jacobian_regular;
  float scale_k, scale_m;
  realw* u0;         // [nglob][E][3]
  realw* u1;
  realw* floc0;      // [nloc][E][3]
  realw* floc1;
  double* mu_kl;     // [E][nloc]
  double* kappa_kl;
  float* mu_part;    // [E][nloc] fp32 partial sums (null: straight into fp64)
  float* kappa_part;
  int kflush;        // this step closes a partial-sum block
  double dt;
  int active_mask;  // bit e set: event e advances this step
  const realw* cij;  // [nspec][21][125] (anisotropic)
  double* cijkl_kl;  // [E][nspec][21][125]
  float* cij_part;   // [E][nspec][21][125] fp32 partial sums
  // element-interior points (k_batch_elem<..., true>): a point >= n_shared of an
  // element that holds none of the batch's sources has this element's term only,
  // so the pass finishes it (Newmark of every active event, density kernel)
  int n_shared;
  int n_src;
  int src_el[EMAX];  // the batch's source elements (their points are gathered)
  int last_mask;     // bit e set: this is event e's last step of the run
  const realw* rm;   // [nglob][3] inverse mass
  realw *v0, *a0, *v1, *a1;  // [nglob][E][3] (field 1 runs backwards in time)
  double* rho_kl;    // [nglob][E]
  float* rho_part;   // [nglob][E]
  realw fdt;         // field time step
};

__device__ __forceinline__ float bf16f(uint16_t x) { return __uint_as_float((uint32_t)x << 16); }
__device__ __forceinline__ float h16f(uint16_t x) {
  _Float16 h;
  __builtin_memcpy(&h, &x, 2);
  return (float)h;
}

// contiguous 3E floats of one point (16-byte vectors when E is a multiple of 4)
template <int E>
__device__ __forceinline__ void load_pt(const realw* __restrict__ p, realw (&x)[3 * E]) {
  if constexpr (E % 4 == 0) {
    const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
    for (int k = 0; k < 3 * E / 4; ++k) {
      const float4 v = q[k];
      x[4 * k] = v.x; x[4 * k + 1] = v.y; x[4 * k + 2] = v.z; x[4 * k + 3] = v.w;
    }
  } else if constexpr (E == 2) {
    const float2* q = reinterpret_cast<const float2*>(p);
#pragma unroll
    for (int k = 0; k < 3; ++k) { const float2 v = q[k]; x[2 * k] = v.x; x[2 * k + 1] = v.y; }
  } else {
#pragma unroll
    for (int k = 0; k < 3 * E; ++k) x[k] = p[k];
  }
}

template <int E>
__device__ __forceinline__ void store_pt(realw* __restrict__ p, const realw (&x)[3 * E]) {
  if constexpr (E % 4 == 0) {
    float4* q = reinterpret_cast<float4*>(p);
#pragma unroll
    for (int k = 0; k < 3 * E / 4; ++k) q[k] = make_float4(x[4 * k], x[4 * k + 1], x[4 * k + 2], x[4 * k + 3]);
  } else if constexpr (E == 2) {
    float2* q = reinterpret_cast<float2*>(p);
#pragma unroll
    for (int k = 0; k < 3; ++k) q[k] = make_float2(x[2 * k], x[2 * k + 1]);
  } else {
#pragma unroll
    for (int k = 0; k < 3 * E; ++k) p[k] = x[k];
  }
}

// Newmark finish of event ev at a point whose only term is one element's: the
// element pass holds the point's forces f0, f1 and displacements u0, u1; its inverse
// mass r, velocities v0, v1 and density-kernel partial sum were loaded ahead.  Same
// expressions and rounding as k_batch_asm (the gather sums such a row as 0 + term).
template <int E, bool TWO>
__device__ __forceinline__ void batch_point_newmark(const BatchElemArgs& A, int g, int ev, const realw (&r)[3],
                                                    const realw (&f0)[3], const realw (&f1)[3],
                                                    const realw (&u0)[3], const realw (&u1)[3],
                                                    const realw (&v0)[3], const realw (&v1)[3], float part) {
  const size_t o = (size_t)g * (3 * E) + 3 * ev;
  const bool last = (A.last_mask >> ev) & 1;
  const realw z = 0.0f, d0 = A.fdt, h0 = 0.5f * A.fdt, s0 = 0.5f * A.fdt * A.fdt;
  realw Ac[3], V[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    Ac[c] = (z + f0[c]) * r[c];
    V[c] = fmaf(h0, Ac[c], v0[c]);
  }
  if (TWO) {
    const realw d1 = -A.fdt, h1 = 0.5f * d1, s1 = 0.5f * d1 * d1;
    realw BA[3], BV[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      BA[c] = (z + f1[c]) * r[c];
      BV[c] = fmaf(h1, BA[c], v1[c]);
    }
    // kernel_acc with the partial sum loaded ahead
    const float x = fmaf(Ac[2], u1[2], fmaf(Ac[1], u1[1], Ac[0] * u1[0]));
    const size_t q = (size_t)g * E + ev;
    if (!A.rho_part) {
      A.rho_kl[q] = fma(A.dt, (double)x, A.rho_kl[q]);
    } else if (A.kflush) {
      A.rho_kl[q] = fma(A.dt, (double)(part + x), A.rho_kl[q]);
      A.rho_part[q] = 0.0f;
    } else {
      A.rho_part[q] = part + x;
    }
    if (last) {
      store3(A.a1 + o, BA[0], BA[1], BA[2]);
      store3(A.v1 + o, BV[0], BV[1], BV[2]);
    } else {
      store3(A.u1 + o, fmaf(s1, BA[0], fmaf(d1, BV[0], u1[0])), fmaf(s1, BA[1], fmaf(d1, BV[1], u1[1])),
             fmaf(s1, BA[2], fmaf(d1, BV[2], u1[2])));
      store3(A.v1 + o, fmaf(h1, BA[0], BV[0]), fmaf(h1, BA[1], BV[1]), fmaf(h1, BA[2], BV[2]));
    }
  }
  if (last) {
    store3(A.a0 + o, Ac[0], Ac[1], Ac[2]);
    store3(A.v0 + o, V[0], V[1], V[2]);
  } else {
    store3(A.u0 + o, fmaf(s0, Ac[0], fmaf(d0, V[0], u0[0])), fmaf(s0, Ac[1], fmaf(d0, V[1], u0[1])),
           fmaf(s0, Ac[2], fmaf(d0, V[2], u0[2])));
    store3(A.v0 + o, fmaf(h0, Ac[0], V[0]), fmaf(h0, Ac[1], V[1]), fmaf(h0, Ac[2], V[2]));
  }
}

// Element pass of E events (NF = 1 forward, NF = 2 adjoint + reconstruction
// with in-pass mu/kappa kernels), regular iso meshes, folded tables,
// persistent XCD-chunked schedule.  The fields of all events are loaded
// together (3E contiguous floats per lane), then the events are processed one
// after the other out of LDS; each event's forces overwrite its own LDS slot
// and are stored together at the end (3E contiguous floats per lane).
// INT: the element-interior points are finished here (batch_point_newmark) and
// their rows, the last 27 of each element row, are not stored.
// (the waves-per-SIMD floor is what the pass without the interior points reaches,
// so that their Newmark tail does not cost occupancy)
constexpr int batch_elem_waves(int E, int NF) { return NF == 2 ? (E >= 8 ? 3 : 4) : (E >= 8 ? 5 : E == 4 ? 6 : 7); }

template <int E, int NF, int SP, bool INT>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(INT ? 1 : batch_elem_waves(E, NF))))
void k_batch_elem(BatchElemArgs A) {
  __shared__ realw s_tab[128];
  __shared__ realw s_u[NF][E][3][128];
  __shared__ realw s_s[NF][6][128];
  const int t = threadIdx.x;
  if (t < 125) s_tab[t] = A.tables[t];
  const bool act = t < 125;
  const int tt = act ? t : 124;
  const int I = tt % 5, J = (tt / 5) % 5, K = tt / 25;
  __syncthreads();
  const realw xr = A.xix_regular;
  const realw jx = A.jacobian_regular * A.xix_regular;
  realw hpI[5], hpJ[5], hpK[5], hwI[5], hwJ[5], hwK[5];
  {
    const realw f1 = -jx * s_tab[50 + J + 5 * K];
    const realw f2 = -jx * s_tab[75 + I + 5 * K];
    const realw f3 = -jx * s_tab[100 + I + 5 * J];
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      hpI[l] = s_tab[I + 5 * l] * xr;
      hpJ[l] = s_tab[J + 5 * l] * xr;
      hpK[l] = s_tab[K + 5 * l] * xr;
      hwI[l] = s_tab[25 + l + 5 * I] * f1;
      hwJ[l] = s_tab[25 + l + 5 * J] * f2;
      hwK[l] = s_tab[25 + l + 5 * K] * f3;
    }
  }
  const int La0 = 5 * J + 25 * K, Lb0 = I + 25 * K, Lc0 = I + 5 * J;
  const int nb = gridDim.x;
  const int x = blockIdx.x & 7, per_xcd = nb >> 3, slot = blockIdx.x >> 3;
  const int chunk = (A.nspec + 7) >> 3;
  const int e_end = min(A.nspec, (x + 1) * chunk);
  const size_t nloc = (size_t)A.nspec * 125;
  const int mask = A.active_mask;
  for (int e = x * chunk + slot; e < e_end; e += per_xcd) {
    const size_t Lg = (size_t)e * 125 + tt;
    const int iglob = A.ibool[Lg];
    bool src = false;  // uniform over the workgroup
    if (INT)
      for (int k = 0; k < A.n_src; ++k) src |= e == A.src_el[k];
    realw kl, ml;
    if (SP == SEM_STATIC_FP32) {
      kl = A.kappa[Lg];
      ml = A.mu[Lg];
    } else {
      const uint32_t pk = reinterpret_cast<const uint32_t*>(A.kmu_lp)[Lg];
      const float2 kb = reinterpret_cast<const float2*>(A.kmu_base)[e];
      kl = kb.x + (SP == SEM_STATIC_BF16 ? bf16f((uint16_t)(pk & 0xffffu)) : h16f((uint16_t)(pk & 0xffffu))) * A.scale_k;
      ml = kb.y + (SP == SEM_STATIC_BF16 ? bf16f((uint16_t)(pk >> 16)) : h16f((uint16_t)(pk >> 16))) * A.scale_m;
    }
    const realw lam = kl - (2.0f / 3.0f) * ml;
    const realw m2 = 2.0f * ml;
    __syncthreads();  // previous element done with s_u
    if (act) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        realw x3[3 * E];
        load_pt<E>((f == 0 ? A.u0 : A.u1) + (size_t)iglob * (3 * E), x3);
#pragma unroll
        for (int ev = 0; ev < E; ++ev) {
          s_u[f][ev][0][t] = x3[3 * ev];
          s_u[f][ev][1][t] = x3[3 * ev + 1];
          s_u[f][ev][2][t] = x3[3 * ev + 2];
        }
      }
    }
    const bool inner = INT && act && !src && iglob >= A.n_shared;
    realw rmi[3] = {0.0f, 0.0f, 0.0f};
    if (INT && inner) {
#pragma unroll
      for (int c = 0; c < 3; ++c) rmi[c] = A.rm[3 * (size_t)iglob + c];
    }
    __syncthreads();
#pragma unroll 1
    for (int ev = 0; ev < E; ++ev) {
      if (!((mask >> ev) & 1)) continue;  // uniform over the grid
      // an interior point's velocities and density-kernel partial sum for this event,
      // loaded now so that they arrive during the contractions
      realw pv0[3] = {0.0f, 0.0f, 0.0f}, pv1[3] = {0.0f, 0.0f, 0.0f};
      float ppart = 0.0f;
      if (INT && inner) {
        const size_t o = (size_t)iglob * (3 * E) + 3 * ev;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          pv0[c] = A.v0[o + c];
          if (NF == 2) pv1[c] = A.v1[o + c];
        }
        if (NF == 2 && A.rho_part) ppart = A.rho_part[(size_t)iglob * E + ev];
      }
      realw ea0 = 0, ea1 = 0, ea2 = 0, ea3 = 0, ea4 = 0, ea5 = 0;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        realw duxdx = 0, duydx = 0, duzdx = 0, duxdy = 0, duydy = 0, duzdy = 0, duxdz = 0, duydz = 0, duzdz = 0;
#pragma unroll
        for (int l = 0; l < 5; ++l) {
          const int La = La0 + l, Lb = Lb0 + 5 * l, Lc = Lc0 + 25 * l;
          duxdx += s_u[f][ev][0][La] * hpI[l]; duydx += s_u[f][ev][1][La] * hpI[l]; duzdx += s_u[f][ev][2][La] * hpI[l];
          duxdy += s_u[f][ev][0][Lb] * hpJ[l]; duydy += s_u[f][ev][1][Lb] * hpJ[l]; duzdy += s_u[f][ev][2][Lb] * hpJ[l];
          duxdz += s_u[f][ev][0][Lc] * hpK[l]; duydz += s_u[f][ev][1][Lc] * hpK[l]; duzdz += s_u[f][ev][2][Lc] * hpK[l];
        }
        const realw exy2 = duxdy + duydx, exz2 = duzdx + duxdz, eyz2 = duzdy + duydz;
        if (NF == 2) {
          if (f == 0) {
            ea0 = duxdx; ea1 = duydy; ea2 = duzdz;
            ea3 = 0.5f * exy2; ea4 = 0.5f * exz2; ea5 = 0.5f * eyz2;
          } else if (act) {
            const realw eb0 = duxdx, eb1 = duydy, eb2 = duzdz;
            const realw eb3 = 0.5f * exy2, eb4 = 0.5f * exz2, eb5 = 0.5f * eyz2;
            const realw tra = ea0 + ea1 + ea2, trb = eb0 + eb1 + eb2;
            const realw t3a = tra / 3.0f, t3b = trb / 3.0f;
            const realw prod = (ea0 - t3a) * (eb0 - t3b) + (ea1 - t3a) * (eb1 - t3b) + (ea2 - t3a) * (eb2 - t3b) +
                               2.0f * (ea3 * eb3 + ea4 * eb4 + ea5 * eb5);
            const size_t kq = (size_t)ev * nloc + Lg;
            kernel_acc(A.mu_kl, A.mu_part, kq, prod, A.dt, A.kflush != 0);
            kernel_acc(A.kappa_kl, A.kappa_part, kq, tra * trb, A.dt, A.kflush != 0);
          }
        }
        const realw ltr = lam * (duxdx + duydy + duzdz);
        s_s[f][0][t] = ltr + m2 * duxdx; s_s[f][1][t] = ltr + m2 * duydy; s_s[f][2][t] = ltr + m2 * duzdz;
        s_s[f][3][t] = ml * exy2; s_s[f][4][t] = ml * exz2; s_s[f][5][t] = ml * eyz2;
      }
      __syncthreads();  // stresses complete; every lane is done reading s_u[*][ev]
      realw ff[NF][3];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        realw fx = 0, fy = 0, fz = 0;
#pragma unroll
        for (int l = 0; l < 5; ++l) {
          const int La = La0 + l, Lb = Lb0 + 5 * l, Lc = Lc0 + 25 * l;
          fx += s_s[f][0][La] * hwI[l]; fy += s_s[f][3][La] * hwI[l]; fz += s_s[f][4][La] * hwI[l];
          fx += s_s[f][3][Lb] * hwJ[l]; fy += s_s[f][1][Lb] * hwJ[l]; fz += s_s[f][5][Lb] * hwJ[l];
          fx += s_s[f][4][Lc] * hwK[l]; fy += s_s[f][5][Lc] * hwK[l]; fz += s_s[f][2][Lc] * hwK[l];
        }
        ff[f][0] = fx; ff[f][1] = fy; ff[f][2] = fz;
      }
      if (INT && inner) {
        // own slot: this event's displacements, before the forces replace them
        const realw u0[3] = {s_u[0][ev][0][t], s_u[0][ev][1][t], s_u[0][ev][2][t]};
        const realw u1[3] = {s_u[NF - 1][ev][0][t], s_u[NF - 1][ev][1][t], s_u[NF - 1][ev][2][t]};
        batch_point_newmark<E, NF == 2>(A, iglob, ev, rmi, ff[0], ff[NF - 1], u0, u1, pv0, pv1, ppart);
      }
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        s_u[f][ev][0][t] = ff[f][0]; s_u[f][ev][1][t] = ff[f][1]; s_u[f][ev][2][t] = ff[f][2];  // own slot only
      }
      __syncthreads();  // s_s free for the next event
    }
    const int fs = sem_floc_slot(tt);
    // the 27 inner slots sit at the end of the element row, so leaving the rows of the
    // points finished here out leaves the rest whole (a free-surface element's top
    // face points that have no other term lie mid-row and are stored anyway)
    if (act && !(inner && fs >= 98)) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        realw out[3 * E];
#pragma unroll
        for (int ev = 0; ev < E; ++ev) {
          out[3 * ev] = s_u[f][ev][0][t];
          out[3 * ev + 1] = s_u[f][ev][1][t];
          out[3 * ev + 2] = s_u[f][ev][2][t];
        }
        store_pt<E>((f == 0 ? A.floc0 : A.floc1) + (Lg - tt + fs) * (3 * E), out);
      }
    }
  }
}

// Anisotropic element pass of E events: as k_batch_elem, with the stress of
// each event from the 21 c_ij of the point (read per event: the element's
// 21 x 125 values come from HBM once and from the caches for the other
// events) and, in the adjoint pass, the 21 c_ij kernel integrands
// (e_I' e_J + e_J' e_I) / (1 + delta_IJ) of engineering Voigt strains
// accumulated per event.
template <int E, int NF>
__global__ __launch_bounds__(128) void k_batch_elem_aniso(BatchElemArgs A) {
  __shared__ realw s_tab[128];
  __shared__ realw s_u[NF][E][3][128];
  __shared__ realw s_s[NF][6][128];
  const int t = threadIdx.x;
  if (t < 125) s_tab[t] = A.tables[t];
  const bool act = t < 125;
  const int tt = act ? t : 124;
  const int I = tt % 5, J = (tt / 5) % 5, K = tt / 25;
  __syncthreads();
  const realw xr = A.xix_regular;
  const realw jx = A.jacobian_regular * A.xix_regular;
  realw hpI[5], hpJ[5], hpK[5], hwI[5], hwJ[5], hwK[5];
  {
    const realw f1 = -jx * s_tab[50 + J + 5 * K];
    const realw f2 = -jx * s_tab[75 + I + 5 * K];
    const realw f3 = -jx * s_tab[100 + I + 5 * J];
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      hpI[l] = s_tab[I + 5 * l] * xr;
      hpJ[l] = s_tab[J + 5 * l] * xr;
      hpK[l] = s_tab[K + 5 * l] * xr;
      hwI[l] = s_tab[25 + l + 5 * I] * f1;
      hwJ[l] = s_tab[25 + l + 5 * J] * f2;
      hwK[l] = s_tab[25 + l + 5 * K] * f3;
    }
  }
  const int La0 = 5 * J + 25 * K, Lb0 = I + 25 * K, Lc0 = I + 5 * J;
  const int nb = gridDim.x;
  const int x = blockIdx.x & 7, per_xcd = nb >> 3, slot = blockIdx.x >> 3;
  const int chunk = (A.nspec + 7) >> 3;
  const int e_end = min(A.nspec, (x + 1) * chunk);
  const size_t nloc21 = (size_t)A.nspec * 21 * 125;
  const int mask = A.active_mask;
  for (int e = x * chunk + slot; e < e_end; e += per_xcd) {
    const size_t Lg = (size_t)e * 125 + tt;
    const int iglob = A.ibool[Lg];
    const realw* C = A.cij + (size_t)e * 21 * 125 + tt;
    __syncthreads();  // previous element done with s_u
    if (act) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        realw x3[3 * E];
        load_pt<E>((f == 0 ? A.u0 : A.u1) + (size_t)iglob * (3 * E), x3);
#pragma unroll
        for (int ev = 0; ev < E; ++ev) {
          s_u[f][ev][0][t] = x3[3 * ev];
          s_u[f][ev][1][t] = x3[3 * ev + 1];
          s_u[f][ev][2][t] = x3[3 * ev + 2];
        }
      }
    }
    __syncthreads();
#pragma unroll 1
    for (int ev = 0; ev < E; ++ev) {
      if (!((mask >> ev) & 1)) continue;  // uniform over the grid
      realw va[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        realw duxdx = 0, duydx = 0, duzdx = 0, duxdy = 0, duydy = 0, duzdy = 0, duxdz = 0, duydz = 0, duzdz = 0;
#pragma unroll
        for (int l = 0; l < 5; ++l) {
          const int La = La0 + l, Lb = Lb0 + 5 * l, Lc = Lc0 + 25 * l;
          duxdx += s_u[f][ev][0][La] * hpI[l]; duydx += s_u[f][ev][1][La] * hpI[l]; duzdx += s_u[f][ev][2][La] * hpI[l];
          duxdy += s_u[f][ev][0][Lb] * hpJ[l]; duydy += s_u[f][ev][1][Lb] * hpJ[l]; duzdy += s_u[f][ev][2][Lb] * hpJ[l];
          duxdz += s_u[f][ev][0][Lc] * hpK[l]; duydz += s_u[f][ev][1][Lc] * hpK[l]; duzdz += s_u[f][ev][2][Lc] * hpK[l];
        }
        const realw exy2 = duxdy + duydx, exz2 = duzdx + duxdz, eyz2 = duzdy + duydz;
        // engineering Voigt strains 1..6 = xx yy zz yz xz xy
        const realw e1 = duxdx, e2 = duydy, e3 = duzdz, e4 = eyz2, e5 = exz2, e6 = exy2;
        if (NF == 2) {
          if (f == 0) {
            va[0] = e1; va[1] = e2; va[2] = e3; va[3] = 2.0f * (0.5f * e4); va[4] = 2.0f * (0.5f * e5);
            va[5] = 2.0f * (0.5f * e6);
          } else if (act) {
            const realw vb[6] = {e1, e2, e3, 2.0f * (0.5f * e4), 2.0f * (0.5f * e5), 2.0f * (0.5f * e6)};
            const size_t k0 = (size_t)ev * nloc21 + (size_t)e * 21 * 125 + t;
            int q = 0;
#pragma unroll
            for (int P = 0; P < 6; ++P)
#pragma unroll
              for (int Q = P; Q < 6; ++Q) {
                realw pr = va[P] * vb[Q];
                if (Q > P) pr += va[Q] * vb[P];
                kernel_acc(A.cijkl_kl, A.cij_part, k0 + (size_t)q * 125, pr, A.dt, A.kflush != 0);
                ++q;
              }
          }
        }
        asm volatile("" ::: "memory");
        const realw c11 = C[0], c12 = C[125], c13 = C[250], c14 = C[375], c15 = C[500], c16 = C[625];
        const realw c22 = C[750], c23 = C[875], c24 = C[1000], c25 = C[1125], c26 = C[1250];
        const realw c33 = C[1375], c34 = C[1500], c35 = C[1625], c36 = C[1750];
        const realw c44 = C[1875], c45 = C[2000], c46 = C[2125];
        const realw c55 = C[2250], c56 = C[2375], c66 = C[2500];
        s_s[f][0][t] = c11 * e1 + c12 * e2 + c13 * e3 + c14 * e4 + c15 * e5 + c16 * e6;
        s_s[f][1][t] = c12 * e1 + c22 * e2 + c23 * e3 + c24 * e4 + c25 * e5 + c26 * e6;
        s_s[f][2][t] = c13 * e1 + c23 * e2 + c33 * e3 + c34 * e4 + c35 * e5 + c36 * e6;
        s_s[f][5][t] = c14 * e1 + c24 * e2 + c34 * e3 + c44 * e4 + c45 * e5 + c46 * e6;  // syz
        s_s[f][4][t] = c15 * e1 + c25 * e2 + c35 * e3 + c45 * e4 + c55 * e5 + c56 * e6;  // sxz
        s_s[f][3][t] = c16 * e1 + c26 * e2 + c36 * e3 + c46 * e4 + c56 * e5 + c66 * e6;  // sxy
      }
      __syncthreads();  // stresses complete; every lane is done reading s_u[*][ev]
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        realw fx = 0, fy = 0, fz = 0;
#pragma unroll
        for (int l = 0; l < 5; ++l) {
          const int La = La0 + l, Lb = Lb0 + 5 * l, Lc = Lc0 + 25 * l;
          fx += s_s[f][0][La] * hwI[l]; fy += s_s[f][3][La] * hwI[l]; fz += s_s[f][4][La] * hwI[l];
          fx += s_s[f][3][Lb] * hwJ[l]; fy += s_s[f][1][Lb] * hwJ[l]; fz += s_s[f][5][Lb] * hwJ[l];
          fx += s_s[f][4][Lc] * hwK[l]; fy += s_s[f][5][Lc] * hwK[l]; fz += s_s[f][2][Lc] * hwK[l];
        }
        s_u[f][ev][0][t] = fx; s_u[f][ev][1][t] = fy; s_u[f][ev][2][t] = fz;  // own slot only
      }
      __syncthreads();  // s_s free for the next event
    }
    if (act) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        realw out[3 * E];
#pragma unroll
        for (int ev = 0; ev < E; ++ev) {
          out[3 * ev] = s_u[f][ev][0][t];
          out[3 * ev + 1] = s_u[f][ev][1][t];
          out[3 * ev + 2] = s_u[f][ev][2][t];
        }
        store_pt<E>((f == 0 ? A.floc0 : A.floc1) + (Lg - tt + sem_floc_slot(tt)) * (3 * E), out);
      }
    }
  }
}

struct BatchAsmArgs {
  int nglob, nloc, nfp, nrec, nmax, it;
  const int* off;
  const int* idx;
  const realw* sarr;   // [E][125][3]
  const realw* rech;   // [nrec][125]
  const realw* rm;     // [nglob][3]
  const realw* floc0;  // [nloc][E][3]
  const realw* floc1;
  realw *u0, *v0, *a0, *u1, *v1, *a1;  // [nglob][E][3]
  double* rho_kl;      // [nglob][E]
  float* rho_part;     // [nglob][E] fp32 partial sums (null: straight into fp64)
  int kflush;
  const int* nstep;    // [E]
  const realw* face0;  // [nfp][E][3] traction of field 0 (forward: store row it-1; adjoint: scratch)
  const realw* face1;  // [nfp][E][3] replayed store row NSTEP_max - it (adjoint)
  const double* stf;   // [E][NSTEP_max]
  const realw* adj;    // [E][NSTEP_max][nrec][3]
  realw dt;
  double dtk;
  int n_main;          // points [0, n_main) (all of them, or the shared ones when the
  const int* extra;    // element pass finishes the rest), then the n_extra listed ones
  int n_extra;
};

// Gather assembly + Newmark, one lane per (point, event).  TWO: adjoint field
// (dt > 0) and reconstructed field (dt < 0) with the fp64 density kernel.
// CSR segments: element-local points, face points, E*125 source points
// (event = p / 125), receiver points; summed in that fixed order.
template <int E, bool TWO>
__global__ void k_batch_asm(BatchAsmArgs A) {
  const size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nm = (size_t)A.n_main * E;
  if (l >= nm + (size_t)A.n_extra * E) return;
  const int g = l < nm ? (int)(l / E) : A.extra[(l - nm) / E];
  const int ev = (int)(l % E);  // nm is a multiple of E
  const size_t q = (size_t)g * E + ev;
  const int ne = A.nstep[ev];
  const bool active = TWO ? (A.it > A.nmax - ne) : (A.it <= ne);
  if (!active) return;
  const bool last = TWO ? (A.it == A.nmax) : (A.it == ne);
  const int row = TWO ? A.nmax - A.it : A.it - 1;  // stf / adjoint row of this step
  realw f0x = 0, f0y = 0, f0z = 0, f1x = 0, f1y = 0, f1z = 0;
  const int nsp = E * 125;
  const int q1 = A.off[g + 1];
  int k = A.off[g];
  // pairs of element-local contributions: both rows loaded before either is
  // added (two loads in flight per lane), same summation order
  for (; k + 1 < q1; k += 2) {
    const int c0 = A.idx[k], c1 = A.idx[k + 1];
    if (c1 >= A.nloc) break;  // element terms lead each row: c0 is one too
    const realw* P0 = A.floc0 + ((size_t)c0 * E + ev) * 3;
    const realw* P1 = A.floc0 + ((size_t)c1 * E + ev) * 3;
    const realw a0 = P0[0], a1 = P0[1], a2 = P0[2], b0 = P1[0], b1 = P1[1], b2 = P1[2];
    if (TWO) {
      const realw* B0 = A.floc1 + ((size_t)c0 * E + ev) * 3;
      const realw* B1 = A.floc1 + ((size_t)c1 * E + ev) * 3;
      const realw d0 = B0[0], d1 = B0[1], d2 = B0[2], e0 = B1[0], e1 = B1[1], e2 = B1[2];
      f1x += d0; f1y += d1; f1z += d2;
      f1x += e0; f1y += e1; f1z += e2;
    }
    f0x += a0; f0y += a1; f0z += a2;
    f0x += b0; f0y += b1; f0z += b2;
  }
  for (; k < q1; ++k) {
    const int c = A.idx[k];
    if (c < A.nloc) {
      const realw* P = A.floc0 + ((size_t)c * E + ev) * 3;
      f0x += P[0]; f0y += P[1]; f0z += P[2];
      if (TWO) {
        const realw* B = A.floc1 + ((size_t)c * E + ev) * 3;
        f1x += B[0]; f1y += B[1]; f1z += B[2];
      }
    } else if (c < A.nloc + A.nfp) {
      const size_t p = ((size_t)(c - A.nloc) * E + ev) * 3;
      f0x -= A.face0[p]; f0y -= A.face0[p + 1]; f0z -= A.face0[p + 2];
      if (TWO) { f1x -= A.face1[p]; f1y -= A.face1[p + 1]; f1z -= A.face1[p + 2]; }
    } else if (c < A.nloc + A.nfp + nsp) {
      const int p = c - A.nloc - A.nfp;
      if (p / 125 == ev) {
        const realw* S = A.sarr + 3 * (size_t)p;
        const realw w = (realw)A.stf[(size_t)ev * A.nmax + row];
        if (TWO) { f1x += S[0] * w; f1y += S[1] * w; f1z += S[2] * w; }
        else { f0x += S[0] * w; f0y += S[1] * w; f0z += S[2] * w; }
      }
    } else if (TWO) {
      const int p = c - A.nloc - A.nfp - nsp;
      const realw h = A.rech[p];
      const realw* R = A.adj + (((size_t)ev * A.nmax + row) * A.nrec + p / 125) * 3;
      f0x += h * R[0]; f0y += h * R[1]; f0z += h * R[2];
    }
  }
  // explicit fmas: the element pass finishes its points with the same expressions
  // (batch_point_newmark), so both give the same bits whatever the compiler contracts
  const size_t b = 3 * (size_t)g;
  const realw r0 = A.rm[b], r1 = A.rm[b + 1], r2 = A.rm[b + 2];
  const realw dt = A.dt, h0 = 0.5f * dt, s0 = 0.5f * dt * dt;
  const size_t o = 3 * q;
  const realw Ac[3] = {f0x * r0, f0y * r1, f0z * r2};
  realw V[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) V[c] = fmaf(h0, Ac[c], A.v0[o + c]);
  if (TWO) {
    const realw d1 = -dt, h1 = 0.5f * d1, s1 = 0.5f * d1 * d1;
    const realw BA[3] = {f1x * r0, f1y * r1, f1z * r2};
    realw BV[3], BU[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) { BV[c] = fmaf(h1, BA[c], A.v1[o + c]); BU[c] = A.u1[o + c]; }
    const float x = fmaf(Ac[2], BU[2], fmaf(Ac[1], BU[1], Ac[0] * BU[0]));
    kernel_acc(A.rho_kl, A.rho_part, q, x, A.dtk, A.kflush != 0);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (last) { A.a1[o + c] = BA[c]; A.v1[o + c] = BV[c]; }
      else { A.u1[o + c] = fmaf(s1, BA[c], fmaf(d1, BV[c], BU[c])); A.v1[o + c] = fmaf(h1, BA[c], BV[c]); }
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (last) { A.a0[o + c] = Ac[c]; A.v0[o + c] = V[c]; }
    else { A.u0[o + c] = fmaf(s0, Ac[c], fmaf(dt, V[c], A.u0[o + c])); A.v0[o + c] = fmaf(h0, Ac[c], V[c]); }
  }
}

// Stacey traction of field 0, one lane per (face point, event), written to
// out [nfp][E][3] (forward: the store row of this step; adjoint: scratch)
__global__ void k_batch_stacey(int nfp, int E, int it, int nmax, int adjoint, const int* __restrict__ nstep,
                               const int* __restrict__ figlob, const realw* __restrict__ fcoef,
                               const realw* __restrict__ v, realw* __restrict__ out) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)nfp * E) return;
  const int ev = (int)(q % E), p = (int)(q / E);
  const int ne = nstep[ev];
  if (adjoint ? (it <= nmax - ne) : (it > ne)) return;
  const size_t g = (size_t)figlob[p];
  const realw* C = fcoef + 8 * (size_t)p;
  const realw nx = C[0], ny = C[1], nz = C[2], jw = C[3], rvp = C[4], rvs = C[5];
  const realw* V = v + (g * E + ev) * 3;
  const realw vx = V[0], vy = V[1], vz = V[2];
  const realw vn = vx * nx + vy * ny + vz * nz;
  realw* o = out + 3 * q;
  o[0] = (rvp * vn * nx + rvs * (vx - vn * nx)) * jw;
  o[1] = (rvp * vn * ny + rvs * (vy - vn * ny)) * jw;
  o[2] = (rvp * vn * nz + rvs * (vz - vn * nz)) * jw;
}

// Compact Stacey store: a record (one step of one event, nfp*3 values) keeps
// fp16 values v/s with s = max|v| / 16384 (the record's largest value maps to
// 16384, well inside fp16's range, so the rounding error of every value is at
// most 2^-11 of its own size above s/16384's subnormal floor).  The scale is
// the record's max|v| held as float bits (non-negative floats order like their
// bits, so an integer atomicMax finds it).
constexpr float kStoreMaxMapped = 16384.0f;

// max|v| of each event's traction in this step's face scratch [nfp][E][3],
// reduced over the lanes of a wavefront that hold the same event (lanes q and
// q' hold one event when q - q' is a multiple of E) and folded into
// maxbits[ev] with one integer atomicMax per wavefront and event
__global__ void k_batch_face_absmax(int nfp, int E, const realw* __restrict__ face, unsigned int* __restrict__ maxbits) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  float m = 0.0f;
  if (q < (size_t)nfp * E) {
    const realw* o = face + 3 * q;
    m = fmaxf(fabsf(o[0]), fmaxf(fabsf(o[1]), fabsf(o[2])));
  }
  for (int d = E; d < 64; d <<= 1) m = fmaxf(m, __shfl_xor(m, d));
  // lane ev < E of every wavefront that holds any face point reports event ev
  const int lane = threadIdx.x & 63;
  if (lane < E && q - lane < (size_t)nfp * E) atomicMax(maxbits + lane, __float_as_uint(m));
}

__device__ __forceinline__ float store_scale_of(unsigned int maxbits) { return __uint_as_float(maxbits) / kStoreMaxMapped; }

// face scratch [nfp][E][3] -> fp16 record with the scales of this step
__global__ void k_batch_store_pack(int nfp, int E, const realw* __restrict__ face, const unsigned int* __restrict__ maxbits,
                                   uint16_t* __restrict__ rec) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)nfp * E) return;
  const float mx = __uint_as_float(maxbits[q % E]);
  const float inv = mx > 0.0f ? kStoreMaxMapped / mx : 0.0f;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const _Float16 h = (_Float16)(face[3 * q + c] * inv);
    uint16_t b;
    __builtin_memcpy(&b, &h, 2);
    rec[3 * q + c] = b;
  }
}

// fp16 record -> face scratch [nfp][E][3] for the replay
__global__ void k_batch_store_unpack(int nfp, int E, const uint16_t* __restrict__ rec,
                                     const unsigned int* __restrict__ maxbits, realw* __restrict__ face) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)nfp * E) return;
  const float sc = store_scale_of(maxbits[q % E]);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    _Float16 h;
    const uint16_t b = rec[3 * q + c];
    __builtin_memcpy(&h, &b, 2);
    face[3 * q + c] = (float)h * sc;
  }
}

__global__ void k_fill_u32(unsigned int* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// seismograms of the forward field, one block per (receiver, event), into
// seis [E][NSTEP_max][nrec][3] row it-1
__global__ __launch_bounds__(128) void k_batch_seismo(int nrec, int E, int it, int nmax,
                                                      const int* __restrict__ nstep, const int* __restrict__ riglob,
                                                      const realw* __restrict__ rh, const realw* __restrict__ u,
                                                      realw* __restrict__ seis) {
  __shared__ realw s[3][128];
  const int r = blockIdx.x / E, ev = blockIdx.x % E, t = threadIdx.x;
  if (it > nstep[ev]) return;  // block-uniform
  realw x = 0, y = 0, z = 0;
  if (t < 125) {
    const realw h = rh[(size_t)r * 125 + t];
    const size_t g = (size_t)riglob[(size_t)r * 125 + t];
    const realw* U = u + (g * E + ev) * 3;
    x = h * U[0]; y = h * U[1]; z = h * U[2];
  }
  s[0][t] = x; s[1][t] = y; s[2][t] = z;
  __syncthreads();
  for (int w = 64; w > 0; w >>= 1) {
    if (t < w) { s[0][t] += s[0][t + w]; s[1][t] += s[1][t + w]; s[2][t] += s[2][t + w]; }
    __syncthreads();
  }
  if (t == 0) {
    realw* o = seis + (((size_t)ev * nmax + (it - 1)) * nrec + r) * 3;
    o[0] = s[0][0]; o[1] = s[1][0]; o[2] = s[2][0];
  }
}

__global__ void k_batch_vshift(size_t n, realw* __restrict__ v, const realw* __restrict__ a, realw dto2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] += dto2 * a[i];
}

__global__ void k_fill(realw* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

__global__ void k_fill_d(double* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

struct SemBatch {
  SemMesh* m = nullptr;
  int E = 0, nmax = 0;
  size_t nfp = 0;
  realw dt = 0;
  std::vector<int> nstep;
  int* d_nstep = nullptr;
  int* d_off = nullptr;
  int* d_idx = nullptr;
  realw* d_sarr = nullptr;                           // [E][125][3]
  double* d_stf = nullptr;                           // [E][nmax]
  realw *u = nullptr, *v = nullptr, *a = nullptr;     // forward / adjoint field [nglob][E][3]
  realw *bu = nullptr, *bv = nullptr, *ba = nullptr;  // reconstructed field
  realw *floc0 = nullptr, *floc1 = nullptr;           // [nloc][E][3]
  realw* face_scratch = nullptr;                      // [nfp][E][3]
  realw* store = nullptr;                             // [nmax][nfp][E][3] fp32 store
  int store_precision = 0;                            // 0 fp32, 1 fp16 with a scale per record and event
  uint16_t* store16 = nullptr;                        // [nmax][nfp][E][3] fp16 store
  unsigned int* store_max = nullptr;                  // [nmax][E] max|v| per record and event (float bits)
  realw* face_replay = nullptr;                       // [nfp][E][3] unpacked replay row (fp16 store)
  realw* seis = nullptr;                              // [E][nmax][nrec][3]
  realw* adj = nullptr;                               // [E][nmax][nrec][3]
  double *mu_kl = nullptr, *kappa_kl = nullptr;       // [E][nloc]
  bool aniso = false;                                 // 21 c_ij model and kernels
  double* cijkl_kl = nullptr;                         // [E][nspec][21][125]
  float* cij_part = nullptr;                          // [E][nspec][21][125] fp32 partial sums
  float *mu_part = nullptr, *kappa_part = nullptr;    // [E][nloc] fp32 partial sums
  float* rho_part = nullptr;                          // [nglob][E]
  int kflush = 1;
  double* rho_kl = nullptr;                           // [nglob][E]
  double elem_ms = 0, asm_ms = 0;
  int elem_n = 0, asm_n = 0;
  int blocks_per_cu = 8;
  int interior = 0;                                   // element pass finishes the interior points
  std::vector<int> src_el;                            // distinct source elements of the batch
  int* d_extra = nullptr;                             // interior points of those elements (gathered)
  int n_extra = 0;
};

// true when the element pass finishes the points [n_shared, nglob) outside the
// source elements (isotropic batches on a mesh numbered with interior points)
bool batch_interior(const SemBatch* B) {
  return B->interior && !B->aniso && B->m->n_shared < B->m->nglob;
}

SemBatch* batch_of(long* p) {
  SemBatch* b = reinterpret_cast<SemBatch*>(*p);
  if (!b) SEM_FAIL("Batch_pointer is null");
  return b;
}

template <int E>
void launch_elem(SemBatch* B, int nf, const int* active, const int* last) {
  SemMesh* m = B->m;
  BatchElemArgs A{};
  A.ibool = m->d_ibool; A.kappa = m->d_kappa; A.mu = m->d_mu;
  A.kmu_lp = reinterpret_cast<const uint16_t*>(m->d_kmu_lp);
  A.kmu_base = m->d_kmu_base;
  A.tables = m->d_tables; A.nspec = m->nspec;
  A.xix_regular = m->xix_regular; A.jacobian_regular = m->jacobian_regular;
  A.scale_k = m->static_scale_kappa; A.scale_m = m->static_scale_mu;
  A.u0 = B->u; A.u1 = B->bu; A.floc0 = B->floc0; A.floc1 = B->floc1;
  A.mu_kl = B->mu_kl; A.kappa_kl = B->kappa_kl; A.dt = (double)B->dt;
  A.mu_part = B->mu_part; A.kappa_part = B->kappa_part; A.kflush = B->kflush;
  A.cij = m->d_cij; A.cijkl_kl = B->cijkl_kl; A.cij_part = B->cij_part;
  A.active_mask = A.last_mask = 0;
  for (int e = 0; e < E; ++e) {
    A.active_mask |= active[e] ? 1 << e : 0;
    A.last_mask |= last[e] ? 1 << e : 0;
  }
  const bool in = batch_interior(B);
  if (in) {
    A.n_shared = m->n_shared;
    A.n_src = (int)B->src_el.size();
    for (int k = 0; k < A.n_src; ++k) A.src_el[k] = B->src_el[k];
    A.rm = m->d_rmass_inv;
    A.v0 = B->v; A.a0 = B->a; A.v1 = B->bv; A.a1 = B->ba;
    A.rho_kl = B->rho_kl; A.rho_part = B->rho_part;
    A.fdt = B->dt;
  }
  int dev = 0, ncu = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int nb = std::max(8, std::min((m->nspec + 7) & ~7, (ncu * B->blocks_per_cu + 7) & ~7));
  const int sp = m->static_precision;
  hipStream_t s = m->stream;
#define LB(NF, SP)                                                                \
  if (in) SEM_LAUNCH((k_batch_elem<E, NF, SP, true>), dim3(nb), dim3(128), 0, s, A);  \
  else SEM_LAUNCH((k_batch_elem<E, NF, SP, false>), dim3(nb), dim3(128), 0, s, A)
  if (B->aniso) {
    if (nf == 2) SEM_LAUNCH((k_batch_elem_aniso<E, 2>), dim3(nb), dim3(128), 0, s, A);
    else SEM_LAUNCH((k_batch_elem_aniso<E, 1>), dim3(nb), dim3(128), 0, s, A);
  } else if (nf == 2) {
    if (sp == SEM_STATIC_BF16) LB(2, SEM_STATIC_BF16);
    else if (sp == SEM_STATIC_FP16) LB(2, SEM_STATIC_FP16);
    else LB(2, SEM_STATIC_FP32);
  } else {
    if (sp == SEM_STATIC_BF16) LB(1, SEM_STATIC_BF16);
    else if (sp == SEM_STATIC_FP16) LB(1, SEM_STATIC_FP16);
    else LB(1, SEM_STATIC_FP32);
  }
#undef LB
  HIP_CHECK(hipGetLastError());
}

void elem_pass(SemBatch* B, int nf, const int* active, const int* last) {
  switch (B->E) {
    case 1: launch_elem<1>(B, nf, active, last); break;
    case 2: launch_elem<2>(B, nf, active, last); break;
    case 4: launch_elem<4>(B, nf, active, last); break;
    case 8: launch_elem<8>(B, nf, active, last); break;
    default: SEM_FAIL("batch size %d not supported (1, 2, 4, 8)", B->E);
  }
}

void asm_pass(SemBatch* B, bool two, int it) {
  SemMesh* m = B->m;
  BatchAsmArgs A{};
  A.nglob = m->nglob; A.nloc = m->nspec * 125; A.nfp = (int)B->nfp; A.nrec = m->nrec;
  A.nmax = B->nmax; A.it = it;
  A.off = B->d_off; A.idx = B->d_idx; A.sarr = B->d_sarr; A.rech = m->d_rec_h; A.rm = m->d_rmass_inv;
  A.floc0 = B->floc0; A.floc1 = B->floc1;
  A.u0 = B->u; A.v0 = B->v; A.a0 = B->a; A.u1 = B->bu; A.v1 = B->bv; A.a1 = B->ba;
  A.rho_kl = B->rho_kl; A.nstep = B->d_nstep; A.stf = B->d_stf; A.adj = B->adj;
  A.rho_part = B->rho_part; A.kflush = B->kflush;
  A.dt = B->dt; A.dtk = (double)B->dt;
  const bool in = batch_interior(B);
  A.n_main = in ? m->n_shared : m->nglob;
  A.extra = in ? B->d_extra : nullptr;
  A.n_extra = in ? B->n_extra : 0;
  const size_t row = B->nfp * B->E * 3;
  if (two) {
    A.face0 = B->face_scratch;
    A.face1 = B->store_precision ? B->face_replay : B->store + (size_t)(B->nmax - it) * row;
  } else {
    A.face0 = B->store_precision ? B->face_scratch : B->store + (size_t)(it - 1) * row;
  }
  const dim3 gr(nblk(((size_t)A.n_main + A.n_extra) * B->E)), bl(TPB);
  hipStream_t s = m->stream;
#define LA(EE)                                                           \
  if (two) SEM_LAUNCH((k_batch_asm<EE, true>), gr, bl, 0, s, A);       \
  else SEM_LAUNCH((k_batch_asm<EE, false>), gr, bl, 0, s, A);
  switch (B->E) {
    case 1: LA(1); break;
    case 2: LA(2); break;
    case 4: LA(4); break;
    case 8: LA(8); break;
    default: SEM_FAIL("batch size %d not supported (1, 2, 4, 8)", B->E);
  }
#undef LA
  HIP_CHECK(hipGetLastError());
}

// Stacey terms of this step.  Forward: the traction of field 0 goes to the
// store row it-1 (fp32 store) or to the scratch, then into the fp16 record
// it-1 with its scales; adjoint: the adjoint field's traction to the scratch
// and, with the fp16 store, record NSTEP_max - it unpacked for the replay.
void stacey_pass(SemBatch* B, int it, bool adjoint) {
  if (!B->nfp) return;
  SemMesh* m = B->m;
  hipStream_t s = m->stream;
  const int nfp = (int)B->nfp, E = B->E;
  const size_t row = B->nfp * E * 3;
  const bool compact = B->store_precision != 0;
  realw* out = (adjoint || compact) ? B->face_scratch : B->store + (size_t)(it - 1) * row;
  hipLaunchKernelGGL(k_batch_stacey, dim3(nblk(B->nfp * E)), dim3(TPB), 0, s, nfp, E, it, B->nmax, (int)adjoint,
                     B->d_nstep, m->d_face_iglob, m->d_face_coef, B->v, out);
  if (compact && !adjoint) {
    unsigned int* mx = B->store_max + (size_t)(it - 1) * E;
    hipLaunchKernelGGL(k_batch_face_absmax, dim3(nblk(B->nfp * E)), dim3(TPB), 0, s, nfp, E, B->face_scratch, mx);
    hipLaunchKernelGGL(k_batch_store_pack, dim3(nblk(B->nfp * E)), dim3(TPB), 0, s, nfp, E, B->face_scratch, mx,
                       B->store16 + (size_t)(it - 1) * row);
  } else if (compact && adjoint) {
    const size_t r = (size_t)(B->nmax - it);
    hipLaunchKernelGGL(k_batch_store_unpack, dim3(nblk(B->nfp * E)), dim3(TPB), 0, s, nfp, E, B->store16 + r * row,
                       B->store_max + r * E, B->face_replay);
  }
  HIP_CHECK(hipGetLastError());
}

// the Stacey store in the precision asked for, allocated before a forward run
void ensure_store(SemBatch* B) {
  if (!B->nfp) return;
  const size_t n = (size_t)B->nmax * B->nfp * B->E * 3;
  hipStream_t s = B->m->stream;
  if (B->store_precision == 0) {
    dev_free(B->store16);
    dev_free(B->store_max);
    dev_free(B->face_replay);
    if (!B->store) B->store = dev_alloc<realw>(n);
  } else {
    dev_free(B->store);
    if (!B->store16) B->store16 = dev_alloc<uint16_t>(n);
    if (!B->store_max) B->store_max = dev_alloc<unsigned int>((size_t)B->nmax * B->E);
    if (!B->face_replay) B->face_replay = dev_alloc<realw>(B->nfp * B->E * 3);
    const size_t ns = (size_t)B->nmax * B->E;
    hipLaunchKernelGGL(k_fill_u32, dim3(nblk(ns)), dim3(TPB), 0, s, B->store_max, ns);
  }
}

// element / assembly pass times summed over the sampled steps of a timed run
// (every stride-th step, at most about 256 of them): four events per sampled
// step, whatever NSTEP is
struct PassTimer {
  std::vector<hipEvent_t> ev;
  int stride = 1, samples = 0;
  explicit PassTimer(int steps) {
    stride = std::max(1, steps / 256);
    ev.resize(4 * (size_t)((steps + stride - 1) / stride));
    for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
  }
  bool sampled(int it) const { return (it - 1) % stride == 0; }
  // k = 0 / 2 arm the element / assembly launch's start and end events, k = 1 / 3
  // check that the launch took them
  void mark(int it, int k, hipStream_t) {
    if (!sampled(it)) return;
    const size_t q = 4 * (size_t)((it - 1) / stride) + k;
    if (k % 2 == 0) sem_arm_launch_events(ev[q], ev[q + 1]);
    else if (!sem_launch_events_consumed()) SEM_FAIL("timed batch pass without a kernel launch");
  }
  // sums over the sampled steps and their count
  int sum(int steps, double& elem, double& asmb) {
    elem = asmb = 0;
    const int n = (steps + stride - 1) / stride;
    for (int k = 0; k < n; ++k) {
      float a = 0, b = 0;
      HIP_CHECK(hipEventElapsedTime(&a, ev[4 * (size_t)k], ev[4 * (size_t)k + 1]));
      HIP_CHECK(hipEventElapsedTime(&b, ev[4 * (size_t)k + 2], ev[4 * (size_t)k + 3]));
      elem += a; asmb += b;
    }
    return n;
  }
  ~PassTimer() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};

}  // namespace

extern "C" {

// Create a batch of `nevents` events on a mesh prepared with
// prepare_constants_device_ / prepare_fields_elastic_device_ (its own source
// list is not used; its receivers and Stacey faces are shared by every event).
//   nstep (nevents): NSTEP of each event; h_sourcearrays (NDIM,NGLLX,NGLLY,NGLLZ,nevents);
//   h_ispec_selected_source (nevents); h_stf (NSTEP_max, nevents), steps 1..NSTEP_e used
void fwat_batch_create_(long* Mesh_pointer, long* Batch_pointer, int* nevents, int* nstep, realw* DT,
                        realw* h_sourcearrays, int* h_ispec_selected_source, double* h_stf) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  const int E = *nevents;
  if (E != 1 && E != 2 && E != 4 && E != 8) SEM_FAIL("batch size %d not supported (1, 2, 4, 8)", E);
  if (m->nspec_irregular != 0) SEM_FAIL("batches need a regular (all-cube) mesh");
  if (m->anisotropy && m->static_precision != SEM_STATIC_FP32)
    SEM_FAIL("anisotropic batches keep their c_ij in fp32");
  for (int e = 0; e < E; ++e)
    if (nstep[e] < 1) SEM_FAIL("event %d has NSTEP < 1", e + 1);
  SemBatch* B = new SemBatch();
  *Batch_pointer = reinterpret_cast<long>(B);
  B->m = m;
  B->E = E;
  B->aniso = m->anisotropy != 0;
  B->dt = *DT;
  B->nstep.assign(nstep, nstep + E);
  B->nmax = *std::max_element(B->nstep.begin(), B->nstep.end());
  B->nfp = m->absorbing ? (size_t)m->nfaces * 25 : 0;
  hipStream_t s = m->stream;
  const size_t nloc = (size_t)m->nspec * 125, ng = (size_t)m->nglob, nfp = B->nfp;
  // CSR with the batch's E sources in the source segment (E*125 source points)
  std::vector<int> ib(nloc), fi(nfp), ri((size_t)m->nrec * 125);
  HIP_CHECK(hipMemcpyAsync(ib.data(), m->d_ibool, nloc * sizeof(int), hipMemcpyDeviceToHost, s));
  if (nfp) HIP_CHECK(hipMemcpyAsync(fi.data(), m->d_face_iglob, nfp * sizeof(int), hipMemcpyDeviceToHost, s));
  if (m->nrec) HIP_CHECK(hipMemcpyAsync(ri.data(), m->d_rec_iglob, ri.size() * sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  const size_t nsp = (size_t)E * 125;
  std::vector<int> owner(nloc + nfp + nsp + ri.size());
  for (size_t L = 0; L < nloc; ++L) owner[L] = ib[L];
  for (size_t q = 0; q < nfp; ++q) owner[nloc + q] = fi[q];
  for (int e = 0; e < E; ++e) {
    const int el = h_ispec_selected_source[e] - 1;
    if (el < 0 || el >= m->nspec) SEM_FAIL("ispec_selected_source(%d) = %d out of range", e + 1, el + 1);
    for (int L = 0; L < 125; ++L) owner[nloc + nfp + (size_t)e * 125 + L] = ib[(size_t)el * 125 + L];
  }
  for (size_t q = 0; q < ri.size(); ++q) owner[nloc + nfp + nsp + q] = ri[q];
  // the source elements' points have terms beyond their element's: the gather
  // assembles the ones the element pass would otherwise finish
  for (int e = 0; e < E; ++e) {
    const int el = h_ispec_selected_source[e] - 1;
    if (std::find(B->src_el.begin(), B->src_el.end(), el) == B->src_el.end()) B->src_el.push_back(el);
  }
  std::vector<int> extra;
  for (int el : B->src_el)
    for (int L = 0; L < 125; ++L)
      if (ib[(size_t)el * 125 + L] >= m->n_shared) extra.push_back(ib[(size_t)el * 125 + L]);
  B->n_extra = (int)extra.size();
  if (B->n_extra) B->d_extra = dev_upload(extra.data(), extra.size(), s);
  // every point's element terms in the mesh's phase-loop order, as the single-event path
  std::vector<int> visit = m->h_visit, off, idx;
  if (visit.empty()) {
    visit.resize(m->nspec);
    for (int e = 0; e < m->nspec; ++e) visit[e] = e;
  }
  sem_build_assembly_csr(m->nglob, m->nspec, owner, visit, off, idx);
  // element-local terms addressed by their place in the grouped element rows (sem_floc_slot)
  for (int& q : idx)
    if ((size_t)q < nloc) q = q - q % 125 + sem_floc_slot(q % 125);
  B->d_off = dev_upload(off.data(), off.size(), s);
  B->d_idx = dev_upload(idx.data(), idx.size(), s);
  B->d_sarr = dev_upload(h_sourcearrays, (size_t)E * 375, s);
  B->d_stf = dev_upload(h_stf, (size_t)E * B->nmax, s);
  B->d_nstep = dev_upload(nstep, (size_t)E, s);
  const size_t nf3 = ng * E * 3;
  B->u = dev_alloc<realw>(nf3); B->v = dev_alloc<realw>(nf3); B->a = dev_alloc<realw>(nf3);
  B->floc0 = dev_alloc<realw>(nloc * E * 3);
  if (nfp) B->face_scratch = dev_alloc<realw>(nfp * E * 3);  // the store: ensure_store, at the forward run
  if (m->nrec) {
    const size_t ns = (size_t)E * B->nmax * m->nrec * 3;
    B->seis = dev_alloc<realw>(ns);
    hipLaunchKernelGGL(k_fill, dim3(nblk(ns)), dim3(TPB), 0, s, B->seis, ns);
  }
  for (realw* p : {B->u, B->v, B->a}) hipLaunchKernelGGL(k_fill, dim3(nblk(nf3)), dim3(TPB), 0, s, p, nf3);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(s));
}

// 1: isotropic batches finish the element-interior points (one element term, no
// source) in the element pass and assemble the rest; 0 (default): every point in
// the assembly.  Same results either way; on 64x64x32 elements with 2 and 4 events
// per launch the adjoint pass is slower with 1 (each event's 12-byte point pieces
// are strided by the event rows), the forward pass the same.
void fwat_batch_set_interior_(long* Batch_pointer, int* on) {
  SemBatch* B = batch_of(Batch_pointer);
  if (*on != 0 && *on != 1) SEM_FAIL("interior must be 0 or 1");
  B->interior = *on;
}

void fwat_batch_set_tuning_(long* Batch_pointer, int* blocks_per_cu) {
  SemBatch* B = batch_of(Batch_pointer);
  if (*blocks_per_cu < 1 || *blocks_per_cu > 64) SEM_FAIL("blocks per CU must be in 1..64");
  B->blocks_per_cu = *blocks_per_cu;
}

// Precision of the batch's Stacey store, for the next forward run: 0 fp32
// (12 bytes per face point, event and step), 1 fp16 with one scale per record
// and event (6 bytes).  The replayed traction of the reconstruction carries the
// fp16 rounding (at most 2^-11 of a value above the record's max / 2^28).
void fwat_batch_set_store_precision_(long* Batch_pointer, int* precision) {
  SemBatch* B = batch_of(Batch_pointer);
  if (*precision != 0 && *precision != 1) SEM_FAIL("store precision must be 0 (fp32) or 1 (fp16 + scale)");
  B->store_precision = *precision;
}

// Forward runs of all events from rest (SAVE_FORWARD: Stacey rows kept in HBM).
void fwat_batch_forward_(long* Batch_pointer, int* timed) {
  SemBatch* B = batch_of(Batch_pointer);
  SemMesh* m = B->m;
  hipStream_t s = m->stream;
  const int E = B->E, N = B->nmax;
  ensure_store(B);
  // every call is a run from rest
  const size_t nf3 = (size_t)m->nglob * E * 3;
  for (realw* p : {B->u, B->v, B->a}) hipLaunchKernelGGL(k_fill, dim3(nblk(nf3)), dim3(TPB), 0, s, p, nf3);
  if (m->nrec) {
    const size_t ns = (size_t)E * N * m->nrec * 3;
    hipLaunchKernelGGL(k_fill, dim3(nblk(ns)), dim3(TPB), 0, s, B->seis, ns);
  }
  HIP_CHECK(hipGetLastError());
  B->elem_ms = B->asm_ms = 0;
  B->elem_n = B->asm_n = 0;
  PassTimer* tm = *timed ? new PassTimer(N) : nullptr;
  int active[EMAX] = {0}, last[EMAX] = {0};
  for (int it = 1; it <= N; ++it) {
    for (int e = 0; e < E; ++e) {
      active[e] = it <= B->nstep[e];
      last[e] = it == B->nstep[e];
    }
    if (tm) tm->mark(it, 0, s);
    elem_pass(B, 1, active, last);
    if (tm) tm->mark(it, 1, s);
    stacey_pass(B, it, false);
    if (m->nrec)
      hipLaunchKernelGGL(k_batch_seismo, dim3(m->nrec * E), dim3(128), 0, s, m->nrec, E, it, N, B->d_nstep,
                         m->d_rec_iglob, m->d_rec_h, B->u, B->seis);
    if (tm) tm->mark(it, 2, s);
    asm_pass(B, false, it);
    if (tm) tm->mark(it, 3, s);
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(s));
  if (tm) {
    B->elem_n = B->asm_n = tm->sum(N, B->elem_ms, B->asm_ms);
    delete tm;
  }
}

// Adjoint traces (NDIM, nrec, NSTEP_max, nevents): event e uses its first NSTEP_e steps.
void fwat_batch_set_adjoint_sources_(long* Batch_pointer, realw* h_adj) {
  SemBatch* B = batch_of(Batch_pointer);
  SemMesh* m = B->m;
  if (!m->nrec) return;
  dev_free(B->adj);
  B->adj = dev_upload(h_adj, (size_t)B->E * B->nmax * m->nrec * 3, m->stream);
  HIP_CHECK(hipStreamSynchronize(m->stream));
}

// Adjoint + reconstruction + kernels of all events, aligned at the end.
void fwat_batch_adjoint_(long* Batch_pointer, int* timed) {
  SemBatch* B = batch_of(Batch_pointer);
  SemMesh* m = B->m;
  hipStream_t s = m->stream;
  const int E = B->E, N = B->nmax;
  const size_t nloc = (size_t)m->nspec * 125, ng = (size_t)m->nglob, nf3 = ng * E * 3;
  if (m->nrec && !B->adj) SEM_FAIL("fwat_batch_set_adjoint_sources_ not called");
  if (B->nfp && !(B->store_precision ? (void*)B->store16 : (void*)B->store)) SEM_FAIL("fwat_batch_forward_ has not run");
  if (!B->bu) {
    B->bu = dev_alloc<realw>(nf3); B->bv = dev_alloc<realw>(nf3); B->ba = dev_alloc<realw>(nf3);
    B->floc1 = dev_alloc<realw>(nloc * E * 3);
    if (B->aniso) {
      B->cijkl_kl = dev_alloc<double>((size_t)E * nloc * 21);
    } else {
      B->mu_kl = dev_alloc<double>((size_t)E * nloc);
      B->kappa_kl = dev_alloc<double>((size_t)E * nloc);
    }
    B->rho_kl = dev_alloc<double>((size_t)E * ng);
  }
  // reconstructed field <- final forward states (v shifted by dt/2 a); adjoint field from rest
  HIP_CHECK(hipMemcpyAsync(B->bu, B->u, nf3 * 4, hipMemcpyDeviceToDevice, s));
  HIP_CHECK(hipMemcpyAsync(B->bv, B->v, nf3 * 4, hipMemcpyDeviceToDevice, s));
  HIP_CHECK(hipMemcpyAsync(B->ba, B->a, nf3 * 4, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_batch_vshift, dim3(nblk(nf3)), dim3(TPB), 0, s, nf3, B->bv, B->ba, 0.5f * B->dt);
  for (realw* p : {B->u, B->v, B->a}) hipLaunchKernelGGL(k_fill, dim3(nblk(nf3)), dim3(TPB), 0, s, p, nf3);
  if (B->aniso) {
    hipLaunchKernelGGL(k_fill_d, dim3(nblk((size_t)E * nloc * 21)), dim3(TPB), 0, s, B->cijkl_kl, (size_t)E * nloc * 21);
  } else {
    hipLaunchKernelGGL(k_fill_d, dim3(nblk((size_t)E * nloc)), dim3(TPB), 0, s, B->mu_kl, (size_t)E * nloc);
    hipLaunchKernelGGL(k_fill_d, dim3(nblk((size_t)E * nloc)), dim3(TPB), 0, s, B->kappa_kl, (size_t)E * nloc);
  }
  hipLaunchKernelGGL(k_fill_d, dim3(nblk((size_t)E * ng)), dim3(TPB), 0, s, B->rho_kl, (size_t)E * ng);
  // fp32 partial sums of mu/kappa over blocks of the mesh handle's kernel_block steps
  const int kb = m->kernel_block;
  dev_free(B->mu_part);
  dev_free(B->kappa_part);
  dev_free(B->cij_part);
  dev_free(B->rho_part);
  if (kb > 1) {
    if (B->aniso) {
      B->cij_part = dev_alloc<float>((size_t)E * nloc * 21);
      hipLaunchKernelGGL(k_fill, dim3(nblk((size_t)E * nloc * 21)), dim3(TPB), 0, s, B->cij_part, (size_t)E * nloc * 21);
    } else {
      B->mu_part = dev_alloc<float>((size_t)E * nloc);
      B->kappa_part = dev_alloc<float>((size_t)E * nloc);
      hipLaunchKernelGGL(k_fill, dim3(nblk((size_t)E * nloc)), dim3(TPB), 0, s, B->mu_part, (size_t)E * nloc);
      hipLaunchKernelGGL(k_fill, dim3(nblk((size_t)E * nloc)), dim3(TPB), 0, s, B->kappa_part, (size_t)E * nloc);
    }
    B->rho_part = dev_alloc<float>((size_t)E * ng);
    hipLaunchKernelGGL(k_fill, dim3(nblk((size_t)E * ng)), dim3(TPB), 0, s, B->rho_part, (size_t)E * ng);
  }
  HIP_CHECK(hipGetLastError());
  B->elem_ms = B->asm_ms = 0;
  B->elem_n = B->asm_n = 0;
  PassTimer* tm = *timed ? new PassTimer(N) : nullptr;
  int active[EMAX] = {0}, last[EMAX] = {0};
  for (int it = 1; it <= N; ++it) {
    for (int e = 0; e < E; ++e) {
      active[e] = it > N - B->nstep[e];
      last[e] = it == N;
    }
    B->kflush = kb <= 1 || it % kb == 0 || it == N;  // every event is active at it == N
    if (tm) tm->mark(it, 0, s);
    elem_pass(B, 2, active, last);
    if (tm) tm->mark(it, 1, s);
    stacey_pass(B, it, true);
    if (tm) tm->mark(it, 2, s);
    asm_pass(B, true, it);
    if (tm) tm->mark(it, 3, s);
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(s));
  if (tm) {
    B->elem_n = B->asm_n = tm->sum(N, B->elem_ms, B->asm_ms);
    delete tm;
  }
}

// seismograms (NDIM, nrec, NSTEP_max, nevents); rows past NSTEP_e are zero
void fwat_batch_seismograms_(long* Batch_pointer, realw* h_seis) {
  SemBatch* B = batch_of(Batch_pointer);
  SemMesh* m = B->m;
  if (!m->nrec) return;
  HIP_CHECK(hipMemcpyAsync(h_seis, B->seis, sizeof(realw) * (size_t)B->E * B->nmax * m->nrec * 3,
                           hipMemcpyDeviceToHost, m->stream));
  HIP_CHECK(hipStreamSynchronize(m->stream));
}

// Scaled kernels per event: rho_kl, mu_kl, kappa_kl (NGLLX,NGLLY,NGLLZ,NSPEC_AB,nevents), fp64,
// with save_adjoint_kernels' scaling by rhostore, mustore and kappastore.
void fwat_batch_kernels_(long* Batch_pointer, realw* h_rhostore, double* h_rho_kl, double* h_mu_kl,
                         double* h_kappa_kl) {
  SemBatch* B = batch_of(Batch_pointer);
  SemMesh* m = B->m;
  if (B->aniso) SEM_FAIL("anisotropic batch: its kernels come from fwat_batch_kernels_ani_");
  if (!B->mu_kl) SEM_FAIL("fwat_batch_adjoint_ has not run");
  const int E = B->E;
  const size_t nloc = (size_t)m->nspec * 125, ng = (size_t)m->nglob;
  std::vector<double> rg((size_t)E * ng);
  std::vector<int> ib(nloc);
  std::vector<float> kap(nloc), mu(nloc);
  hipStream_t s = m->stream;
  HIP_CHECK(hipMemcpyAsync(rg.data(), B->rho_kl, rg.size() * 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(h_mu_kl, B->mu_kl, (size_t)E * nloc * 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(h_kappa_kl, B->kappa_kl, (size_t)E * nloc * 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(ib.data(), m->d_ibool, nloc * 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(kap.data(), m->d_kappa, nloc * 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(mu.data(), m->d_mu, nloc * 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  for (int e = 0; e < E; ++e)
    for (size_t L = 0; L < nloc; ++L) {
      const size_t q = (size_t)e * nloc + L;
      h_rho_kl[q] = -(double)h_rhostore[L] * rg[(size_t)ib[L] * E + e];
      h_mu_kl[q] *= -2.0 * (double)mu[L];
      h_kappa_kl[q] *= -(double)kap[L];
    }
}

// Kernels of an anisotropic batch per event: rho_kl (NGLLX,NGLLY,NGLLZ,NSPEC_AB,nevents) scaled by
// rhostore, and cijkl_kl (21,NGLLX,NGLLY,NGLLZ,NSPEC_AB,nevents) in the order kernel_io.f90:100-156
// reads (c11..c66), scaled by -1 as the single-event transfer_kernels_ani_to_host_dp_.
void fwat_batch_kernels_ani_(long* Batch_pointer, realw* h_rhostore, double* h_rho_kl, double* h_cijkl_kl) {
  SemBatch* B = batch_of(Batch_pointer);
  SemMesh* m = B->m;
  if (!B->aniso) SEM_FAIL("isotropic batch: its kernels come from fwat_batch_kernels_");
  if (!B->cijkl_kl) SEM_FAIL("fwat_batch_adjoint_ has not run");
  const int E = B->E;
  const size_t nloc = (size_t)m->nspec * 125, ng = (size_t)m->nglob;
  std::vector<double> rg((size_t)E * ng), ck(nloc * 21);
  std::vector<int> ib(nloc);
  hipStream_t s = m->stream;
  HIP_CHECK(hipMemcpyAsync(rg.data(), B->rho_kl, rg.size() * 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(ib.data(), m->d_ibool, nloc * 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  for (int e = 0; e < E; ++e) {
    HIP_CHECK(hipMemcpyAsync(ck.data(), B->cijkl_kl + (size_t)e * nloc * 21, ck.size() * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    double* out = h_cijkl_kl + (size_t)e * nloc * 21;
    for (size_t el = 0; el < (size_t)m->nspec; ++el)
      for (int q = 0; q < 21; ++q)
        for (int L = 0; L < 125; ++L) out[(el * 125 + L) * 21 + q] = -ck[(el * 21 + q) * 125 + L];
    for (size_t L = 0; L < nloc; ++L)
      h_rho_kl[(size_t)e * nloc + L] = -(double)h_rhostore[L] * rg[(size_t)ib[L] * E + e];
  }
}

// element / assembly pass time of the last timed batch run (ms) and step counts
void fwat_batch_last_time_(long* Batch_pointer, double* elem_ms, int* elem_n, double* asm_ms, int* asm_n) {
  SemBatch* B = batch_of(Batch_pointer);
  *elem_ms = B->elem_ms; *elem_n = B->elem_n;
  *asm_ms = B->asm_ms; *asm_n = B->asm_n;
}

void fwat_batch_destroy_(long* Batch_pointer) {
  SemBatch* B = reinterpret_cast<SemBatch*>(*Batch_pointer);
  if (!B) return;
  HIP_CHECK(hipStreamSynchronize(B->m->stream));
  dev_free(B->d_extra);
  dev_free(B->d_nstep); dev_free(B->d_off); dev_free(B->d_idx); dev_free(B->d_sarr); dev_free(B->d_stf);
  dev_free(B->u); dev_free(B->v); dev_free(B->a); dev_free(B->bu); dev_free(B->bv); dev_free(B->ba);
  dev_free(B->floc0); dev_free(B->floc1); dev_free(B->face_scratch); dev_free(B->store);
  dev_free(B->store16); dev_free(B->store_max); dev_free(B->face_replay);
  dev_free(B->seis); dev_free(B->adj);
  dev_free(B->mu_kl); dev_free(B->kappa_kl); dev_free(B->rho_kl); dev_free(B->cijkl_kl);
  dev_free(B->mu_part); dev_free(B->kappa_part); dev_free(B->rho_part); dev_free(B->cij_part);
  delete B;
  *Batch_pointer = 0;
}

}  // extern "C"
