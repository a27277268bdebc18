// Multi-event batches: E events (own source, source time function, NSTEP
// and adjoint sources) on one mesh and model, advanced together.
//
// No counterpart in the reference: SpecFWAT loops over events one at a time
// (src/preproc/fwat_fwd_measure_adj.f90:43-51) and SPECFEM's batched-run
// option is acoustic forward only (read_mesh_databases.f90:261-268).  The
// per-event arithmetic is the single-event fused path's; what batching
// changes is the data layout and what is read once per batch:
//   fields, element-local forces   [nglob or nloc][E][3]  (events innermost)
//   statics (ibool, kappa/mu), CSR, rmass_inv, face geometry: read once per
//   element / point for all E events
// NSTEP may differ per event: the forward run advances event e for its first
// NSTEP_e steps (its state is frozen afterwards); the adjoint run is aligned
// at the end (event e idles for the first NSTEP_max - NSTEP_e steps), so every
// event pairs adjoint step it_e with its own forward state NSTEP_e + 1 - it_e.
// Regular (all-cube) isotropic meshes only; other meshes are refused.
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
  const realw* tables;
  int nspec;
  realw xix_regular, jacobian_regular;
  float scale_k, scale_m;
  const realw* u0;   // [nglob][E][3]
  const realw* u1;
  realw* floc0;      // [nloc][E][3]
  realw* floc1;
  double* mu_kl;     // [E][nloc]
  double* kappa_kl;
  double dt;
  int active[EMAX];
};

__device__ __forceinline__ float bf16f(uint16_t x) { return __uint_as_float((uint32_t)x << 16); }
__device__ __forceinline__ float h16f(uint16_t x) {
  _Float16 h;
  __builtin_memcpy(&h, &x, 2);
  return (float)h;
}

// Element pass of E events (NF = 1 forward, NF = 2 adjoint + reconstruction
// with in-pass mu/kappa kernels), regular iso meshes, folded tables.
template <int E, int NF, int SP>
__global__ __launch_bounds__(128) void k_batch_elem(BatchElemArgs A) {
  __shared__ realw s_tab[128];
  __shared__ realw s_u[3][128];
  __shared__ realw s_s[6][128];
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
  for (int e = x * chunk + slot; e < e_end; e += per_xcd) {
    const size_t Lg = (size_t)e * 125 + tt;
    const int iglob = A.ibool[Lg];
    realw kl, ml;
    if (SP == SEM_STATIC_FP32) {
      kl = A.kappa[Lg];
      ml = A.mu[Lg];
    } else {
      const uint32_t pk = reinterpret_cast<const uint32_t*>(A.kmu_lp)[Lg];
      kl = (SP == SEM_STATIC_BF16 ? bf16f((uint16_t)(pk & 0xffffu)) : h16f((uint16_t)(pk & 0xffffu))) * A.scale_k;
      ml = (SP == SEM_STATIC_BF16 ? bf16f((uint16_t)(pk >> 16)) : h16f((uint16_t)(pk >> 16))) * A.scale_m;
    }
    const realw lam = kl - (2.0f / 3.0f) * ml;
    const realw m2 = 2.0f * ml;
#pragma unroll
    for (int ev = 0; ev < E; ++ev) {
      if (!A.active[ev]) continue;  // uniform over the grid
      realw ea[6];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const realw* U = (f == 0 ? A.u0 : A.u1) + ((size_t)iglob * E + ev) * 3;
        __syncthreads();
        if (act) { s_u[0][t] = U[0]; s_u[1][t] = U[1]; s_u[2][t] = U[2]; }
        __syncthreads();
        realw duxdx = 0, duydx = 0, duzdx = 0, duxdy = 0, duydy = 0, duzdy = 0, duxdz = 0, duydz = 0, duzdz = 0;
#pragma unroll
        for (int l = 0; l < 5; ++l) {
          const int La = La0 + l, Lb = Lb0 + 5 * l, Lc = Lc0 + 25 * l;
          duxdx += s_u[0][La] * hpI[l]; duydx += s_u[1][La] * hpI[l]; duzdx += s_u[2][La] * hpI[l];
          duxdy += s_u[0][Lb] * hpJ[l]; duydy += s_u[1][Lb] * hpJ[l]; duzdy += s_u[2][Lb] * hpJ[l];
          duxdz += s_u[0][Lc] * hpK[l]; duydz += s_u[1][Lc] * hpK[l]; duzdz += s_u[2][Lc] * hpK[l];
        }
        const realw exy2 = duxdy + duydx, exz2 = duzdx + duxdz, eyz2 = duzdy + duydz;
        if (NF == 2) {
          if (f == 0) {
            ea[0] = duxdx; ea[1] = duydy; ea[2] = duzdz;
            ea[3] = 0.5f * exy2; ea[4] = 0.5f * exz2; ea[5] = 0.5f * eyz2;
          } else if (act) {
            const realw eb0 = duxdx, eb1 = duydy, eb2 = duzdz;
            const realw eb3 = 0.5f * exy2, eb4 = 0.5f * exz2, eb5 = 0.5f * eyz2;
            const realw tra = ea[0] + ea[1] + ea[2], trb = eb0 + eb1 + eb2;
            const realw t3a = tra / 3.0f, t3b = trb / 3.0f;
            const realw prod = (ea[0] - t3a) * (eb0 - t3b) + (ea[1] - t3a) * (eb1 - t3b) +
                               (ea[2] - t3a) * (eb2 - t3b) +
                               2.0f * (ea[3] * eb3 + ea[4] * eb4 + ea[5] * eb5);
            const size_t kq = (size_t)ev * nloc + Lg;
            A.mu_kl[kq] += A.dt * (double)prod;
            A.kappa_kl[kq] += A.dt * (double)(tra * trb);
          }
        }
        const realw ltr = lam * (duxdx + duydy + duzdz);
        if (act) {
          s_s[0][t] = ltr + m2 * duxdx; s_s[1][t] = ltr + m2 * duydy; s_s[2][t] = ltr + m2 * duzdz;
          s_s[3][t] = ml * exy2; s_s[4][t] = ml * exz2; s_s[5][t] = ml * eyz2;
        }
        __syncthreads();
        realw fx = 0, fy = 0, fz = 0;
#pragma unroll
        for (int l = 0; l < 5; ++l) {
          const int La = La0 + l, Lb = Lb0 + 5 * l, Lc = Lc0 + 25 * l;
          fx += s_s[0][La] * hwI[l]; fy += s_s[3][La] * hwI[l]; fz += s_s[4][La] * hwI[l];
          fx += s_s[3][Lb] * hwJ[l]; fy += s_s[1][Lb] * hwJ[l]; fz += s_s[5][Lb] * hwJ[l];
          fx += s_s[4][Lc] * hwK[l]; fy += s_s[5][Lc] * hwK[l]; fz += s_s[2][Lc] * hwK[l];
        }
        if (act) {
          realw* F = (f == 0 ? A.floc0 : A.floc1) + (Lg * E + ev) * 3;
          F[0] = fx; F[1] = fy; F[2] = fz;
        }
      }
    }
  }
}

struct BatchStep {
  int active[EMAX];         // event advances this step
  int last[EMAX];           // corrector only (end of the event's run)
  const realw* face0[EMAX];  // [nfp][3] traction of field 0 to subtract, or null
  const realw* face1[EMAX];  // replayed record of the reconstructed field, or null
  double stf0[EMAX];        // source time function value for field 0 (0 when none)
  double stf1[EMAX];        // for the reconstructed field
  const realw* adj[EMAX];   // [nrec][3] adjoint source row of field 0, or null
};

// Gather assembly + Newmark for E events (TWO: adjoint + reconstruction,
// with the fp64 density kernel).  CSR segments: element-local points, face
// points, E*125 source points (event = p/125), receiver points.
template <int E, bool TWO>
__global__ void k_batch_asm(int nglob, const int* __restrict__ off, const int* __restrict__ idx, int nloc,
                            int nfp, const realw* __restrict__ sarr, const realw* __restrict__ rech,
                            const realw* __restrict__ rm, const realw* __restrict__ floc0,
                            const realw* __restrict__ floc1, realw* __restrict__ u0, realw* __restrict__ v0,
                            realw* __restrict__ a0, realw* __restrict__ u1, realw* __restrict__ v1,
                            realw* __restrict__ a1, double* __restrict__ rho_kl, realw dt, double dtk,
                            BatchStep S) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nglob) return;
  realw f0[E][3], f1[E][3];
#pragma unroll
  for (int ev = 0; ev < E; ++ev)
#pragma unroll
    for (int c = 0; c < 3; ++c) { f0[ev][c] = 0.0f; f1[ev][c] = 0.0f; }
  const int q1 = off[g + 1];
  const int nsp = E * 125;
  for (int q = off[g]; q < q1; ++q) {
    const int c = idx[q];
    if (c < nloc) {
      const realw* P = floc0 + (size_t)c * (3 * E);
#pragma unroll
      for (int ev = 0; ev < E; ++ev) { f0[ev][0] += P[3 * ev]; f0[ev][1] += P[3 * ev + 1]; f0[ev][2] += P[3 * ev + 2]; }
      if (TWO) {
        const realw* B = floc1 + (size_t)c * (3 * E);
#pragma unroll
        for (int ev = 0; ev < E; ++ev) { f1[ev][0] += B[3 * ev]; f1[ev][1] += B[3 * ev + 1]; f1[ev][2] += B[3 * ev + 2]; }
      }
    } else if (c < nloc + nfp) {
      const size_t p = 3 * (size_t)(c - nloc);
#pragma unroll
      for (int ev = 0; ev < E; ++ev) {
        if (S.face0[ev]) { f0[ev][0] -= S.face0[ev][p]; f0[ev][1] -= S.face0[ev][p + 1]; f0[ev][2] -= S.face0[ev][p + 2]; }
        if (TWO && S.face1[ev]) { f1[ev][0] -= S.face1[ev][p]; f1[ev][1] -= S.face1[ev][p + 1]; f1[ev][2] -= S.face1[ev][p + 2]; }
      }
    } else if (c < nloc + nfp + nsp) {
      const int p = c - nloc - nfp;
      const int es = p / 125;
      const realw* Sv = sarr + 3 * (size_t)p;
#pragma unroll
      for (int ev = 0; ev < E; ++ev) {
        if (ev != es) continue;
        const realw w0 = (realw)S.stf0[ev];
        f0[ev][0] += Sv[0] * w0; f0[ev][1] += Sv[1] * w0; f0[ev][2] += Sv[2] * w0;
        if (TWO) {
          const realw w1 = (realw)S.stf1[ev];
          f1[ev][0] += Sv[0] * w1; f1[ev][1] += Sv[1] * w1; f1[ev][2] += Sv[2] * w1;
        }
      }
    } else {
      const int p = c - nloc - nfp - nsp;
      const realw h = rech[p];
      const size_t r = 3 * (size_t)(p / 125);
#pragma unroll
      for (int ev = 0; ev < E; ++ev)
        if (S.adj[ev]) { f0[ev][0] += h * S.adj[ev][r]; f0[ev][1] += h * S.adj[ev][r + 1]; f0[ev][2] += h * S.adj[ev][r + 2]; }
    }
  }
  const size_t b = 3 * (size_t)g;
  const realw r0 = rm[b], r1 = rm[b + 1], r2 = rm[b + 2];
  const realw h0 = 0.5f * dt, s0 = 0.5f * dt * dt;
#pragma unroll
  for (int ev = 0; ev < E; ++ev) {
    if (!S.active[ev]) continue;
    const size_t o = (size_t)g * (3 * E) + 3 * ev;
    realw A_[3] = {f0[ev][0] * r0, f0[ev][1] * r1, f0[ev][2] * r2};
    realw V[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) V[c] = v0[o + c] + h0 * A_[c];
    if (TWO) {
      realw BA[3] = {f1[ev][0] * r0, f1[ev][1] * r1, f1[ev][2] * r2};
      realw BV[3], BU[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) { BV[c] = v1[o + c] - h0 * BA[c]; BU[c] = u1[o + c]; }
      rho_kl[(size_t)ev * nglob + g] += dtk * (double)(A_[0] * BU[0] + A_[1] * BU[1] + A_[2] * BU[2]);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (S.last[ev]) { a1[o + c] = BA[c]; v1[o + c] = BV[c]; }
        else { u1[o + c] = BU[c] - dt * BV[c] + s0 * BA[c]; v1[o + c] = BV[c] - h0 * BA[c]; }
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (S.last[ev]) { a0[o + c] = A_[c]; v0[o + c] = V[c]; }
      else { u0[o + c] = u0[o + c] + dt * V[c] + s0 * A_[c]; v0[o + c] = V[c] + h0 * A_[c]; }
    }
  }
}

struct FaceOut {
  realw* out[EMAX];
  int active[EMAX];
};

// Stacey traction of field 0 for every active event, written per face point
__global__ void k_batch_stacey(int nfp, int E, const int* __restrict__ figlob, const realw* __restrict__ fcoef,
                               const realw* __restrict__ v, FaceOut F) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nfp * E) return;
  const int ev = q % E, p = q / E;
  if (!F.active[ev] || !F.out[ev]) return;
  const size_t g = (size_t)figlob[p];
  const realw* C = fcoef + 8 * (size_t)p;
  const realw nx = C[0], ny = C[1], nz = C[2], jw = C[3], rvp = C[4], rvs = C[5];
  const realw* V = v + (g * E + ev) * 3;
  const realw vx = V[0], vy = V[1], vz = V[2];
  const realw vn = vx * nx + vy * ny + vz * nz;
  realw* o = F.out[ev] + 3 * (size_t)p;
  o[0] = (rvp * vn * nx + rvs * (vx - vn * nx)) * jw;
  o[1] = (rvp * vn * ny + rvs * (vy - vn * ny)) * jw;
  o[2] = (rvp * vn * nz + rvs * (vz - vn * nz)) * jw;
}

struct SeisOut {
  realw* row[EMAX];
};

__global__ __launch_bounds__(128) void k_batch_seismo(int nrec, int E, const int* __restrict__ riglob,
                                                      const realw* __restrict__ rh, const realw* __restrict__ u,
                                                      SeisOut O) {
  __shared__ realw s[3][128];
  const int r = blockIdx.x / E, ev = blockIdx.x % E, t = threadIdx.x;
  if (!O.row[ev]) return;  // block-uniform
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
  if (t == 0) { O.row[ev][3 * r] = s[0][0]; O.row[ev][3 * r + 1] = s[1][0]; O.row[ev][3 * r + 2] = s[2][0]; }
}

// u += dt v + dt^2/2 a ; v += dt/2 a (field with sign of dt); events with
// mask 0 are left untouched
__global__ void k_batch_predictor(size_t npts, int E, realw* __restrict__ u, realw* __restrict__ v,
                                  const realw* __restrict__ a, realw dt, realw dt2o2, FaceOut M) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts * E) return;
  const int ev = (int)(q % E);
  if (!M.active[ev]) return;
  for (int c = 0; c < 3; ++c) {
    const realw ac = a[3 * q + c];
    u[3 * q + c] += dt * v[3 * q + c] + dt2o2 * ac;
    v[3 * q + c] += 0.5f * dt * ac;
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
  realw dt = 0;
  std::vector<int> nstep;
  int* d_off = nullptr;
  int* d_idx = nullptr;
  realw* d_sarr = nullptr;       // [E][125][3]
  std::vector<double> stf;       // host [E][nmax]
  realw *u = nullptr, *v = nullptr, *a = nullptr;        // forward / adjoint field [nglob][E][3]
  realw *bu = nullptr, *bv = nullptr, *ba = nullptr;     // reconstructed field
  realw *floc0 = nullptr, *floc1 = nullptr;              // [nloc][E][3]
  realw* face_scratch = nullptr;                         // [E][nfp][3]
  std::vector<realw*> store;                             // per event [nstep_e][nfp][3]
  realw* seis = nullptr;                                 // [E][nmax][nrec][3]
  realw* adj = nullptr;                                  // [E][nmax][nrec][3]
  double *mu_kl = nullptr, *kappa_kl = nullptr, *rho_kl = nullptr;  // [E][nloc], [E][nglob]
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double elem_ms = 0, asm_ms = 0;
  int elem_n = 0, asm_n = 0;
  int blocks_per_cu = 8;
};

SemBatch* batch_of(long* p) {
  SemBatch* b = reinterpret_cast<SemBatch*>(*p);
  if (!b) SEM_FAIL("Batch_pointer is null");
  return b;
}

template <int E>
void launch_elem(SemBatch* B, int nf, const int* active) {
  SemMesh* m = B->m;
  BatchElemArgs A{};
  A.ibool = m->d_ibool; A.kappa = m->d_kappa; A.mu = m->d_mu;
  A.kmu_lp = reinterpret_cast<const uint16_t*>(m->d_kmu_lp);
  A.tables = m->d_tables; A.nspec = m->nspec;
  A.xix_regular = m->xix_regular; A.jacobian_regular = m->jacobian_regular;
  A.scale_k = m->static_scale_kappa; A.scale_m = m->static_scale_mu;
  A.u0 = B->u; A.u1 = B->bu; A.floc0 = B->floc0; A.floc1 = B->floc1;
  A.mu_kl = B->mu_kl; A.kappa_kl = B->kappa_kl; A.dt = (double)B->dt;
  for (int e = 0; e < EMAX; ++e) A.active[e] = e < E ? active[e] : 0;
  int dev = 0, ncu = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  int nb = std::max(8, std::min((m->nspec + 7) & ~7, (ncu * B->blocks_per_cu + 7) & ~7));
  const int sp = m->static_precision;
  hipStream_t s = m->stream;
#define LB(NF, SP) hipLaunchKernelGGL((k_batch_elem<E, NF, SP>), dim3(nb), dim3(128), 0, s, A)
  if (nf == 2) {
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

void elem_pass(SemBatch* B, int nf, const int* active) {
  switch (B->E) {
    case 1: launch_elem<1>(B, nf, active); break;
    case 2: launch_elem<2>(B, nf, active); break;
    case 4: launch_elem<4>(B, nf, active); break;
    case 8: launch_elem<8>(B, nf, active); break;
    default: SEM_FAIL("batch size %d not supported (1, 2, 4, 8)", B->E);
  }
}

template <int E>
void launch_asm(SemBatch* B, bool two, const BatchStep& S) {
  SemMesh* m = B->m;
  const int nfp = m->absorbing ? m->nfaces * 25 : 0;
  const dim3 gr(nblk(m->nglob)), bl(TPB);
  if (two)
    hipLaunchKernelGGL((k_batch_asm<E, true>), gr, bl, 0, m->stream, m->nglob, B->d_off, B->d_idx, m->nspec * 125, nfp,
                       B->d_sarr, m->d_rec_h, m->d_rmass_inv, B->floc0, B->floc1, B->u, B->v, B->a, B->bu, B->bv,
                       B->ba, B->rho_kl, B->dt, (double)B->dt, S);
  else
    hipLaunchKernelGGL((k_batch_asm<E, false>), gr, bl, 0, m->stream, m->nglob, B->d_off, B->d_idx, m->nspec * 125, nfp,
                       B->d_sarr, m->d_rec_h, m->d_rmass_inv, B->floc0, B->floc1, B->u, B->v, B->a, B->bu, B->bv,
                       B->ba, B->rho_kl, B->dt, (double)B->dt, S);
  HIP_CHECK(hipGetLastError());
}

void asm_pass(SemBatch* B, bool two, const BatchStep& S) {
  switch (B->E) {
    case 1: launch_asm<1>(B, two, S); break;
    case 2: launch_asm<2>(B, two, S); break;
    case 4: launch_asm<4>(B, two, S); break;
    case 8: launch_asm<8>(B, two, S); break;
    default: SEM_FAIL("batch size %d not supported (1, 2, 4, 8)", B->E);
  }
}

}  // namespace

extern "C" {

// Create a batch of `nevents` events on a mesh prepared with
// prepare_constants_device_ / prepare_fields_elastic_device_ (its own source
// list is not used; its receivers are shared by every event).
//   nstep (nevents): NSTEP of each event; h_sourcearrays (NDIM,NGLLX,NGLLY,NGLLZ,nevents);
//   h_ispec_selected_source (nevents); h_stf (NSTEP_max, nevents) for steps 1..NSTEP_e
void fwat_batch_create_(long* Mesh_pointer, long* Batch_pointer, int* nevents, int* nstep, realw* DT,
                        realw* h_sourcearrays, int* h_ispec_selected_source, double* h_stf) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  const int E = *nevents;
  if (E != 1 && E != 2 && E != 4 && E != 8) SEM_FAIL("batch size %d not supported (1, 2, 4, 8)", E);
  if (m->nspec_irregular != 0 || m->anisotropy) SEM_FAIL("batches need a regular isotropic mesh");
  SemBatch* B = new SemBatch();
  *Batch_pointer = reinterpret_cast<long>(B);
  B->m = m;
  B->E = E;
  B->dt = *DT;
  B->nstep.assign(nstep, nstep + E);
  B->nmax = *std::max_element(B->nstep.begin(), B->nstep.end());
  for (int e = 0; e < E; ++e)
    if (B->nstep[e] < 1) SEM_FAIL("event %d has NSTEP < 1", e + 1);
  hipStream_t s = m->stream;
  const size_t nloc = (size_t)m->nspec * 125, ng = (size_t)m->nglob;
  const size_t nfp = m->absorbing ? (size_t)m->nfaces * 25 : 0;
  B->stf.assign(h_stf, h_stf + (size_t)E * B->nmax);
  // CSR with the batch's E sources
  std::vector<int> ib(nloc);
  HIP_CHECK(hipMemcpyAsync(ib.data(), m->d_ibool, nloc * sizeof(int), hipMemcpyDeviceToHost, s));
  std::vector<int> fi(nfp), ri((size_t)m->nrec * 125);
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
  if (owner.size() >= (size_t)INT32_MAX) SEM_FAIL("too many contributions for 32-bit CSR indices");
  std::vector<int> off(ng + 1, 0), idx(owner.size());
  for (size_t q = 0; q < owner.size(); ++q) off[(size_t)owner[q] + 1]++;
  for (size_t g = 0; g < ng; ++g) off[g + 1] += off[g];
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (size_t q = 0; q < owner.size(); ++q) idx[fill[owner[q]]++] = (int)q;
  B->d_off = dev_upload(off.data(), off.size(), s);
  B->d_idx = dev_upload(idx.data(), idx.size(), s);
  B->d_sarr = dev_upload(h_sourcearrays, (size_t)E * 375, s);
  const size_t nf3 = ng * E * 3;
  B->u = dev_alloc<realw>(nf3); B->v = dev_alloc<realw>(nf3); B->a = dev_alloc<realw>(nf3);
  B->floc0 = dev_alloc<realw>(nloc * E * 3);
  if (nfp) B->face_scratch = dev_alloc<realw>((size_t)E * nfp * 3);
  B->store.assign(E, nullptr);
  for (int e = 0; e < E; ++e)
    if (nfp) B->store[e] = dev_alloc<realw>((size_t)B->nstep[e] * nfp * 3);
  if (m->nrec) {
    B->seis = dev_alloc<realw>((size_t)E * B->nmax * m->nrec * 3);
    hipLaunchKernelGGL(k_fill, dim3(nblk((size_t)E * B->nmax * m->nrec * 3)), dim3(TPB), 0, s, B->seis,
                       (size_t)E * B->nmax * m->nrec * 3);
  }
  for (realw* p : {B->u, B->v, B->a}) hipLaunchKernelGGL(k_fill, dim3(nblk(nf3)), dim3(TPB), 0, s, p, nf3);
  HIP_CHECK(hipEventCreate(&B->ev0));
  HIP_CHECK(hipEventCreate(&B->ev1));
  HIP_CHECK(hipStreamSynchronize(s));
}

void fwat_batch_set_tuning_(long* Batch_pointer, int* blocks_per_cu) {
  SemBatch* B = batch_of(Batch_pointer);
  if (*blocks_per_cu < 1 || *blocks_per_cu > 64) SEM_FAIL("blocks per CU must be in 1..64");
  B->blocks_per_cu = *blocks_per_cu;
}

// Forward runs of all events (SAVE_FORWARD: Stacey stores kept per event).
void fwat_batch_forward_(long* Batch_pointer, int* timed) {
  SemBatch* B = batch_of(Batch_pointer);
  SemMesh* m = B->m;
  hipStream_t s = m->stream;
  const int E = B->E, N = B->nmax;
  const size_t nfp = m->absorbing ? (size_t)m->nfaces * 25 : 0;
  B->elem_ms = B->asm_ms = 0;
  B->elem_n = B->asm_n = 0;
  std::vector<hipEvent_t> evs;
  if (*timed) {
    evs.resize(4 * (size_t)N);
    for (auto& e : evs) HIP_CHECK(hipEventCreate(&e));
  }
  int active[EMAX] = {0};
  for (int it = 1; it <= N; ++it) {
    BatchStep S{};
    FaceOut F{};
    SeisOut O{};
    for (int e = 0; e < E; ++e) {
      active[e] = it <= B->nstep[e];
      S.active[e] = F.active[e] = active[e];
      S.last[e] = it == B->nstep[e];
      S.stf0[e] = active[e] ? B->stf[(size_t)e * N + (it - 1)] : 0.0;
      S.face0[e] = (active[e] && nfp) ? B->store[e] + (size_t)(it - 1) * nfp * 3 : nullptr;
      F.out[e] = const_cast<realw*>(S.face0[e]);
      O.row[e] = (active[e] && m->nrec) ? B->seis + (((size_t)e * N + (it - 1)) * m->nrec) * 3 : nullptr;
    }
    if (*timed) HIP_CHECK(hipEventRecord(evs[4 * (it - 1)], s));
    elem_pass(B, 1, active);
    if (*timed) HIP_CHECK(hipEventRecord(evs[4 * (it - 1) + 1], s));
    if (nfp)
      hipLaunchKernelGGL(k_batch_stacey, dim3(nblk(nfp * E)), dim3(TPB), 0, s, (int)nfp, E, m->d_face_iglob,
                         m->d_face_coef, B->v, F);
    if (m->nrec) hipLaunchKernelGGL(k_batch_seismo, dim3(m->nrec * E), dim3(128), 0, s, m->nrec, E, m->d_rec_iglob,
                                    m->d_rec_h, B->u, O);
    if (*timed) HIP_CHECK(hipEventRecord(evs[4 * (it - 1) + 2], s));
    asm_pass(B, false, S);
    if (*timed) HIP_CHECK(hipEventRecord(evs[4 * (it - 1) + 3], s));
  }
  HIP_CHECK(hipStreamSynchronize(s));
  if (*timed) {
    for (int it = 0; it < N; ++it) {
      float a = 0, b = 0;
      HIP_CHECK(hipEventElapsedTime(&a, evs[4 * it], evs[4 * it + 1]));
      HIP_CHECK(hipEventElapsedTime(&b, evs[4 * it + 2], evs[4 * it + 3]));
      B->elem_ms += a; B->asm_ms += b;
    }
    B->elem_n = B->asm_n = N;
    for (auto e : evs) HIP_CHECK(hipEventDestroy(e));
  }
}

// Adjoint traces (NDIM, nrec, NSTEP_max, nevents): event e uses its first NSTEP_e steps.
void fwat_batch_set_adjoint_sources_(long* Batch_pointer, realw* h_adj) {
  SemBatch* B = batch_of(Batch_pointer);
  SemMesh* m = B->m;
  if (!m->nrec) return;
  const size_t n = (size_t)B->E * B->nmax * m->nrec * 3;
  dev_free(B->adj);
  B->adj = dev_upload(h_adj, n, m->stream);
  HIP_CHECK(hipStreamSynchronize(m->stream));
}

// Adjoint + reconstruction + kernels of all events, aligned at the end.
void fwat_batch_adjoint_(long* Batch_pointer, int* timed) {
  SemBatch* B = batch_of(Batch_pointer);
  SemMesh* m = B->m;
  hipStream_t s = m->stream;
  const int E = B->E, N = B->nmax;
  const size_t nloc = (size_t)m->nspec * 125, ng = (size_t)m->nglob, nf3 = ng * E * 3;
  const size_t nfp = m->absorbing ? (size_t)m->nfaces * 25 : 0;
  if (m->nrec && !B->adj) SEM_FAIL("fwat_batch_set_adjoint_sources_ not called");
  // reconstructed field <- final forward states; adjoint field from rest
  if (!B->bu) {
    B->bu = dev_alloc<realw>(nf3); B->bv = dev_alloc<realw>(nf3); B->ba = dev_alloc<realw>(nf3);
    B->floc1 = dev_alloc<realw>(nloc * E * 3);
    B->mu_kl = dev_alloc<double>((size_t)E * nloc);
    B->kappa_kl = dev_alloc<double>((size_t)E * nloc);
    B->rho_kl = dev_alloc<double>((size_t)E * ng);
  }
  HIP_CHECK(hipMemcpyAsync(B->bu, B->u, nf3 * 4, hipMemcpyDeviceToDevice, s));
  HIP_CHECK(hipMemcpyAsync(B->bv, B->v, nf3 * 4, hipMemcpyDeviceToDevice, s));
  HIP_CHECK(hipMemcpyAsync(B->ba, B->a, nf3 * 4, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_batch_vshift, dim3(nblk(nf3)), dim3(TPB), 0, s, nf3, B->bv, B->ba, 0.5f * B->dt);
  for (realw* p : {B->u, B->v, B->a}) hipLaunchKernelGGL(k_fill, dim3(nblk(nf3)), dim3(TPB), 0, s, p, nf3);
  hipLaunchKernelGGL(k_fill_d, dim3(nblk((size_t)E * nloc)), dim3(TPB), 0, s, B->mu_kl, (size_t)E * nloc);
  hipLaunchKernelGGL(k_fill_d, dim3(nblk((size_t)E * nloc)), dim3(TPB), 0, s, B->kappa_kl, (size_t)E * nloc);
  hipLaunchKernelGGL(k_fill_d, dim3(nblk((size_t)E * ng)), dim3(TPB), 0, s, B->rho_kl, (size_t)E * ng);
  B->elem_ms = B->asm_ms = 0;
  B->elem_n = B->asm_n = 0;
  std::vector<hipEvent_t> evs;
  if (*timed) {
    evs.resize(4 * (size_t)N);
    for (auto& e : evs) HIP_CHECK(hipEventCreate(&e));
  }
  int active[EMAX] = {0};
  for (int it = 1; it <= N; ++it) {
    BatchStep S{};
    FaceOut F{};
    for (int e = 0; e < E; ++e) {
      const int ite = it - (N - B->nstep[e]);  // event-local adjoint step
      active[e] = ite >= 1;
      S.active[e] = F.active[e] = active[e];
      S.last[e] = it == N;
      if (!active[e]) continue;
      const int rec0 = B->nstep[e] - ite;  // forward step N_e + 1 - it_e, 0-based
      S.face0[e] = nfp ? B->face_scratch + (size_t)e * nfp * 3 : nullptr;
      F.out[e] = const_cast<realw*>(S.face0[e]);
      S.face1[e] = nfp ? B->store[e] + (size_t)rec0 * nfp * 3 : nullptr;
      S.stf1[e] = B->stf[(size_t)e * N + rec0];
      S.adj[e] = m->nrec ? B->adj + (((size_t)e * N + rec0) * m->nrec) * 3 : nullptr;
    }
    if (*timed) HIP_CHECK(hipEventRecord(evs[4 * (it - 1)], s));
    elem_pass(B, 2, active);
    if (*timed) HIP_CHECK(hipEventRecord(evs[4 * (it - 1) + 1], s));
    if (nfp)
      hipLaunchKernelGGL(k_batch_stacey, dim3(nblk(nfp * E)), dim3(TPB), 0, s, (int)nfp, E, m->d_face_iglob,
                         m->d_face_coef, B->v, F);
    if (*timed) HIP_CHECK(hipEventRecord(evs[4 * (it - 1) + 2], s));
    asm_pass(B, true, S);
    if (*timed) HIP_CHECK(hipEventRecord(evs[4 * (it - 1) + 3], s));
  }
  HIP_CHECK(hipStreamSynchronize(s));
  if (*timed) {
    for (int it = 0; it < N; ++it) {
      float a = 0, b = 0;
      HIP_CHECK(hipEventElapsedTime(&a, evs[4 * it], evs[4 * it + 1]));
      HIP_CHECK(hipEventElapsedTime(&b, evs[4 * it + 2], evs[4 * it + 3]));
      B->elem_ms += a; B->asm_ms += b;
    }
    B->elem_n = B->asm_n = N;
    for (auto e : evs) HIP_CHECK(hipEventDestroy(e));
  }
}

// seismograms (NDIM, nrec, NSTEP_max, nevents)
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
  if (!B->mu_kl) SEM_FAIL("fwat_batch_adjoint_ has not run");
  const size_t nloc = (size_t)m->nspec * 125, ng = (size_t)m->nglob;
  std::vector<double> rg((size_t)B->E * ng);
  std::vector<int> ib(nloc);
  std::vector<float> kap(nloc), mu(nloc);
  hipStream_t s = m->stream;
  HIP_CHECK(hipMemcpyAsync(rg.data(), B->rho_kl, rg.size() * 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(h_mu_kl, B->mu_kl, (size_t)B->E * nloc * 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(h_kappa_kl, B->kappa_kl, (size_t)B->E * nloc * 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(ib.data(), m->d_ibool, nloc * 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(kap.data(), m->d_kappa, nloc * 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(mu.data(), m->d_mu, nloc * 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  for (int e = 0; e < B->E; ++e)
    for (size_t L = 0; L < nloc; ++L) {
      const size_t q = (size_t)e * nloc + L;
      h_rho_kl[q] = -(double)h_rhostore[L] * rg[(size_t)e * ng + ib[L]];
      h_mu_kl[q] *= -2.0 * (double)mu[L];
      h_kappa_kl[q] *= -(double)kap[L];
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
  dev_free(B->d_off); dev_free(B->d_idx); dev_free(B->d_sarr);
  dev_free(B->u); dev_free(B->v); dev_free(B->a); dev_free(B->bu); dev_free(B->bv); dev_free(B->ba);
  dev_free(B->floc0); dev_free(B->floc1); dev_free(B->face_scratch);
  for (auto& p : B->store) dev_free(p);
  dev_free(B->seis); dev_free(B->adj);
  dev_free(B->mu_kl); dev_free(B->kappa_kl); dev_free(B->rho_kl);
  HIP_CHECK(hipEventDestroy(B->ev0));
  HIP_CHECK(hipEventDestroy(B->ev1));
  delete B;
  *Batch_pointer = 0;
}

}  // extern "C"
