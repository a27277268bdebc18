// Brick-fused element + assembly pass for regular lattice meshes.
//
// On an all-regular isotropic mesh whose elements form a lattice (element
// (ix, iy, iz) = ix + nex*(iy + ney*iz), neighbours sharing faces, checked by
// fwat_set_lattice_), a workgroup owns a brick of BX x BY x BZ elements.  It
// computes the element forces of the brick one element after the other and
// adds them into LDS over the brick's (4BX+1)(4BY+1)(4BZ+1) lattice points, in
// a fixed element order.  Points strictly inside the brick receive all their
// element contributions there, so the workgroup finishes them at once with
// the fused Newmark update (newmark_point).  Points on the brick boundary
// (shared with other bricks), and every point of a brick holding a source or
// receiver element, write their partial sums; the halo pass then gathers, per
// point and in brick order, the partials plus the Stacey, source and receiver
// terms and applies the same Newmark update.  The element-local force array
// of the gather path (written and read once per local point and step) is
// replaced by partial sums of the brick-boundary points only.  No atomics:
// every sum has a fixed order and runs are reproducible.
//
// Scheduling: the Stacey traction and the seismograms of a step read only
// halo points (mesh-boundary faces lie on brick boundaries; source and
// receiver elements sit in bricks whose points all go through the halo
// pass), so the per-step order is  brick pass -> Stacey face terms,
// seismograms -> halo pass.
#include "sem_device.h"

#include <algorithm>
#include <cstring>

namespace {

constexpr int TPB = 256;

struct BrickArgs {
  const int* ibool;
  const realw* kappa;
  const realw* mu;
  const uint16_t* kmu_lp;
  const float* kmu_base;
  const realw* tables;
  int nex, ney, nbx, nby, nbricks;
  realw xix_regular, jacobian_regular;
  float scale_k, scale_m;
  const uint8_t* special;
  realw* part0;
  realw* part1;
  const realw* rm;
  AsmField F0, F1;
  double* mu_kl;
  double* kappa_kl;
  double* rho_kl;
  double* hess_kl;
  float* mu_part;     // fp32 partial sums of mu/kappa/rho (null: straight into fp64)
  float* kappa_part;
  float* rho_part;
  int kflush;
  double dt;   // kernel time weight of the element pass (mu/kappa)
  double dtk;  // kernel time weight of the point update (rho, Hessian)
  int last;
};

__device__ __forceinline__ float bk_bf16(uint16_t x) { return __uint_as_float((uint32_t)x << 16); }
__device__ __forceinline__ float bk_h16(uint16_t x) {
  _Float16 h;
  __builtin_memcpy(&h, &x, 2);
  return (float)h;
}

template <int BX, int BY, int BZ, int NF, int SP>
__global__ __launch_bounds__(128) void k_brick(BrickArgs A) {
  constexpr int LX = 4 * BX + 1, LY = 4 * BY + 1, LZ = 4 * BZ + 1, NPTS = LX * LY * LZ;
  __shared__ realw s_tab[128];
  __shared__ realw s_u[3][128];
  __shared__ realw s_s[6][128];
  __shared__ realw s_acc[NF][3][NPTS];
  __shared__ int s_g[NPTS];
  const int t = threadIdx.x;
  if (t < 125) s_tab[t] = A.tables[t];
  for (int q = t; q < NF * 3 * NPTS; q += 128) (&s_acc[0][0][0])[q] = 0.0f;
  const bool act = t < 125;
  const int tt = act ? t : 124;
  const int I = tt % 5, J = (tt / 5) % 5, K = tt / 25;
  __syncthreads();
  // folded tables (as the regular-mesh element pass): hprime * xix_regular and
  // -jacobian * xix_regular * hprimewgll * wgllwgll
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
  // XCD-aware persistent schedule over bricks (lattice order)
  const int nb = gridDim.x;
  const int x = blockIdx.x & 7, per_xcd = nb >> 3, slot = blockIdx.x >> 3;
  const int chunk = (A.nbricks + 7) >> 3;
  const int b_end = min(A.nbricks, (x + 1) * chunk);
  for (int b = x * chunk + slot; b < b_end; b += per_xcd) {
    const int bx = b % A.nbx, by = (b / A.nbx) % A.nby, bz = b / (A.nbx * A.nby);
#pragma unroll 1
    for (int el = 0; el < BX * BY * BZ; ++el) {
      const int ex = el % BX, ey = (el / BX) % BY, ez = el / (BX * BY);
      const int ispec = (bx * BX + ex) + A.nex * ((by * BY + ey) + A.ney * (bz * BZ + ez));
      const size_t Lg = (size_t)ispec * 125 + tt;
      const int iglob = A.ibool[Lg];
      const int bl = (4 * ex + I) + LX * ((4 * ey + J) + LY * (4 * ez + K));
      // each lattice point has one owner element: the last one along each axis
      const bool owner = act && (I < 4 || ex == BX - 1) && (J < 4 || ey == BY - 1) && (K < 4 || ez == BZ - 1);
      realw kl, ml;
      if (SP == SEM_STATIC_FP32) {
        kl = A.kappa[Lg];
        ml = A.mu[Lg];
      } else {
        const uint32_t pk = reinterpret_cast<const uint32_t*>(A.kmu_lp)[Lg];
        const float2 kb = reinterpret_cast<const float2*>(A.kmu_base)[ispec];
        kl = kb.x + (SP == SEM_STATIC_BF16 ? bk_bf16((uint16_t)(pk & 0xffffu)) : bk_h16((uint16_t)(pk & 0xffffu))) * A.scale_k;
        ml = kb.y + (SP == SEM_STATIC_BF16 ? bk_bf16((uint16_t)(pk >> 16)) : bk_h16((uint16_t)(pk >> 16))) * A.scale_m;
      }
      const realw lam = kl - (2.0f / 3.0f) * ml;
      const realw m2 = 2.0f * ml;
      realw ea[6];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const realw* U = f == 0 ? A.F0.u : A.F1.u;
        __syncthreads();  // previous element / field (or brick update) done with s_u, s_s, s_g
        if (f == 0 && owner) s_g[bl] = iglob;
        if (act) {
          s_u[0][t] = U[3 * (size_t)iglob];
          s_u[1][t] = U[3 * (size_t)iglob + 1];
          s_u[2][t] = U[3 * (size_t)iglob + 2];
        }
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
            kernel_acc(A.mu_kl, A.mu_part, Lg, prod, A.dt, A.kflush != 0);
            kernel_acc(A.kappa_kl, A.kappa_part, Lg, tra * trb, A.dt, A.kflush != 0);
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
        // distinct lattice points within an element; elements are ordered by the barriers above
        if (act) { s_acc[f][0][bl] += fx; s_acc[f][1][bl] += fy; s_acc[f][2][bl] += fz; }
      }
    }
    __syncthreads();  // all element sums and owners of this brick are in LDS
    const bool whole_halo = A.special[b] != 0;
    for (int q = t; q < NPTS; q += 128) {
      const int px = q % LX, py = (q / LX) % LY, pz = q / (LX * LY);
      const bool inner = !whole_halo && px > 0 && px < LX - 1 && py > 0 && py < LY - 1 && pz > 0 && pz < LZ - 1;
      const realw f0x = s_acc[0][0][q], f0y = s_acc[0][1][q], f0z = s_acc[0][2][q];
      const realw f1x = NF == 2 ? s_acc[NF - 1][0][q] : 0.0f;
      const realw f1y = NF == 2 ? s_acc[NF - 1][1][q] : 0.0f;
      const realw f1z = NF == 2 ? s_acc[NF - 1][2][q] : 0.0f;
      if (inner) {
        newmark_point<NF == 2>((size_t)s_g[q], A.rm, f0x, f0y, f0z, f1x, f1y, f1z, A.F0, A.F1, A.last != 0,
                               A.rho_kl, A.rho_part, A.kflush != 0, A.hess_kl, A.dtk);
      } else {
        const size_t o = 3 * ((size_t)b * NPTS + q);
        A.part0[o] = f0x; A.part0[o + 1] = f0y; A.part0[o + 2] = f0z;
        if (NF == 2) { A.part1[o] = f1x; A.part1[o + 1] = f1y; A.part1[o + 2] = f1z; }
      }
#pragma unroll
      for (int f = 0; f < NF; ++f) { s_acc[f][0][q] = 0.0f; s_acc[f][1][q] = 0.0f; s_acc[f][2][q] = 0.0f; }
    }
    // the next brick's first element barrier orders these clears and s_g reads before reuse
  }
}

// Halo points: brick partial sums (slot b*npts + bl, brick order), then Stacey
// face terms, source points and receiver points, then the Newmark update.
template <bool TWO>
__global__ void k_halo(int nhalo, const int* __restrict__ halo_g, const int* __restrict__ off,
                       const int* __restrict__ idx, int nslot, int nfp, int nsp, const realw* __restrict__ sarr,
                       const realw* __restrict__ rech, const realw* __restrict__ rm, const realw* __restrict__ part0,
                       const realw* __restrict__ part1, AsmField F0, AsmField F1, int last,
                       double* __restrict__ rho_kl, float* __restrict__ rho_part, int flush,
                       double* __restrict__ hess_kl, double dtk) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nhalo) return;
  realw f0x = 0, f0y = 0, f0z = 0, f1x = 0, f1y = 0, f1z = 0;
  const int q1 = off[h + 1];
  for (int q = off[h]; q < q1; ++q) {
    const int c = idx[q];
    if (c < nslot) {
      const realw* P = part0 + 3 * (size_t)c;
      f0x += P[0]; f0y += P[1]; f0z += P[2];
      if (TWO) {
        const realw* B = part1 + 3 * (size_t)c;
        f1x += B[0]; f1y += B[1]; f1z += B[2];
      }
    } else if (c < nslot + nfp) {
      const size_t p = 3 * (size_t)(c - nslot);
      if (F0.face) { f0x -= F0.face[p]; f0y -= F0.face[p + 1]; f0z -= F0.face[p + 2]; }
      if (TWO && F1.face) { f1x -= F1.face[p]; f1y -= F1.face[p + 1]; f1z -= F1.face[p + 2]; }
    } else if (c < nslot + nfp + nsp) {
      const int p = c - nslot - nfp;
      const realw* S = sarr + 3 * (size_t)p;
      if (F0.stf) { const realw w = (realw)F0.stf[p / 125]; f0x = fmaf(S[0], w, f0x); f0y = fmaf(S[1], w, f0y); f0z = fmaf(S[2], w, f0z); }
      if (TWO && F1.stf) { const realw w = (realw)F1.stf[p / 125]; f1x = fmaf(S[0], w, f1x); f1y = fmaf(S[1], w, f1y); f1z = fmaf(S[2], w, f1z); }
    } else {
      const int p = c - nslot - nfp - nsp;
      const realw hh = rech[p];
      const size_t r = 3 * (size_t)(p / 125);
      if (F0.adj) { f0x = fmaf(hh, F0.adj[r], f0x); f0y = fmaf(hh, F0.adj[r + 1], f0y); f0z = fmaf(hh, F0.adj[r + 2], f0z); }
      if (TWO && F1.adj) { f1x = fmaf(hh, F1.adj[r], f1x); f1y = fmaf(hh, F1.adj[r + 1], f1y); f1z = fmaf(hh, F1.adj[r + 2], f1z); }
    }
  }
  newmark_point<TWO>((size_t)halo_g[h], rm, f0x, f0y, f0z, f1x, f1y, f1z, F0, F1, last != 0, rho_kl, rho_part,
                     flush != 0, hess_kl, dtk);
}

void brick_free(SemMesh* m) {
  dev_free(m->d_brick_special);
  dev_free(m->d_brick_part0);
  dev_free(m->d_brick_part1);
  dev_free(m->d_halo_g);
  dev_free(m->d_halo_off);
  dev_free(m->d_halo_idx);
  m->brick_x = m->brick_y = m->brick_z = 0;
  m->nbricks = m->nhalo = 0;
}

bool brick_shape_supported(int bx, int by, int bz) {
  return (bx == 2 && by == 2 && bz == 2) || (bx == 4 && by == 2 && bz == 2) || (bx == 4 && by == 4 && bz == 2);
}

template <int BX, int BY, int BZ>
void launch_brick_shape(SemMesh* m, int nf, const BrickArgs& a) {
  int nb = m->nbricks;
  if (m->brick_bpc > 0) {
    int dev = 0, ncu = 0;
    HIP_CHECK(hipGetDevice(&dev));
    HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    nb = std::min(nb, ncu * m->brick_bpc);
  }
  nb = std::max(8, (nb + 7) & ~7);
  const int sp = m->static_precision;
  hipStream_t s = m->stream;
#define LK(NF, SP) SEM_LAUNCH((k_brick<BX, BY, BZ, NF, SP>), dim3(nb), dim3(128), 0, s, a)
  if (nf == 2) {
    if (sp == SEM_STATIC_BF16) LK(2, SEM_STATIC_BF16);
    else if (sp == SEM_STATIC_FP16) LK(2, SEM_STATIC_FP16);
    else LK(2, SEM_STATIC_FP32);
  } else {
    if (sp == SEM_STATIC_BF16) LK(1, SEM_STATIC_BF16);
    else if (sp == SEM_STATIC_FP16) LK(1, SEM_STATIC_FP16);
    else LK(1, SEM_STATIC_FP32);
  }
#undef LK
  HIP_CHECK(hipGetLastError());
}

}  // namespace

bool sem_brick_active(const SemMesh* m, int nfields) {
  if (m->brick_x == 0 || m->nspec_irregular != 0 || m->anisotropy) return false;
  return nfields == 1 ? m->d_brick_part0 != nullptr : m->d_brick_part1 != nullptr;
}

void sem_launch_brick(SemMesh* m, int nfields, const AsmField& F0, const AsmField& F1, bool last, double dtk) {
  if (!sem_brick_active(m, nfields)) SEM_FAIL("brick pass not prepared for %d field(s)", nfields);
  if (m->static_precision != SEM_STATIC_FP32 && !m->d_kmu_lp) SEM_FAIL("reduced-precision statics not prepared");
  if (nfields == 2 && (!m->d_mu_kl || !m->d_kappa_kl || !m->d_rho_kl_glob))
    SEM_FAIL("brick adjoint pass needs the kernel accumulators");
  BrickArgs a{};
  a.ibool = m->d_ibool; a.kappa = m->d_kappa; a.mu = m->d_mu;
  a.kmu_lp = reinterpret_cast<const uint16_t*>(m->d_kmu_lp);
  a.kmu_base = m->d_kmu_base;
  a.tables = m->d_tables;
  a.nex = m->lat_nex; a.ney = m->lat_ney;
  a.nbx = m->lat_nex / m->brick_x; a.nby = m->lat_ney / m->brick_y; a.nbricks = m->nbricks;
  a.xix_regular = m->xix_regular; a.jacobian_regular = m->jacobian_regular;
  a.scale_k = m->static_scale_kappa; a.scale_m = m->static_scale_mu;
  a.special = m->d_brick_special;
  a.part0 = m->d_brick_part0; a.part1 = m->d_brick_part1;
  a.rm = m->d_rmass_inv;
  a.F0 = F0; a.F1 = F1;
  a.mu_kl = m->d_mu_kl; a.kappa_kl = m->d_kappa_kl;
  a.rho_kl = m->d_rho_kl_glob; a.hess_kl = m->d_hess_kl_glob;
  if (nfields == 2 && m->kernel_block > 1) {
    a.mu_part = m->d_mu_part;
    a.kappa_part = m->d_kappa_part;
    a.rho_part = m->d_rho_part;
    if (!a.mu_part || !a.kappa_part || !a.rho_part) SEM_FAIL("kernel partial sums not allocated");
  }
  a.kflush = m->kernel_flush;
  a.dt = (double)F0.dt;
  a.dtk = dtk;
  a.last = last ? 1 : 0;
  const int bx = m->brick_x, by = m->brick_y, bz = m->brick_z;
  if (bx == 2 && by == 2 && bz == 2) launch_brick_shape<2, 2, 2>(m, nfields, a);
  else if (bx == 4 && by == 2 && bz == 2) launch_brick_shape<4, 2, 2>(m, nfields, a);
  else if (bx == 4 && by == 4 && bz == 2) launch_brick_shape<4, 4, 2>(m, nfields, a);
  else SEM_FAIL("brick shape %dx%dx%d not built", bx, by, bz);
}

void sem_launch_halo(SemMesh* m, int nfields, const AsmField& F0, const AsmField& F1, bool last, double dtk) {
  if (!sem_brick_active(m, nfields)) SEM_FAIL("brick pass not prepared for %d field(s)", nfields);
  if (m->nhalo == 0) return;
  const int npts = (4 * m->brick_x + 1) * (4 * m->brick_y + 1) * (4 * m->brick_z + 1);
  const int nslot = m->nbricks * npts;
  const int nfp = m->absorbing ? m->nfaces * 25 : 0;
  const int nsp = m->nsources * 125;
  const dim3 gr((m->nhalo + TPB - 1) / TPB), bl(TPB);
  if (nfields == 2)
    SEM_LAUNCH(k_halo<true>, gr, bl, 0, m->stream, m->nhalo, m->d_halo_g, m->d_halo_off, m->d_halo_idx, nslot,
                       nfp, nsp, m->d_sourcearrays, m->d_rec_h, m->d_rmass_inv, m->d_brick_part0, m->d_brick_part1,
                       F0, F1, last ? 1 : 0, m->d_rho_kl_glob, m->kernel_block > 1 ? m->d_rho_part : nullptr,
                       m->kernel_flush, m->d_hess_kl_glob, dtk);
  else
    SEM_LAUNCH(k_halo<false>, gr, bl, 0, m->stream, m->nhalo, m->d_halo_g, m->d_halo_off, m->d_halo_idx, nslot,
                       nfp, nsp, m->d_sourcearrays, m->d_rec_h, m->d_rmass_inv, m->d_brick_part0, m->d_brick_part1,
                       F0, F1, last ? 1 : 0, (double*)nullptr, (float*)nullptr, 1, (double*)nullptr, dtk);
  HIP_CHECK(hipGetLastError());
}

extern "C" {

// Declare that the mesh's elements form an nex x ney x nez lattice with element
// (ix, iy, iz) = ix + nex*(iy + ney*iz).  Checked against ibool: neighbouring
// elements share their 25 face points, and every point appears in exactly as
// many elements as its lattice position implies (1 inside an element, 2 on an
// inner face, 4 on an inner edge, 8 at an inner corner).  Aborts otherwise.
void fwat_set_lattice_(long* Mesh_pointer, int* nex, int* ney, int* nez) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  const int NX = *nex, NY = *ney, NZ = *nez;
  if (NX < 1 || NY < 1 || NZ < 1 || (long)NX * NY * NZ != m->nspec)
    SEM_FAIL("fwat_set_lattice_: %d x %d x %d does not match NSPEC_AB = %d", NX, NY, NZ, m->nspec);
  const size_t nloc = (size_t)m->nspec * 125;
  std::vector<int> ib(nloc);
  HIP_CHECK(hipMemcpyAsync(ib.data(), m->d_ibool, nloc * sizeof(int), hipMemcpyDeviceToHost, m->stream));
  HIP_CHECK(hipStreamSynchronize(m->stream));
  const auto I = [&](int e, int i, int j, int k) { return ib[(size_t)e * 125 + i + 5 * j + 25 * k]; };
  for (int iz = 0; iz < NZ; ++iz)
    for (int iy = 0; iy < NY; ++iy)
      for (int ix = 0; ix < NX; ++ix) {
        const int e = ix + NX * (iy + NY * iz);
        for (int a = 0; a < 5; ++a)
          for (int c = 0; c < 5; ++c) {
            if (ix + 1 < NX && I(e, 4, a, c) != I(e + 1, 0, a, c))
              SEM_FAIL("fwat_set_lattice_: element %d and its +x neighbour do not share a face", e + 1);
            if (iy + 1 < NY && I(e, a, 4, c) != I(e + NX, a, 0, c))
              SEM_FAIL("fwat_set_lattice_: element %d and its +y neighbour do not share a face", e + 1);
            if (iz + 1 < NZ && I(e, a, c, 4) != I(e + NX * NY, a, c, 0))
              SEM_FAIL("fwat_set_lattice_: element %d and its +z neighbour do not share a face", e + 1);
          }
      }
  std::vector<int> cnt(m->nglob, 0);
  for (size_t L = 0; L < nloc; ++L) cnt[ib[L]]++;
  const auto mult = [](int p, int n) { return (p % 4 == 0 && p > 0 && p < 4 * n) ? 2 : 1; };
  for (int iz = 0; iz < NZ; ++iz)
    for (int iy = 0; iy < NY; ++iy)
      for (int ix = 0; ix < NX; ++ix) {
        const int e = ix + NX * (iy + NY * iz);
        for (int k = 0; k < 5; ++k)
          for (int j = 0; j < 5; ++j)
            for (int i = 0; i < 5; ++i) {
              const int want = mult(4 * ix + i, NX) * mult(4 * iy + j, NY) * mult(4 * iz + k, NZ);
              if (cnt[I(e, i, j, k)] != want)
                SEM_FAIL("fwat_set_lattice_: point (%d,%d,%d) of element %d is shared by %d elements, the lattice implies %d",
                         i + 1, j + 1, k + 1, e + 1, cnt[I(e, i, j, k)], want);
            }
      }
  brick_free(m);
  m->lat_nex = NX; m->lat_ney = NY; m->lat_nez = NZ;
}

// Brick pass on or off: bricks of bx x by x bz elements (2x2x2, 4x2x2, 4x4x2;
// bx = 0 turns it off and frees its buffers), blocks_per_cu persistent
// workgroups per CU (0: one workgroup per brick).  Needs fwat_set_lattice_, an
// all-regular isotropic mesh and lattice dimensions divisible by the brick.
void fwat_set_brick_(long* Mesh_pointer, int* bx, int* by, int* bz, int* blocks_per_cu) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  brick_free(m);
  if (*bx == 0) return;
  const int BX = *bx, BY = *by, BZ = *bz;
  if (!brick_shape_supported(BX, BY, BZ)) SEM_FAIL("fwat_set_brick_: brick %dx%dx%d not built", BX, BY, BZ);
  if (*blocks_per_cu < 0 || *blocks_per_cu > 64) SEM_FAIL("fwat_set_brick_: blocks per CU must be in 0..64");
  if (m->lat_nex == 0) SEM_FAIL("fwat_set_brick_: call fwat_set_lattice_ first");
  if (m->nspec_irregular != 0 || m->anisotropy) SEM_FAIL("fwat_set_brick_: needs an all-regular isotropic mesh");
  if (m->lat_nex % BX || m->lat_ney % BY || m->lat_nez % BZ)
    SEM_FAIL("fwat_set_brick_: lattice %d x %d x %d not divisible by %d x %d x %d", m->lat_nex, m->lat_ney,
             m->lat_nez, BX, BY, BZ);
  const int nbx = m->lat_nex / BX, nby = m->lat_ney / BY, nbz = m->lat_nez / BZ;
  const int LX = 4 * BX + 1, LY = 4 * BY + 1, LZ = 4 * BZ + 1, npts = LX * LY * LZ;
  const long nbricks = (long)nbx * nby * nbz;
  if (nbricks * npts >= (long)INT32_MAX) SEM_FAIL("fwat_set_brick_: too many slots for 32-bit indices");
  hipStream_t s = m->stream;
  const size_t nloc = (size_t)m->nspec * 125;
  const size_t nfp = m->absorbing ? (size_t)m->nfaces * 25 : 0;
  std::vector<int> ib(nloc), fi(nfp), si((size_t)m->nsources * 125), ri((size_t)m->nrec * 125);
  HIP_CHECK(hipMemcpyAsync(ib.data(), m->d_ibool, nloc * sizeof(int), hipMemcpyDeviceToHost, s));
  if (nfp) HIP_CHECK(hipMemcpyAsync(fi.data(), m->d_face_iglob, nfp * sizeof(int), hipMemcpyDeviceToHost, s));
  if (!si.empty()) HIP_CHECK(hipMemcpyAsync(si.data(), m->d_src_iglob, si.size() * sizeof(int), hipMemcpyDeviceToHost, s));
  if (!ri.empty()) HIP_CHECK(hipMemcpyAsync(ri.data(), m->d_rec_iglob, ri.size() * sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  // bricks holding a source or receiver point: all their points go to the halo pass
  std::vector<uint8_t> special(nbricks, 0);
  std::vector<int> brick_of_glob(m->nglob, -1);  // brick owning an inner point
  const auto brick_of = [&](int e) {
    const int ix = e % m->lat_nex, iy = (e / m->lat_nex) % m->lat_ney, iz = e / (m->lat_nex * m->lat_ney);
    return (ix / BX) + nbx * ((iy / BY) + nby * (iz / BZ));
  };
  // (Stacey face points need no marking: they lie on the mesh boundary, hence on brick boundaries)
  std::vector<uint8_t> touched(m->nglob, 0);
  for (int g : si) touched[g] = 1;
  for (int g : ri) touched[g] = 1;
  for (int e = 0; e < m->nspec; ++e)
    for (int L = 0; L < 125; ++L)
      if (touched[ib[(size_t)e * 125 + L]]) special[brick_of(e)] = 1;
  // lattice point (b, q) -> global point, through its owner element
  const auto glob_of = [&](long b, int q) {
    const int bxi = (int)(b % nbx), byi = (int)((b / nbx) % nby), bzi = (int)(b / ((long)nbx * nby));
    const int px = q % LX, py = (q / LX) % LY, pz = q / (LX * LY);
    const int ex = std::min(px / 4, BX - 1), ey = std::min(py / 4, BY - 1), ez = std::min(pz / 4, BZ - 1);
    const int e = (bxi * BX + ex) + m->lat_nex * ((byi * BY + ey) + m->lat_ney * (bzi * BZ + ez));
    return ib[(size_t)e * 125 + (px - 4 * ex) + 5 * (py - 4 * ey) + 25 * (pz - 4 * ez)];
  };
  const auto is_inner = [&](long b, int q) {
    const int px = q % LX, py = (q / LX) % LY, pz = q / (LX * LY);
    return !special[b] && px > 0 && px < LX - 1 && py > 0 && py < LY - 1 && pz > 0 && pz < LZ - 1;
  };
  // halo CSR: per point, its slots in brick order, then face, source, receiver terms
  std::vector<int> cnt(m->nglob, 0);
  std::vector<uint8_t> inner_done(m->nglob, 0);
  for (long b = 0; b < nbricks; ++b)
    for (int q = 0; q < npts; ++q) {
      const int g = glob_of(b, q);
      if (is_inner(b, q)) {
        if (inner_done[g]) SEM_FAIL("fwat_set_brick_: point %d is inner to two bricks", g + 1);
        inner_done[g] = 1;
      } else {
        cnt[g]++;
      }
    }
  for (int g : fi) cnt[g]++;
  for (int g : si) cnt[g]++;
  for (int g : ri) cnt[g]++;
  std::vector<int> hidx(m->nglob, -1), halo_g;
  for (int g = 0; g < m->nglob; ++g) {
    if (inner_done[g] && cnt[g]) SEM_FAIL("fwat_set_brick_: inner point %d also has halo terms", g + 1);
    if (!inner_done[g] && !cnt[g]) SEM_FAIL("fwat_set_brick_: point %d is in no brick", g + 1);
    if (cnt[g]) { hidx[g] = (int)halo_g.size(); halo_g.push_back(g); }
  }
  const int nh = (int)halo_g.size();
  std::vector<int> off(nh + 1, 0);
  for (int h = 0; h < nh; ++h) off[h + 1] = off[h] + cnt[halo_g[h]];
  std::vector<int> fill(off.begin(), off.end() - 1), idx(off[nh]);
  const int nslot = (int)(nbricks * npts);
  for (long b = 0; b < nbricks; ++b)
    for (int q = 0; q < npts; ++q)
      if (!is_inner(b, q)) {
        const int h = hidx[glob_of(b, q)];
        idx[fill[h]++] = (int)(b * npts + q);
      }
  for (size_t q = 0; q < nfp; ++q) idx[fill[hidx[fi[q]]]++] = nslot + (int)q;
  for (size_t q = 0; q < si.size(); ++q) idx[fill[hidx[si[q]]]++] = nslot + (int)nfp + (int)q;
  for (size_t q = 0; q < ri.size(); ++q) idx[fill[hidx[ri[q]]]++] = nslot + (int)nfp + (int)si.size() + (int)q;
  m->brick_x = BX; m->brick_y = BY; m->brick_z = BZ;
  m->brick_bpc = *blocks_per_cu;
  m->nbricks = (int)nbricks;
  m->nhalo = nh;
  m->d_brick_special = dev_upload(special.data(), special.size(), s);
  m->d_halo_g = dev_upload(halo_g.data(), halo_g.size(), s);
  m->d_halo_off = dev_upload(off.data(), off.size(), s);
  m->d_halo_idx = dev_upload(idx.data(), idx.size(), s);
  m->d_brick_part0 = dev_alloc<realw>((size_t)nslot * 3);
  if (m->simulation_type == 3) m->d_brick_part1 = dev_alloc<realw>((size_t)nslot * 3);
  HIP_CHECK(hipStreamSynchronize(s));
}

// Sizes of the brick decomposition (for traffic models): points assembled by the
// halo pass, brick partial-sum slots written per step and field, halo CSR length.
void fwat_brick_info_(long* Mesh_pointer, int* nhalo, long* nslots, long* ncontrib) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  *nhalo = m->nhalo;
  *nslots = 0;
  *ncontrib = 0;
  if (!m->brick_x) return;
  std::vector<int> off(m->nhalo + 1);
  HIP_CHECK(hipMemcpyAsync(off.data(), m->d_halo_off, off.size() * sizeof(int), hipMemcpyDeviceToHost, m->stream));
  HIP_CHECK(hipStreamSynchronize(m->stream));
  const long nfp = m->absorbing ? (long)m->nfaces * 25 : 0;
  *ncontrib = off[m->nhalo];
  *nslots = *ncontrib - nfp - 125L * (m->nsources + m->nrec);
}

}  // extern "C"
