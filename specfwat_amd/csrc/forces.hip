// Elastic internal forces, accel -= K u, for SPECFEM-style NGLL=5 hexahedra on
// CDNA4 (gfx950).
//
// Replaces the solver's internal-force kernel that SpecFWAT reaches through
// iterate_time (src/preproc/prepare_fwd.f90:364; SPECFEM3D
// compute_forces_viscoelastic, source absent from the tree).  The arithmetic
// is the GLL weak form:
//   gradient     du/dxi_l  = sum_m u(m) hprime(l,m)              (3 directions x 3 components)
//   strain/stress (isotropic kappa,mu or 21 c_ij, Voigt 1..6 = xx yy zz yz xz xy)
//   flux         t_d = jacobian * sigma . grad(xi_d)
//   divergence   f = wgllwgll_yz sum t_1 hprimewgll + wgllwgll_xz sum t_2 .. + wgllwgll_xy sum t_3 ..
//   scatter      accel(iglob) -= f
//
// Work mapping: one element per 128-lane workgroup (two wave64s, 125 active
// lanes, lane L = i + 5j + 25k).  The element's displacement (3x125 fp32) and
// its nine flux fields (9x125) are staged in LDS; the 5x5 derivative tables
// live in registers.  Workgroups are remapped so each XCD owns a contiguous
// run of elements (shared GLL points then meet in that XCD's L2).  Assembly is
// an fp32 atomic add to HBM (global_atomic_add_f32).
//
// Template switches:
//   NF    1 = one field, 2 = adjoint + reconstructed field in the same pass
//   ANISO 21 c_ij instead of kappa/mu
//   KMODE 0 = forces only, 1 = also store the strain (SPECFEM-style kernel
//         path), 2 = accumulate the mu/kappa (or c_ij) kernels in fp64 in the
//         same pass (requires NF == 2)
//   SP    storage precision of kappa/mu/Jacobian: fp32, bf16, fp16 (scaled)
//   ASM   0 = fp32 atomic add into accel (SPECFEM's scatter), 1 = write the
//         element-local forces to floc[ispec][sem_floc_slot(L)][3]; a per-point gather over
//         a static CSR then assembles them in a fixed order (no atomics)
#include "sem_device.h"

#include <algorithm>

namespace {

struct ForceArgs {
  const int* ibool;
  const int* irregular;
  const realw* jinv10;
  const uint16_t* jinv10_lp;
  const realw* kappa;
  const realw* mu;
  const uint16_t* kmu_lp;
  const float* kmu_base;     // [nspec][2] per-element means (reduced-precision statics)
  const float* jinv10_base;  // [nspec_irr][10]
  const realw* cij;
  const realw* tables;
  const int* elem_list;
  // pipelined regular pass in column blocks (regular_path 5): columns of the element lattice
  // (ncol = nex * ney) and elements per column; col_nez = 0: element order
  int col_ncol, col_nez;
  int grouped;  // element-local force rows grouped by reader (sem_floc_slot), else in local-point order
  int elem_begin, elem_count;
  int nblocks_padded;
  realw xix_regular, jacobian_regular;
  float scale_k, scale_m, scale_jinv, scale_jac;
  const realw* u0;
  realw* a0;
  const realw* u1;
  realw* a1;
  realw* eps0;
  realw* floc0;
  realw* floc1;
  double* mu_kl;
  double* kappa_kl;
  double* cijkl_kl;
  float* mu_part;     // fp32 partial sums of the kernels (null: straight into fp64)
  float* kappa_part;
  float* cij_part;
  int kflush;         // this step closes a partial-sum block
  double dt;
  // interior-point Newmark (pipelined pass of a fused loop): points >= n_shared have
  // this element's term only and are finished here (newmark_point); every lane's
  // element term still goes to floc
  int n_shared;
  int last;
  const realw* rm;
  AsmField F0, F1;
  double* rho_kl;
  float* rho_part;
  double* hess_kl;
  double dtk;
};

__device__ __forceinline__ float bf16_to_f(uint16_t x) { return __uint_as_float((uint32_t)x << 16); }
__device__ __forceinline__ float h16_to_f(uint16_t x) {
  _Float16 h;
  __builtin_memcpy(&h, &x, 2);
  return (float)h;
}
template <int SP>
__device__ __forceinline__ float lp_to_f(uint16_t x) {
  return SP == SEM_STATIC_BF16 ? bf16_to_f(x) : h16_to_f(x);
}

template <int NF, bool ANISO, int KMODE, int SP, int ASM>
__global__ __launch_bounds__(128) void k_forces(ForceArgs A) {
  __shared__ realw s_tab[128];
  __shared__ realw s_u[3][128];
  __shared__ realw s_t[9][128];

  const int t = threadIdx.x;
  if (t < 125) s_tab[t] = A.tables[t];
  const int lb = xcd_remap(blockIdx.x, A.nblocks_padded);
  if (lb >= A.elem_count) return;  // block-uniform
  const int ispec = A.elem_list ? A.elem_list[A.elem_begin + lb] : A.elem_begin + lb;
  const bool act = t < 125;
  const int tt = act ? t : 124;
  const int I = tt % 5, J = (tt / 5) % 5, K = tt / 25;
  __syncthreads();

  realw hpI[5], hpJ[5], hpK[5], hwI[5], hwJ[5], hwK[5];
#pragma unroll
  for (int l = 0; l < 5; ++l) {
    hpI[l] = s_tab[I + 5 * l];
    hpJ[l] = s_tab[J + 5 * l];
    hpK[l] = s_tab[K + 5 * l];
    hwI[l] = s_tab[25 + l + 5 * I];
    hwJ[l] = s_tab[25 + l + 5 * J];
    hwK[l] = s_tab[25 + l + 5 * K];
  }
  const realw fac1 = s_tab[50 + J + 5 * K];
  const realw fac2 = s_tab[75 + I + 5 * K];
  const realw fac3 = s_tab[100 + I + 5 * J];

  const size_t Lg = (size_t)ispec * 125 + tt;
  const int iglob = A.ibool[Lg];

  // geometry
  const int ir = A.irregular[ispec];
  realw xix, xiy, xiz, etx, ety, etz, gax, gay, gaz, jl;
  if (ir) {
    const size_t base = (size_t)(ir - 1) * 1250 + tt;
    if (SP == SEM_STATIC_FP32) {
      const realw* G = A.jinv10 + base;
      xix = G[0]; xiy = G[125]; xiz = G[250];
      etx = G[375]; ety = G[500]; etz = G[625];
      gax = G[750]; gay = G[875]; gaz = G[1000];
      jl = G[1125];
    } else {
      // per-element fp32 mean + reduced-precision deviation
      const uint16_t* G = A.jinv10_lp + base;
      const float* B = A.jinv10_base + (size_t)(ir - 1) * 10;
      const float sj = A.scale_jinv;
      xix = B[0] + lp_to_f<SP>(G[0]) * sj; xiy = B[1] + lp_to_f<SP>(G[125]) * sj;
      xiz = B[2] + lp_to_f<SP>(G[250]) * sj; etx = B[3] + lp_to_f<SP>(G[375]) * sj;
      ety = B[4] + lp_to_f<SP>(G[500]) * sj; etz = B[5] + lp_to_f<SP>(G[625]) * sj;
      gax = B[6] + lp_to_f<SP>(G[750]) * sj; gay = B[7] + lp_to_f<SP>(G[875]) * sj;
      gaz = B[8] + lp_to_f<SP>(G[1000]) * sj;
      jl = B[9] + lp_to_f<SP>(G[1125]) * A.scale_jac;
    }
  } else {
    xix = A.xix_regular; ety = A.xix_regular; gaz = A.xix_regular;
    xiy = xiz = etx = etz = gax = gay = 0.0f;
    jl = A.jacobian_regular;
  }

  // material
  realw kl = 0, ml = 0;
  const realw* C = nullptr;
  if (!ANISO) {
    if (SP == SEM_STATIC_FP32) {
      kl = A.kappa[Lg];
      ml = A.mu[Lg];
    } else {
      const uint32_t pk = reinterpret_cast<const uint32_t*>(A.kmu_lp)[Lg];
      const float2 kb = reinterpret_cast<const float2*>(A.kmu_base)[ispec];
      kl = kb.x + lp_to_f<SP>((uint16_t)(pk & 0xffffu)) * A.scale_k;
      ml = kb.y + lp_to_f<SP>((uint16_t)(pk >> 16)) * A.scale_m;
    }
  } else {
    C = A.cij + (size_t)ispec * 21 * 125 + tt;
  }

  realw ea[6];  // strain of field 0 (KMODE 2): exx eyy ezz exy exz eyz

#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const realw* U = f == 0 ? A.u0 : A.u1;
    realw* Acc = f == 0 ? A.a0 : A.a1;
    realw* Floc = f == 0 ? A.floc0 : A.floc1;
    if (act) {
      s_u[0][t] = U[3 * (size_t)iglob];
      s_u[1][t] = U[3 * (size_t)iglob + 1];
      s_u[2][t] = U[3 * (size_t)iglob + 2];
    }
    __syncthreads();
    realw ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0, cx = 0, cy = 0, cz = 0;
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      const int La = l + 5 * J + 25 * K, Lb = I + 5 * l + 25 * K, Lc = I + 5 * J + 25 * l;
      ax += s_u[0][La] * hpI[l]; ay += s_u[1][La] * hpI[l]; az += s_u[2][La] * hpI[l];
      bx += s_u[0][Lb] * hpJ[l]; by += s_u[1][Lb] * hpJ[l]; bz += s_u[2][Lb] * hpJ[l];
      cx += s_u[0][Lc] * hpK[l]; cy += s_u[1][Lc] * hpK[l]; cz += s_u[2][Lc] * hpK[l];
    }
    const realw duxdx = xix * ax + etx * bx + gax * cx;
    const realw duxdy = xiy * ax + ety * bx + gay * cx;
    const realw duxdz = xiz * ax + etz * bx + gaz * cx;
    const realw duydx = xix * ay + etx * by + gax * cy;
    const realw duydy = xiy * ay + ety * by + gay * cy;
    const realw duydz = xiz * ay + etz * by + gaz * cy;
    const realw duzdx = xix * az + etx * bz + gax * cz;
    const realw duzdy = xiy * az + ety * bz + gay * cz;
    const realw duzdz = xiz * az + etz * bz + gaz * cz;
    const realw exy2 = duxdy + duydx, exz2 = duzdx + duxdz, eyz2 = duzdy + duydz;

    if (KMODE == 1 && act) {
      realw* E = A.eps0 + (size_t)ispec * 750 + t;
      E[0] = duxdx; E[125] = duydy; E[250] = duzdz;
      E[375] = 0.5f * exy2; E[500] = 0.5f * exz2; E[625] = 0.5f * eyz2;
    }
    if (KMODE == 2) {
      if (f == 0) {
        ea[0] = duxdx; ea[1] = duydy; ea[2] = duzdz;
        ea[3] = 0.5f * exy2; ea[4] = 0.5f * exz2; ea[5] = 0.5f * eyz2;
      } else if (act) {
        const realw eb0 = duxdx, eb1 = duydy, eb2 = duzdz;
        const realw eb3 = 0.5f * exy2, eb4 = 0.5f * exz2, eb5 = 0.5f * eyz2;
        if (!ANISO) {
          const realw tra = ea[0] + ea[1] + ea[2], trb = eb0 + eb1 + eb2;
          const realw t3a = tra / 3.0f, t3b = trb / 3.0f;
          const realw prod = (ea[0] - t3a) * (eb0 - t3b) + (ea[1] - t3a) * (eb1 - t3b) +
                             (ea[2] - t3a) * (eb2 - t3b) +
                             2.0f * (ea[3] * eb3 + ea[4] * eb4 + ea[5] * eb5);
          kernel_acc(A.mu_kl, A.mu_part, Lg, prod, A.dt, A.kflush != 0);
          kernel_acc(A.kappa_kl, A.kappa_part, Lg, tra * trb, A.dt, A.kflush != 0);
        } else {
          const realw va[6] = {ea[0], ea[1], ea[2], 2.0f * ea[5], 2.0f * ea[4], 2.0f * ea[3]};
          const realw vb[6] = {eb0, eb1, eb2, 2.0f * eb5, 2.0f * eb4, 2.0f * eb3};
          const size_t k0 = (size_t)ispec * 21 * 125 + t;
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
    }

    realw sxx, syy, szz, sxy, sxz, syz;
    if (!ANISO) {
      const realw lp2m = kl + (4.0f / 3.0f) * ml;
      const realw lam = lp2m - 2.0f * ml;
      sxx = lp2m * duxdx + lam * (duydy + duzdz);
      syy = lp2m * duydy + lam * (duxdx + duzdz);
      szz = lp2m * duzdz + lam * (duxdx + duydy);
      sxy = ml * exy2;
      sxz = ml * exz2;
      syz = ml * eyz2;
    } else {
      // c index (P,Q), P<=Q, upper-triangle row-major: 11 12 13 14 15 16 22 .. 66
      const realw e1 = duxdx, e2 = duydy, e3 = duzdz, e4 = eyz2, e5 = exz2, e6 = exy2;
      asm volatile("" ::: "memory");  // reload c_ij per field rather than keep 21 registers live
      const realw c11 = C[0], c12 = C[125], c13 = C[250], c14 = C[375], c15 = C[500], c16 = C[625];
      const realw c22 = C[750], c23 = C[875], c24 = C[1000], c25 = C[1125], c26 = C[1250];
      const realw c33 = C[1375], c34 = C[1500], c35 = C[1625], c36 = C[1750];
      const realw c44 = C[1875], c45 = C[2000], c46 = C[2125];
      const realw c55 = C[2250], c56 = C[2375], c66 = C[2500];
      sxx = c11 * e1 + c12 * e2 + c13 * e3 + c14 * e4 + c15 * e5 + c16 * e6;
      syy = c12 * e1 + c22 * e2 + c23 * e3 + c24 * e4 + c25 * e5 + c26 * e6;
      szz = c13 * e1 + c23 * e2 + c33 * e3 + c34 * e4 + c35 * e5 + c36 * e6;
      syz = c14 * e1 + c24 * e2 + c34 * e3 + c44 * e4 + c45 * e5 + c46 * e6;
      sxz = c15 * e1 + c25 * e2 + c35 * e3 + c45 * e4 + c55 * e5 + c56 * e6;
      sxy = c16 * e1 + c26 * e2 + c36 * e3 + c46 * e4 + c56 * e5 + c66 * e6;
    }
    if (act) {
      s_t[0][t] = jl * (sxx * xix + sxy * xiy + sxz * xiz);
      s_t[1][t] = jl * (sxy * xix + syy * xiy + syz * xiz);
      s_t[2][t] = jl * (sxz * xix + syz * xiy + szz * xiz);
      s_t[3][t] = jl * (sxx * etx + sxy * ety + sxz * etz);
      s_t[4][t] = jl * (sxy * etx + syy * ety + syz * etz);
      s_t[5][t] = jl * (sxz * etx + syz * ety + szz * etz);
      s_t[6][t] = jl * (sxx * gax + sxy * gay + sxz * gaz);
      s_t[7][t] = jl * (sxy * gax + syy * gay + syz * gaz);
      s_t[8][t] = jl * (sxz * gax + syz * gay + szz * gaz);
    }
    __syncthreads();
    realw nx1 = 0, ny1 = 0, nz1 = 0, nx2 = 0, ny2 = 0, nz2 = 0, nx3 = 0, ny3 = 0, nz3 = 0;
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      const int La = l + 5 * J + 25 * K, Lb = I + 5 * l + 25 * K, Lc = I + 5 * J + 25 * l;
      nx1 += s_t[0][La] * hwI[l]; ny1 += s_t[1][La] * hwI[l]; nz1 += s_t[2][La] * hwI[l];
      nx2 += s_t[3][Lb] * hwJ[l]; ny2 += s_t[4][Lb] * hwJ[l]; nz2 += s_t[5][Lb] * hwJ[l];
      nx3 += s_t[6][Lc] * hwK[l]; ny3 += s_t[7][Lc] * hwK[l]; nz3 += s_t[8][Lc] * hwK[l];
    }
    if (act) {
      const realw fx = fac1 * nx1 + fac2 * nx2 + fac3 * nx3;
      const realw fy = fac1 * ny1 + fac2 * ny2 + fac3 * ny3;
      const realw fz = fac1 * nz1 + fac2 * nz2 + fac3 * nz3;
      if (ASM == 0) {
        unsafeAtomicAdd(Acc + 3 * (size_t)iglob, -fx);
        unsafeAtomicAdd(Acc + 3 * (size_t)iglob + 1, -fy);
        unsafeAtomicAdd(Acc + 3 * (size_t)iglob + 2, -fz);
      } else {
        realw* F = Floc + 3 * (Lg - tt + (A.grouped ? sem_floc_slot(tt) : tt));
        F[0] = -fx;
        F[1] = -fy;
        F[2] = -fz;
      }
    }
  }
}

// Regular-mesh element pass (every element an axis-aligned cube of the
// mesh size, NSPEC_IRREGULAR = 0), gather-assembly output.  The inverse
// Jacobian is xix_regular * I, so:
//   * the derivative tables carry xix_regular: the contraction yields
//     du/dx directly;
//   * the divergence tables carry -jacobian_regular * xix_regular *
//     wgllwgll(..) of the output point, and LDS holds the six stress
//     components only (the flux of direction d is row d of sigma);
//   * the three directions accumulate into one register per component.
// Workgroups are persistent: each loops over its share of an XCD-contiguous
// run of elements, so the table set-up is paid once per workgroup.
template <int NF, bool ANISO, int KMODE, int SP>
__device__ __forceinline__ void forces_reg_body(const ForceArgs& A) {
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
  // XCD-aware persistent schedule: XCD x = blockIdx % 8 owns elements
  // [x*chunk, (x+1)*chunk); its workgroups stride through that run together.
  const int nblk = gridDim.x;
  const int x = blockIdx.x & 7, per_xcd = nblk >> 3, slot = blockIdx.x >> 3;
  const int chunk = (A.elem_count + 7) >> 3;
  const int e_end = min(A.elem_count, (x + 1) * chunk);
  for (int le = x * chunk + slot; le < e_end; le += per_xcd) {
    const int e = A.elem_begin + le;  // elements [elem_begin, elem_begin + elem_count)
    const size_t Lg = (size_t)e * 125 + tt;
    const int iglob = A.ibool[Lg];
    realw kl = 0, ml = 0;
    const realw* C = nullptr;
    if (!ANISO) {
      if (SP == SEM_STATIC_FP32) {
        kl = A.kappa[Lg];
        ml = A.mu[Lg];
      } else {
        const uint32_t pk = reinterpret_cast<const uint32_t*>(A.kmu_lp)[Lg];
        const float2 kb = reinterpret_cast<const float2*>(A.kmu_base)[e];
        kl = kb.x + lp_to_f<SP>((uint16_t)(pk & 0xffffu)) * A.scale_k;
        ml = kb.y + lp_to_f<SP>((uint16_t)(pk >> 16)) * A.scale_m;
      }
    } else {
      C = A.cij + (size_t)e * 21 * 125 + tt;
    }
    realw ea[6];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const realw* U = f == 0 ? A.u0 : A.u1;
      realw* Floc = f == 0 ? A.floc0 : A.floc1;
      __syncthreads();  // previous element / field done with s_u and s_s
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
      if (KMODE == 2) {
        if (f == 0) {
          ea[0] = duxdx; ea[1] = duydy; ea[2] = duzdz;
          ea[3] = 0.5f * exy2; ea[4] = 0.5f * exz2; ea[5] = 0.5f * eyz2;
        } else if (act) {
          const realw eb0 = duxdx, eb1 = duydy, eb2 = duzdz;
          const realw eb3 = 0.5f * exy2, eb4 = 0.5f * exz2, eb5 = 0.5f * eyz2;
          if (!ANISO) {
            const realw tra = ea[0] + ea[1] + ea[2], trb = eb0 + eb1 + eb2;
            const realw t3a = tra / 3.0f, t3b = trb / 3.0f;
            const realw prod = (ea[0] - t3a) * (eb0 - t3b) + (ea[1] - t3a) * (eb1 - t3b) +
                               (ea[2] - t3a) * (eb2 - t3b) +
                               2.0f * (ea[3] * eb3 + ea[4] * eb4 + ea[5] * eb5);
            kernel_acc(A.mu_kl, A.mu_part, Lg, prod, A.dt, A.kflush != 0);
            kernel_acc(A.kappa_kl, A.kappa_part, Lg, tra * trb, A.dt, A.kflush != 0);
          } else {
            const realw va[6] = {ea[0], ea[1], ea[2], 2.0f * ea[5], 2.0f * ea[4], 2.0f * ea[3]};
            const realw vb[6] = {eb0, eb1, eb2, 2.0f * eb5, 2.0f * eb4, 2.0f * eb3};
            const size_t k0 = (size_t)e * 21 * 125 + t;
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
      }
      realw sxx, syy, szz, sxy, sxz, syz;
      if (!ANISO) {
        const realw lam = kl - (2.0f / 3.0f) * ml;
        const realw m2 = 2.0f * ml;
        const realw ltr = lam * (duxdx + duydy + duzdz);
        sxx = ltr + m2 * duxdx;
        syy = ltr + m2 * duydy;
        szz = ltr + m2 * duzdz;
        sxy = ml * exy2;
        sxz = ml * exz2;
        syz = ml * eyz2;
      } else {
        const realw e1 = duxdx, e2 = duydy, e3 = duzdz, e4 = eyz2, e5 = exz2, e6 = exy2;
        // load the 21 c_ij here for each field (cache hits the second time) instead of
        // keeping them live across both fields: 144 -> fewer VGPRs, more waves per SIMD
        asm volatile("" ::: "memory");
        const realw c11 = C[0], c12 = C[125], c13 = C[250], c14 = C[375], c15 = C[500], c16 = C[625];
        const realw c22 = C[750], c23 = C[875], c24 = C[1000], c25 = C[1125], c26 = C[1250];
        const realw c33 = C[1375], c34 = C[1500], c35 = C[1625], c36 = C[1750];
        const realw c44 = C[1875], c45 = C[2000], c46 = C[2125];
        const realw c55 = C[2250], c56 = C[2375], c66 = C[2500];
        sxx = c11 * e1 + c12 * e2 + c13 * e3 + c14 * e4 + c15 * e5 + c16 * e6;
        syy = c12 * e1 + c22 * e2 + c23 * e3 + c24 * e4 + c25 * e5 + c26 * e6;
        szz = c13 * e1 + c23 * e2 + c33 * e3 + c34 * e4 + c35 * e5 + c36 * e6;
        syz = c14 * e1 + c24 * e2 + c34 * e3 + c44 * e4 + c45 * e5 + c46 * e6;
        sxz = c15 * e1 + c25 * e2 + c35 * e3 + c45 * e4 + c55 * e5 + c56 * e6;
        sxy = c16 * e1 + c26 * e2 + c36 * e3 + c46 * e4 + c56 * e5 + c66 * e6;
      }
      if (act) {
        s_s[0][t] = sxx; s_s[1][t] = syy; s_s[2][t] = szz;
        s_s[3][t] = sxy; s_s[4][t] = sxz; s_s[5][t] = syz;
      }
      __syncthreads();
      realw fx = 0, fy = 0, fz = 0;
#pragma unroll
      for (int l = 0; l < 5; ++l) {
        const int La = La0 + l, Lb = Lb0 + 5 * l, Lc = Lc0 + 25 * l;
        // direction x: row x of sigma at (l,J,K); y: row y at (I,l,K); z: row z at (I,J,l)
        fx += s_s[0][La] * hwI[l]; fy += s_s[3][La] * hwI[l]; fz += s_s[4][La] * hwI[l];
        fx += s_s[3][Lb] * hwJ[l]; fy += s_s[1][Lb] * hwJ[l]; fz += s_s[5][Lb] * hwJ[l];
        fx += s_s[4][Lc] * hwK[l]; fy += s_s[5][Lc] * hwK[l]; fz += s_s[2][Lc] * hwK[l];
      }
      if (act) {
        realw* F = Floc + 3 * (Lg - tt + (A.grouped ? sem_floc_slot(tt) : tt));
        F[0] = fx;
        F[1] = fy;
        F[2] = fz;
      }
    }
  }
}

// Software-pipelined variant of the isotropic regular-mesh pass: each
// workgroup walks its elements with the next element's displacements (both
// fields) and moduli loaded one element ahead and the point indices two
// elements ahead, so the dependent ibool -> displacement gather of element
// n+1 is in flight while element n's contractions run; the kernel partial
// sums of element n are read at its start.  Arithmetic and results are those
// of forces_reg_body.
template <int NF, int KMODE, int SP>
__device__ __forceinline__ void forces_reg_pipe_body(const ForceArgs& A) {
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
  const int nblk = gridDim.x;
  const int x = blockIdx.x & 7, per_xcd = nblk >> 3, slot = blockIdx.x >> 3;
  const int chunk = (A.elem_count + 7) >> 3;
  const int e_end = min(A.elem_count, (x + 1) * chunk);
  const bool flush = A.kflush != 0;

  const auto load_moduli = [&](size_t Lg, int e, realw& kl, realw& ml) {
    if (SP == SEM_STATIC_FP32) {
      kl = A.kappa[Lg];
      ml = A.mu[Lg];
    } else {
      const uint32_t pk = reinterpret_cast<const uint32_t*>(A.kmu_lp)[Lg];
      const float2 kb = reinterpret_cast<const float2*>(A.kmu_base)[e];
      kl = kb.x + lp_to_f<SP>((uint16_t)(pk & 0xffffu)) * A.scale_k;
      ml = kb.y + lp_to_f<SP>((uint16_t)(pk >> 16)) * A.scale_m;
    }
  };
  // prologue: element le's displacements and moduli, element le + per_xcd's indices
  int le = x * chunk + slot;
  realw un[NF][3];
  realw kn = 0, mn = 0;
  int ig_next = 0;
  if (le < e_end) {
    const int e = A.elem_begin + le;
    const size_t Lg = (size_t)e * 125 + tt;
    const int ig = A.ibool[Lg];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const realw* U = f == 0 ? A.u0 : A.u1;
#pragma unroll
      for (int c = 0; c < 3; ++c) un[f][c] = U[3 * (size_t)ig + c];
    }
    load_moduli(Lg, e, kn, mn);
    if (le + per_xcd < e_end) ig_next = A.ibool[(size_t)(e + per_xcd) * 125 + tt];
  }
  for (; le < e_end; le += per_xcd) {
    const int e = A.elem_begin + le;
    const size_t Lg = (size_t)e * 125 + tt;
    realw uc[NF][3];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int c = 0; c < 3; ++c) uc[f][c] = un[f][c];
    const realw kl = kn, ml = mn;
    // loads of the next element (its indices were read one element ago)
    if (le + per_xcd < e_end) {
      const int e1 = e + per_xcd;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const realw* U = f == 0 ? A.u0 : A.u1;
#pragma unroll
        for (int c = 0; c < 3; ++c) un[f][c] = U[3 * (size_t)ig_next + c];
      }
      load_moduli((size_t)e1 * 125 + tt, e1, kn, mn);
      if (le + 2 * per_xcd < e_end) ig_next = A.ibool[(size_t)(e1 + per_xcd) * 125 + tt];
    }
    // this element's kernel partial sums
    float pm = 0.0f, pk = 0.0f;
    if (KMODE == 2 && A.mu_part && act) {
      pm = A.mu_part[Lg];
      pk = A.kappa_part[Lg];
    }
    realw ea[6];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      realw* Floc = f == 0 ? A.floc0 : A.floc1;
      __syncthreads();  // previous element / field done with s_u and s_s
      if (act) {
        s_u[0][t] = uc[f][0];
        s_u[1][t] = uc[f][1];
        s_u[2][t] = uc[f][2];
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
      if (KMODE == 2) {
        if (f == 0) {
          ea[0] = duxdx; ea[1] = duydy; ea[2] = duzdz;
          ea[3] = 0.5f * exy2; ea[4] = 0.5f * exz2; ea[5] = 0.5f * eyz2;
        } else if (act) {
          const realw eb0 = duxdx, eb1 = duydy, eb2 = duzdz;
          const realw eb3 = 0.5f * exy2, eb4 = 0.5f * exz2, eb5 = 0.5f * eyz2;
          const realw tra = ea[0] + ea[1] + ea[2], trb = eb0 + eb1 + eb2;
          const realw t3a = tra / 3.0f, t3b = trb / 3.0f;
          const realw prod = (ea[0] - t3a) * (eb0 - t3b) + (ea[1] - t3a) * (eb1 - t3b) +
                             (ea[2] - t3a) * (eb2 - t3b) + 2.0f * (ea[3] * eb3 + ea[4] * eb4 + ea[5] * eb5);
          const realw kx = tra * trb;
          if (!A.mu_part) {
            kernel_acc(A.mu_kl, nullptr, Lg, prod, A.dt, true);
            kernel_acc(A.kappa_kl, nullptr, Lg, kx, A.dt, true);
          } else if (flush) {  // kernel_acc with the partial sums read at the element's start
            A.mu_kl[Lg] += A.dt * (double)(pm + prod);
            A.kappa_kl[Lg] += A.dt * (double)(pk + kx);
            A.mu_part[Lg] = 0.0f;
            A.kappa_part[Lg] = 0.0f;
          } else {
            A.mu_part[Lg] = pm + prod;
            A.kappa_part[Lg] = pk + kx;
          }
        }
      }
      const realw lam = kl - (2.0f / 3.0f) * ml;
      const realw m2 = 2.0f * ml;
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
        realw* F = Floc + 3 * (Lg - tt + (A.grouped ? sem_floc_slot(tt) : tt));
        F[0] = fx;
        F[1] = fy;
        F[2] = fz;
      }
    }
  }
}

// The pipelined pass with two barriers per element instead of three per
// field: both fields' displacements go to LDS at once, both fields' stresses
// after them, and consecutive elements alternate between two LDS buffers, so
// an element's writes never wait for the previous element's readers (those
// finished before the previous element's second barrier).  Arithmetic and
// results are forces_reg_pipe_body's.
template <int NF, int KMODE, int SP, bool NT, bool INT>
__device__ __forceinline__ void forces_reg_pp_body(const ForceArgs& A) {
  // NT: the streamed per-point arrays (indices, moduli, kernel partial sums) are read and
  // written with non-temporal hints, leaving the caches to the displacement gathers
  const auto ld = [](const auto* p) { if constexpr (NT) return __builtin_nontemporal_load(p); else return *p; };
  __shared__ realw s_tab[128];
  __shared__ realw s_u[2][NF][3][128];
  __shared__ realw s_s[2][NF][6][128];
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
  const int nblk = gridDim.x;
  const int x = blockIdx.x & 7, per_xcd = nblk >> 3, slot = blockIdx.x >> 3;
  const int chunk = (A.elem_count + 7) >> 3;
  const int e_end = min(A.elem_count, (x + 1) * chunk);
  const bool flush = A.kflush != 0;
  const auto load_moduli = [&](size_t Lg, int e, realw& kl, realw& ml) {
    if (SP == SEM_STATIC_FP32) {
      kl = ld(&A.kappa[Lg]);
      ml = ld(&A.mu[Lg]);
    } else {
      const uint32_t pk = ld(&reinterpret_cast<const uint32_t*>(A.kmu_lp)[Lg]);
      const float2 kb = reinterpret_cast<const float2*>(A.kmu_base)[e];
      kl = kb.x + lp_to_f<SP>((uint16_t)(pk & 0xffffu)) * A.scale_k;
      ml = kb.y + lp_to_f<SP>((uint16_t)(pk >> 16)) * A.scale_m;
    }
  };
  // element at traversal index i (uniform over the workgroup): element order, or the column
  // blocks of sem_column_block_order (XCD x's share [x*chunk, (x+1)*chunk) holds whole
  // columns; groups of per_xcd columns are visited layer by layer), by integer arithmetic
  const int cz = A.col_nez;
  const int c_first = cz ? (x * chunk) / cz : 0;
  const int c_count = cz ? (e_end - x * chunk) / cz : 0;
  const auto elem_at = [&](int i) {
    if (!cz) return A.elem_begin + i;
    const int t = i - x * chunk;
    const int g = t / (per_xcd * cz);
    const int r = t - g * per_xcd * cz;
    const int pg = min(per_xcd, c_count - g * per_xcd);
    const int z = r / pg;
    return c_first + g * per_xcd + (r - z * pg) + A.col_ncol * z;
  };
  int le = x * chunk + slot;
  realw un[NF][3];
  realw kn = 0, mn = 0;
  int ig_next = 0, ig_cur = 0;
  int e_cur = 0, e_nxt = 0;
  if (le < e_end) {
    e_cur = elem_at(le);
    const size_t Lg = (size_t)e_cur * 125 + tt;
    const int ig = ld(&A.ibool[Lg]);
    ig_cur = ig;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const realw* U = f == 0 ? A.u0 : A.u1;
#pragma unroll
      for (int c = 0; c < 3; ++c) un[f][c] = U[3 * (size_t)ig + c];
    }
    load_moduli(Lg, e_cur, kn, mn);
    if (le + per_xcd < e_end) {
      e_nxt = elem_at(le + per_xcd);
      ig_next = ld(&A.ibool[(size_t)e_nxt * 125 + tt]);
    }
  }
  int b = 0;
  for (; le < e_end; le += per_xcd, b ^= 1) {
    const int e = e_cur;
    const int ig = ig_cur;  // this lane's point of element e
    const size_t Lg = (size_t)e * 125 + tt;
    if (act) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        s_u[b][f][0][t] = un[f][0];
        s_u[b][f][1][t] = un[f][1];
        s_u[b][f][2][t] = un[f][2];
      }
    }
    const realw kl = kn, ml = mn;
    if (le + per_xcd < e_end) {
      const int e1 = e_nxt;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const realw* U = f == 0 ? A.u0 : A.u1;
#pragma unroll
        for (int c = 0; c < 3; ++c) un[f][c] = U[3 * (size_t)ig_next + c];
      }
      load_moduli((size_t)e1 * 125 + tt, e1, kn, mn);
      e_cur = e1;
      ig_cur = ig_next;
      if (le + 2 * per_xcd < e_end) {
        e_nxt = elem_at(le + 2 * per_xcd);
        ig_next = ld(&A.ibool[(size_t)e_nxt * 125 + tt]);
      }
    }
    float pm = 0.0f, pk = 0.0f;
    if (KMODE == 2 && A.mu_part && act) {
      pm = ld(&A.mu_part[Lg]);
      pk = ld(&A.kappa_part[Lg]);
    }
    const bool inner = INT && act && ig >= A.n_shared;
    __syncthreads();  // displacements of both fields in buffer b
    realw ea[6];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      realw duxdx = 0, duydx = 0, duzdx = 0, duxdy = 0, duydy = 0, duzdy = 0, duxdz = 0, duydz = 0, duzdz = 0;
#pragma unroll
      for (int l = 0; l < 5; ++l) {
        const int La = La0 + l, Lb = Lb0 + 5 * l, Lc = Lc0 + 25 * l;
        duxdx += s_u[b][f][0][La] * hpI[l]; duydx += s_u[b][f][1][La] * hpI[l]; duzdx += s_u[b][f][2][La] * hpI[l];
        duxdy += s_u[b][f][0][Lb] * hpJ[l]; duydy += s_u[b][f][1][Lb] * hpJ[l]; duzdy += s_u[b][f][2][Lb] * hpJ[l];
        duxdz += s_u[b][f][0][Lc] * hpK[l]; duydz += s_u[b][f][1][Lc] * hpK[l]; duzdz += s_u[b][f][2][Lc] * hpK[l];
      }
      const realw exy2 = duxdy + duydx, exz2 = duzdx + duxdz, eyz2 = duzdy + duydz;
      if (KMODE == 2) {
        if (f == 0) {
          ea[0] = duxdx; ea[1] = duydy; ea[2] = duzdz;
          ea[3] = 0.5f * exy2; ea[4] = 0.5f * exz2; ea[5] = 0.5f * eyz2;
        } else if (act) {
          const realw eb0 = duxdx, eb1 = duydy, eb2 = duzdz;
          const realw eb3 = 0.5f * exy2, eb4 = 0.5f * exz2, eb5 = 0.5f * eyz2;
          const realw tra = ea[0] + ea[1] + ea[2], trb = eb0 + eb1 + eb2;
          const realw t3a = tra / 3.0f, t3b = trb / 3.0f;
          const realw prod = (ea[0] - t3a) * (eb0 - t3b) + (ea[1] - t3a) * (eb1 - t3b) +
                             (ea[2] - t3a) * (eb2 - t3b) + 2.0f * (ea[3] * eb3 + ea[4] * eb4 + ea[5] * eb5);
          const realw kx = tra * trb;
          if (!A.mu_part) {
            kernel_acc(A.mu_kl, nullptr, Lg, prod, A.dt, true);
            kernel_acc(A.kappa_kl, nullptr, Lg, kx, A.dt, true);
          } else if (flush) {
            A.mu_kl[Lg] += A.dt * (double)(pm + prod);
            A.kappa_kl[Lg] += A.dt * (double)(pk + kx);
            if constexpr (NT) {
              __builtin_nontemporal_store(0.0f, &A.mu_part[Lg]);
              __builtin_nontemporal_store(0.0f, &A.kappa_part[Lg]);
            } else {
              A.mu_part[Lg] = 0.0f;
              A.kappa_part[Lg] = 0.0f;
            }
          } else if constexpr (NT) {
            __builtin_nontemporal_store(pm + prod, &A.mu_part[Lg]);
            __builtin_nontemporal_store(pk + kx, &A.kappa_part[Lg]);
          } else {
            A.mu_part[Lg] = pm + prod;
            A.kappa_part[Lg] = pk + kx;
          }
        }
      }
      const realw lam = kl - (2.0f / 3.0f) * ml;
      const realw m2 = 2.0f * ml;
      const realw ltr = lam * (duxdx + duydy + duzdz);
      if (act) {
        s_s[b][f][0][t] = ltr + m2 * duxdx; s_s[b][f][1][t] = ltr + m2 * duydy; s_s[b][f][2][t] = ltr + m2 * duzdz;
        s_s[b][f][3][t] = ml * exy2; s_s[b][f][4][t] = ml * exz2; s_s[b][f][5][t] = ml * eyz2;
      }
    }
    // an interior point's inverse mass, velocities and density-kernel partial sum,
    // loaded now so that they arrive during the divergence contraction
    PointState P;
    if (INT && inner) {
      const size_t pb = 3 * (size_t)ig;
      P.part = (NF == 2 && A.rho_part) ? A.rho_part[ig] : 0.0f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        P.r[c] = A.rm[pb + c];
        P.v0[c] = A.F0.v[pb + c];
        if (NF == 2) P.v1[c] = A.F1.v[pb + c];
      }
    }
    __syncthreads();  // stresses of both fields in buffer b
    realw ff[NF][3];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      realw fx = 0, fy = 0, fz = 0;
#pragma unroll
      for (int l = 0; l < 5; ++l) {
        const int La = La0 + l, Lb = Lb0 + 5 * l, Lc = Lc0 + 25 * l;
        fx += s_s[b][f][0][La] * hwI[l]; fy += s_s[b][f][3][La] * hwI[l]; fz += s_s[b][f][4][La] * hwI[l];
        fx += s_s[b][f][3][Lb] * hwJ[l]; fy += s_s[b][f][1][Lb] * hwJ[l]; fz += s_s[b][f][5][Lb] * hwJ[l];
        fx += s_s[b][f][4][Lc] * hwK[l]; fy += s_s[b][f][5][Lc] * hwK[l]; fz += s_s[b][f][2][Lc] * hwK[l];
      }
      ff[f][0] = fx; ff[f][1] = fy; ff[f][2] = fz;
      // element-local forces; a point this pass finishes needs none, and the 27 inner
      // slots sit at the end of the row (sem_floc_slot), so leaving them out leaves the
      // rest of the row whole (the inner points of a free-surface element's top face
      // lie mid-row and are stored anyway)
      const int fs = A.grouped ? sem_floc_slot(tt) : tt;
      if (act && !(INT && inner && A.grouped && fs >= 98)) {
        realw* F = (f == 0 ? A.floc0 : A.floc1) + 3 * (Lg - tt + fs);
        F[0] = fx;
        F[1] = fy;
        F[2] = fz;
      }
    }
    if (INT && inner) {
      // the point's one element term, summed as the gather sums a row (0 + term),
      // then the assembly's Newmark finish (same expressions, same rounding)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        P.u0[c] = s_u[b][0][c][t];
        if (NF == 2) P.u1[c] = s_u[b][NF - 1][c][t];
      }
      const realw z = 0.0f;
      if constexpr (NF == 2)
        newmark_point<true>((size_t)ig, P, z + ff[0][0], z + ff[0][1], z + ff[0][2], z + ff[NF - 1][0],
                            z + ff[NF - 1][1], z + ff[NF - 1][2], A.F0, A.F1, A.last != 0, A.rho_kl, A.rho_part,
                            flush, A.hess_kl, A.dtk);
      else
        newmark_point<false>((size_t)ig, P, z + ff[0][0], z + ff[0][1], z + ff[0][2], z, z, z, A.F0, A.F1,
                             A.last != 0, nullptr, nullptr, true, nullptr, A.dtk);
    }
  }
}

template <int NF, int KMODE, int SP, bool NT = false, bool INT = false>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(NF == 2 ? 4 : 5))) void k_forces_reg_pp(
    ForceArgs A) {
  forces_reg_pp_body<NF, KMODE, SP, NT, INT>(A);
}

// Two fields hold 6 more registers of look-ahead: 4 waves per SIMD (128 VGPRs)
// keeps them out of scratch.
template <int NF, int KMODE, int SP>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(NF == 2 ? 4 : 5))) void k_forces_reg_pipe(
    ForceArgs A) {
  forces_reg_pipe_body<NF, KMODE, SP>(A);
}

// Kernel entries of the regular-mesh pass.  The isotropic one asks for five
// waves per SIMD (fits in 96 VGPRs without spills); the anisotropic one keeps
// the compiler's allocation (96 would spill the c_ij products).
template <int NF, bool ANISO, int KMODE, int SP>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(5))) void k_forces_reg(ForceArgs A) {
  forces_reg_body<NF, ANISO, KMODE, SP>(A);
}

template <int NF, bool ANISO, int KMODE, int SP>
__global__ __launch_bounds__(128) void k_forces_reg_aniso(ForceArgs A) {
  forces_reg_body<NF, ANISO, KMODE, SP>(A);
}

// Workgroups of a regular-pass launch: a multiple of 8 (one share per XCD);
// blocks_per_cu = 0: one workgroup per element.
int reg_grid(int elem_count, int blocks_per_cu) {
  int dev = 0, ncu = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int all = (elem_count + 7) & ~7;
  const int nb = blocks_per_cu > 0 ? std::min(all, (ncu * blocks_per_cu + 7) & ~7) : all;
  return std::max(8, nb);
}

template <int NF, bool ANISO, int KMODE, int SP>
void launch_reg(const ForceArgs& a, hipStream_t s, int blocks_per_cu, int variant = 1) {
  const int nb = reg_grid(a.elem_count, blocks_per_cu);
  if constexpr (ANISO) SEM_LAUNCH((k_forces_reg_aniso<NF, ANISO, KMODE, SP>), dim3(nb), dim3(128), 0, s, a);
  else if (variant == 2) SEM_LAUNCH((k_forces_reg_pipe<NF, KMODE, SP>), dim3(nb), dim3(128), 0, s, a);
  else if (variant == 3 && a.rm) SEM_LAUNCH((k_forces_reg_pp<NF, KMODE, SP, false, true>), dim3(nb), dim3(128), 0, s, a);
  else if (variant == 3) SEM_LAUNCH((k_forces_reg_pp<NF, KMODE, SP>), dim3(nb), dim3(128), 0, s, a);
  else if (variant == 4) SEM_LAUNCH((k_forces_reg_pp<NF, KMODE, SP, true>), dim3(nb), dim3(128), 0, s, a);
  else SEM_LAUNCH((k_forces_reg<NF, ANISO, KMODE, SP>), dim3(nb), dim3(128), 0, s, a);
}

template <int NF, bool ANISO, int KMODE, int SP>
void launch_t(const ForceArgs& a, hipStream_t s) {
  const int nb = a.nblocks_padded;
  if (a.floc0)
    SEM_LAUNCH((k_forces<NF, ANISO, KMODE, SP, 1>), dim3(nb), dim3(128), 0, s, a);
  else
    SEM_LAUNCH((k_forces<NF, ANISO, KMODE, SP, 0>), dim3(nb), dim3(128), 0, s, a);
}

}  // namespace

// Column-block traversal of a lattice mesh (element (ix,iy,iz) = ix + nex*(iy + ney*iz))
// for the pipelined regular pass with P workgroups per XCD.  The pass gives XCD x the
// traversal indices [x*chunk, (x+1)*chunk) and its P workgroups take P consecutive
// indices per iteration.  Elements are listed column by column (a column = the nez
// elements above one (ix,iy)), so each XCD owns whole columns; inside an XCD's share
// groups of P consecutive columns are visited layer by layer.  One iteration then
// covers a patch of P columns at one depth and the next one the patch above it: the
// face an element shares with the element below was read one iteration earlier, when
// it is still in the XCD's L2, where the element-order traversal reaches it
// nex*ney elements later.  Every element appears once.  When chunk is not a multiple
// of nez the shares would split columns, and the order is the element order.  The
// pass computes the same map by integer arithmetic (forces_reg_pp_body, elem_at).
std::vector<int> sem_column_block_order(int nex, int ney, int nez, int per_xcd) {
  const int ncol = nex * ney, n = ncol * nez;
  const int chunk = (n + 7) >> 3;
  if (chunk % nez != 0) {  // the shares would split columns: element order (as the pass keeps it)
    std::vector<int> id(n);
    for (int e = 0; e < n; ++e) id[e] = e;
    return id;
  }
  std::vector<int> cm(n);  // column-major list
  for (int c = 0, q = 0; c < ncol; ++c)
    for (int z = 0; z < nez; ++z) cm[q++] = c + ncol * z;
  std::vector<int> order(n);
  const int P = std::max(1, per_xcd);
  for (int x = 0; x < 8; ++x) {
    const int b = std::min(n, x * chunk), end = std::min(n, (x + 1) * chunk);
    if (b >= end) continue;
    // columns of this share in list order, and each element's group / depth / column
    const int c0 = cm[b] % ncol;
    std::vector<long long> key(end - b);
    std::vector<int> pos(end - b);
    for (int q = b; q < end; ++q) {
      const int e = cm[q], c = e % ncol, z = e / ncol;
      const long long g = (c - c0) / P;
      key[q - b] = (g * nez + z) * (long long)ncol + c;
      pos[q - b] = q;
    }
    std::stable_sort(pos.begin(), pos.end(), [&](int u, int v) { return key[u - b] < key[v - b]; });
    for (int q = b; q < end; ++q) order[q] = cm[pos[q - b]];
  }
  return order;
}

extern "C" {
// The column-block element order of the pipelined regular pass (regular_path 5)
// for a nex x ney x nez lattice and per_xcd workgroups per XCD, 0-based
// elements in traversal order (host only; no counterpart in the reference).
void fwat_column_block_order_(int* nex, int* ney, int* nez, int* per_xcd, int* order) {
  if (*nex <= 0 || *ney <= 0 || *nez <= 0 || *per_xcd <= 0) SEM_FAIL("fwat_column_block_order_: bad arguments");
  const std::vector<int> o = sem_column_block_order(*nex, *ney, *nez, *per_xcd);
  std::copy(o.begin(), o.end(), order);
}
}  // extern "C"

bool sem_interior_active(const SemMesh* m, int nfields) {
  return m->interior_newmark && m->n_shared < m->nglob && m->nspec_irregular == 0 && !m->anisotropy &&
         (m->regular_fast_path == 3 || m->regular_fast_path == 5) && m->nslabs == 0 &&
         !sem_brick_active(m, nfields) && m->d_floc != nullptr;
}

void sem_launch_forces(SemMesh* m, int nfields, int kernel_mode, int elem_begin, int elem_count,
                       realw deltat, const int* elem_list, bool gather, const SemInterior* in) {
  if (elem_count <= 0) return;
  ForceArgs a{};
  a.ibool = m->d_ibool;
  a.irregular = m->d_irregular;
  a.jinv10 = m->d_jinv10;
  a.jinv10_lp = reinterpret_cast<const uint16_t*>(m->d_jinv10_lp);
  a.kappa = m->d_kappa;
  a.mu = m->d_mu;
  a.kmu_lp = reinterpret_cast<const uint16_t*>(m->d_kmu_lp);
  a.kmu_base = m->d_kmu_base;
  a.jinv10_base = m->d_jinv10_base;
  a.cij = m->d_cij;
  a.tables = m->d_tables;
  a.elem_list = elem_list;
  a.grouped = m->floc_grouped;
  a.elem_begin = elem_begin;
  a.elem_count = elem_count;
  a.nblocks_padded = xcd_pad(elem_count);
  a.xix_regular = m->xix_regular;
  a.jacobian_regular = m->jacobian_regular;
  a.scale_k = m->static_scale_kappa;
  a.scale_m = m->static_scale_mu;
  a.scale_jinv = m->static_scale_jinv;
  a.scale_jac = m->static_scale_jac;
  a.dt = (double)deltat;
  if (nfields == 2) {
    a.u0 = m->d_displ; a.a0 = m->d_accel;
    a.u1 = m->d_b_displ; a.a1 = m->d_b_accel;
  } else if (nfields == 1) {
    a.u0 = m->d_displ; a.a0 = m->d_accel;
  } else {  // nfields == 3: the reconstructed (b_) field alone
    a.u0 = m->d_b_displ; a.a0 = m->d_b_accel;
    a.eps0 = m->d_b_epsilon;
  }
  if (nfields != 3) a.eps0 = m->d_epsilon;
  if (gather) {
    if (!m->d_floc) SEM_FAIL("gather assembly buffers not allocated");
    if (elem_list) SEM_FAIL("gather assembly takes a contiguous element range");
    if (elem_begin < 0 || elem_begin + elem_count > m->nspec) SEM_FAIL("element range outside the mesh");
    a.floc0 = m->d_floc;
    a.floc1 = nfields == 2 ? m->d_floc_b : nullptr;
    if (nfields == 2 && !a.floc1) SEM_FAIL("second element-local force buffer not allocated");
  }
  if (kernel_mode == 1 && !a.eps0) SEM_FAIL("strain store requested but not allocated");
  if (a.u0 == nullptr || a.a0 == nullptr) SEM_FAIL("force pass on unallocated fields");
  if (m->static_precision != SEM_STATIC_FP32 && !m->d_kmu_lp) SEM_FAIL("reduced-precision statics not prepared");
  a.mu_kl = m->d_mu_kl;
  a.kappa_kl = m->d_kappa_kl;
  a.cijkl_kl = m->d_cijkl_kl;
  if (kernel_mode == 2 && m->kernel_block > 1) {
    a.mu_part = m->d_mu_part;
    a.kappa_part = m->d_kappa_part;
    a.cij_part = m->d_cij_part;
    if (m->anisotropic_kl ? !a.cij_part : (!a.mu_part || !a.kappa_part))
      SEM_FAIL("kernel partial sums not allocated (fwat_iterate_time_ allocates them)");
  }
  a.kflush = m->kernel_flush;
  const int sp = m->static_precision;
  const bool an = m->anisotropy != 0;
  hipStream_t s = m->stream;
  if (gather && m->nspec_irregular == 0 && m->regular_fast_path && kernel_mode != 1) {
    // regular_fast_path 2: the software-pipelined isotropic pass, 3: the same with two
    // barriers per element (ping-pong LDS buffers)
    int pipe = m->regular_fast_path;
    const int bpc = nfields == 2 ? m->blocks_per_cu_adj : m->blocks_per_cu_fwd;
    if (pipe == 5) {
      // 5: the ping-pong pass in column-block order on a lattice mesh over the whole mesh,
      // when every XCD's share is whole columns (anything else keeps element order)
      pipe = 3;
      const int chunk = (elem_count + 7) >> 3;
      if (!an && m->lat_nex > 0 && elem_begin == 0 && elem_count == m->nspec && chunk % m->lat_nez == 0) {
        a.col_ncol = m->lat_nex * m->lat_ney;
        a.col_nez = m->lat_nez;
      }
    }
    if (in && pipe == 3 && !an && kernel_mode != 1 && elem_begin == 0 && elem_count == m->nspec &&
        sem_interior_active(m, nfields)) {
      // the pass finishes the points [n_shared, nglob) (Newmark, density kernel)
      a.n_shared = m->n_shared;
      a.last = in->last ? 1 : 0;
      a.rm = m->d_rmass_inv;
      a.F0 = in->F0;
      a.F1 = in->F1;
      a.dtk = in->dtk;
      if (nfields == 2) {
        a.rho_kl = m->d_rho_kl_glob;
        a.rho_part = m->kernel_block > 1 ? m->d_rho_part : nullptr;
        a.hess_kl = m->d_hess_kl_glob;
        if (!a.rho_kl) SEM_FAIL("two-field pass needs the density kernel accumulator");
      }
    }
    if (nfields == 2) {
      if (an ? !m->d_cijkl_kl : (!m->d_mu_kl || !m->d_kappa_kl))
        SEM_FAIL("two-field force pass needs the %s kernel accumulators", an ? "cijkl" : "mu/kappa");
      if (an) launch_reg<2, true, 2, SEM_STATIC_FP32>(a, s, bpc);
      else if (sp == SEM_STATIC_BF16) launch_reg<2, false, 2, SEM_STATIC_BF16>(a, s, bpc, pipe);
      else if (sp == SEM_STATIC_FP16) launch_reg<2, false, 2, SEM_STATIC_FP16>(a, s, bpc, pipe);
      else launch_reg<2, false, 2, SEM_STATIC_FP32>(a, s, bpc, pipe);
    } else {
      if (an) launch_reg<1, true, 0, SEM_STATIC_FP32>(a, s, bpc);
      else if (sp == SEM_STATIC_BF16) launch_reg<1, false, 0, SEM_STATIC_BF16>(a, s, bpc, pipe);
      else if (sp == SEM_STATIC_FP16) launch_reg<1, false, 0, SEM_STATIC_FP16>(a, s, bpc, pipe);
      else launch_reg<1, false, 0, SEM_STATIC_FP32>(a, s, bpc, pipe);
    }
    HIP_CHECK(hipGetLastError());
    return;
  }
  if (nfields == 2) {
    if (kernel_mode != 2) SEM_FAIL("two-field force pass requires kernel mode 2");
    if (!m->d_b_displ || !m->d_b_accel) SEM_FAIL("two-field force pass needs the b_ fields");
    if (an ? !m->d_cijkl_kl : (!m->d_mu_kl || !m->d_kappa_kl))
      SEM_FAIL("two-field force pass needs the %s kernel accumulators", an ? "cijkl" : "mu/kappa");
    if (an) launch_t<2, true, 2, SEM_STATIC_FP32>(a, s);
    else if (sp == SEM_STATIC_BF16) launch_t<2, false, 2, SEM_STATIC_BF16>(a, s);
    else if (sp == SEM_STATIC_FP16) launch_t<2, false, 2, SEM_STATIC_FP16>(a, s);
    else launch_t<2, false, 2, SEM_STATIC_FP32>(a, s);
  } else {
    if (kernel_mode == 2) SEM_FAIL("kernel mode 2 needs both fields");
    if (an) {
      if (kernel_mode == 1) launch_t<1, true, 1, SEM_STATIC_FP32>(a, s);
      else launch_t<1, true, 0, SEM_STATIC_FP32>(a, s);
    } else if (sp == SEM_STATIC_BF16) {
      if (kernel_mode == 1) launch_t<1, false, 1, SEM_STATIC_BF16>(a, s);
      else launch_t<1, false, 0, SEM_STATIC_BF16>(a, s);
    } else if (sp == SEM_STATIC_FP16) {
      if (kernel_mode == 1) launch_t<1, false, 1, SEM_STATIC_FP16>(a, s);
      else launch_t<1, false, 0, SEM_STATIC_FP16>(a, s);
    } else {
      if (kernel_mode == 1) launch_t<1, false, 1, SEM_STATIC_FP32>(a, s);
      else launch_t<1, false, 0, SEM_STATIC_FP32>(a, s);
    }
  }
  HIP_CHECK(hipGetLastError());
}
