// PDE (explicit diffusion) kernel smoothing on the device.
//
// Drop-in for smooth_pde_cuda_ (src/postproc/smooth_sem_pde.cu:298-479,
// called from smooth_sem_pde, src/postproc/smooth.f90:392-402) with the same
// 42 by-reference arguments.  One diffusion step is
//   ddat = - sum_e  W (hprimewgll^T) [ J c(grad dat) ]      (element pass, scatter-add)
//   dat += rvol * ddat ;  ddat = 0                          (one streaming pass)
//   dat(PML points) = dat_bak                               (only when PML elements exist)
// for ntstep steps.  Geometry per element follows the reference: irregular
// elements read their nine inverse-Jacobian terms, regular elements use
// xix_regular; faithful mode (the default, as the reference's CUDA and CPU
// paths) puts xix_regular in all nine slots, the alternative uses the
// diagonal map of an axis-aligned cube.
//
// Differences in mechanism (not in result): the reference scatters with
// atomics into ddat, launches 4-5 kernels per diffusion step (memset, two
// phase launches, apply_mass, keep_pml) and exchanges halos over MPI; here
// the element pass writes element-local terms and one gather launch per step
// sums them per point in element order (the reference CPU loop's order, no
// atomics, run-to-run reproducible) fused with dat += rvol * ddat, the
// element phases are merged (single GPU), and up to 24 fields share one
// geometry read with the fields innermost in HBM ([nglob][fields]).
#include "sem_device.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

constexpr int SM_TPB = 256;
constexpr int SM_NFMAX = 24;  // fields per device run (a larger batch runs in chunks)

inline unsigned sm_blocks(size_t n) { return (unsigned)((n + SM_TPB - 1) / SM_TPB); }

struct SmoothArgs {
  const int* ibool;       // 0-based [nspec][125]
  const int* irregular;   // [nspec]
  const realw* jinv9;     // [nirr][9][125]
  const realw* jac;       // [nspec][125]  (jacobian_all)
  const realw* rot;       // [nspec][3][125] when spherical
  const realw* tables;    // hprimeT[25] hprimewgll[25] wgllwgll_yz[25] _xz[25] _xy[25]
  const realw* dat;       // [nglob][NFP]
  realw* dloc;            // [nspec][125][NFP] element-local -(weak-form stiffness term)
  int nspec, nfield;
  int nblocks_padded;
  realw cv, ch, xix_regular;
  int is_sph;
  int faithful;
};

// Element pass: one element per 128-lane workgroup; every lane loads the NFP
// field values of its point in one contiguous read, the fields are then
// processed one after the other out of LDS (geometry read once for all), and
// each lane writes its NFP element-local results in one contiguous write.
template <int NFP>
__global__ __launch_bounds__(128) void k_smooth_elem(SmoothArgs A) {
  __shared__ realw s_tab[128];
  __shared__ realw s_u[NFP][128];
  __shared__ realw s_t[3][128];
  const int t = threadIdx.x;
  if (t < 125) s_tab[t] = A.tables[t];
  const int e = xcd_remap(blockIdx.x, A.nblocks_padded);
  if (e >= A.nspec) return;  // block-uniform
  const bool act = t < 125;
  const int tt = act ? t : 124;
  const int I = tt % 5, J = (tt / 5) % 5, K = tt / 25;
  __syncthreads();
  // mxm_opt{x,y,z} of the reference: u(l,J,K) * hprimeT[I*5+l], etc.
  realw hpI[5], hpJ[5], hpK[5], hwI[5], hwJ[5], hwK[5];
#pragma unroll
  for (int l = 0; l < 5; ++l) {
    hpI[l] = s_tab[I * 5 + l];
    hpJ[l] = s_tab[J * 5 + l];
    hpK[l] = s_tab[K * 5 + l];
    hwI[l] = s_tab[25 + I * 5 + l];
    hwJ[l] = s_tab[25 + J * 5 + l];
    hwK[l] = s_tab[25 + K * 5 + l];
  }
  const realw fac1 = s_tab[50 + K * 5 + J];
  const realw fac2 = s_tab[75 + K * 5 + I];
  const realw fac3 = s_tab[100 + J * 5 + I];
  const size_t Lg = (size_t)e * 125 + tt;
  const int iglob = A.ibool[Lg];
  const int ir = A.irregular[e];
  realw xix, xiy, xiz, etx, ety, etz, gax, gay, gaz;
  if (ir) {
    const realw* G = A.jinv9 + (size_t)(ir - 1) * 1125 + tt;
    xix = G[0]; xiy = G[125]; xiz = G[250];
    etx = G[375]; ety = G[500]; etz = G[625];
    gax = G[750]; gay = G[875]; gaz = G[1000];
  } else if (A.faithful) {
    xix = xiy = xiz = etx = ety = etz = gax = gay = gaz = A.xix_regular;
  } else {
    xix = ety = gaz = A.xix_regular;
    xiy = xiz = etx = etz = gax = gay = 0.0f;
  }
  const realw jl = A.jac[Lg];
  realw rx = 0, ry = 0, rz = 0;
  if (A.is_sph) {
    rx = A.rot[(size_t)e * 375 + tt];
    ry = A.rot[(size_t)e * 375 + 125 + tt];
    rz = A.rot[(size_t)e * 375 + 250 + tt];
  }
  if (act) {
    const realw* D = A.dat + (size_t)iglob * NFP;
    if constexpr (NFP % 4 == 0) {
#pragma unroll
      for (int k = 0; k < NFP / 4; ++k) {
        const float4 v = reinterpret_cast<const float4*>(D)[k];
        s_u[4 * k][t] = v.x; s_u[4 * k + 1][t] = v.y; s_u[4 * k + 2][t] = v.z; s_u[4 * k + 3][t] = v.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < NFP; ++k) s_u[k][t] = D[k];
    }
  }
  __syncthreads();
  const realw cv = A.cv, ch = A.ch;
#pragma unroll 1
  for (int f = 0; f < A.nfield; ++f) {
    realw t1 = 0, t2 = 0, t3 = 0;
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      t1 += s_u[f][l + 5 * J + 25 * K] * hpI[l];
      t2 += s_u[f][I + 5 * l + 25 * K] * hpJ[l];
      t3 += s_u[f][I + 5 * J + 25 * l] * hpK[l];
    }
    const realw dx = t1 * xix + t2 * etx + t3 * gax;
    const realw dy = t1 * xiy + t2 * ety + t3 * gay;
    const realw dz = t1 * xiz + t2 * etz + t3 * gaz;
    realw f1, f2, f3;
    if (A.is_sph) {
      const realw rd = rx * dx + ry * dy + rz * dz;
      f1 = ((cv - ch) * (rx * xix + ry * xiy + rz * xiz) * rd + ch * (xix * dx + xiy * dy + xiz * dz)) * jl;
      f2 = ((cv - ch) * (rx * etx + ry * ety + rz * etz) * rd + ch * (etx * dx + ety * dy + etz * dz)) * jl;
      f3 = ((cv - ch) * (rx * gax + ry * gay + rz * gaz) * rd + ch * (gax * dx + gay * dy + gaz * dz)) * jl;
    } else {
      f1 = ((cv - ch) * xiz * dz + ch * (xix * dx + xiy * dy + xiz * dz)) * jl;
      f2 = ((cv - ch) * etz * dz + ch * (etx * dx + ety * dy + etz * dz)) * jl;
      f3 = ((cv - ch) * gaz * dz + ch * (gax * dx + gay * dy + gaz * dz)) * jl;
    }
    if (act) { s_t[0][t] = f1; s_t[1][t] = f2; s_t[2][t] = f3; }
    __syncthreads();  // fluxes complete; every lane is done reading s_u[f]
    realw n1 = 0, n2 = 0, n3 = 0;
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      n1 += s_t[0][l + 5 * J + 25 * K] * hwI[l];
      n2 += s_t[1][I + 5 * l + 25 * K] * hwJ[l];
      n3 += s_t[2][I + 5 * J + 25 * l] * hwK[l];
    }
    if (act) s_u[f][t] = -(fac1 * n1 + fac2 * n2 + fac3 * n3);  // own slot only
    __syncthreads();  // s_t free for the next field
  }
  if (act) {
    realw* O = A.dloc + Lg * NFP;
    if constexpr (NFP % 4 == 0) {
#pragma unroll
      for (int k = 0; k < NFP / 4; ++k)
        reinterpret_cast<float4*>(O)[k] = make_float4(s_u[4 * k][t], s_u[4 * k + 1][t], s_u[4 * k + 2][t], s_u[4 * k + 3][t]);
    } else {
#pragma unroll
      for (int k = 0; k < NFP; ++k) O[k] = s_u[k][t];
    }
  }
}

// ddat(g) = sum of the element-local terms of g in element order (the
// reference CPU loop's accumulation order); dat += rvol * ddat.  One lane per
// point carrying all NFP fields: each contribution is one contiguous
// NFP-float row (16-byte vectors), and the dat update is a contiguous
// read-modify-write of the point's row.
template <int NFP>
__global__ void k_smooth_gather(int nglob, const int* __restrict__ off, const int* __restrict__ idx,
                                const realw* __restrict__ dloc, const realw* __restrict__ rvol,
                                realw* __restrict__ dat) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nglob) return;
  realw d[NFP];
#pragma unroll
  for (int f = 0; f < NFP; ++f) d[f] = 0.0f;
  const int k1 = off[g + 1];
  for (int k = off[g]; k < k1; ++k) {
    const realw* P = dloc + (size_t)idx[k] * NFP;
    if constexpr (NFP % 4 == 0) {
#pragma unroll
      for (int q = 0; q < NFP / 4; ++q) {
        const float4 v = reinterpret_cast<const float4*>(P)[q];
        d[4 * q] += v.x; d[4 * q + 1] += v.y; d[4 * q + 2] += v.z; d[4 * q + 3] += v.w;
      }
    } else {
#pragma unroll
      for (int f = 0; f < NFP; ++f) d[f] += P[f];
    }
  }
  const realw r = rvol[g];
  realw* D = dat + (size_t)g * NFP;
  if constexpr (NFP % 4 == 0) {
#pragma unroll
    for (int q = 0; q < NFP / 4; ++q) {
      float4 v = reinterpret_cast<const float4*>(D)[q];
      v.x += r * d[4 * q]; v.y += r * d[4 * q + 1]; v.z += r * d[4 * q + 2]; v.w += r * d[4 * q + 3];
      reinterpret_cast<float4*>(D)[q] = v;
    }
  } else {
#pragma unroll
    for (int f = 0; f < NFP; ++f) D[f] += r * d[f];
  }
}

// pin PML points: dat(iglob) = dat_bak(local), in element/point order
__global__ void k_keep_pml(int npml, int nfield, int nfp, const int* __restrict__ pml_iglob,
                           const realw* __restrict__ pml_val, realw* __restrict__ dat) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npml) return;
  for (int f = 0; f < nfield; ++f) dat[(size_t)pml_iglob[q] * nfp + f] = pml_val[(size_t)f * npml + q];
}

// [nfield][nglob] <-> [nglob][nfp]
__global__ void k_to_point_major(int nglob, int nfield, int nfp, const realw* __restrict__ a, realw* __restrict__ b) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)nglob * nfp) return;
  const int g = (int)(q / nfp), f = (int)(q % nfp);
  b[q] = f < nfield ? a[(size_t)f * nglob + g] : 0.0f;
}

__global__ void k_to_field_major(int nglob, int nfield, int nfp, const realw* __restrict__ b, realw* __restrict__ a) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)nglob * nfield) return;
  const int f = (int)(q / nglob), g = (int)(q % nglob);
  a[q] = b[(size_t)g * nfp + f];
}

int g_smooth_faithful = 1;

// fields per point row in HBM: 1-3 as they are (rows of 4, 8 or 12 bytes),
// more padded to a multiple of 4 for 16-byte vector access
int padded_fields(int nf) { return nf <= 3 ? nf : ((nf + 3) / 4) * 4; }

struct SmoothPlan {
  int nspec = 0, nglob = 0, nirr = 0, npml = 0, nfield = 0, nfp = 1;
  int* d_ibool = nullptr;
  int* d_irr = nullptr;
  realw* d_jinv9 = nullptr;
  realw* d_jac = nullptr;
  realw* d_rot = nullptr;
  realw* d_tables = nullptr;
  realw* d_rvol = nullptr;
  int* d_off = nullptr;     // CSR point -> local points, element order
  int* d_idx = nullptr;
  realw* d_dat = nullptr;   // [nglob][nfp]
  realw* d_dloc = nullptr;  // [nloc][nfp]
  realw* d_io = nullptr;    // [nfield][nglob] staging of the caller's layout
  int* d_pml_iglob = nullptr;
  realw* d_pml_val = nullptr;
  hipStream_t stream = nullptr;
  double elem_ms = 0, update_ms = 0;
};

void plan_free(SmoothPlan& p) {
  dev_free(p.d_ibool); dev_free(p.d_irr); dev_free(p.d_jinv9); dev_free(p.d_jac); dev_free(p.d_rot);
  dev_free(p.d_tables); dev_free(p.d_rvol); dev_free(p.d_off); dev_free(p.d_idx); dev_free(p.d_dat);
  dev_free(p.d_dloc); dev_free(p.d_io); dev_free(p.d_pml_iglob); dev_free(p.d_pml_val);
  if (p.stream) HIP_CHECK(hipStreamDestroy(p.stream));
  p = SmoothPlan();
}

// elem_ms / update_ms of a timed run: the element and gather launches of every
// stride-th step (at most about 256 steps) timed by their dispatch timestamps,
// scaled to all nstep steps
template <int NFP>
void launch_steps(SmoothPlan& p, const SmoothArgs& a, int nstep, bool timed) {
  const int stride = std::max(1, nstep / 256), nsamp = (nstep + stride - 1) / stride;
  std::vector<hipEvent_t> ev;
  if (timed) {
    ev.resize(4 * (size_t)nsamp);
    for (auto& x : ev) HIP_CHECK(hipEventCreate(&x));
  }
  for (int it = 0; it < nstep; ++it) {
    const bool tm = timed && it % stride == 0;
    const size_t q = 4 * (size_t)(it / stride);
    if (tm) sem_arm_launch_events(ev[q], ev[q + 1]);
    SEM_LAUNCH(k_smooth_elem<NFP>, dim3(a.nblocks_padded), dim3(128), 0, p.stream, a);
    if (tm) sem_arm_launch_events(ev[q + 2], ev[q + 3]);
    SEM_LAUNCH(k_smooth_gather<NFP>, dim3(sm_blocks((size_t)p.nglob)), dim3(SM_TPB), 0, p.stream, p.nglob,
               p.d_off, p.d_idx, p.d_dloc, p.d_rvol, p.d_dat);
    if (p.npml > 0)
      hipLaunchKernelGGL(k_keep_pml, dim3((p.npml + SM_TPB - 1) / SM_TPB), dim3(SM_TPB), 0, p.stream, p.npml,
                         p.nfield, p.nfp, p.d_pml_iglob, p.d_pml_val, p.d_dat);
  }
  HIP_CHECK(hipGetLastError());
  if (timed) {
    HIP_CHECK(hipStreamSynchronize(p.stream));
    double e = 0.0, u = 0.0;
    for (int k = 0; k < nsamp; ++k) {
      float a1 = 0, a2 = 0;
      HIP_CHECK(hipEventElapsedTime(&a1, ev[4 * (size_t)k], ev[4 * (size_t)k + 1]));
      HIP_CHECK(hipEventElapsedTime(&a2, ev[4 * (size_t)k + 2], ev[4 * (size_t)k + 3]));
      e += a1;
      u += a2;
    }
    p.elem_ms += e * nstep / nsamp;
    p.update_ms += u * nstep / nsamp;
    for (auto x : ev) HIP_CHECK(hipEventDestroy(x));
  }
}

void run_steps(SmoothPlan& p, int nstep, realw cv, realw ch, realw xix_regular, int is_sph, bool timed) {
  SmoothArgs a{};
  a.ibool = p.d_ibool; a.irregular = p.d_irr; a.jinv9 = p.d_jinv9; a.jac = p.d_jac; a.rot = p.d_rot;
  a.tables = p.d_tables; a.dat = p.d_dat; a.dloc = p.d_dloc;
  a.nspec = p.nspec; a.nfield = p.nfield;
  a.nblocks_padded = xcd_pad(p.nspec);
  a.cv = cv; a.ch = ch; a.xix_regular = xix_regular; a.is_sph = is_sph; a.faithful = g_smooth_faithful;
  p.elem_ms = p.update_ms = 0;
  switch (p.nfp) {
    case 1: launch_steps<1>(p, a, nstep, timed); break;
    case 2: launch_steps<2>(p, a, nstep, timed); break;
    case 3: launch_steps<3>(p, a, nstep, timed); break;
    case 4: launch_steps<4>(p, a, nstep, timed); break;
    case 8: launch_steps<8>(p, a, nstep, timed); break;
    case 12: launch_steps<12>(p, a, nstep, timed); break;
    case 16: launch_steps<16>(p, a, nstep, timed); break;
    case 20: launch_steps<20>(p, a, nstep, timed); break;
    case 24: launch_steps<24>(p, a, nstep, timed); break;
    default: SEM_FAIL("smooth: %d padded fields not supported", p.nfp);
  }
}

// Build the device plan from the reference's host arrays.
void plan_build(SmoothPlan& p, int nspec, int nglob, int nirr, int is_sph, const realw* const* jin,
                const realw* jaco, const realw* rotate0, const realw* wxy, const realw* wxz,
                const realw* wyz, const realw* hprimeT, const realw* hprime_wgll, const int* ibool,
                const int* is_CPML, const int* irregular, const realw* rvol, int nfield) {
  if (nfield < 1 || nfield > SM_NFMAX) SEM_FAIL("smooth: %d fields per run (1..%d)", nfield, SM_NFMAX);
  HIP_CHECK(hipStreamCreateWithFlags(&p.stream, hipStreamNonBlocking));
  hipStream_t s = p.stream;
  p.nspec = nspec; p.nglob = nglob; p.nirr = nirr; p.nfield = nfield; p.nfp = padded_fields(nfield);
  const size_t nloc = (size_t)nspec * 125;
  std::vector<int> ib(nloc);
  for (size_t L = 0; L < nloc; ++L) {
    int g = ibool[L] - 1;
    if (g < 0 || g >= nglob) SEM_FAIL("smooth_pde: ibool(%zu)=%d outside 1..%d", L, g + 1, nglob);
    ib[L] = g;
  }
  p.d_ibool = dev_upload(ib.data(), nloc, s);
  // CSR point -> local points in increasing local index (element order)
  if (nloc >= (size_t)INT32_MAX) SEM_FAIL("smooth: too many local points for 32-bit CSR indices");
  {
    std::vector<int> off((size_t)nglob + 1, 0), idx(nloc);
    for (size_t L = 0; L < nloc; ++L) off[(size_t)ib[L] + 1]++;
    for (int g = 0; g < nglob; ++g) off[g + 1] += off[g];
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (size_t L = 0; L < nloc; ++L) idx[fill[ib[L]]++] = (int)L;
    p.d_off = dev_upload(off.data(), off.size(), s);
    p.d_idx = dev_upload(idx.data(), idx.size(), s);
  }
  std::vector<int> irr(irregular, irregular + nspec);
  for (int e = 0; e < nspec; ++e)
    if (irr[e] < 0 || irr[e] > nirr) SEM_FAIL("smooth_pde: irregular_element_number(%d)=%d outside 0..%d", e + 1, irr[e], nirr);
  p.d_irr = dev_upload(irr.data(), (size_t)nspec, s);
  if (nirr > 0) {
    std::vector<realw> j9((size_t)nirr * 1125);
    for (size_t e = 0; e < (size_t)nirr; ++e)
      for (int c = 0; c < 9; ++c) std::memcpy(&j9[(e * 9 + c) * 125], &jin[c][e * 125], 125 * sizeof(realw));
    p.d_jinv9 = dev_upload(j9.data(), j9.size(), s);
  }
  p.d_jac = dev_upload(jaco, nloc, s);
  if (is_sph) {
    // rotate_r(NDIM,NGLLX,NGLLY,NGLLZ,NSPEC) -> [nspec][3][125]
    std::vector<realw> r(nloc * 3);
    for (size_t e = 0; e < (size_t)nspec; ++e)
      for (int c = 0; c < 3; ++c)
        for (int L = 0; L < 125; ++L) r[(e * 3 + c) * 125 + L] = rotate0[((e * 125) + L) * 3 + c];
    p.d_rot = dev_upload(r.data(), r.size(), s);
  }
  realw tab[125];
  for (int q = 0; q < 25; ++q) {
    tab[q] = hprimeT[q];
    tab[25 + q] = hprime_wgll[q];
    tab[50 + q] = wyz[q];
    tab[75 + q] = wxz[q];
    tab[100 + q] = wxy[q];
  }
  p.d_tables = dev_upload(tab, 125, s);
  p.d_rvol = dev_upload(rvol, (size_t)nglob, s);
  p.d_dat = dev_alloc<realw>((size_t)nglob * p.nfp);
  p.d_dloc = dev_alloc<realw>(nloc * p.nfp);
  p.d_io = dev_alloc<realw>((size_t)nglob * nfield);
  // PML points in element order (last writer of a shared point wins, as the
  // reference's sequential CPU loop)
  std::vector<int> piglob;
  if (is_CPML) {
    std::vector<int> last(nglob, -1);
    for (int e = 0; e < nspec; ++e)
      if (is_CPML[e])
        for (int L = 0; L < 125; ++L) last[ib[(size_t)e * 125 + L]] = e * 125 + L;
    for (int g = 0; g < nglob; ++g)
      if (last[g] >= 0) piglob.push_back(g);
  }
  p.npml = (int)piglob.size();
  if (p.npml) {
    p.d_pml_iglob = dev_upload(piglob.data(), piglob.size(), s);
    std::vector<realw> zero((size_t)nfield * p.npml, 0.0f);
    p.d_pml_val = dev_upload(zero.data(), zero.size(), s);
  }
  HIP_CHECK(hipStreamSynchronize(s));
}

// caller layout (NGLOB_AB, nfield) <-> device layout [nglob][nfp]
void upload_fields(SmoothPlan& p, const realw* dat_glob) {
  const size_t n = (size_t)p.nglob * p.nfield;
  HIP_CHECK(hipMemcpyAsync(p.d_io, dat_glob, n * sizeof(realw), hipMemcpyHostToDevice, p.stream));
  hipLaunchKernelGGL(k_to_point_major, dim3(sm_blocks((size_t)p.nglob * p.nfp)), dim3(SM_TPB), 0, p.stream, p.nglob,
                     p.nfield, p.nfp, p.d_io, p.d_dat);
  HIP_CHECK(hipGetLastError());
}

void download_fields(SmoothPlan& p, realw* dat_glob) {
  const size_t n = (size_t)p.nglob * p.nfield;
  hipLaunchKernelGGL(k_to_field_major, dim3(sm_blocks(n)), dim3(SM_TPB), 0, p.stream, p.nglob, p.nfield, p.nfp,
                     p.d_dat, p.d_io);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(dat_glob, p.d_io, n * sizeof(realw), hipMemcpyDeviceToHost, p.stream));
  HIP_CHECK(hipStreamSynchronize(p.stream));
}

double g_last_elem_ms = 0.0, g_last_update_ms = 0.0;

}  // namespace

extern "C" {

// smooth_sem_pde.cu:298-316 signature.  dat_bak is the ZERO_PML pinned value
// (all zero in the reference, smooth.f90:21,184-188).
void smooth_pde_cuda_(int* h_is_sph, int* h_nspec, int* h_nglob, int* nprocs, int* h_nstep,
                      int* nspec_max, int* phase_ispec_inner_elastic, int* h_nspec_outer,
                      int* h_nspec_inner, const realw* xix, const realw* xiy, const realw* xiz,
                      const realw* etax, const realw* etay, const realw* etaz, const realw* gamx,
                      const realw* gamy, const realw* gamz, const realw* jaco, const realw* rotate0,
                      const realw* wgllwgll_xy, const realw* wgllwgll_xz, const realw* wgllwgll_yz,
                      const realw* hprimeT, const realw* hprime_wgll, const int* ibool,
                      const int* is_CPML, const realw* h_cv, const realw* h_ch, int* h_num_intfs,
                      int* h_max_nibool, const int* my_neighbors, const int* nibool_intf,
                      const int* ibool_intf, const realw* dat_bak, realw* dat, const realw* rvol,
                      realw* dat_glob, realw* ddat_glob, const int* irregular_element_number,
                      const realw* h_xix_regular, const int* h_nspec_irregular) {
  (void)nspec_max; (void)phase_ispec_inner_elastic; (void)h_num_intfs; (void)h_max_nibool;
  (void)my_neighbors; (void)nibool_intf; (void)ibool_intf; (void)ddat_glob;
  if (*nprocs != 1)
    SEM_FAIL("smooth_pde_cuda_: multi-rank halo assembly is not part of this single-GPU build (nprocs=%d)", *nprocs);
  if (*h_nspec_outer + *h_nspec_inner != *h_nspec)
    SEM_FAIL("smooth_pde_cuda_: phases do not cover all elements");
  const int nspec = *h_nspec, nglob = *h_nglob;
  const realw* jin[9] = {xix, xiy, xiz, etax, etay, etaz, gamx, gamy, gamz};
  SmoothPlan p;
  plan_build(p, nspec, nglob, *h_nspec_irregular, *h_is_sph, jin, jaco, rotate0, wgllwgll_xy,
             wgllwgll_xz, wgllwgll_yz, hprimeT, hprime_wgll, ibool, is_CPML,
             irregular_element_number, rvol, 1);
  if (p.npml) {
    // pinned values from dat_bak at the last-writer local point
    std::vector<int> last(nglob, -1);
    for (int e = 0; e < nspec; ++e)
      if (is_CPML[e])
        for (int L = 0; L < 125; ++L) last[ibool[(size_t)e * 125 + L] - 1] = e * 125 + L;
    std::vector<realw> v;
    for (int g = 0; g < nglob; ++g)
      if (last[g] >= 0) v.push_back(dat_bak[last[g]]);
    HIP_CHECK(hipMemcpyAsync(p.d_pml_val, v.data(), v.size() * sizeof(realw), hipMemcpyHostToDevice, p.stream));
  }
  upload_fields(p, dat_glob);
  run_steps(p, *h_nstep, *h_cv, *h_ch, *h_xix_regular, *h_is_sph, false);
  download_fields(p, dat_glob);
  for (size_t L = 0; L < (size_t)nspec * 125; ++L) dat[L] = dat_glob[ibool[L] - 1];
  plan_free(p);
}

// 1 = reference arithmetic on regular elements (xix_regular in all nine
// inverse-Jacobian slots, smooth_sem_pde.cu:161-171), 0 = diagonal map.
void fwat_smooth_set_faithful_(int* faithful) { g_smooth_faithful = *faithful != 0; }

// Several fields smoothed together, sharing one geometry read per element:
// dat_glob (NGLOB_AB, nfield) in/out, already projected and rvol-scaled as
// smooth_sem_pde does before calling smooth_pde_cuda_.  No PML, Cartesian or
// spherical.  timed = 1 records the element-pass time (fwat_smooth_last_time_).
void fwat_smooth_pde_batch_(int* h_is_sph, int* h_nspec, int* h_nglob, int* h_nstep, int* h_nfield,
                            const realw* xix, const realw* xiy, const realw* xiz, const realw* etax,
                            const realw* etay, const realw* etaz, const realw* gamx, const realw* gamy,
                            const realw* gamz, const realw* jaco, const realw* rotate0,
                            const realw* wgllwgll_xy, const realw* wgllwgll_xz, const realw* wgllwgll_yz,
                            const realw* hprimeT, const realw* hprime_wgll, const int* ibool,
                            const realw* h_cv, const realw* h_ch, const realw* rvol, realw* dat_glob,
                            const int* irregular_element_number, const realw* h_xix_regular,
                            const int* h_nspec_irregular, int* timed) {
  const int nspec = *h_nspec, nglob = *h_nglob, nf = *h_nfield;
  if (nf <= 0) SEM_FAIL("fwat_smooth_pde_batch_: nfield must be positive");
  const realw* jin[9] = {xix, xiy, xiz, etax, etay, etaz, gamx, gamy, gamz};
  g_last_elem_ms = g_last_update_ms = 0;
  for (int f0 = 0; f0 < nf; f0 += SM_NFMAX) {  // chunks of at most SM_NFMAX fields
    const int nc = std::min(SM_NFMAX, nf - f0);
    SmoothPlan p;
    plan_build(p, nspec, nglob, *h_nspec_irregular, *h_is_sph, jin, jaco, rotate0, wgllwgll_xy,
               wgllwgll_xz, wgllwgll_yz, hprimeT, hprime_wgll, ibool, nullptr,
               irregular_element_number, rvol, nc);
    realw* chunk = dat_glob + (size_t)f0 * nglob;
    upload_fields(p, chunk);
    HIP_CHECK(hipStreamSynchronize(p.stream));
    run_steps(p, *h_nstep, *h_cv, *h_ch, *h_xix_regular, *h_is_sph, *timed != 0);
    download_fields(p, chunk);
    g_last_elem_ms += p.elem_ms;
    g_last_update_ms += p.update_ms;
    plan_free(p);
  }
}

void fwat_smooth_last_time_(double* elem_ms, double* update_ms) {
  *elem_ms = g_last_elem_ms;
  *update_ms = g_last_update_ms;
}

}  // extern "C"
