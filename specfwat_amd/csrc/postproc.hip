// Kernel post-processing on the device: event sums of the sensitivity
// kernels, the diagonal-Hessian preconditioner, projection of GLL fields onto
// the regular inversion grid and the grid taper (SpecFWAT's xfwat_post_proc,
// src/postproc/fwat_post_proc.f90:22-47 driving post_processing.f90).
//
// HBM layout:
//   post sums      fp64 [nker][nspec][125]   (alpha, beta, rhop[, gcp, gsp])
//   Hessian sum    fp64 [nspec][125]
//   HTI model      fp32 [6][nspec][125]      c44, c45, c55, rho_vp, rho_vs, rho
//   grid owner     int32 [nz][ny][nx]        element claiming the grid point, INT_MAX if none
//   grid ref       fp64 [nz][ny][nx][3]      reference coordinates in that element
// Event kernels enter either straight from a finished adjoint run's device
// arrays (no host round trip) or as the single-precision arrays of the
// EKERNEL files.  Every pass is one lane per point, coalesced, HBM-bound.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

#include "sem_device.h"

namespace {

constexpr int PTPB = 256;
inline unsigned pblk(size_t n) { return (unsigned)((n + PTPB - 1) / PTPB); }

enum { POST_ISO = 1, POST_HTI = 2 };

struct PostState {
  int nspec = 0;
  size_t nloc = 0;
  int model_type = POST_ISO;
  int nker = 3;
  int use_rho_scaling = 0;
  hipStream_t stream = nullptr;
  double* d_ker = nullptr;
  double* d_hess = nullptr;
  float* d_model = nullptr;
  float* d_stage = nullptr;
  size_t stage_n = 0;
  unsigned long long* d_max = nullptr;
  int nevents = 0, nhess = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;  // device time of the last add / invert / apply call
};

void post_begin(PostState* p) { HIP_CHECK(hipEventRecord(p->ev0, p->stream)); }

void post_end(PostState* p) {
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipEventRecord(p->ev1, p->stream));
  HIP_CHECK(hipStreamSynchronize(p->stream));
  float ms = 0.0f;
  HIP_CHECK(hipEventElapsedTime(&ms, p->ev0, p->ev1));
  p->last_ms = ms;
}

PostState* post_state(long* pp) {
  PostState* p = reinterpret_cast<PostState*>(*pp);
  if (!p) SEM_FAIL("post-processing handle is null (fwat_post_create_ not called?)");
  return p;
}

// ------------------------------------------------------------ event sums

// save_adjoint_kernels scaling and the alpha/beta/rhop combination of one
// isotropic event, straight from the adjoint run's accumulators
__global__ void k_post_add_iso_run(size_t nloc, const int* __restrict__ ibool, const double* __restrict__ rho_glob,
                                   const double* __restrict__ mu_kl, const double* __restrict__ kappa_kl,
                                   const realw* __restrict__ rho, const realw* __restrict__ kappa,
                                   const realw* __restrict__ mu, const double* __restrict__ hess_glob,
                                   int hess_abs, int use_rhop, double* __restrict__ ker, double* __restrict__ hess) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nloc) return;
  const int g = ibool[i];
  const double k = (double)kappa[i], m = (double)mu[i];
  const double rs = -(double)rho[i] * rho_glob[g];
  const double ms = -2.0 * m * mu_kl[i];
  const double ks = -k * kappa_kl[i];
  const double r = 4.0 * m / (3.0 * k);
  ker[i] += 2.0 * (1.0 + r) * ks;
  ker[nloc + i] += 2.0 * (ms - r * ks);
  if (use_rhop) ker[2 * nloc + i] += rs + ks + ms;
  if (hess_glob) {
    const double h = hess_glob[g];
    hess[i] += hess_abs ? fabs(h) : h;
  }
}

// kernel_cijkl2hti: alpha, beta, rho (relative) and gcp, gsp kernels from the
// rho kernel and the c11 c12 c13 c22 c23 c33 c44 c45 c55 c66 kernels K[10];
// the model coefficients are formed in single precision as the reference does
__device__ __forceinline__ void hti5(const float* __restrict__ mdl, size_t nloc, size_t i, double rk,
                                     const double* K, double* out) {
  const float c44 = mdl[i], c45 = mdl[nloc + i], c55 = mdl[2 * nloc + i];
  const float rvp = mdl[3 * nloc + i], rvs = mdl[4 * nloc + i], rho = mdl[5 * nloc + i];
  const float L = 0.5f * (c44 + c55), gc = 0.5f * (c44 - c55);
  const float gcp = gc / L, gsp = c45 / L;
  const float vp = rvp / rho, vs = rvs / rho;
  const double a2 = 2.0 * (double)rvp, b2 = 2.0 * (double)rvs, b4 = -4.0 * (double)rvs;
  const double gcb = 2.0 * (double)gcp * (double)rvs, gsb = 2.0 * (double)gsp * (double)rvs;
  const double vp2 = (double)vp * vp, vs2 = (double)vs * vs, d = vp2 - 2.0 * vs2;
  const double a = (K[0] + K[1] + K[2] + K[3] + K[4] + K[5]) * a2;
  const double b = (K[1] + K[2] + K[4]) * b4 + K[6] * (gcb + b2) + K[7] * gsb + K[8] * (b2 - gcb) + K[9] * b2;
  const double r = rk + (K[0] + K[3] + K[5]) * vp2 + (K[1] + K[2] + K[4]) * d + K[6] * ((double)gcp * vs2 + vs2) +
                   K[7] * ((double)gsp * vs2) + K[8] * (vs2 - (double)gcp * vs2) + K[9] * vs2;
  const double lv = (double)rvs * (double)vs;
  out[0] = a * (double)vp;
  out[1] = b * (double)vs;
  out[2] = r * (double)rho;
  out[3] = (K[6] - K[8]) * lv;
  out[4] = K[7] * lv;
}

// Voigt positions of c11 c12 c13 c22 c23 c33 c44 c45 c55 c66 in the 21-term order
__constant__ int c_hti_q[10] = {0, 1, 2, 6, 7, 11, 15, 16, 18, 20};

__global__ void k_post_add_hti_run(size_t nloc, const int* __restrict__ ibool, const double* __restrict__ rho_glob,
                                   const realw* __restrict__ rho_run, const double* __restrict__ cijkl_kl,
                                   const float* __restrict__ mdl, const double* __restrict__ hess_glob,
                                   int hess_abs, double* __restrict__ ker, double* __restrict__ hess) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nloc) return;
  const size_t e = i / 125, L = i - e * 125;
  double K[10];
#pragma unroll
  for (int q = 0; q < 10; ++q) K[q] = -cijkl_kl[(e * 21 + c_hti_q[q]) * 125 + L];
  const int g = ibool[i];
  double o[5];
  hti5(mdl, nloc, i, -(double)rho_run[i] * rho_glob[g], K, o);
#pragma unroll
  for (int q = 0; q < 5; ++q) ker[q * nloc + i] += o[q];
  if (hess_glob) {
    const double h = hess_glob[g];
    hess[i] += hess_abs ? fabs(h) : h;
  }
}

// event kernels as stored in the EKERNEL files (single precision)
//   iso: [3][nloc] alpha, beta, rhop    HTI: [11][nloc] rho, c11 ... c66
__global__ void k_post_add_files(size_t nloc, int model_type, int use_rhop, const float* __restrict__ in,
                                 const float* __restrict__ mdl, const float* __restrict__ hin, int hess_abs,
                                 double* __restrict__ ker, double* __restrict__ hess) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nloc) return;
  if (model_type == POST_ISO) {
    ker[i] += (double)in[i];
    ker[nloc + i] += (double)in[nloc + i];
    if (use_rhop) ker[2 * nloc + i] += (double)in[2 * nloc + i];
  } else {
    double K[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) K[q] = (double)in[(q + 1) * nloc + i];
    double o[5];
    hti5(mdl, nloc, i, (double)in[i], K, o);
#pragma unroll
    for (int q = 0; q < 5; ++q) ker[q * nloc + i] += o[q];
  }
  if (hin) {
    const double h = (double)hin[i];
    hess[i] += hess_abs ? fabs(h) : h;
  }
}

// ------------------------------------------------------------ preconditioner

__global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = v;
}

// max |x| over n values; non-negative doubles order like their bit patterns,
// so the cross-workgroup step is one integer atomicMax per workgroup
__global__ void __launch_bounds__(PTPB) k_maxabs(size_t n, const double* __restrict__ x,
                                                 unsigned long long* __restrict__ out) {
  __shared__ double s[PTPB / 64];
  double m = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    m = fmax(m, fabs(x[i]));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < PTPB / 64; ++w) m = fmax(m, s[w]);
    atomicMax(out, (unsigned long long)__double_as_longlong(m));
  }
}

// invert_hess, first half: normalise by max|H| (all ones if it is below
// 1e-18), invert above the 1e-3 threshold
__global__ void k_hess_invert(size_t n, double* __restrict__ h, const unsigned long long* __restrict__ mx,
                              int take_abs) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double m = __longlong_as_double((long long)mx[0]);
  double v = take_abs ? fabs(h[i]) : h[i];
  v = m < 1e-18 ? 1.0 : v / m;
  h[i] = fabs(v) > 1e-3 ? 1.0 / v : 1.0 / 1e-3;
}

__global__ void k_scale_by_inv_max(size_t n, double* __restrict__ h, const unsigned long long* __restrict__ mx) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  h[i] = h[i] / __longlong_as_double((long long)mx[0]);
}

__global__ void k_apply_precond(size_t nloc, int nker, double* __restrict__ ker, const double* __restrict__ h) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nloc) return;
  const double w = h[i];
  for (int q = 0; q < nker; ++q) ker[q * nloc + i] *= w;
}

void device_maxabs(PostState* p, const double* x, size_t n) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, p->stream, p->d_max, 1, 0ull);
  const unsigned nb = (unsigned)std::min<size_t>(pblk(n), 2048);
  hipLaunchKernelGGL(k_maxabs, dim3(nb), dim3(PTPB), 0, p->stream, n, x, p->d_max);
}

// ------------------------------------------------------------ SEM -> grid

struct GridState {
  int nspec = 0, nglob = 0;
  int nx = 0, ny = 0, nz = 0;
  float ox = 0, oy = 0, oz = 0, hx = 1, hy = 1, hz = 1;
  int num_iter = 5;
  long nfound = 0;
  hipStream_t stream = nullptr;
  int* d_ibool = nullptr;
  float* d_xyz = nullptr;  // [3][nglob]
  int* d_owner = nullptr;
  double* d_ref = nullptr;
  int* d_count = nullptr;
  float* d_in = nullptr;
  float* d_out = nullptr;
  size_t in_n = 0, out_n = 0;
  double last_ms = 0.0;
};

GridState* grid_state(long* gp) {
  GridState* g = reinterpret_cast<GridState*>(*gp);
  if (!g) SEM_FAIL("grid projection handle is null (fwat_gll2grid_prepare_ not called?)");
  return g;
}

__constant__ double c_xg[5];

#pragma clang fp contract(off)
// 8-node hexahedron: anchors at GLL indices 0/4, SPECFEM's anchor order
__device__ void hex8_map(const double* xe, const double* ye, const double* ze, double xi, double eta, double gam,
                         double* p, double* J) {
  const double ra1 = 1.0 + xi, ra2 = 1.0 - xi, rb1 = 1.0 + eta, rb2 = 1.0 - eta;
  const double rc1 = 1.0 + gam, rc2 = 1.0 - gam;
  const double e8 = 0.125;
  const double sh[8] = {e8 * ra2 * rb2 * rc2, e8 * ra1 * rb2 * rc2, e8 * ra1 * rb1 * rc2, e8 * ra2 * rb1 * rc2,
                        e8 * ra2 * rb2 * rc1, e8 * ra1 * rb2 * rc1, e8 * ra1 * rb1 * rc1, e8 * ra2 * rb1 * rc1};
  const double dxi[8] = {-e8 * rb2 * rc2, e8 * rb2 * rc2, e8 * rb1 * rc2, -e8 * rb1 * rc2,
                         -e8 * rb2 * rc1, e8 * rb2 * rc1, e8 * rb1 * rc1, -e8 * rb1 * rc1};
  const double deta[8] = {-e8 * ra2 * rc2, -e8 * ra1 * rc2, e8 * ra1 * rc2, e8 * ra2 * rc2,
                          -e8 * ra2 * rc1, -e8 * ra1 * rc1, e8 * ra1 * rc1, e8 * ra2 * rc1};
  const double dgam[8] = {-e8 * ra2 * rb2, -e8 * ra1 * rb2, -e8 * ra1 * rb1, -e8 * ra2 * rb1,
                          e8 * ra2 * rb2, e8 * ra1 * rb2, e8 * ra1 * rb1, e8 * ra2 * rb1};
  double x = 0, y = 0, z = 0, xxi = 0, xeta = 0, xgam = 0, yxi = 0, yeta = 0, ygam = 0, zxi = 0, zeta = 0,
         zgam = 0;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    x = x + sh[a] * xe[a]; y = y + sh[a] * ye[a]; z = z + sh[a] * ze[a];
    xxi = xxi + dxi[a] * xe[a]; xeta = xeta + deta[a] * xe[a]; xgam = xgam + dgam[a] * xe[a];
    yxi = yxi + dxi[a] * ye[a]; yeta = yeta + deta[a] * ye[a]; ygam = ygam + dgam[a] * ye[a];
    zxi = zxi + dxi[a] * ze[a]; zeta = zeta + deta[a] * ze[a]; zgam = zgam + dgam[a] * ze[a];
  }
  const double jac = xxi * (yeta * zgam - ygam * zeta) - xeta * (yxi * zgam - ygam * zxi) +
                     xgam * (yxi * zeta - yeta * zxi);
  p[0] = x; p[1] = y; p[2] = z;
  J[0] = (yeta * zgam - ygam * zeta) / jac;
  J[1] = (xgam * zeta - xeta * zgam) / jac;
  J[2] = (xeta * ygam - xgam * yeta) / jac;
  J[3] = (ygam * zxi - yxi * zgam) / jac;
  J[4] = (xxi * zgam - xgam * zxi) / jac;
  J[5] = (xgam * yxi - xxi * ygam) / jac;
  J[6] = (yxi * zeta - yeta * zxi) / jac;
  J[7] = (xeta * zxi - xxi * zeta) / jac;
  J[8] = (xxi * yeta - xeta * yxi) / jac;
}

// locate_point_in_element: first guess at the closest interior GLL point,
// num_iter Newton steps, each clamped to the reference cube
__device__ void locate_in_element(const int* __restrict__ ib, const float* __restrict__ xs,
                                  const float* __restrict__ ys, const float* __restrict__ zs, double xt, double yt,
                                  double zt, int num_iter, double* ref) {
  double dmin = 1e30;
  int gi = 0, gj = 0, gk = 0;
  for (int k = 1; k < 4; ++k)
    for (int j = 1; j < 4; ++j)
      for (int i = 1; i < 4; ++i) {
        const int g = ib[i + 5 * j + 25 * k];
        const double dx = xt - (double)xs[g], dy = yt - (double)ys[g], dz = zt - (double)zs[g];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < dmin) { dmin = d2; gi = i; gj = j; gk = k; }
      }
  double xi = c_xg[gi], eta = c_xg[gj], gam = c_xg[gk];
  double xe[8], ye[8], ze[8];
  const int corner[8] = {0, 4, 24, 20, 100, 104, 124, 120};
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int g = ib[corner[a]];
    xe[a] = (double)xs[g]; ye[a] = (double)ys[g]; ze[a] = (double)zs[g];
  }
  double p[3], J[9];
  for (int it = 0; it < num_iter; ++it) {
    hex8_map(xe, ye, ze, xi, eta, gam, p, J);
    const double dx = -(p[0] - xt), dy = -(p[1] - yt), dz = -(p[2] - zt);
    xi = xi + (J[0] * dx + J[1] * dy + J[2] * dz);
    eta = eta + (J[3] * dx + J[4] * dy + J[5] * dz);
    gam = gam + (J[6] * dx + J[7] * dy + J[8] * dz);
    xi = fmin(fmax(xi, -1.0), 1.0);
    eta = fmin(fmax(eta, -1.0), 1.0);
    gam = fmin(fmax(gam, -1.0), 1.0);
  }
  ref[0] = xi; ref[1] = eta; ref[2] = gam;
}

__device__ void lagrange5(double x, double* h) {
#pragma unroll
  for (int d = 0; d < 5; ++d) {
    double p1 = 1.0, p2 = 1.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      if (i == d) continue;
      p1 = p1 * (x - c_xg[i]);
      p2 = p2 * (c_xg[d] - c_xg[i]);
    }
    h[d] = p1 / p2;
  }
}

// One workgroup per element: its bounding box in grid-index space (single
// precision, truncated like the Fortran integer conversion), then every grid
// point in that box is located in the element; the lowest element index that
// accepts a point owns it (the reference keeps the first one in element order)
__global__ void __launch_bounds__(128) k_grid_owner(int nspec, const int* __restrict__ ibool,
                                                    const float* __restrict__ xyz, int nglob, float ox, float oy,
                                                    float oz, float hx, float hy, float hz, int nx, int ny, int nz,
                                                    int num_iter, int* __restrict__ owner) {
  const int e = blockIdx.x;
  if (e >= nspec) return;
  const int* ib = ibool + (size_t)e * 125;
  const float* xs = xyz;
  const float* ys = xyz + nglob;
  const float* zs = xyz + 2 * (size_t)nglob;
  __shared__ float s_lo[3][2], s_hi[3][2];
  float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
  if (threadIdx.x < 125) {
    const int g = ib[threadIdx.x];
    const float v[3] = {xs[g], ys[g], zs[g]};
    for (int c = 0; c < 3; ++c) { lo[c] = v[c]; hi[c] = v[c]; }
  }
  for (int c = 0; c < 3; ++c)
    for (int o = 32; o > 0; o >>= 1) {
      lo[c] = fminf(lo[c], __shfl_xor(lo[c], o, 64));
      hi[c] = fmaxf(hi[c], __shfl_xor(hi[c], o, 64));
    }
  if ((threadIdx.x & 63) == 0)
    for (int c = 0; c < 3; ++c) { s_lo[c][threadIdx.x >> 6] = lo[c]; s_hi[c][threadIdx.x >> 6] = hi[c]; }
  __syncthreads();
  float bl[3], bh[3];
  for (int c = 0; c < 3; ++c) { bl[c] = fminf(s_lo[c][0], s_lo[c][1]); bh[c] = fmaxf(s_hi[c][0], s_hi[c][1]); }
  // values are clamped to [-1, n + 1] before the conversion; indices outside
  // 1..n are skipped either way
  auto idx = [](float t, int n) { return (int)fminf(fmaxf(t, -1.0f), (float)(n + 1)); };
  const int imin = idx(1.0f + (bl[0] - ox) / hx, nx), imax = idx(1.0f + (bh[0] - ox) / hx, nx);
  const int jmin = idx(1.0f + (bl[1] - oy) / hy, ny), jmax = idx(1.0f + (bh[1] - oy) / hy, ny);
  const int kmin = idx(1.0f + (bl[2] - oz) / hz, nz), kmax = idx(1.0f + (bh[2] - oz) / hz, nz);
  const int ni = imax - imin + 1, nj = jmax - jmin + 1, nk = kmax - kmin + 1;
  if (ni <= 0 || nj <= 0 || nk <= 0) return;
  const long ncand = (long)ni * nj * nk;
  for (long c = threadIdx.x; c < ncand; c += blockDim.x) {
    const int i_f = imin + (int)(c % ni), jf = jmin + (int)((c / ni) % nj), kf = kmin + (int)(c / ((long)ni * nj));
    if (i_f < 1 || i_f > nx || jf < 1 || jf > ny || kf < 1 || kf > nz) continue;
    const size_t p = (size_t)(i_f - 1) + (size_t)nx * ((size_t)(jf - 1) + (size_t)ny * (kf - 1));
    if (owner[p] <= e) continue;  // a lower element has it already
    const float xf = ox + (float)(i_f - 1) * hx, yf = oy + (float)(jf - 1) * hy, zf = oz + (float)(kf - 1) * hz;
    double r[3];
    locate_in_element(ib, xs, ys, zs, (double)xf, (double)yf, (double)zf, num_iter, r);
    if (fabs(r[0]) < 1.05 && fabs(r[1]) < 1.05 && fabs(r[2]) < 1.05) atomicMin(&owner[p], e);
  }
}

// reference coordinates of each owned grid point in its owner; count them
__global__ void k_grid_ref(int nx, int ny, int nz, const int* __restrict__ owner, const int* __restrict__ ibool,
                           const float* __restrict__ xyz, int nglob, float ox, float oy, float oz, float hx, float hy,
                           float hz, int num_iter, double* __restrict__ ref, int* __restrict__ count) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t npt = (size_t)nx * ny * nz;
  int own = 0;
  if (p < npt && owner[p] != INT_MAX) {
    own = 1;
    const int i_f = (int)(p % nx) + 1, jf = (int)((p / nx) % ny) + 1, kf = (int)(p / ((size_t)nx * ny)) + 1;
    const float xf = ox + (float)(i_f - 1) * hx, yf = oy + (float)(jf - 1) * hy, zf = oz + (float)(kf - 1) * hz;
    double r[3];
    locate_in_element(ibool + (size_t)owner[p] * 125, xyz, xyz + nglob, xyz + 2 * (size_t)nglob, (double)xf,
                      (double)yf, (double)zf, num_iter, r);
    ref[3 * p] = r[0]; ref[3 * p + 1] = r[1]; ref[3 * p + 2] = r[2];
  }
  for (int o = 32; o > 0; o >>= 1) own += __shfl_xor(own, o, 64);
  if ((threadIdx.x & 63) == 0 && own) atomicAdd(count, own);
}

// Project_model_SEM2FD_grid: Lagrange interpolation of the owner element's
// 125 values at the point's reference coordinates (double), stored single;
// points no element owns are 0
__global__ void k_grid_project(size_t npt, int nfield, size_t nloc, const int* __restrict__ owner,
                               const double* __restrict__ ref, const float* __restrict__ model,
                               float* __restrict__ out) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npt) return;
  const int e = owner[p];
  if (e == INT_MAX) {
    for (int f = 0; f < nfield; ++f) out[f * npt + p] = 0.0f;
    return;
  }
  double hxi[5], het[5], hga[5];
  lagrange5(ref[3 * p], hxi);
  lagrange5(ref[3 * p + 1], het);
  lagrange5(ref[3 * p + 2], hga);
  for (int f = 0; f < nfield; ++f) {
    const float* m = model + f * nloc + (size_t)e * 125;
    double v = 0.0;
    for (int k = 0; k < 5; ++k)
      for (int j = 0; j < 5; ++j)
#pragma unroll
        for (int i = 0; i < 5; ++i) v = v + hxi[i] * het[j] * hga[k] * (double)m[i + 5 * j + 25 * k];
    out[f * npt + p] = (float)v;
  }
}

// ------------------------------------------------------------ grid taper

// Hann half-windows of taper_3D_mod.f90 create_taper / coswin along one axis
__device__ float coswin_at(int n2, int opt, int i) {
  const float pi = 3.1415926535f;
  const int nn = 2 * n2 + 1;
  const int w = opt == 1 ? i + 1 : n2 + 2 + i;  // 1-based window index
  return 0.5f - 0.5f * cosf(2.0f * pi * (float)(w - 1) / (float)(nn - 1));
}

// lateral window value at 1-based taper-grid index t along an axis of n points
__device__ float lateral_win(int t, int n, int n1, int n2) {
  const int x1 = n1 + 1, x2 = n1 + n2 + 1, x3 = n - n1 - n2, x4 = n - n1;
  float v = 0.0f;
  if (t >= x1 && t <= x2 - 1) v = coswin_at(n2, 1, t - x1);
  if (t >= x3 + 1 && t <= x4) v = coswin_at(n2, 0, t - x3 - 1);
  if (t >= x2 && t <= x3) v = 1.0f;
  return v;
}

__device__ float bottom_win(int t, int n1, int n2, int nz) {
  const int z1 = n1 + n2 + 1;
  float v = 0.0f;
  if (t >= n1 + 1 && t <= z1 - 1) v = coswin_at(n2, 1, t - n1 - 1);
  if (t >= z1 && t <= nz) v = 1.0f;
  return v;
}

// locate_bissection on the regular taper coordinates c0 + i dh (utils.f90:1490-1520)
__device__ int bissect(float c0, float dh, int n, double x) {
  auto c = [&](int i1) { return (double)(c0 + (float)(i1 - 1) * dh); };
  if (x < c(1) || x > c(n)) return -1;
  int jl = 0, ju = n + 1;
  while (ju - jl > 1) {
    const int jm = (ju + jl) / 2;
    if (x >= c(jm)) jl = jm; else ju = jm;
  }
  if (x == c(1)) return 1;
  if (x == c(n)) return n - 1;
  return jl;
}

struct TaperAxis {
  float c0, dh;
  int n, n1, n2, bottom;
};

__device__ float taper_val(const TaperAxis& A, int t) {
  return A.bottom ? bottom_win(t, A.n1, A.n2, A.n) : lateral_win(t, A.n, A.n1, A.n2);
}

// taper_cls%Interp: interp3 of the separable taper grid (values formed as
// (xt * yt) * zt like the stored array) at one grid point
__device__ float taper_interp(const TaperAxis& X, const TaperAxis& Y, const TaperAxis& Z, float x, float y,
                              float z) {
  auto crd = [](const TaperAxis& A, int i1) { return A.c0 + (float)(i1 - 1) * A.dh; };
  if (x < crd(X, 1) || x > crd(X, X.n) || y < crd(Y, 1) || y > crd(Y, Y.n) || z < crd(Z, 1) || z > crd(Z, Z.n))
    return 0.0f;
  const int ix = bissect(X.c0, X.dh, X.n, (double)x), iy = bissect(Y.c0, Y.dh, Y.n, (double)y),
            iz = bissect(Z.c0, Z.dh, Z.n, (double)z);
  const int kx = min(max(ix, 1), X.n - 1), ky = min(max(iy, 1), Y.n - 1), kz = min(max(iz, 1), Z.n - 1);
  auto V = [&](int a, int b, int c) { return taper_val(X, a) * taper_val(Y, b) * taper_val(Z, c); };
  const float xix1 = x - crd(X, kx), yiy1 = y - crd(Y, ky), ziz1 = z - crd(Z, kz);
  const float xix2 = x - crd(X, kx + 1), yiy2 = y - crd(Y, ky + 1), ziz2 = z - crd(Z, kz + 1);
  const float x1x2 = 1.0f / (crd(X, kx) - crd(X, kx + 1)), y1y2 = 1.0f / (crd(Y, ky) - crd(Y, ky + 1)),
              z1z2 = 1.0f / (crd(Z, kz) - crd(Z, kz + 1));
  const float x2x1 = 1.0f / (crd(X, kx + 1) - crd(X, kx)), y2y1 = 1.0f / (crd(Y, ky + 1) - crd(Y, ky)),
              z2z1 = 1.0f / (crd(Z, kz + 1) - crd(Z, kz));
  const float f111 = xix2 * yiy2 * ziz2 * x1x2 * y1y2 * z1z2 * V(kx, ky, kz);
  const float f112 = xix2 * yiy2 * ziz1 * x1x2 * y1y2 * z2z1 * V(kx, ky, kz + 1);
  const float f121 = xix2 * yiy1 * ziz2 * x1x2 * y2y1 * z1z2 * V(kx, ky + 1, kz);
  const float f122 = xix2 * yiy1 * ziz1 * x1x2 * y2y1 * z2z1 * V(kx, ky + 1, kz + 1);
  const float f211 = xix1 * yiy2 * ziz2 * x2x1 * y1y2 * z1z2 * V(kx + 1, ky, kz);
  const float f212 = xix1 * yiy2 * ziz1 * x2x1 * y1y2 * z2z1 * V(kx + 1, ky, kz + 1);
  const float f221 = xix1 * yiy1 * ziz2 * x2x1 * y2y1 * z1z2 * V(kx + 1, ky + 1, kz);
  const float f222 = xix1 * yiy1 * ziz1 * x2x1 * y2y1 * z2z1 * V(kx + 1, ky + 1, kz + 1);
  return f111 + f112 + f121 + f122 + f211 + f212 + f221 + f222;
}
#pragma clang fp contract(fast)

__global__ void k_taper(int gnx, int gny, int gnz, float ox, float oy, float oz, float hx, float hy, float hz,
                        TaperAxis X, TaperAxis Y, TaperAxis Z, int nfield, float* __restrict__ data,
                        float* __restrict__ mask) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t npt = (size_t)gnx * gny * gnz;
  if (p >= npt) return;
  const int i = (int)(p % gnx), j = (int)((p / gnx) % gny), k = (int)(p / ((size_t)gnx * gny));
  const float t = taper_interp(X, Y, Z, ox + (float)i * hx, oy + (float)j * hy, oz + (float)k * hz);
  if (mask) mask[p] = t;
  for (int f = 0; f < nfield; ++f) data[f * npt + p] *= t;
}

}  // namespace

// ================================================================ C ABI
extern "C" {

void fwat_post_create_(long* pp, int* nspec, int* model_type, int* use_rho_scaling) {
  if (*model_type != POST_ISO && *model_type != POST_HTI) SEM_FAIL("fwat_post_create_: MODEL_TYPE must be 1 or 2");
  if (*nspec <= 0) SEM_FAIL("fwat_post_create_: nspec must be positive");
  PostState* p = new PostState();
  p->nspec = *nspec;
  p->nloc = (size_t)*nspec * 125;
  p->model_type = *model_type;
  p->nker = *model_type == POST_ISO ? 3 : 5;
  p->use_rho_scaling = *use_rho_scaling != 0;
  HIP_CHECK(hipStreamCreate(&p->stream));
  HIP_CHECK(hipEventCreate(&p->ev0));
  HIP_CHECK(hipEventCreate(&p->ev1));
  p->d_ker = dev_alloc<double>(p->nker * p->nloc);
  p->d_hess = dev_alloc<double>(p->nloc);
  p->d_max = dev_alloc<unsigned long long>(1);
  HIP_CHECK(hipMemsetAsync(p->d_ker, 0, p->nker * p->nloc * sizeof(double), p->stream));
  HIP_CHECK(hipMemsetAsync(p->d_hess, 0, p->nloc * sizeof(double), p->stream));
  HIP_CHECK(hipStreamSynchronize(p->stream));
  *pp = reinterpret_cast<long>(p);
}

void fwat_post_set_model_(long* pp, realw* c44store, realw* c45store, realw* c55store, realw* rho_vp,
                          realw* rho_vs, realw* rhostore) {
  PostState* p = post_state(pp);
  if (p->model_type != POST_HTI) SEM_FAIL("fwat_post_set_model_: only the HTI conversion needs the model");
  std::vector<float> m(6 * p->nloc);
  const realw* src[6] = {c44store, c45store, c55store, rho_vp, rho_vs, rhostore};
  for (int q = 0; q < 6; ++q) std::memcpy(&m[q * p->nloc], src[q], p->nloc * sizeof(float));
  dev_free(p->d_model);
  p->d_model = dev_upload(m.data(), m.size(), p->stream);
  HIP_CHECK(hipStreamSynchronize(p->stream));
}

void fwat_post_add_event_(long* pp, long* Mesh_pointer, int* precond_type) {
  PostState* p = post_state(pp);
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (m->nspec != p->nspec) SEM_FAIL("fwat_post_add_event_: NSPEC mismatch");
  if (!m->d_rho_kl_glob) SEM_FAIL("fwat_post_add_event_: the run holds no kernels (SIMULATION_TYPE 3 only)");
  HIP_CHECK(hipStreamSynchronize(m->stream));
  const double* hg = (*precond_type == 0 || *precond_type == 1) ? m->d_hess_kl_glob : nullptr;
  if ((*precond_type == 0 || *precond_type == 1) && !hg)
    SEM_FAIL("fwat_post_add_event_: PRECOND_TYPE %d needs the run's Hessian (APPROXIMATE_HESS_KL)", *precond_type);
  const int habs = *precond_type == 1;
  post_begin(p);
  if (p->model_type == POST_ISO) {
    if (m->anisotropic_kl) SEM_FAIL("fwat_post_add_event_: isotropic sums from an anisotropic-kernel run");
    hipLaunchKernelGGL(k_post_add_iso_run, dim3(pblk(p->nloc)), dim3(PTPB), 0, p->stream, p->nloc, m->d_ibool,
                       m->d_rho_kl_glob, m->d_mu_kl, m->d_kappa_kl, m->d_rho, m->d_kappa, m->d_mu, hg, habs,
                       p->use_rho_scaling ? 0 : 1, p->d_ker, p->d_hess);
  } else {
    if (!m->d_cijkl_kl) SEM_FAIL("fwat_post_add_event_: HTI sums need the run's c_ij kernels (ANISOTROPIC_KL)");
    if (!p->d_model) SEM_FAIL("fwat_post_add_event_: call fwat_post_set_model_ first");
    hipLaunchKernelGGL(k_post_add_hti_run, dim3(pblk(p->nloc)), dim3(PTPB), 0, p->stream, p->nloc, m->d_ibool,
                       m->d_rho_kl_glob, m->d_rho, m->d_cijkl_kl, p->d_model, hg, habs, p->d_ker, p->d_hess);
  }
  post_end(p);
  p->nevents++;
  if (hg) p->nhess++;
}

void fwat_post_add_event_files_(long* pp, realw* ker, realw* hess, int* precond_type) {
  PostState* p = post_state(pp);
  if (p->model_type == POST_HTI && !p->d_model) SEM_FAIL("fwat_post_add_event_files_: call fwat_post_set_model_ first");
  const int nin = p->model_type == POST_ISO ? 3 : 11;
  const bool with_h = hess && (*precond_type == 0 || *precond_type == 1);
  const size_t n = (size_t)(nin + (with_h ? 1 : 0)) * p->nloc;
  if (n > p->stage_n) {
    dev_free(p->d_stage);
    p->d_stage = dev_alloc<float>(n);
    p->stage_n = n;
  }
  HIP_CHECK(hipMemcpyAsync(p->d_stage, ker, (size_t)nin * p->nloc * sizeof(float), hipMemcpyHostToDevice, p->stream));
  if (with_h)
    HIP_CHECK(hipMemcpyAsync(p->d_stage + (size_t)nin * p->nloc, hess, p->nloc * sizeof(float),
                             hipMemcpyHostToDevice, p->stream));
  post_begin(p);
  hipLaunchKernelGGL(k_post_add_files, dim3(pblk(p->nloc)), dim3(PTPB), 0, p->stream, p->nloc, p->model_type,
                     p->use_rho_scaling ? 0 : 1, p->d_stage, p->d_model,
                     with_h ? p->d_stage + (size_t)nin * p->nloc : nullptr, *precond_type == 1 ? 1 : 0, p->d_ker,
                     p->d_hess);
  post_end(p);
  p->nevents++;
  if (with_h) p->nhess++;
}

void fwat_post_get_kernels_(long* pp, double* h_ker) {
  PostState* p = post_state(pp);
  HIP_CHECK(hipMemcpyAsync(h_ker, p->d_ker, p->nker * p->nloc * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  HIP_CHECK(hipStreamSynchronize(p->stream));
}

void fwat_post_get_hess_(long* pp, double* h_hess) {
  PostState* p = post_state(pp);
  HIP_CHECK(hipMemcpyAsync(h_hess, p->d_hess, p->nloc * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  HIP_CHECK(hipStreamSynchronize(p->stream));
}

void fwat_post_set_hess_(long* pp, double* h_hess) {
  PostState* p = post_state(pp);
  HIP_CHECK(hipMemcpyAsync(p->d_hess, h_hess, p->nloc * sizeof(double), hipMemcpyHostToDevice, p->stream));
  HIP_CHECK(hipStreamSynchronize(p->stream));
}

void fwat_post_invert_hess_(long* pp, int* precond_type) {
  PostState* p = post_state(pp);
  post_begin(p);
  device_maxabs(p, p->d_hess, p->nloc);
  hipLaunchKernelGGL(k_hess_invert, dim3(pblk(p->nloc)), dim3(PTPB), 0, p->stream, p->nloc, p->d_hess, p->d_max,
                     *precond_type == 0 ? 1 : 0);
  device_maxabs(p, p->d_hess, p->nloc);
  hipLaunchKernelGGL(k_scale_by_inv_max, dim3(pblk(p->nloc)), dim3(PTPB), 0, p->stream, p->nloc, p->d_hess,
                     p->d_max);
  post_end(p);
}

void fwat_post_apply_precond_(long* pp) {
  PostState* p = post_state(pp);
  post_begin(p);
  hipLaunchKernelGGL(k_apply_precond, dim3(pblk(p->nloc)), dim3(PTPB), 0, p->stream, p->nloc, p->nker, p->d_ker,
                     p->d_hess);
  post_end(p);
}

void fwat_post_info_(long* pp, int* nker, int* nevents, int* nhess) {
  PostState* p = post_state(pp);
  *nker = p->nker;
  *nevents = p->nevents;
  *nhess = p->nhess;
}

void fwat_post_last_time_(long* pp, double* ms) { *ms = post_state(pp)->last_ms; }

void fwat_post_destroy_(long* pp) {
  if (!*pp) return;
  PostState* p = post_state(pp);
  HIP_CHECK(hipStreamSynchronize(p->stream));
  HIP_CHECK(hipEventDestroy(p->ev0));
  HIP_CHECK(hipEventDestroy(p->ev1));
  dev_free(p->d_ker); dev_free(p->d_hess); dev_free(p->d_model); dev_free(p->d_stage); dev_free(p->d_max);
  HIP_CHECK(hipStreamDestroy(p->stream));
  delete p;
  *pp = 0;
}

void fwat_gll2grid_prepare_(long* gp, int* nspec, int* nglob, int* ibool, realw* xstore, realw* ystore,
                            realw* zstore, realw* ox, realw* oy, realw* oz, realw* hx, realw* hy, realw* hz, int* nx,
                            int* ny, int* nz, int* num_iter, long* nb_fd_point) {
  if (*nspec <= 0 || *nglob <= 0 || *nx <= 0 || *ny <= 0 || *nz <= 0)
    SEM_FAIL("fwat_gll2grid_prepare_: sizes must be positive");
  if (!(*hx > 0.0f && *hy > 0.0f && *hz > 0.0f)) SEM_FAIL("fwat_gll2grid_prepare_: grid spacing must be positive");
  if (*num_iter < 1 || *num_iter > 64) SEM_FAIL("fwat_gll2grid_prepare_: num_iter must be in 1..64");
  const size_t npt = (size_t)*nx * *ny * *nz;
  if (npt > (size_t)INT_MAX) SEM_FAIL("fwat_gll2grid_prepare_: grid larger than 2^31 points");
  const size_t nloc = (size_t)*nspec * 125;
  std::vector<int> ib(nloc);
  for (size_t q = 0; q < nloc; ++q) {
    ib[q] = ibool[q] - 1;
    if (ib[q] < 0 || ib[q] >= *nglob) SEM_FAIL("fwat_gll2grid_prepare_: ibool entry %d out of 1..%d", ibool[q], *nglob);
  }
  GridState* g = new GridState();
  g->nspec = *nspec; g->nglob = *nglob;
  g->nx = *nx; g->ny = *ny; g->nz = *nz;
  g->ox = *ox; g->oy = *oy; g->oz = *oz; g->hx = *hx; g->hy = *hy; g->hz = *hz;
  g->num_iter = *num_iter;
  HIP_CHECK(hipStreamCreate(&g->stream));
  double xg[5];
  const double s = std::sqrt(3.0 / 7.0);
  xg[0] = -1.0; xg[1] = -s; xg[2] = 0.0; xg[3] = s; xg[4] = 1.0;
  HIP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_xg), xg, sizeof(xg), 0, hipMemcpyHostToDevice, g->stream));
  std::vector<float> xyz(3 * (size_t)*nglob);
  std::memcpy(xyz.data(), xstore, *nglob * sizeof(float));
  std::memcpy(xyz.data() + *nglob, ystore, *nglob * sizeof(float));
  std::memcpy(xyz.data() + 2 * (size_t)*nglob, zstore, *nglob * sizeof(float));
  g->d_ibool = dev_upload(ib.data(), nloc, g->stream);
  g->d_xyz = dev_upload(xyz.data(), xyz.size(), g->stream);
  g->d_owner = dev_alloc<int>(npt);
  g->d_ref = dev_alloc<double>(3 * npt);
  g->d_count = dev_alloc<int>(1);
  std::vector<int> init(npt, INT_MAX);
  HIP_CHECK(hipMemcpyAsync(g->d_owner, init.data(), npt * sizeof(int), hipMemcpyHostToDevice, g->stream));
  HIP_CHECK(hipMemsetAsync(g->d_count, 0, sizeof(int), g->stream));
  hipEvent_t a, b;
  HIP_CHECK(hipEventCreate(&a));
  HIP_CHECK(hipEventCreate(&b));
  HIP_CHECK(hipEventRecord(a, g->stream));
  hipLaunchKernelGGL(k_grid_owner, dim3(g->nspec), dim3(128), 0, g->stream, g->nspec, g->d_ibool, g->d_xyz,
                     g->nglob, g->ox, g->oy, g->oz, g->hx, g->hy, g->hz, g->nx, g->ny, g->nz, g->num_iter,
                     g->d_owner);
  hipLaunchKernelGGL(k_grid_ref, dim3(pblk(npt)), dim3(PTPB), 0, g->stream, g->nx, g->ny, g->nz, g->d_owner,
                     g->d_ibool, g->d_xyz, g->nglob, g->ox, g->oy, g->oz, g->hx, g->hy, g->hz, g->num_iter, g->d_ref,
                     g->d_count);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipEventRecord(b, g->stream));
  int cnt = 0;
  HIP_CHECK(hipMemcpyAsync(&cnt, g->d_count, sizeof(int), hipMemcpyDeviceToHost, g->stream));
  HIP_CHECK(hipStreamSynchronize(g->stream));
  float ms = 0.0f;
  HIP_CHECK(hipEventElapsedTime(&ms, a, b));
  HIP_CHECK(hipEventDestroy(a));
  HIP_CHECK(hipEventDestroy(b));
  g->last_ms = ms;
  g->nfound = cnt;
  *nb_fd_point = cnt;
  *gp = reinterpret_cast<long>(g);
}

void fwat_gll2grid_(long* gp, int* nfield, realw* gll_data, realw* grid_data) {
  GridState* g = grid_state(gp);
  if (*nfield < 1) SEM_FAIL("fwat_gll2grid_: nfield must be positive");
  const size_t nloc = (size_t)g->nspec * 125, npt = (size_t)g->nx * g->ny * g->nz;
  const size_t nin = (size_t)*nfield * nloc, nout = (size_t)*nfield * npt;
  if (nin > g->in_n) { dev_free(g->d_in); g->d_in = dev_alloc<float>(nin); g->in_n = nin; }
  if (nout > g->out_n) { dev_free(g->d_out); g->d_out = dev_alloc<float>(nout); g->out_n = nout; }
  HIP_CHECK(hipMemcpyAsync(g->d_in, gll_data, nin * sizeof(float), hipMemcpyHostToDevice, g->stream));
  hipEvent_t a, b;
  HIP_CHECK(hipEventCreate(&a));
  HIP_CHECK(hipEventCreate(&b));
  HIP_CHECK(hipEventRecord(a, g->stream));
  hipLaunchKernelGGL(k_grid_project, dim3(pblk(npt)), dim3(PTPB), 0, g->stream, npt, *nfield, nloc, g->d_owner,
                     g->d_ref, g->d_in, g->d_out);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipEventRecord(b, g->stream));
  HIP_CHECK(hipMemcpyAsync(grid_data, g->d_out, nout * sizeof(float), hipMemcpyDeviceToHost, g->stream));
  HIP_CHECK(hipStreamSynchronize(g->stream));
  float ms = 0.0f;
  HIP_CHECK(hipEventElapsedTime(&ms, a, b));
  HIP_CHECK(hipEventDestroy(a));
  HIP_CHECK(hipEventDestroy(b));
  g->last_ms = ms;
}

void fwat_gll2grid_owner_(long* gp, int* owner, double* ref) {
  GridState* g = grid_state(gp);
  const size_t npt = (size_t)g->nx * g->ny * g->nz;
  std::vector<int> o(npt);
  HIP_CHECK(hipMemcpyAsync(o.data(), g->d_owner, npt * sizeof(int), hipMemcpyDeviceToHost, g->stream));
  if (ref) HIP_CHECK(hipMemcpyAsync(ref, g->d_ref, 3 * npt * sizeof(double), hipMemcpyDeviceToHost, g->stream));
  HIP_CHECK(hipStreamSynchronize(g->stream));
  for (size_t q = 0; q < npt; ++q) owner[q] = o[q] == INT_MAX ? 0 : o[q] + 1;
}

void fwat_gll2grid_last_time_(long* gp, double* ms) { *ms = grid_state(gp)->last_ms; }

void fwat_gll2grid_destroy_(long* gp) {
  if (!*gp) return;
  GridState* g = grid_state(gp);
  HIP_CHECK(hipStreamSynchronize(g->stream));
  dev_free(g->d_ibool); dev_free(g->d_xyz); dev_free(g->d_owner); dev_free(g->d_ref); dev_free(g->d_count);
  dev_free(g->d_in); dev_free(g->d_out);
  HIP_CHECK(hipStreamDestroy(g->stream));
  delete g;
  *gp = 0;
}

void fwat_taper_grid_(realw* x_min_glob, realw* x_max_glob, realw* y_min_glob, realw* y_max_glob, realw* z_min_glob,
                      realw* z_max_glob, realw* distance_min_glob, realw* distance_max_glob, realw* h_suppress,
                      realw* h_buffer, realw* v_suppress, realw* v_buffer, realw* ox, realw* oy, realw* oz,
                      realw* hx, realw* hy, realw* hz, int* nx, int* ny, int* nz, int* nfield, realw* data,
                      realw* mask) {
  const float dh = (*distance_max_glob + *distance_min_glob) / 2.0f;
  if (!(dh > 0.0f)) SEM_FAIL("fwat_taper_grid_: taper spacing must be positive");
  if (*nx <= 0 || *ny <= 0 || *nz <= 0 || *nfield < 0) SEM_FAIL("fwat_taper_grid_: sizes must be positive");
  TaperAxis X, Y, Z;
  X.c0 = *x_min_glob; Y.c0 = *y_min_glob; Z.c0 = *z_min_glob;
  X.dh = Y.dh = Z.dh = dh;
  X.n = (int)std::floor((*x_max_glob - *x_min_glob) / dh + 1.0f);
  Y.n = (int)std::floor((*y_max_glob - *y_min_glob) / dh + 1.0f);
  Z.n = (int)std::floor((*z_max_glob - *z_min_glob) / dh + 1.0f);
  X.n1 = Y.n1 = (int)std::floor(*h_suppress / dh);
  X.n2 = Y.n2 = (int)std::floor(*h_buffer / dh);
  Z.n1 = (int)std::floor(*v_suppress / dh);
  Z.n2 = (int)std::floor(*v_buffer / dh);
  X.bottom = Y.bottom = 0;
  Z.bottom = 1;
  for (const TaperAxis* A : {&X, &Y}) {
    const int x2 = A->n1 + A->n2 + 1, x3 = A->n - A->n1 - A->n2, x4 = A->n - A->n1;
    if (A->n < 2 || x3 < 0 || x4 > A->n || x2 - 1 > A->n || x3 + 1 < 1)
      SEM_FAIL("fwat_taper_grid_: lateral taper widths do not fit the mesh");
  }
  if (Z.n < 2 || Z.n1 + Z.n2 > Z.n) SEM_FAIL("fwat_taper_grid_: vertical taper widths do not fit the mesh");
  const size_t npt = (size_t)*nx * *ny * *nz;
  hipStream_t s;
  HIP_CHECK(hipStreamCreate(&s));
  float* d = dev_alloc<float>(std::max<size_t>(1, (size_t)*nfield * npt));
  float* dm = mask ? dev_alloc<float>(npt) : nullptr;
  if (*nfield) HIP_CHECK(hipMemcpyAsync(d, data, (size_t)*nfield * npt * sizeof(float), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_taper, dim3(pblk(npt)), dim3(PTPB), 0, s, *nx, *ny, *nz, *ox, *oy, *oz, *hx, *hy, *hz, X, Y,
                     Z, *nfield, d, dm);
  HIP_CHECK(hipGetLastError());
  if (*nfield) HIP_CHECK(hipMemcpyAsync(data, d, (size_t)*nfield * npt * sizeof(float), hipMemcpyDeviceToHost, s));
  if (mask) HIP_CHECK(hipMemcpyAsync(mask, dm, npt * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  dev_free(d);
  dev_free(dm);
  HIP_CHECK(hipStreamDestroy(s));
}

}  // extern "C"
