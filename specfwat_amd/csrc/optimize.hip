// Grid optimiser on the device: steepest-descent, L-BFGS and conjugate-gradient
// search directions and the model update (SpecFWAT's xfwat_optimize,
// src/optimization/optimize_grid.f90).  The reference forms every inner
// product as a single-precision sum (inner_product, src/shared/utils.f90:
// 1527-1534); here all vectors derived from the inputs (model and gradient
// differences, q, r) are fp64 and every reduction is an fp64 two-stage sum in a
// fixed order (per-workgroup partials, then one workgroup), so a direction is
// reproducible bit for bit.  The L-BFGS coefficients stay on the device: each
// axpy reads the dot products it needs from a small scalar array, so the two-
// loop recursion runs without host round trips.
//
// HBM layout: grids (nx,ny,nz,nkernel) flattened, npg = nx*ny*nz points per
// parameter; history [nstore+1][n] fp32 as uploaded, s_k / y_k [nstore][n] fp64.
#include <algorithm>
#include <cmath>
#include <vector>

#include "sem_device.h"

namespace {

constexpr int OTPB = 256;
constexpr int ONB = 512;  // workgroups of a reduction (partials summed by one workgroup)

inline unsigned oblk(size_t n) { return (unsigned)((n + OTPB - 1) / OTPB); }

double g_last_ms = 0.0;

// scalar slots
enum { S_DYS = 0, S_DSQ = 64, S_DYR = 128, S_DYY = 192, S_MAX = 193, S_GD = 194, S_GG = 195, S_DD = 196,
       S_A = 197, S_B = 198, S_NSLOT = 256 };
constexpr int MAX_STORE = 64;

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
  return t;
}

__device__ __forceinline__ double block_max(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t = fmax(t, sh[w]);
  return t;
}

// partial sums of a . b (a, b fp64); b == nullptr: a . a
__global__ void __launch_bounds__(OTPB) k_dot_part(size_t n, const double* __restrict__ a,
                                                   const double* __restrict__ b, double* __restrict__ part) {
  __shared__ double sh[OTPB / 64];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    s += a[i] * (b ? b[i] : a[i]);
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void __launch_bounds__(OTPB) k_maxabs_part(size_t n, const double* __restrict__ a,
                                                      double* __restrict__ part) {
  __shared__ double sh[OTPB / 64];
  double m = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    m = fmax(m, fabs(a[i]));
  const double t = block_max(m, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// one workgroup: partials -> scal[slot] (sum or max), fixed order
__global__ void __launch_bounds__(ONB) k_reduce_final(int np, const double* __restrict__ part,
                                                      double* __restrict__ scal, int slot, int is_max) {
  __shared__ double sh[ONB / 64];
  double v = (int)threadIdx.x < np ? part[threadIdx.x] : 0.0;
  const double t = is_max ? block_max(fabs(v), sh) : block_sum(v, sh);
  if (threadIdx.x == 0) scal[slot] = t;
}

// s = model difference in the inversion variables (log for the first three
// parameters), y = gradient difference; both fp64
__global__ void k_pair(size_t n, size_t npg, const float* __restrict__ m1, const float* __restrict__ m0,
                       const float* __restrict__ g1, const float* __restrict__ g0, double* __restrict__ s,
                       double* __restrict__ y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  s[i] = i / npg < 3 ? log((double)m1[i]) - log((double)m0[i]) : (double)m1[i] - (double)m0[i];
  y[i] = (double)g1[i] - (double)g0[i];
}

__global__ void k_to_f64(size_t n, const float* __restrict__ a, double* __restrict__ b) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

// first loop: q -= a_k y_k with a_k = <s_k, q> / <y_k, s_k>
__global__ void k_q_update(size_t n, double* __restrict__ q, const double* __restrict__ y,
                           const double* __restrict__ scal, int k) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  q[i] -= (scal[S_DSQ + k] / scal[S_DYS + k]) * y[i];
}

// r = p_k hess q with p_k = <y, s> / <y, y> of the latest pair
__global__ void k_r_init(size_t n, double* __restrict__ r, const double* __restrict__ q,
                         const float* __restrict__ hess, const double* __restrict__ scal, int last) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  r[i] = (scal[S_DYS + last] / scal[S_DYY]) * (double)hess[i] * q[i];
}

// second loop: r += (a_k - b_k) s_k with b_k = <y_k, r> / <y_k, s_k>
__global__ void k_r_update(size_t n, double* __restrict__ r, const double* __restrict__ s,
                           const double* __restrict__ scal, int k) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dys = scal[S_DYS + k];
  r[i] += (scal[S_DSQ + k] / dys - scal[S_DYR + k] / dys) * s[i];
}

__global__ void k_negate(size_t n, const double* __restrict__ r, double* __restrict__ d) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = -r[i];
}

// d / max|d| -> fp32 output, and the normalised direction kept in fp64
__global__ void k_normalise(size_t n, double* __restrict__ d, const double* __restrict__ scal,
                            float* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = d[i] / scal[S_MAX];
  d[i] = v;
  out[i] = (float)v;
}

__global__ void k_sd(size_t n, const float* __restrict__ g, const float* __restrict__ h, double* __restrict__ d) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = -(double)g[i] * (double)h[i];
}

// hess * (-g1 - alpha g0) with alpha = <g1 - g0, g1> / <g0, g0> (sums over all parameters)
__global__ void k_cg_dir(size_t n, const float* __restrict__ g1, const float* __restrict__ g0,
                         const float* __restrict__ h, const double* __restrict__ scal, double* __restrict__ d) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double alpha = scal[S_B] / scal[S_A];
  d[i] = (double)h[i] * (-(double)g1[i] - alpha * (double)g0[i]);
}

__global__ void k_diff_f64(size_t n, const float* __restrict__ a, const float* __restrict__ b,
                           double* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (double)a[i] - (double)b[i];
}

// model_update + alpha_scaling (single precision like the stored model)
__global__ void k_model_update(size_t npg, int niso, int do_iso, int do_ani, float step, float r0, float r1,
                               float* __restrict__ model, const float* __restrict__ dir) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npg) return;
  if (do_iso) {
    for (int k = 0; k < niso; ++k) model[k * npg + i] = model[k * npg + i] * expf(step * dir[k * npg + i]);
    const float vs = model[npg + i];
    const float ratio = model[i] / vs;
    if (ratio < r0) model[i] = r0 * vs;
    if (ratio > r1) model[i] = r1 * vs;
  }
  if (do_ani)
    for (int k = 3; k < 5; ++k) model[k * npg + i] = model[k * npg + i] + step * dir[k * npg + i];
}

struct OptRun {
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double* part = nullptr;
  double* scal = nullptr;
  OptRun() {
    HIP_CHECK(hipStreamCreate(&s));
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    part = dev_alloc<double>(ONB);
    scal = dev_alloc<double>(S_NSLOT);
    HIP_CHECK(hipEventRecord(e0, s));
  }
  void dot(size_t n, const double* a, const double* b, int slot) {
    hipLaunchKernelGGL(k_dot_part, dim3(ONB), dim3(OTPB), 0, s, n, a, b, part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(ONB), 0, s, ONB, part, scal, slot, 0);
  }
  void maxabs(size_t n, const double* a, int slot) {
    hipLaunchKernelGGL(k_maxabs_part, dim3(ONB), dim3(OTPB), 0, s, n, a, part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(ONB), 0, s, ONB, part, scal, slot, 1);
  }
  void finish(std::vector<double>* h_scal) {
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(e1, s));
    if (h_scal) {
      h_scal->resize(S_NSLOT);
      HIP_CHECK(hipMemcpyAsync(h_scal->data(), scal, S_NSLOT * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    g_last_ms = ms;
  }
  ~OptRun() {
    dev_free(part);
    dev_free(scal);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
  }
};

void check_n(long n, long npg, const char* who) {
  if (n <= 0 || npg <= 0 || n % npg) SEM_FAIL("%s: n must be a positive multiple of nx*ny*nz", who);
}

}  // namespace

// ================================================================ C ABI
extern "C" {

void fwat_opt_sd_direction_(long* n, realw* gradient, realw* hess, realw* direction) {
  if (*n <= 0) SEM_FAIL("fwat_opt_sd_direction_: n must be positive");
  const size_t N = (size_t)*n;
  float* dg = dev_alloc<float>(2 * N);
  double* d = dev_alloc<double>(N);
  float* out = dev_alloc<float>(N);
  {
    OptRun R;
    HIP_CHECK(hipMemcpyAsync(dg, gradient, N * 4, hipMemcpyHostToDevice, R.s));
    HIP_CHECK(hipMemcpyAsync(dg + N, hess, N * 4, hipMemcpyHostToDevice, R.s));
    HIP_CHECK(hipEventRecord(R.e0, R.s));
    hipLaunchKernelGGL(k_sd, dim3(oblk(N)), dim3(OTPB), 0, R.s, N, dg, dg + N, d);
    R.maxabs(N, d, S_MAX);
    hipLaunchKernelGGL(k_normalise, dim3(oblk(N)), dim3(OTPB), 0, R.s, N, d, R.scal, out);
    R.finish(nullptr);
    HIP_CHECK(hipMemcpy(direction, out, N * 4, hipMemcpyDeviceToHost));
  }
  dev_free(dg); dev_free(d); dev_free(out);
}

void fwat_opt_lbfgs_direction_(long* n, long* npg, int* nstore, realw* grads, realw* models, realw* hess,
                               realw* dx, realw* dy, realw* dz, realw* direction, double* angle) {
  check_n(*n, *npg, "fwat_opt_lbfgs_direction_");
  const int m = *nstore;
  if (m < 1 || m > MAX_STORE) SEM_FAIL("fwat_opt_lbfgs_direction_: nstore must be in 1..%d", MAX_STORE);
  if (!(*dx > 0.0f && *dy > 0.0f && *dz > 0.0f)) SEM_FAIL("fwat_opt_lbfgs_direction_: grid spacing must be positive");
  const size_t N = (size_t)*n, NPG = (size_t)*npg;
  float* dgm = dev_alloc<float>(2 * (size_t)(m + 1) * N + N);  // grads, models, hess
  float* dG = dgm;
  float* dM = dgm + (size_t)(m + 1) * N;
  float* dH = dgm + 2 * (size_t)(m + 1) * N;
  double* S = dev_alloc<double>((size_t)m * N);
  double* Y = dev_alloc<double>((size_t)m * N);
  double* q = dev_alloc<double>(N);
  double* r = dev_alloc<double>(N);
  double* g = dev_alloc<double>(N);
  float* out = dev_alloc<float>(N);
  std::vector<double> sc;
  {
    OptRun R;
    HIP_CHECK(hipMemcpyAsync(dG, grads, (size_t)(m + 1) * N * 4, hipMemcpyHostToDevice, R.s));
    HIP_CHECK(hipMemcpyAsync(dM, models, (size_t)(m + 1) * N * 4, hipMemcpyHostToDevice, R.s));
    HIP_CHECK(hipMemcpyAsync(dH, hess, N * 4, hipMemcpyHostToDevice, R.s));
    HIP_CHECK(hipEventRecord(R.e0, R.s));
    for (int k = 0; k < m; ++k)
      hipLaunchKernelGGL(k_pair, dim3(oblk(N)), dim3(OTPB), 0, R.s, N, NPG, dM + (size_t)(k + 1) * N,
                         dM + (size_t)k * N, dG + (size_t)(k + 1) * N, dG + (size_t)k * N, S + (size_t)k * N,
                         Y + (size_t)k * N);
    hipLaunchKernelGGL(k_to_f64, dim3(oblk(N)), dim3(OTPB), 0, R.s, N, dG + (size_t)m * N, q);
    hipLaunchKernelGGL(k_to_f64, dim3(oblk(N)), dim3(OTPB), 0, R.s, N, dG + (size_t)m * N, g);
    for (int k = m - 1; k >= 0; --k) {
      R.dot(N, Y + (size_t)k * N, S + (size_t)k * N, S_DYS + k);
      R.dot(N, S + (size_t)k * N, q, S_DSQ + k);
      hipLaunchKernelGGL(k_q_update, dim3(oblk(N)), dim3(OTPB), 0, R.s, N, q, Y + (size_t)k * N, R.scal, k);
    }
    R.dot(N, Y + (size_t)(m - 1) * N, nullptr, S_DYY);
    hipLaunchKernelGGL(k_r_init, dim3(oblk(N)), dim3(OTPB), 0, R.s, N, r, q, dH, R.scal, m - 1);
    for (int k = 0; k < m; ++k) {
      R.dot(N, Y + (size_t)k * N, r, S_DYR + k);
      hipLaunchKernelGGL(k_r_update, dim3(oblk(N)), dim3(OTPB), 0, R.s, N, r, S + (size_t)k * N, R.scal, k);
    }
    hipLaunchKernelGGL(k_negate, dim3(oblk(N)), dim3(OTPB), 0, R.s, N, r, q);  // q <- d = -r
    R.maxabs(N, q, S_MAX);
    hipLaunchKernelGGL(k_normalise, dim3(oblk(N)), dim3(OTPB), 0, R.s, N, q, R.scal, out);
    R.dot(N, g, q, S_GD);
    R.dot(N, g, nullptr, S_GG);
    R.dot(N, q, nullptr, S_DD);
    R.finish(&sc);
    HIP_CHECK(hipMemcpy(direction, out, N * 4, hipMemcpyDeviceToHost));
  }
  // angle between the direction and the negative gradient (degrees)
  *angle = std::acos(-sc[S_GD] / (std::sqrt(sc[S_GG]) * std::sqrt(sc[S_DD]))) * (180.0 / M_PI);
  dev_free(dgm); dev_free(S); dev_free(Y); dev_free(q); dev_free(r); dev_free(g); dev_free(out);
}

void fwat_opt_cg_direction_(long* n, long* npg, realw* g1, realw* g0, realw* hess, realw* direction,
                            double* alpha) {
  check_n(*n, *npg, "fwat_opt_cg_direction_");
  const size_t N = (size_t)*n;
  float* d3 = dev_alloc<float>(3 * N);
  double* a = dev_alloc<double>(N);
  double* b = dev_alloc<double>(N);
  float* out = dev_alloc<float>(N);
  std::vector<double> sc;
  {
    OptRun R;
    HIP_CHECK(hipMemcpyAsync(d3, g1, N * 4, hipMemcpyHostToDevice, R.s));
    HIP_CHECK(hipMemcpyAsync(d3 + N, g0, N * 4, hipMemcpyHostToDevice, R.s));
    HIP_CHECK(hipMemcpyAsync(d3 + 2 * N, hess, N * 4, hipMemcpyHostToDevice, R.s));
    HIP_CHECK(hipEventRecord(R.e0, R.s));
    // sum over parameters of the per-parameter sums = one sum over all points
    hipLaunchKernelGGL(k_to_f64, dim3(oblk(N)), dim3(OTPB), 0, R.s, N, d3 + N, a);            // g0
    hipLaunchKernelGGL(k_diff_f64, dim3(oblk(N)), dim3(OTPB), 0, R.s, N, d3, d3 + N, b);      // g1 - g0
    R.dot(N, a, nullptr, S_A);
    hipLaunchKernelGGL(k_to_f64, dim3(oblk(N)), dim3(OTPB), 0, R.s, N, d3, a);                // g1
    R.dot(N, b, a, S_B);
    hipLaunchKernelGGL(k_cg_dir, dim3(oblk(N)), dim3(OTPB), 0, R.s, N, d3, d3 + N, d3 + 2 * N, R.scal, a);
    R.maxabs(N, a, S_MAX);
    hipLaunchKernelGGL(k_normalise, dim3(oblk(N)), dim3(OTPB), 0, R.s, N, a, R.scal, out);
    R.finish(&sc);
    HIP_CHECK(hipMemcpy(direction, out, N * 4, hipMemcpyDeviceToHost));
  }
  *alpha = sc[S_B] / sc[S_A];
  dev_free(d3); dev_free(a); dev_free(b); dev_free(out);
}

void fwat_opt_model_update_(long* npg, int* nkernel, int* model_type, int* update_iso, int* update_ani,
                            realw* step_len, realw* vpvs_ratio_range, realw* model, realw* direction) {
  if (*npg <= 0) SEM_FAIL("fwat_opt_model_update_: npg must be positive");
  if (*model_type == 1 && *nkernel < 2) SEM_FAIL("fwat_opt_model_update_: model type 1 needs vp and vs");
  if (*model_type == 2 && *nkernel != 5) SEM_FAIL("fwat_opt_model_update_: model type 2 has 5 parameters");
  if (*model_type != 1 && *model_type != 2) SEM_FAIL("fwat_opt_model_update_: unknown model type %d", *model_type);
  const size_t NPG = (size_t)*npg, N = NPG * *nkernel;
  float* dm = dev_alloc<float>(2 * N);
  {
    OptRun R;
    HIP_CHECK(hipMemcpyAsync(dm, model, N * 4, hipMemcpyHostToDevice, R.s));
    HIP_CHECK(hipMemcpyAsync(dm + N, direction, N * 4, hipMemcpyHostToDevice, R.s));
    HIP_CHECK(hipEventRecord(R.e0, R.s));
    const int niso = *model_type == 1 ? *nkernel : 3;
    const int do_iso = *model_type == 1 || *update_iso;
    const int do_ani = *model_type == 2 && *update_ani;
    hipLaunchKernelGGL(k_model_update, dim3(oblk(NPG)), dim3(OTPB), 0, R.s, NPG, niso, do_iso, do_ani, *step_len,
                       vpvs_ratio_range[0], vpvs_ratio_range[1], dm, dm + N);
    R.finish(nullptr);
    HIP_CHECK(hipMemcpy(model, dm, N * 4, hipMemcpyDeviceToHost));
  }
  dev_free(dm);
}

void fwat_opt_last_time_(double* ms) { *ms = g_last_ms; }

}  // extern "C"
