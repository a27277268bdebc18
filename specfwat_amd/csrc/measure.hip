// Waveform misfit and adjoint sources of all receiver components of an event
// in one device pass (SpecFWAT's per-trace loop of leq_data.f90:148 over
// WaveformMisfit%calc_adjoint_source, waveform_misfit.f90:15-64, and
// filter_adj, adjoint_source.f90:92-114).
//
// Layout: traces in the seismogram order [nstep][ntrace] (ntrace = 3 * nrec),
// one lane per trace, so every sample step of a wavefront reads 64 adjacent
// floats.  Per window: tapered synthetics and data, their difference tapered
// again as the adjoint source, misfit 0.5 * simpson((s - d)^2) and the mean
// residual in double precision (as the reference).  The adjoint source is then
// band-passed in single precision by the second-order sections the caller
// designs (zero phase: forward, then reversed, direct form I as SAC's APPLY)
// and tapered over [first window start, last window end].  The per-index taper
// weights are the same for every trace and come precomputed in double.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sem_device.h"

#include <mutex>

namespace {

constexpr int MTPB = 64;
constexpr int MAX_WIN = 64;

struct WinInfo {
  int nb[MAX_WIN];  // 0-based first sample
  int nl[MAX_WIN];  // samples
  int woff[MAX_WIN];  // offset of the window's taper weights
};

#pragma clang fp contract(off)
__global__ void __launch_bounds__(MTPB) k_waveform_adj(int nstep, int ntrace, double dt, int nwin, WinInfo W,
                                                       const double* __restrict__ wts, const float* __restrict__ obs,
                                                       const float* __restrict__ syn, float* __restrict__ adj,
                                                       double* __restrict__ misfit, double* __restrict__ resid) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntrace) return;
  const size_t T = (size_t)ntrace;
  for (int i = 0; i < nstep; ++i) adj[(size_t)i * T + t] = 0.0f;
  for (int w = 0; w < nwin; ++w) {
    const int nb = W.nb[w], nl = W.nl[w];
    const double* tw = wts + W.woff[w];
    // (s - d) of sample k of the window: tapered synthetics minus tapered data
    auto r_at = [&](int k) {
      const size_t a = (size_t)(nb + k) * T + t;
      return (double)syn[a] * tw[k] - (double)obs[a] * tw[k];
    };
    double rsum = 0.0;
    for (int k = 0; k < nl; ++k) {
      const double r = r_at(k);
      adj[(size_t)(nb + k) * T + t] = (float)(r * tw[k]);
      rsum = rsum + r;
    }
    // simpson (utils.f90:1630-1663) of (s - d)^2 with the reference's loops
    auto y = [&](int k) { const double r = r_at(k); return r * r; };
    double integ = 0.0;
    if (nl >= 2) {
      if (nl % 2 == 0) {
        double sum = y(0) + y(nl - 2);
        for (int k = 1; k <= nl - 3; k += 2) sum = sum + 4.0 * y(k) + 2.0 * y(k + 1);
        integ = (dt / 3.0) * sum + 0.5 * dt * (y(nl - 2) + y(nl - 1));
      } else {
        double sum = y(0) + y(nl - 1);
        for (int k = 1; k <= nl - 2; k += 2) sum = sum + 4.0 * y(k);
        for (int k = 2; k <= nl - 3; k += 2) sum = sum + 2.0 * y(k);
        integ = (dt / 3.0) * sum;
      }
    }
    misfit[(size_t)w * T + t] = 0.5 * integ;
    resid[(size_t)w * T + t] = rsum / nl;
  }
}

// zero-phase cascade of second-order sections over each trace, in place;
// samples are fetched 16 ahead into registers, the recursion is SAC's
constexpr int SOS_AHEAD = 16;
__global__ void __launch_bounds__(MTPB) k_sos_filter(int nstep, int ntrace, int nsect, const float* __restrict__ sos,
                                                     float* __restrict__ x) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntrace) return;
  const size_t T = (size_t)ntrace;
  for (int pass = 0; pass < 2; ++pass)
    for (int j = 0; j < nsect; ++j) {
      const float b0 = sos[6 * j], b1 = sos[6 * j + 1], b2 = sos[6 * j + 2];
      const float a1 = sos[6 * j + 4], a2 = sos[6 * j + 5];
      float x1 = 0.0f, x2 = 0.0f, y1 = 0.0f, y2 = 0.0f;
      for (int k0 = 0; k0 < nstep; k0 += SOS_AHEAD) {
        float buf[SOS_AHEAD];
#pragma unroll
        for (int u = 0; u < SOS_AHEAD; ++u) {
          const int k = k0 + u;
          buf[u] = k < nstep ? x[(size_t)(pass == 0 ? k : nstep - 1 - k) * T + t] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < SOS_AHEAD; ++u) {
          const float xi = buf[u];
          float out = b0 * xi + b1 * x1 + b2 * x2;
          out = out - (a1 * y1 + a2 * y2);
          y2 = y1;
          y1 = out;
          x2 = x1;
          x1 = xi;
          buf[u] = out;
        }
#pragma unroll
        for (int u = 0; u < SOS_AHEAD; ++u) {
          const int k = k0 + u;
          if (k < nstep) x[(size_t)(pass == 0 ? k : nstep - 1 - k) * T + t] = buf[u];
        }
      }
    }
}

// The order-2 bandpass (two sections) on a lane pair per trace: the even lane
// runs section 0 on sample n while the odd lane runs section 1 on sample
// n - 1, fed section 0's previous output across the pair; each lane's
// instruction stream is one section's, and each section sees the same inputs
// in the same order as section by section.
__global__ void __launch_bounds__(MTPB) k_sos_filter2(int nstep, int ntrace, const float* __restrict__ sos,
                                                      float* __restrict__ x) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x, t = gid >> 1, sec = gid & 1;
  if (t >= ntrace) return;  // both lanes of a pair leave together
  const int src = (threadIdx.x & ~1);  // the pair's section-0 lane
  const size_t T = (size_t)ntrace;
  const float* c = sos + 6 * sec;
  const float b0 = c[0], b1 = c[1], b2 = c[2], a1 = c[4], a2 = c[5];
  for (int pass = 0; pass < 2; ++pass) {
    auto at = [&](int k) -> size_t { return (size_t)(pass == 0 ? k : nstep - 1 - k) * T + t; };
    float x1 = 0.0f, x2 = 0.0f, y1 = 0.0f, y2 = 0.0f, prev = 0.0f;
    // step n: section 0 filters sample n (n < nstep), section 1 sample n - 1 (n >= 1)
    auto step = [&](int n, float xin) {
      const float from0 = __shfl(prev, src, 64);
      const float in = sec == 0 ? xin : from0;
      const bool valid = sec == 0 ? n < nstep : n >= 1;
      if (valid) {
        float out = b0 * in + b1 * x1 + b2 * x2;
        out = out - (a1 * y1 + a2 * y2);
        y2 = y1; y1 = out; x2 = x1; x1 = in;
        if (sec == 0) prev = out;
        else x[at(n - 1)] = out;
      }
    };
    for (int k0 = 0; k0 < nstep; k0 += SOS_AHEAD) {
      float buf[SOS_AHEAD];
#pragma unroll
      for (int u = 0; u < SOS_AHEAD; ++u) {
        const int k = k0 + u;
        buf[u] = (sec == 0 && k < nstep) ? x[at(k)] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < SOS_AHEAD; ++u) {
        if (k0 + u > nstep) break;
        step(k0 + u, buf[u]);
      }
    }
    if (nstep % SOS_AHEAD == 0) step(nstep, 0.0f);  // section 1 on the last sample
    __threadfence();  // the backward pass reads what the pair's other lane wrote
  }
}

void launch_sos(hipStream_t s, int N, int T, int nsect, const float* d_sos, float* d_x) {
  if (nsect == 2)
    hipLaunchKernelGGL(k_sos_filter2, dim3((unsigned)((2 * (size_t)T + MTPB - 1) / MTPB)), dim3(MTPB), 0, s, N, T,
                       d_sos, d_x);
  else
    hipLaunchKernelGGL(k_sos_filter, dim3((unsigned)((T + MTPB - 1) / MTPB)), dim3(MTPB), 0, s, N, T, nsect, d_sos,
                       d_x);
}

// taper over [nb, nb + nl) in double, zero outside
__global__ void k_window_final(int nstep, int ntrace, int nb, int nl, const double* __restrict__ tw,
                               float* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nstep * ntrace) return;
  const int k = (int)(i / ntrace) - nb;
  x[i] = (k >= 0 && k < nl) ? (float)((double)x[i] * tw[k]) : 0.0f;
}
// ---- cross-correlation family (cc_tt_misfit.f90, utils/cc.f90)

// tapered window of every trace, transposed to trace-major rows [ntrace][nl]
// through 64 x 64 LDS tiles (reads and writes both coalesced)
__global__ void __launch_bounds__(256) k_cc_rows(int ntrace, int nb, int nl, const double* __restrict__ tw,
                                                 const float* __restrict__ obs, const float* __restrict__ syn,
                                                 double* __restrict__ D, double* __restrict__ S) {
  __shared__ double sd[64][65], ss[64][65];
  const int t0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
  for (int r = threadIdx.x / 64; r < 64; r += 4) {
    const int k = k0 + r, t = t0 + (threadIdx.x & 63);
    if (k < nl && t < ntrace) {
      const size_t a = (size_t)(nb + k) * ntrace + t;
      sd[r][threadIdx.x & 63] = (double)obs[a] * tw[k];
      ss[r][threadIdx.x & 63] = (double)syn[a] * tw[k];
    }
  }
  __syncthreads();
  for (int r = threadIdx.x / 64; r < 64; r += 4) {
    const int t = t0 + r, k = k0 + (threadIdx.x & 63);
    if (k < nl && t < ntrace) {
      D[(size_t)t * nl + k] = sd[threadIdx.x & 63][r];
      S[(size_t)t * nl + k] = ss[threadIdx.x & 63][r];
    }
  }
}

// xcorr_shift, first stage: workgroup (trace, block of 1024 lags, 4 consecutive
// lags per lane) of the full correlation; d and the matching stretch of s come
// through LDS in chunks of 1024 samples taken in increasing i1, so each lag's
// sum runs over i1 in increasing order as in myconvolution_dp.  A lane keeps
// its four lags' s samples in a sliding register window (one LDS read per
// step for four products); it runs over the union of its lags' i1 ranges, and
// the staged s is zero outside the trace, so each lag only gains exact zeros.
// Writes the block's first maximum.
constexpr int CC_LPL = 4, CC_LAGS = 256 * CC_LPL, CC_CHUNK = 1024;
__global__ void __launch_bounds__(256) k_cc_corr(int nl, const double* __restrict__ D,
                                                 const double* __restrict__ S, double* __restrict__ bval,
                                                 int* __restrict__ bidx) {
  __shared__ double sd[CC_CHUNK], ss[CC_CHUNK + CC_LAGS];
  const int t = blockIdx.y, nblk = gridDim.x;
  const double* d = D + (size_t)t * nl;
  const double* s = S + (size_t)t * nl;
  const int m0 = blockIdx.x * CC_LAGS, mb = m0 + (int)threadIdx.x * CC_LPL;
  const int nlag = 2 * nl - 1;
  // lag m reads s[i1 + nl - 1 - m] for i1 in [max(0, m - nl + 1), min(nl - 1, m)]; union over mb..mb+3
  const int lo = mb - (nl - 1) > 0 ? mb - (nl - 1) : 0, hi = mb + CC_LPL - 1 < nl - 1 ? mb + CC_LPL - 1 : nl - 1;
  const int blo = m0 - (nl - 1) > 0 ? m0 - (nl - 1) : 0;
  const int mlast = min(m0 + CC_LAGS - 1, nlag - 1);
  const int bhi = mlast < nl - 1 ? mlast : nl - 1;
  double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
  for (int i0 = blo; i0 <= bhi; i0 += CC_CHUNK) {
    const int i1end = min(i0 + CC_CHUNK - 1, bhi);
    // s indices of this chunk over the block's lags: [i0 + nl - 1 - (m0 + CC_LAGS - 1), i1end + nl - 1 - m0]
    const int sbase = i0 + nl - 1 - (m0 + CC_LAGS - 1);
    __syncthreads();
    for (int k = threadIdx.x; k < CC_CHUNK; k += 256) sd[k] = (i0 + k <= i1end) ? d[i0 + k] : 0.0;
    for (int k = threadIdx.x; k < CC_CHUNK + CC_LAGS; k += 256) {
      const int j = sbase + k;
      ss[k] = (j >= 0 && j < nl) ? s[j] : 0.0;
    }
    __syncthreads();
    if (mb < nlag) {
      const int a = max(lo, i0), b = min(hi, i1end);
      // s[i1 + nl - 1 - (mb + q)] sits at ss[i1 + off - q], off in [3, CC_LAGS - 1] - i0
      const int off = m0 + CC_LAGS - 1 - mb - i0;
      if (a <= b) {
        double r0 = ss[a + off], r1 = ss[a + off - 1], r2 = ss[a + off - 2], r3 = ss[a + off - 3];
        for (int i1 = a; i1 <= b; ++i1) {
          const double dv = sd[i1 - i0];
          c0 = c0 + dv * r0;
          c1 = c1 + dv * r1;
          c2 = c2 + dv * r2;
          c3 = c3 + dv * r3;
          r3 = r2;
          r2 = r1;
          r1 = r0;
          r0 = ss[i1 + 1 + off];
        }
      }
    }
  }
  __shared__ double rv[256];
  __shared__ int ri[256];
  // the lane's first maximum over its lags (in increasing lag), then over lanes
  double best = -INFINITY;
  int bm = 0x7fffffff;
  const double cv[CC_LPL] = {c0, c1, c2, c3};
  for (int q = 0; q < CC_LPL; ++q)
    if (mb + q < nlag && cv[q] > best) { best = cv[q]; bm = mb + q; }
  rv[threadIdx.x] = best;
  ri[threadIdx.x] = bm;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      const double v2 = rv[threadIdx.x + o];
      const int m2 = ri[threadIdx.x + o];
      if (v2 > rv[threadIdx.x] || (v2 == rv[threadIdx.x] && m2 < ri[threadIdx.x])) {
        rv[threadIdx.x] = v2;
        ri[threadIdx.x] = m2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    bval[(size_t)t * nblk + blockIdx.x] = rv[0];
    bidx[(size_t)t * nblk + blockIdx.x] = ri[0];
  }
}

// xcorr_shift, second stage: first maximum over the blocks of a trace
__global__ void k_cc_pick(int ntrace, int nblk, int nl, const double* __restrict__ bval,
                          const int* __restrict__ bidx, int* __restrict__ ishift) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntrace) return;
  double best = bval[(size_t)t * nblk];
  int bm = bidx[(size_t)t * nblk];
  for (int b = 1; b < nblk; ++b) {
    const double v = bval[(size_t)t * nblk + b];
    if (v > best) { best = v; bm = bidx[(size_t)t * nblk + b]; }
  }
  ishift[t] = bm + 1 - nl;
}

struct CcCfg {
  double dt_sigma_min, dlna_sigma_min, tshift_min, tshift_max, dlna_min, dlna_max, cc_min;
};

// fixed-order tree sum over a 256-lane workgroup, result on every lane
__device__ double wg_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// per trace (one workgroup on its tapered rows d, s): dlna, calc_cc_error
// (cc_correction + gradient), cc_max_coef, cc_measure_select and
// calc_cc_adjsrc; the sums are fixed-order tree sums.  out[8] per (window,
// trace): misfit, residual, error, cc_max, tshift, dlna, selected, ishift; the
// selected traces' tapered adjoint sources go to rows A [ntrace][nl]
__global__ void __launch_bounds__(256) k_cc_finish(int nl, double dt, int imeas, CcCfg cfg,
                                                   const double* __restrict__ tw, const double* __restrict__ D,
                                                   const double* __restrict__ S, const int* __restrict__ ishift_in,
                                                   double* __restrict__ A, double* __restrict__ out, int ostride) {
  __shared__ double sh[256];
  const int t = blockIdx.x;
  const double* d = D + (size_t)t * nl;
  const double* s = S + (size_t)t * nl;
  const int ish = ishift_in[t];
  const double tshift = ish * dt;
  double p1 = 0.0, p2 = 0.0;
  for (int i = threadIdx.x; i < nl; i += 256) { p1 += d[i] * d[i]; p2 += s[i] * s[i]; }
  const double sd2 = wg_sum(p1, sh), ss2 = wg_sum(p2, sh);
  const double dlna = 0.5 * log(sd2 / ss2);
  const double ea = exp(dlna);
  auto scc = [&](int i) { const int j = i - ish; return (j >= 0 && j < nl) ? s[j] : 0.0; };
  auto sccd = [&](int i) { const int j = i - ish; return (j >= 0 && j < nl) ? ea * s[j] : 0.0; };
  auto grad = [&](auto f, int i) {
    if (nl < 2) return 0.0;
    if (i == 0) return (f(1) - f(0)) / dt;
    if (i == nl - 1) return (f(nl - 1) - f(nl - 2)) / dt;
    return (f(i + 1) - f(i - 1)) / (2.0 * dt);
  };
  auto dsh = [&](int i) {
    if (ish > 0 && i < ish) return 0.0;
    if (ish < 0 && i >= nl + ish) return 0.0;
    return d[i];
  };
  double q1 = 0.0, q2 = 0.0, q3 = 0.0, q4 = 0.0, q5 = 0.0;
  for (int i = threadIdx.x; i < nl; i += 256) {
    const double r = d[i] - sccd(i), v = grad(sccd, i), c = scc(i), e = dsh(i);
    q1 += r * r; q2 += v * v; q3 += c * c; q4 += e * c; q5 += e * e;
  }
  const double num = wg_sum(q1, sh), den = wg_sum(q2, sh), den2 = wg_sum(q3, sh);
  const double a = wg_sum(q4, sh), b = wg_sum(q5, sh);
  double sig_dt = sqrt(num / den), sig_dlna = sqrt(num / den2);
  if (sig_dt < cfg.dt_sigma_min || isnan(sig_dt)) sig_dt = cfg.dt_sigma_min;
  if (sig_dlna < cfg.dlna_sigma_min || isnan(sig_dlna)) sig_dlna = cfg.dlna_sigma_min;
  const double cmax = a / sqrt(b * den2);
  const bool sel = !(cmax < cfg.cc_min || tshift < cfg.tshift_min || tshift > cfg.tshift_max ||
                     dlna < cfg.dlna_min || dlna > cfg.dlna_max);
  if (threadIdx.x == 0) {
    double* o = out + ostride * (size_t)t;
    if (ostride >= 10) {  // both error estimates for the multitaper stage
      o[8] = sig_dt;
      o[9] = sig_dlna;
    }
    o[1] = imeas == 11 ? tshift : dlna;
    o[2] = imeas == 11 ? sig_dt : sig_dlna;
    o[3] = cmax;
    o[4] = tshift;
    o[5] = dlna;
    o[6] = sel ? 1.0 : 0.0;
    o[7] = (double)ish;
    o[0] = !sel ? 0.0
                : (imeas == 11 ? 0.5 * (tshift / sig_dt) * (tshift / sig_dt)
                               : 0.5 * (dlna / sig_dlna) * (dlna / sig_dlna));
  }
  if (!sel) return;
  // calc_cc_adjsrc: simpson (utils.f90:1630-1663) of ds/dt^2 (11) or s^2 (12) as a weighted sum
  auto sf = [&](int i) { return s[i]; };
  double pw = 0.0, tail = 0.0;
  for (int i = threadIdx.x; i < nl; i += 256) {
    const double x = imeas == 11 ? grad(sf, i) : s[i];
    const double y = x * x;
    double w;
    if (nl % 2 == 0) {
      // sum = y(1) + y(n-1) + sum over pairs 4 y(i) + 2 y(i+1), i = 2, 4, .., n-2 (1-based)
      w = (i == 0 ? 1.0 : 0.0) + (i == nl - 2 ? 1.0 : 0.0) + ((i % 2 == 1 && i <= nl - 3) ? 4.0 : 0.0) +
          ((i % 2 == 0 && i >= 2 && i <= nl - 2) ? 2.0 : 0.0);
      if (i >= nl - 2) tail += y;
    } else {
      w = (i == 0 || i == nl - 1) ? 1.0 : (i % 2 == 1 ? 4.0 : 2.0);
    }
    pw += w * y;
  }
  const double sw = wg_sum(pw, sh), st = wg_sum(tail, sh);
  double nn = 0.0;
  if (nl >= 2) nn = (dt / 3.0) * sw + (nl % 2 == 0 ? 0.5 * dt * st : 0.0);
  for (int i = threadIdx.x; i < nl; i += 256) {
    const double v = imeas == 11 ? grad(sf, i) * tshift / nn / (sig_dt * sig_dt)
                                 : -1.0 * s[i] * dlna / nn / (sig_dlna * sig_dlna);
    A[(size_t)t * nl + i] = v * tw[i];
  }
}

// CCCMisfit%calc_adjoint_source (ccc_misfit.f90:15-71) per trace (one
// workgroup on its tapered rows d, s): zero-lag correlation coefficient
// p / sqrt(E_d E_s) (fixed-order tree sums), selection by CC_MIN, misfit
// 1 - coef, adjoint source -(d - (p / E_s) s) / sqrt(E_d E_s) tapered into A.
// out[8] as k_cc_finish: misfit, residual (coef), error 0, cc_max (coef),
// tshift 0, dlna 0, selected, ishift 0
__global__ void __launch_bounds__(256) k_ccc_finish(int nl, double dt, double cc_min, const double* __restrict__ tw,
                                                    const double* __restrict__ D, const double* __restrict__ S,
                                                    double* __restrict__ A, double* __restrict__ out) {
  __shared__ double sh[256];
  const int t = blockIdx.x;
  const double* d = D + (size_t)t * nl;
  const double* s = S + (size_t)t * nl;
  double q1 = 0.0, q2 = 0.0, q3 = 0.0;
  for (int i = threadIdx.x; i < nl; i += 256) { q1 += d[i] * s[i]; q2 += d[i] * d[i]; q3 += s[i] * s[i]; }
  const double p = wg_sum(q1, sh) * dt, de = wg_sum(q2, sh) * dt, se = wg_sum(q3, sh) * dt;
  const double coef = p / sqrt(de * se);
  const bool sel = !(coef < cc_min);
  if (threadIdx.x == 0) {
    double* o = out + 8 * (size_t)t;
    o[0] = sel ? 1 - coef : 0.0;
    o[1] = coef;
    o[2] = 0.0;
    o[3] = coef;
    o[4] = 0.0;
    o[5] = 0.0;
    o[6] = sel ? 1.0 : 0.0;
    o[7] = 0.0;
  }
  if (!sel) return;
  for (int i = threadIdx.x; i < nl; i += 256) A[(size_t)t * nl + i] = -(d[i] - (p / se) * s[i]) / sqrt(de * se) * tw[i];
}

// selected traces' adjoint-source rows back into the sample-major traces
// (64 x 64 LDS tiles); unselected traces keep what earlier windows wrote
__global__ void __launch_bounds__(256) k_cc_put(int ntrace, int nb, int nl, const double* __restrict__ A,
                                                const double* __restrict__ out, float* __restrict__ adj) {
  __shared__ double tile[64][65];
  const int t0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
  for (int r = threadIdx.x / 64; r < 64; r += 4) {
    const int t = t0 + r, k = k0 + (threadIdx.x & 63);
    if (t < ntrace && k < nl) tile[r][threadIdx.x & 63] = A[(size_t)t * nl + k];
  }
  __syncthreads();
  for (int r = threadIdx.x / 64; r < 64; r += 4) {
    const int k = k0 + r, t = t0 + (threadIdx.x & 63);
    if (t < ntrace && k < nl && out[8 * (size_t)t + 6] != 0.0)
      adj[(size_t)(nb + k) * ntrace + t] = (float)tile[threadIdx.x & 63][r];
  }
}


// ---- multitaper family (mt_tt_misfit.f90, utils/dpss.f90)
//
// Real transforms of length nf = 2^p (64 <= nf <= 16384) inside one workgroup:
// the nf real samples are packed as M = nf/2 complex values (even + i odd) in
// LDS at bit-reversed positions, a radix-2 decimation-in-time transform of
// length M runs in place, and the spectrum X_k (k = 0..M) is unpacked in pairs
// (p, M - p), each pair owned by one lane.  Twiddles exp(-2 pi i j / nf),
// j < nf/2, come from a table in device memory.
constexpr int MT_THREADS = 1024;
constexpr int MT_MAXM = 8192;
constexpr int MT_PAIRS = MT_MAXM / 2 / MT_THREADS + 1;  // pair p = lane + 1024 r, p <= M/2

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <int NT, class F>
__device__ void mt_load_real(double2* buf, int M, int logM, int len, F f) {
#pragma unroll 2
  for (int m = threadIdx.x; m < M; m += NT) {
    const int n0 = 2 * m, n1 = 2 * m + 1;
    const double a = n0 < len ? f(n0) : 0.0, b = n1 < len ? f(n1) : 0.0;
    buf[__brev((unsigned)m) >> (32 - logM)] = make_double2(a, b);
  }
  __syncthreads();
}

// R consecutive radix-2 stages (half-spans h, 2h, .., 2^(R-1) h) fused: each
// lane takes groups of 2^R elements i0 + q h through the R stages in
// registers, with the same butterflies and twiddles as stage-by-stage, so the
// result is bit-for-bit that of the radix-2 transform with 1/R of the LDS
// passes and barriers
template <int NT, int R>
__device__ void mt_pass(double2* buf, int M, int lh, int nf, const double2* __restrict__ tw, bool inverse) {
  constexpr int G = 1 << R;
  const int h = 1 << lh;
#pragma unroll 1
  for (int g = threadIdx.x; g < M / G; g += NT) {
    const int pos = g & (h - 1);
    const int i0 = ((g >> lh) << (lh + R)) + pos;
    double2 x[G];
#pragma unroll
    for (int q = 0; q < G; ++q) x[q] = buf[i0 + q * h];
#pragma unroll
    for (int st = 0; st < R; ++st) {
      const int half = 1 << st;
      const int tstep = nf >> (lh + st + 1);  // nf / (2 half h)
#pragma unroll
      for (int q = 0; q < G; ++q) {
        if (q & half) continue;
        double2 w = tw[(pos + (q & (half - 1)) * h) * tstep];
        if (inverse) w.y = -w.y;
        const double2 a = x[q], c = cmul(x[q + half], w);
        x[q] = make_double2(a.x + c.x, a.y + c.y);
        x[q + half] = make_double2(a.x - c.x, a.y - c.y);
      }
    }
#pragma unroll
    for (int q = 0; q < G; ++q) buf[i0 + q * h] = x[q];
  }
  __syncthreads();
}

// in-place transform of bit-reversed buf (decimation in time, radix-2 stages
// fused three at a time); inverse uses the conjugate twiddles (unnormalised)
template <int NT>
__device__ void mt_fft(double2* buf, int M, int logM, int nf, const double2* __restrict__ tw, bool inverse) {
  int lh = 0;
  for (; lh + 3 <= logM; lh += 3) mt_pass<NT, 3>(buf, M, lh, nf, tw, inverse);
  if (logM - lh == 2) mt_pass<NT, 2>(buf, M, lh, nf, tw, inverse);
  else if (logM - lh == 1) mt_pass<NT, 1>(buf, M, lh, nf, tw, inverse);
}

// X_k = E - i W_k O with E = (Z_k + conj Z_{M-k}) / 2, O = (Z_k - conj Z_{M-k}) / 2
__device__ inline double2 mt_half(double2 zk, double2 zm, double2 w) {
  const double2 E = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
  const double2 O = make_double2(0.5 * (zk.x - zm.x), 0.5 * (zk.y + zm.y));
  const double2 wo = cmul(w, O);
  return make_double2(E.x + wo.y, E.y - wo.x);
}

// spectrum of the packed real sequence at k = p and k = M - p
__device__ inline void mt_unpack(const double2* buf, int M, const double2* __restrict__ tw, int p, double2& xa,
                                 double2& xb) {
  if (p == 0) {
    const double2 z = buf[0];
    xa = make_double2(z.x + z.y, 0.0);
    xb = make_double2(z.x - z.y, 0.0);
    return;
  }
  const double2 zp = buf[p], zq = buf[M - p];
  xa = mt_half(zp, zq, tw[p]);
  xb = mt_half(zq, zp, tw[M - p]);
}

// inverse packing: Z'_k = ((H_k + conj H_{M-k}) + i V_k (H_k - conj H_{M-k})) / 2,
// V_k = exp(+2 pi i k / nf), for a Hermitian spectrum H of a real sequence
__device__ inline double2 mt_repack(double2 hk, double2 hm, double2 w) {
  const double2 a = make_double2(hk.x + hm.x, hk.y - hm.y);
  const double2 b = make_double2(hk.x - hm.x, hk.y + hm.y);
  const double2 vb = cmul(make_double2(w.x, -w.y), b);
  return make_double2(0.5 * (a.x - vb.y), 0.5 * (a.y + vb.x));
}

template <int NT>
__device__ double mt_bsum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = NT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

template <int NT>
__device__ double mt_bmax(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = NT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// sample-major traces [nstep][ntrace] to trace-major rows [ntrace][nstep]
// through 64 x 64 LDS tiles (reads and writes coalesced)
__global__ void __launch_bounds__(256) k_rows_f32(int nstep, int ntrace, const float* __restrict__ x,
                                                  float* __restrict__ y) {
  __shared__ float tile[64][65];
  const int t0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  for (int r = threadIdx.x / 64; r < 64; r += 4) {
    const int n = n0 + r, t = t0 + (threadIdx.x & 63);
    if (n < nstep && t < ntrace) tile[r][threadIdx.x & 63] = x[(size_t)n * ntrace + t];
  }
  __syncthreads();
  for (int r = threadIdx.x / 64; r < 64; r += 4) {
    const int t = t0 + r, n = n0 + (threadIdx.x & 63);
    if (n < nstep && t < ntrace) y[(size_t)t * nstep + n] = tile[threadIdx.x & 63][r];
  }
}

// calculate_freq_limits (mt_tt_misfit.f90:615-707) of every trace's whole
// synthetic: amplitude spectrum, its first maximum, the water-level
// "thermostat" search up and down from it, the period and cycle bounds and
// the half-bandwidth test; flim[3 t] = nfreq_min, nfreq_max (1-based), accepted
__global__ void __launch_bounds__(MT_THREADS) k_mt_flim(int nstep, int ntrace, double dt, int nf, int logM,
                                                        const double2* __restrict__ tw, const float* __restrict__ synT,
                                                        double water_threshold, double min_period, double max_period,
                                                        double mt_nw, double min_cycle, int* __restrict__ flim) {
  __shared__ double2 buf[MT_MAXM];
  __shared__ double rv[MT_THREADS];
  __shared__ int ri[MT_THREADS];
  const int t = blockIdx.x, M = nf / 2;
  const float* row = synT + (size_t)t * nstep;
  mt_load_real<MT_THREADS>(buf, M, logM, nstep, [&](int n) { return (double)row[n]; });
  mt_fft<MT_THREADS>(buf, M, logM, nf, tw, false);
  // amplitudes |X_k dt| in place: k < M at buf[k].x, k = M at buf[0].y
  double best = -1.0;
  int bk = 0x7fffffff;
  for (int r = 0; r < MT_PAIRS; ++r) {
    const int p = threadIdx.x + MT_THREADS * r;
    if (p > M / 2) break;
    double2 xa, xb;
    mt_unpack(buf, M, tw, p, xa, xb);
    const double aa = hypot(xa.x * dt, xa.y * dt), ab = hypot(xb.x * dt, xb.y * dt);
    const int kb = M - p;
    if (aa > best || (aa == best && p < bk)) { best = aa; bk = p; }
    if (kb != p && (ab > best || (ab == best && kb < bk))) { best = ab; bk = kb; }
    buf[p].x = aa;
    if (p == 0) buf[0].y = ab;
    else if (kb != p) buf[kb].x = ab;
  }
  rv[threadIdx.x] = best;
  ri[threadIdx.x] = bk;
  __syncthreads();
  for (int o = MT_THREADS / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      const double v2 = rv[threadIdx.x + o];
      const int k2 = ri[threadIdx.x + o];
      if (v2 > rv[threadIdx.x] || (v2 == rv[threadIdx.x] && k2 < ri[threadIdx.x])) {
        rv[threadIdx.x] = v2;
        ri[threadIdx.x] = k2;
      }
    }
    __syncthreads();
  }
  const int fnum = M + 1;
  const double ampmax = rv[0];
  const int iamp = ri[0] + 1;
  __syncthreads();
  auto amp = [&](int i) { return i - 1 < M ? buf[i - 1].x : buf[0].y; };
  const double df = 1.0 / (nf * dt) - 0.0, tlen = nstep * dt;
  const double wl = water_threshold * ampmax;
  const int ifmin = max(1, (int)(1.0 / (max_period * df)));
  const int ifmax = min(fnum, (int)(1.0 / (min_period * df)));
  // block-wide integer min / max through ri
  auto ired = [&](int v, bool mx) {
    ri[threadIdx.x] = v;
    __syncthreads();
    for (int o = MT_THREADS / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) ri[threadIdx.x] = mx ? max(ri[threadIdx.x], ri[threadIdx.x + o])
                                                     : min(ri[threadIdx.x], ri[threadIdx.x + o]);
      __syncthreads();
    }
    const int r = ri[0];
    __syncthreads();
    return r;
  };
  // search_frequency_limit's "thermostat" over i = lo..hi in scan order (up:
  // ascending, else descending), starting in the searching state: a sample
  // below wl stops the search, one above 10 wl restarts it, and the limit is
  // the sample of the last change of state.  That sample is the first
  // decisive sample after the last decisive sample of the other kind (or the
  // first stop, when the scan never restarts); found with block reductions.
  auto thermostat = [&](int lo, int hi, bool up, int init) {
    const int none = up ? -1 : 0x7fffffff;
    auto later = [&](int x, int y) { return up ? max(x, y) : min(x, y); };
    int lastL = none, lastH = none;
    for (int i = lo + (int)threadIdx.x; i <= hi; i += MT_THREADS) {
      const double a = amp(i);
      if (a < wl) lastL = later(lastL, i);
      else if (a > 10 * wl) lastH = later(lastH, i);
    }
    lastL = ired(lastL, up);
    lastH = ired(lastH, up);
    const bool hasL = lastL != none, hasH = lastH != none;
    if (!hasL && !hasH) return init;
    const bool endL = hasL && (!hasH || later(lastL, lastH) == lastL);
    const int q = endL ? lastH : lastL;
    const bool hasQ = endL ? hasH : hasL;
    if (!hasQ && !endL) return init;  // never stopped
    // first sample of the final kind after q (or from the start when there is no q)
    int first = up ? 0x7fffffff : -1;
    for (int i = lo + (int)threadIdx.x; i <= hi; i += MT_THREADS) {
      if (hasQ && (up ? i <= q : i >= q)) continue;
      const double a = amp(i);
      const bool kind = endL ? (a < wl) : (a > 10 * wl);
      if (kind) first = up ? min(first, i) : max(first, i);
    }
    return ired(first, !up);
  };
  int nfmax = fnum, nfmin = 1;
  if (iamp < fnum) nfmax = thermostat(iamp + 1, fnum, true, fnum);
  nfmax = min(min(nfmax, ifmax), (int)(1.0 / (2.0 * dt) / df));
  if (iamp > 1) nfmin = thermostat(1, iamp - 1, false, 1);
  nfmin = max(max(nfmin, ifmin), max(1, (int)(min_cycle / tlen / df)));
  if (threadIdx.x != 0) return;
  flim[3 * t] = nfmin;
  flim[3 * t + 1] = nfmax;
  flim[3 * t + 2] = (nfmax - nfmin) * df < mt_nw / (4.0 * tlen) ? 0 : 1;
}

struct MtCfg {
  double wl_tf, dt_fac, err_fac, dt_max_scale, min_cycle, min_period, water_threshold;
  int ntaper, use_cc_error, use_mt_error;
};

// Per-trace rows of the jackknife errors (calculate_mt_error), used only with
// use_mt_error: the products D_j conj(S_j) and |S_j|^2 of every taper j
// ([ntaper][M + 1] per trace), the leave-one-out (dtau, dlnA) estimates
// ([ntaper][M + 1]) and the resulting (err_dtau, err_dlna) per frequency.
struct MtJack {
  double2* pd;
  double* pb;
  double2* xj;
  double2* er;
};

// Multitaper measurement of one window of every selected trace (one
// workgroup per trace), after the cross-correlation stage has written its
// result and adjoint source: prepare_data_for_mtm, calculate_multitaper (sum
// over the DPSS tapers of D conj(S) and |S|^2, water level, phase with
// process_cycle_skipping, dtau and dlnA per frequency),
// check_mtm_time_shift_acceptability, calculate_freq_domain_taper, the
// Simpson misfits and calculate_mt_adjsrc.  A trace that fails a test keeps
// the cross-correlation result, as the reference falls back to it.
// Frequency k = lane + 1024 r of the M + 1 (k = M: Nyquist, lane 0); the
// per-frequency state lives in three device rows of M + 1 per trace, each
// lane touching only its own frequencies (coalesced, no cross-lane hazard):
//   G0 (double2): the synthetic's spectrum of the current taper, then dlnA,
//                 then the weighted adjoint spectrum H
//   G1 (double2): sum D conj(S), then (dtau, w), then the adjoint factors
//   G2 (double) : sum |S|^2, then the adjoint normalisation sum
// With use_mt_error the jackknife of calculate_mt_error forms, for every
// taper left out in turn, the multitaper estimate of the other tapers from the
// stored per-taper products (summed in taper order from zero, as the
// reference recomputes them), then the pseudo-value mean and the spread per
// frequency; the errors enter the acceptance test and, without use_cc_error,
// weight the misfit and adjoint source (with the water level of
// calculate_freq_domain_taper).
// out[10] per trace: misfit, residual, error, cc_max, tshift, dlna, selected,
// ishift, then (on entry) sigma_dt, sigma_dlna and (on exit) the measurement
// used and the other misfit.
__global__ void __launch_bounds__(MT_THREADS) k_mt_window(int nstep, int ntrace, double dt, int nf, int logM,
                                                          const double2* __restrict__ tw, int nb, int nl, int imeas,
                                                          MtCfg cfg, const double* __restrict__ wtap,
                                                          const double* __restrict__ tap,
                                                          const float* __restrict__ obsT,
                                                          const double* __restrict__ S, const int* __restrict__ flim,
                                                          double2* __restrict__ G0, double2* __restrict__ G1,
                                                          double* __restrict__ G2, double* __restrict__ A,
                                                          double* __restrict__ out, MtJack J) {
  __shared__ double2 buf[MT_MAXM];
  __shared__ double red[MT_THREADS];
  __shared__ double wtr[2];
  const double pi = 3.141592653589793;
  const int t = blockIdx.x, M = nf / 2, lane = threadIdx.x;
  const bool tt = imeas == 13;
  double* o = out + 10 * (size_t)t;
  const bool sel = o[6] != 0.0;
  const double tshift = o[4], dlna = o[5], sig_dt = o[8], sig_dlna = o[9];
  bool mt = sel;
  if (mt && (fabs(tshift) <= dt || cfg.min_cycle * cfg.min_period > nl)) mt = false;
  const int ish = (int)(tshift / dt);
  const int nbd = max(1, nb + 1 + ish), ned = min(nstep, nb + nl + ish);
  if (mt && ned - nbd + 1 != nl) mt = false;
  const int nfmin = flim[3 * t], nfmax = flim[3 * t + 1];
  if (mt && flim[3 * t + 2] == 0) mt = false;
  const double mp_cc = 0.5 * (tshift / sig_dt) * (tshift / sig_dt);
  const double mq_cc = 0.5 * (dlna / sig_dlna) * (dlna / sig_dlna);
  __syncthreads();
  if (!mt || nfmax > M + 1 || nfmin < 1) {
    if (lane == 0) {
      o[8] = sel ? (tt ? 11.0 : 12.0) : 0.0;
      o[9] = sel ? (tt ? mq_cc : mp_cc) : 0.0;
    }
    return;
  }
  const double df = 1.0 / (nf * dt) - 0.0;
  const double em = exp(-dlna);
  const double* s = S + (size_t)t * nl;
  const float* drow = obsT + (size_t)t * nstep + (nbd - 1);  // data window shifted by the CC lag
  double2* g0 = G0 + (size_t)t * (M + 1);
  double2* g1 = G1 + (size_t)t * (M + 1);
  double* g2 = G2 + (size_t)t * (M + 1);
  const bool jk = cfg.use_mt_error != 0;
  const size_t rows = (size_t)cfg.ntaper * (M + 1);
  double2* pd = jk ? J.pd + (size_t)t * rows : nullptr;
  double* pb = jk ? J.pb + (size_t)t * rows : nullptr;
  double2* xj = jk ? J.xj + (size_t)t * rows : nullptr;
  double2* er = jk ? J.er + (size_t)t * (M + 1) : nullptr;
  // spectrum of the packed real sequence at 0 <= k < M, and at k = M
  auto xk = [&](int k) {
    if (k == M) { const double2 z = buf[0]; return make_double2(z.x - z.y, 0.0); }
    if (k == 0) { const double2 z = buf[0]; return make_double2(z.x + z.y, 0.0); }
    return mt_half(buf[k], buf[M - k], tw[k]);
  };
  // ---- calculate_multitaper
  for (int j = 0; j < cfg.ntaper; ++j) {
    const double* tj = tap + (size_t)j * nl;
    mt_load_real<MT_THREADS>(buf, M, logM, nl, [&](int n) { return s[n] * tj[n]; });
    mt_fft<MT_THREADS>(buf, M, logM, nf, tw, false);
    for (int k = lane; k <= M; k += MT_THREADS) g0[k] = xk(k);
    __syncthreads();
    mt_load_real<MT_THREADS>(buf, M, logM, nl, [&](int n) {
      return (double)drow[n] * em * wtap[n] * tj[n];
    });
    mt_fft<MT_THREADS>(buf, M, logM, nf, tw, false);
    for (int k = lane; k <= M; k += MT_THREADS) {
      const double2 x = xk(k), y = g0[k];
      const double a = x.x * dt, b = x.y * dt, c = y.x * dt, d = -(y.y * dt);
      double2 tp = j == 0 ? make_double2(0.0, 0.0) : g1[k];
      double bt = j == 0 ? 0.0 : g2[k];
      const double2 pj = make_double2(a * c - b * d, a * d + b * c);
      const double bj = c * c - d * (-d);
      tp.x = tp.x + pj.x;
      tp.y = tp.y + pj.y;
      bt = bt + bj;
      g1[k] = tp;
      g2[k] = bt;
      if (jk) {
        pd[(size_t)j * (M + 1) + k] = pj;
        pb[(size_t)j * (M + 1) + k] = bj;
      }
    }
    __syncthreads();
  }
  double lm = 0.0;
  for (int k = lane; k <= M; k += MT_THREADS) lm = fmax(lm, fabs(g2[k]));
  const double wlevel = mt_bmax<MT_THREADS>(lm, red) * (cfg.wl_tf * cfg.wl_tf);
  // transfer function in the band: phase to LDS for the cycle-skipping pass, log amplitude + dlnA to G0
  double* phiL = reinterpret_cast<double*>(buf);
  for (int k = lane; k <= M; k += MT_THREADS) {
    double lw = 0.0;
    const int i = k + 1;
    if (i >= nfmin && i <= nfmax) {
      const double bt = g2[k];
      const double2 tp = g1[k];
      const double den = fabs(bt) >= wlevel ? bt : bt + wlevel;
      const double tr = tp.x / den, ti = tp.y / den;
      phiL[k] = atan2(ti, tr);
      lw = log(hypot(tr, ti)) + dlna;
    }
    g0[k] = make_double2(lw, 0.0);
  }
  __syncthreads();
  // process_cycle_skipping (signal.f90:510-598), phase_step 1.5, on one lane
  const auto cycle_skip = [&]() {
    if (lane != 0) return;
    for (int iw = nfmin + 2; iw <= nfmax - 1; ++iw) {
      const double p0 = phiL[iw - 2], p1 = phiL[iw - 1], p2 = phiL[iw];
      const double s0 = fabs(p2 + p0 - 2.0 * p1);
      const double s1 = fabs((p2 + 2.0 * pi) + p0 - 2.0 * p1);
      const double s2 = fabs((p2 - 2.0 * pi) + p0 - 2.0 * p1);
      if (fabs(p1 - p2) > 1.5) {
        if (s1 < s0 && s1 < s2)
          for (int i = iw + 1; i <= nfmax; ++i) phiL[i - 1] = phiL[i - 1] + 2.0 * pi;
        if (s2 < s0 && s2 < s1)
          for (int i = iw + 1; i <= nfmax; ++i) phiL[i - 1] = phiL[i - 1] - 2.0 * pi;
      }
    }
  };
  cycle_skip();
  __syncthreads();
  // dtau, the acceptance test, the frequency-domain taper weights -> G1 = (dtau, w)
  const int L = nfmax - nfmin, frac = (int)(L * 1.0 / 2.0), fden = 2 * frac - 1;
  double bad = 0.0, wsum = 0.0;
  for (int k = lane; k <= M; k += MT_THREADS) {
    const int i = k + 1;
    const double f = i <= nf / 2 ? (double)(i - 1) / (nf * dt) : -(double)(nf - i + 1) / (nf * dt);
    const double wvec = 2.0 * pi * f;
    double dtau = 0.0;
    if (i == 1) dtau = tshift;
    else if (i >= nfmin && i <= nfmax) dtau = -(1.0 / wvec) * phiL[k] + tshift;
    if (i >= nfmin && i <= nfmax && i * df > 0.0) {
      // check_mtm_time_shift_acceptability; the multitaper errors are zero
      const double per = 1.0 / (i * df);
      if (fabs(dtau) > per / cfg.dt_fac || fabs(dtau) > cfg.dt_max_scale * fabs(tshift)) bad = 1.0;
    }
    double w = 0.0;
    if (i >= nfmin && i <= nfmax - 1) {
      // calculate_freq_domain_taper: 100 % cosine taper over [nfreq_min, nfreq_max - 1]
      const int q = i - nfmin;
      w = 1.0;
      if (frac > 0) {
        if (q < frac) w = cos(pi * (double)q / fden - pi / 2.0);
        else if (q >= L - frac) w = cos(pi * (double)(frac + q - (L - frac)) / fden - pi / 2.0);
      }
      wsum += w;
    }
    g1[k] = make_double2(dtau, w);
  }
  if (jk) {
    __syncthreads();  // phiL is reused below
    // calculate_mt_error: the estimate without taper q, for q in order
    for (int q = 0; q < cfg.ntaper; ++q) {
      double lq = 0.0;
      for (int k = lane; k <= M; k += MT_THREADS) {
        double bt = 0.0;
        for (int j = 0; j < cfg.ntaper; ++j)
          if (j != q) bt = bt + pb[(size_t)j * (M + 1) + k];
        lq = fmax(lq, fabs(bt));
      }
      const double wlq = mt_bmax<MT_THREADS>(lq, red) * (cfg.wl_tf * cfg.wl_tf);
      for (int k = lane; k <= M; k += MT_THREADS) {
        const int i = k + 1;
        double lw = 0.0;
        if (i >= nfmin && i <= nfmax) {
          double2 tp = make_double2(0.0, 0.0);
          double bt = 0.0;
          for (int j = 0; j < cfg.ntaper; ++j) {
            if (j == q) continue;
            const double2 p = pd[(size_t)j * (M + 1) + k];
            tp.x = tp.x + p.x;
            tp.y = tp.y + p.y;
            bt = bt + pb[(size_t)j * (M + 1) + k];
          }
          const double den = fabs(bt) >= wlq ? bt : bt + wlq;
          const double tr = tp.x / den, ti = tp.y / den;
          phiL[k] = atan2(ti, tr);
          lw = log(hypot(tr, ti)) + dlna;
        }
        xj[(size_t)q * (M + 1) + k] = make_double2(0.0, lw);
      }
      __syncthreads();
      cycle_skip();
      __syncthreads();
      for (int k = lane; k <= M; k += MT_THREADS) {
        const int i = k + 1;
        const double f = i <= nf / 2 ? (double)(i - 1) / (nf * dt) : -(double)(nf - i + 1) / (nf * dt);
        double dq = 0.0;
        if (i == 1) dq = tshift;
        else if (i >= nfmin && i <= nfmax) dq = -(1.0 / (2.0 * pi * f)) * phiL[k] + tshift;
        xj[(size_t)q * (M + 1) + k].x = dq;
      }
      __syncthreads();
    }
    // pseudo-value means and the spread of the estimates about them
    const double fn = (double)cfg.ntaper, fn1 = (double)(cfg.ntaper - 1);
    const double nn = (double)(cfg.ntaper * (cfg.ntaper - 1));
    for (int k = lane; k <= M; k += MT_THREADS) {
      const int i = k + 1;
      double2 e = make_double2(0.0, 0.0);
      if (i >= nfmin && i <= nfmax) {
        const double xt = g1[k].x, xl = g0[k].x;
        double et = 0.0, el = 0.0;
        for (int q = 0; q < cfg.ntaper; ++q) {
          const double2 x = xj[(size_t)q * (M + 1) + k];
          et = et + fn * xt - fn1 * x.x;
          el = el + fn * xl - fn1 * x.y;
        }
        et = et / fn;
        el = el / fn;
        double st = 0.0, sl = 0.0;
        for (int q = 0; q < cfg.ntaper; ++q) {
          const double2 x = xj[(size_t)q * (M + 1) + k];
          st = st + (x.x - et) * (x.x - et);
          sl = sl + (x.y - el) * (x.y - el);
        }
        e = make_double2(sqrt(st / nn), sqrt(sl / nn));
        if (i * df > 0.0 && e.x > (1.0 / (i * df)) / cfg.err_fac) bad = 1.0;
      }
      er[k] = e;
    }
  }
  if (mt_bmax<MT_THREADS>(bad, red) != 0.0) {
    if (lane == 0) {
      o[8] = tt ? 11.0 : 12.0;
      o[9] = tt ? mq_cc : mp_cc;
    }
    return;
  }
  const double ffac = 2.0 * df * mt_bsum<MT_THREADS>(wsum, red);
  if (jk && !cfg.use_cc_error) {
    // water level of the multitaper errors: water_threshold times the mean
    // |dtau|, |dlnA| over [nfreq_min, nfreq_max - 1], summed in order
    if (lane == 0) {
      double st = 0.0, sl = 0.0;
      for (int i = nfmin; i <= nfmax - 1; ++i) {
        st = st + fabs(g1[i - 1].x);
        sl = sl + fabs(g0[i - 1].x);
      }
      wtr[0] = cfg.water_threshold * st / (double)(nfmax - nfmin);
      wtr[1] = cfg.water_threshold * sl / (double)(nfmax - nfmin);
    }
    __syncthreads();
  }
  // Simpson misfits over the nf frequency samples (only k <= nf/2 < nf - 2 can be
  // non-zero); G1 becomes the adjoint factors (dtau, wp) or (dlnA, wq)
  double sp = 0.0, sq = 0.0;
  for (int k = lane; k <= M; k += MT_THREADS) {
    const double2 dw = g1[k];
    const double lw = g0[k].x;
    double wp = dw.y / ffac, wq = dw.y / ffac;
    if (cfg.use_cc_error) {
      wp = wp / (sig_dt * sig_dt);
      wq = wq / (sig_dlna * sig_dlna);
    } else if (jk && k + 1 >= nfmin && k + 1 <= nfmax) {
      const double2 e = er[k];
      const double et = e.x < wtr[0] ? e.x + wtr[0] : e.x;
      const double el = e.y < wtr[1] ? e.y + wtr[1] : e.y;
      wp = wp / (et * et);
      wq = wq / (el * el);
    }
    const double sw = k == 0 ? 1.0 : (k % 2 == 1 ? 4.0 : 2.0);
    sp += sw * (dw.x * dw.x * wp);
    sq += sw * (lw * lw * wq);
    g1[k] = tt ? make_double2(dw.x, wp) : make_double2(lw, wq);
  }
  const double mp = (df / 3.0) * mt_bsum<MT_THREADS>(sp, red);
  const double mq = (df / 3.0) * mt_bsum<MT_THREADS>(sq, red);
  // ---- calculate_mt_adjsrc for the measurement asked (phase 13: d/dt of the
  // tapered synthetic, amplitude 14: the tapered synthetic)
  auto xin = [&](const double* tj, int n) {
    if (!tt) return s[n] * tj[n];
    if (nl < 2) return 0.0;
    if (n == 0) return (s[1] * tj[1] - s[0] * tj[0]) / dt;
    if (n == nl - 1) return (s[nl - 1] * tj[nl - 1] - s[nl - 2] * tj[nl - 2]) / dt;
    return (s[n + 1] * tj[n + 1] - s[n - 1] * tj[n - 1]) / (2.0 * dt);
  };
  for (int j = 0; j < cfg.ntaper; ++j) {
    const double* tj = tap + (size_t)j * nl;
    mt_load_real<MT_THREADS>(buf, M, logM, nl, [&](int n) { return xin(tj, n); });
    mt_fft<MT_THREADS>(buf, M, logM, nf, tw, false);
    for (int k = lane; k <= M; k += MT_THREADS) {
      const double2 x = xk(k);
      const double a = x.x * dt, b = x.y * dt;
      g2[k] = (j == 0 ? 0.0 : g2[k]) + (a * a - b * (-b));
    }
    __syncthreads();
  }
  double* arow = A + (size_t)t * nl;
  const unsigned shb = 32 - logM;
  for (int j = 0; j < cfg.ntaper; ++j) {
    const double* tj = tap + (size_t)j * nl;
    mt_load_real<MT_THREADS>(buf, M, logM, nl, [&](int n) { return xin(tj, n); });
    mt_fft<MT_THREADS>(buf, M, logM, nf, tw, false);
    // weighted one-sided spectrum P_k and its Hermitian part H_k (Re ifft(P) =
    // ifft(H)) go to G0: the transform in LDS is still being read
    for (int k = lane; k <= M; k += MT_THREADS) {
      double2 P = make_double2(0.0, 0.0);
      const int i = k + 1;
      const double bs = g2[k];
      if (i >= nfmin && i <= nfmax && fabs(bs) > 0.0) {
        const double2 x = xk(k), f = g1[k];
        const double a = x.x * dt, b = x.y * dt;
        P = tt ? make_double2(a / bs, b / bs) : make_double2(-(a / bs), -(b / bs));
        P = make_double2(P.x * f.x * f.y, P.y * f.x * f.y);
      }
      g0[k] = (k == 0 || k == M) ? make_double2(P.x, 0.0) : make_double2(0.5 * P.x, 0.5 * P.y);
    }
    __syncthreads();
    for (int k = lane; k < M; k += MT_THREADS)
      buf[__brev((unsigned)k) >> shb] = mt_repack(g0[k], g0[M - k], tw[k]);
    __syncthreads();
    mt_fft<MT_THREADS>(buf, M, logM, nf, tw, true);
    const bool last = j == cfg.ntaper - 1;
    for (int n = lane; n < nl && n < 2 * M; n += MT_THREADS) {
      const double2 z = buf[n >> 1];
      const double y = ((n & 1) ? z.y : z.x) / M;
      double v = (j == 0 ? 0.0 : arow[n]) + y * 2.0 / dt * tj[n];
      if (last) v = v * wtap[n];
      arow[n] = v;
    }
    __syncthreads();
  }
  if (lane == 0) {
    o[0] = tt ? mp : mq;
    o[8] = (double)imeas;
    o[9] = tt ? mq : mp;
  }
}

// Exponentiated phase (exponentiated_phase_misfit.f90:21-104) of one window
// of every trace, one workgroup per trace on its tapered rows d, s.  hilbert
// (fftpack.f90:135-201) on nfe = 2 ** exponent(real(nl)) points: the real
// transform of the row (LDS, as the multitaper pass), then Re(xa) and Im(xa)
// as two real-output inverse transforms of the Hermitian parts of h X and
// -i h X (h: 1 at DC, 2 on positive frequencies, at the Nyquist 1 for an
// even and 0 for an odd row length).  Envelopes with water level, the real
// and imaginary differences (sums and Simpson misfits by fixed-order tree
// sums), then the adjoint source with the Hilbert transforms of
// dr s H / E^3 and di s^2 / E^3; both parts tapered and added into A.
// Scratch rows per trace: G0, G1 (nfe/2 + 1 double2), R0..R3 (nl doubles).
// out[8]: misfit, residual, misfit_real, misfit_imag, residual_real,
// residual_imag, selected (1), 0.
__global__ void __launch_bounds__(MT_THREADS) k_ep_window(int nl, double dt, int nfe, int logM, double wtr_env,
                                                          const double2* __restrict__ tw,
                                                          const double* __restrict__ wtap,
                                                          const double* __restrict__ D, const double* __restrict__ S,
                                                          double2* __restrict__ G0, double2* __restrict__ G1,
                                                          double* __restrict__ R, double* __restrict__ A,
                                                          double* __restrict__ out) {
  __shared__ double2 buf[MT_MAXM];
  __shared__ double red[MT_THREADS];
  const int t = blockIdx.x, M = nfe / 2, lane = threadIdx.x, T = gridDim.x;
  const unsigned shb = 32 - logM;
  const double* s = S + (size_t)t * nl;
  const double* d = D + (size_t)t * nl;
  double2* g0 = G0 + (size_t)t * (M + 1);
  double2* g1 = G1 + (size_t)t * (M + 1);
  double* r0 = R + (size_t)t * nl;           // Re xa(s), then -(dr H^2 / E^3)
  double* r1 = R + ((size_t)T + t) * nl;     // Im xa(s), then di s H / E^3
  double* r2 = R + (2 * (size_t)T + t) * nl;  // Re xa(d), then dr s H / E^3
  double* r3 = R + (3 * (size_t)T + t) * nl;  // Im xa(d), then di s^2 / E^3
  const double hnyq = (nl % 2 == 0) ? 1.0 : 0.0;
  auto xk = [&](int k) {
    if (k == M) { const double2 z = buf[0]; return make_double2(z.x - z.y, 0.0); }
    if (k == 0) { const double2 z = buf[0]; return make_double2(z.x + z.y, 0.0); }
    return mt_half(buf[k], buf[M - k], tw[k]);
  };
  // real-output inverse of the Hermitian spectrum in g (k = 0..M) into y[0..nl)
  auto inverse_to = [&](double2* g, double* y) {
    for (int k = lane; k < M; k += MT_THREADS)
      buf[__brev((unsigned)k) >> shb] = mt_repack(g[k], g[M - k], tw[k]);
    __syncthreads();
    mt_fft<MT_THREADS>(buf, M, logM, nfe, tw, true);
    for (int n = lane; n < nl; n += MT_THREADS) {
      const double2 z = buf[n >> 1];
      y[n] = ((n & 1) ? z.y : z.x) / M;
    }
    __syncthreads();
  };
  // hilbert of row x: real part to yr (when wanted) and imaginary part to yi
  auto hilbert = [&](const double* x, double* yr, double* yi) {
    mt_load_real<MT_THREADS>(buf, M, logM, nl, [&](int n) { return x[n]; });
    mt_fft<MT_THREADS>(buf, M, logM, nfe, tw, false);
    for (int k = lane; k <= M; k += MT_THREADS) {
      const double2 X = xk(k);
      // Hermitian parts of h X and -i h X
      if (yr) g0[k] = k == 0 ? make_double2(X.x, 0.0) : (k == M ? make_double2(hnyq * X.x, 0.0) : X);
      g1[k] = (k == 0 || k == M) ? make_double2(0.0, 0.0) : make_double2(X.y, -X.x);
    }
    __syncthreads();
    if (yr) inverse_to(g0, yr);
    inverse_to(g1, yi);
  };
  hilbert(s, r0, r1);
  hilbert(d, r2, r3);
  double ms = 0.0, md = 0.0;
  for (int n = lane; n < nl; n += MT_THREADS) {
    ms = fmax(ms, hypot(r0[n], r1[n]));
    md = fmax(md, hypot(r2[n], r3[n]));
  }
  const double ts = wtr_env * mt_bmax<MT_THREADS>(ms, red), td = wtr_env * mt_bmax<MT_THREADS>(md, red);
  double rr = 0.0, ri = 0.0, sr = 0.0, si = 0.0;
  for (int n = lane; n < nl; n += MT_THREADS) {
    const double hs = r1[n], esw = hypot(r0[n], hs) + ts, edw = hypot(r2[n], r3[n]) + td;
    const double dr = d[n] / edw - s[n] / esw;
    const double di = r3[n] / edw - hs / esw;
    rr += dr;
    ri += di;
    // simpson (utils.f90:1630-1663) as a weighted sum over the nl samples
    double w, tail = 0.0;
    if (nl % 2 == 0) {
      w = (n == 0 ? 1.0 : 0.0) + (n == nl - 2 ? 1.0 : 0.0) + ((n % 2 == 1 && n <= nl - 3) ? 4.0 : 0.0) +
          ((n % 2 == 0 && n >= 2 && n <= nl - 2) ? 2.0 : 0.0);
      if (n >= nl - 2) tail = 1.0;
    } else {
      w = (n == 0 || n == nl - 1) ? 1.0 : (n % 2 == 1 ? 4.0 : 2.0);
    }
    sr += (w * (dt / 3.0) + tail * 0.5 * dt) * (dr * dr);
    si += (w * (dt / 3.0) + tail * 0.5 * dt) * (di * di);
    const double cub = esw * esw * esw;
    r0[n] = -1.0 * (dr * (hs * hs) / cub);
    r2[n] = dr * s[n] * hs / cub;
    r1[n] = di * s[n] * hs / cub;
    r3[n] = di * (s[n] * s[n]) / cub;
  }
  rr = mt_bsum<MT_THREADS>(rr, red) / nl;
  ri = mt_bsum<MT_THREADS>(ri, red) / nl;
  const double mr = 0.5 * (nl >= 2 ? mt_bsum<MT_THREADS>(sr, red) : 0.0);
  const double mi = 0.5 * (nl >= 2 ? mt_bsum<MT_THREADS>(si, red) : 0.0);
  if (lane == 0) {
    double* o = out + 8 * (size_t)t;
    o[0] = mr + mi;
    o[1] = rr + ri;
    o[2] = mr;
    o[3] = mi;
    o[4] = rr;
    o[5] = ri;
    o[6] = 1.0;
    o[7] = 0.0;
  }
  // Im hilbert of dr s H / E^3 (r2 -> r2) and of di s^2 / E^3 (r3 -> r3)
  hilbert(r2, nullptr, r2);
  hilbert(r3, nullptr, r3);
  double* a = A + (size_t)t * nl;
  for (int n = lane; n < nl; n += MT_THREADS) {
    const double ar = r0[n] - r2[n], ai = r1[n] + r3[n];
    a[n] = ar * wtap[n] + ai * wtap[n];
  }
}

// ---- cross-convolution (waveform_conv_misfit.f90, IMEAS_WAVEFORM_CONV)

// samples kb[r] + j (j < nout[r]) of the full convolution of row r of X (n1,
// optionally reversed and signed) with row r of Y (n2): each output summed over
// X's index in increasing order as myconvolution_dp; workgroup (row, 1024
// outputs), four consecutive outputs per lane on a sliding register window of
// Y, X and the matching stretch of Y through LDS in 1024-sample chunks (Y
// zero outside the row, so an output only gains exact zeros).  mode 0: O = c
// scale; mode 1: O = (O - c) scale (the difference of two convolutions)
constexpr int CV_OUT = 1024, CV_CHUNK = 1024;
__global__ void __launch_bounds__(256) k_conv_mid(int n1, int n2, const double* __restrict__ X, int xrev,
                                                  double xsign, const double* __restrict__ Y,
                                                  const int* __restrict__ kb, const int* __restrict__ nout,
                                                  int ostride, int mode, double scale, double* __restrict__ O) {
  __shared__ double sx[CV_CHUNK], sy[CV_CHUNK + CV_OUT];
  const int r = blockIdx.y, no = nout[r], j0 = blockIdx.x * CV_OUT;
  if (j0 >= no) return;
  const double* x = X + (size_t)r * n1;
  const double* y = Y + (size_t)r * n2;
  const int kbase = kb[r] + j0, jl = j0 + 4 * (int)threadIdx.x, kl = kb[r] + jl;
  const int kmaxb = kb[r] + min(j0 + CV_OUT, no) - 1;
  const int ablo = max(0, kbase - (n2 - 1)), abhi = min(n1 - 1, kmaxb);
  const int alo = max(0, kl - (n2 - 1)), ahi = min(n1 - 1, kl + 3);
  double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
  for (int a0 = ablo; a0 <= abhi; a0 += CV_CHUNK) {
    const int aend = min(a0 + CV_CHUNK - 1, abhi);
    const int ybase = kbase - (a0 + CV_CHUNK - 1);
    __syncthreads();
    for (int i = threadIdx.x; i < CV_CHUNK; i += 256) {
      const int a = a0 + i;
      sx[i] = a <= aend ? xsign * x[xrev ? n1 - 1 - a : a] : 0.0;
    }
    for (int i = threadIdx.x; i < CV_CHUNK + CV_OUT; i += 256) {
      const int j = ybase + i;
      sy[i] = (j >= 0 && j < n2) ? y[j] : 0.0;
    }
    __syncthreads();
    if (jl < no) {
      const int a = max(alo, a0), b = min(ahi, aend);
      const int off = kl - ybase;  // y[kl + q - a] sits at sy[off + q - a], in [0, CV_CHUNK + CV_OUT - 2]
      if (a <= b) {
        double r0 = sy[off - a], r1 = sy[off + 1 - a], r2 = sy[off + 2 - a], r3 = sy[off + 3 - a];
        for (int i = a; i <= b; ++i) {
          const double xv = sx[i - a0];
          c0 = c0 + xv * r0;
          c1 = c1 + xv * r1;
          c2 = c2 + xv * r2;
          c3 = c3 + xv * r3;
          r3 = r2;
          r2 = r1;
          r1 = r0;
          r0 = i < b ? sy[off - i - 1] : 0.0;
        }
      }
    }
  }
  const double cv[4] = {c0, c1, c2, c3};
  for (int q = 0; q < 4; ++q) {
    const int j = jl + q;
    if (j >= no) break;
    double* o = O + (size_t)r * ostride + j;
    *o = mode == 0 ? cv[q] * scale : (*o - cv[q]) * scale;
  }
}

// window_taper weight of sample i of an npts-sample window (signal.f90:419-491)
__device__ double taper_w(int i, int npts, double pct, int type) {
  const int frac = (pct == 0.0 || pct == 1.0) ? (int)(npts * pct / 2.0) : (int)(npts * pct / 2.0 + 0.5);
  if (frac <= 0) return 1.0;
  int rr;
  if (i < frac) rr = i;
  else if (i >= npts - frac) rr = frac + (i - (npts - frac));
  else return 1.0;
  const double pi = 3.14159265358979323846;
  const int den = 2 * frac - 1;
  switch (type) {
    case 1: return 0.5 - 0.5 * cos(2.0 * pi * rr / den);
    case 2: return 0.54 - 0.46 * cos(2.0 * pi * rr / den);
    case 3: return cos(pi * rr / den - pi / 2.0);
    default: return 1.0 - pow(cos(pi * rr / den), 10.0);
  }
}

// per receiver (one workgroup): the tapered windowed difference, its Simpson
// misfit and mean residual, and both tapered adjoint sources into the
// sample-major outputs
__global__ void __launch_bounds__(256) k_conv_finish(int nstep, int nrec, double dt, const int* __restrict__ nbw,
                                                     const int* __restrict__ nlw, double pct, int type,
                                                     const double* __restrict__ CD, const double* __restrict__ AR,
                                                     const double* __restrict__ AZ, double* __restrict__ out,
                                                     float* __restrict__ adj_r, float* __restrict__ adj_z) {
  __shared__ double sh[256];
  const int r = blockIdx.x, nb = nbw[r], nl = nlw[r];
  const double* cd = CD + (size_t)r * nstep + nb;
  double sq = 0.0, sm = 0.0;
  for (int i = threadIdx.x; i < nl; i += 256) {
    const double w = taper_w(i, nl, pct, type);
    const double v = cd[i] * w;
    double wt, tail = 0.0;
    if (nl % 2 == 0) {
      wt = (i == 0 ? 1.0 : 0.0) + (i == nl - 2 ? 1.0 : 0.0) + ((i % 2 == 1 && i <= nl - 3) ? 4.0 : 0.0) +
           ((i % 2 == 0 && i >= 2 && i <= nl - 2) ? 2.0 : 0.0);
      if (i >= nl - 2) tail = 1.0;
    } else {
      wt = (i == 0 || i == nl - 1) ? 1.0 : (i % 2 == 1 ? 4.0 : 2.0);
    }
    sq += (wt * (dt / 3.0) + tail * 0.5 * dt) * (v * v);
    sm += v;
    adj_r[(size_t)(nb + i) * nrec + r] = (float)(AR[(size_t)r * nstep + i] * w);
    adj_z[(size_t)(nb + i) * nrec + r] = (float)(AZ[(size_t)r * nstep + i] * w);
  }
  const double mf = wg_sum(sq, sh), rs = wg_sum(sm, sh);
  if (threadIdx.x == 0) {
    out[2 * r] = nl >= 2 ? 0.5 * mf : 0.0;
    out[2 * r + 1] = rs / nl;
  }
}

// float sample-major traces [nstep][ntrace] to double trace-major rows
__global__ void __launch_bounds__(256) k_rows_f64(int nstep, int ntrace, const float* __restrict__ x,
                                                  double* __restrict__ y) {
  __shared__ float tile[64][65];
  const int t0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  for (int r = threadIdx.x / 64; r < 64; r += 4) {
    const int n = n0 + r, t = t0 + (threadIdx.x & 63);
    if (n < nstep && t < ntrace) tile[r][threadIdx.x & 63] = x[(size_t)n * ntrace + t];
  }
  __syncthreads();
  for (int r = threadIdx.x / 64; r < 64; r += 4) {
    const int t = t0 + r, n = n0 + (threadIdx.x & 63);
    if (n < nstep && t < ntrace) y[(size_t)t * nstep + n] = (double)tile[threadIdx.x & 63][r];
  }
}

// ---- window selection (window_selector/win_sel.f90, WIN_SELECTOR_TYPE)

struct WsCfg {
  double win_len_fac, min_velocity, thr_corr, thr_shift_fac, jump_fac, min_win_len_fac, min_snr, min_energy_ratio;
  int min_peaks, split;
};

// per trace (one workgroup): max|syn| normalisation of the trace-major rows
// (double), and the noise level sqrt(sum(dat(1:nstart)^2) / nstart) of the
// un-normalised data as initialize computes it (win_sel.f90:56-74)
__global__ void __launch_bounds__(256) k_ws_prep(int nstep, double dt, double t0, double min_period,
                                                 const double* __restrict__ tp, const float* __restrict__ obsT,
                                                 const float* __restrict__ synT, double* __restrict__ Dn,
                                                 double* __restrict__ Sn, double* __restrict__ noise) {
  __shared__ double sh[256];
  const int t = blockIdx.x;
  const float* d = obsT + (size_t)t * nstep;
  const float* s = synT + (size_t)t * nstep;
  const double tstart = tp[t] - min_period * 1.5 + t0;
  const int nstart = min(max((int)(tstart / dt) + 1, 1), nstep);
  double m = 0.0;
  for (int i = threadIdx.x; i < nstep; i += 256) m = fmax(m, fabs((double)s[i]));
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  const double smax = sh[0];
  for (int i = threadIdx.x; i < nstep; i += 256) {
    Dn[(size_t)t * nstep + i] = (double)d[i] / smax;
    Sn[(size_t)t * nstep + i] = (double)s[i] / smax;
  }
  if (threadIdx.x == 0) {  // the sum in the reference's order
    double e = 0.0;
    for (int i = 0; i < nstart; ++i) e = e + (double)d[i] * (double)d[i];
    noise[t] = sqrt(e / (double)nstart);
  }
}

// sliding_cc (win_sel.f90:76-100), correlation part: one wavefront per window
// start i; the Hann tapered windows of d and s in LDS, the full correlation
// with every lag's sum in the reference's order, and its first maximum; the
// lag goes to ish, the coefficient is k_ws_coef's.  Lags come in blocks of
// four consecutive lags (a sliding register window over s, as k_cc_corr,
// unrolled by four so the window needs no register moves).  Block b costs
// about min(4 b, 2 nlen - 4 b) multiply-adds and a wavefront runs as long as
// its slowest lane, so the blocks are dealt to the lanes in descending cost
// (the block holding lag nlen - 1 first, then alternately its neighbours):
// the 64 dearest blocks run together, then the next 64
__device__ __forceinline__ int ws_block_of_rank(int r, int c, int nb) {
  const int L = c, R = nb - 1 - c, m = L < R ? L : R;
  if (r == 0) return c;
  if (r <= 2 * m) { const int k = (r + 1) / 2; return (r & 1) ? c - k : c + k; }
  const int k = r - m;
  return L > R ? c - k : c + k;
}

__global__ void __launch_bounds__(256) k_ws_sliding(int nstep, int nlen, int nw,
                                                             const double* __restrict__ wt,
                                                             const double* __restrict__ Dn,
                                                             const double* __restrict__ Sn, int* __restrict__ ish_out) {
  extern __shared__ double wsh[];
  const int nwv = blockDim.x / 64;  // one window start per wavefront, 2 nlen + 7 doubles of LDS each
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, t = blockIdx.y;
  const int i = blockIdx.x * nwv + wave;
  double* dw = wsh + (size_t)wave * (2 * nlen + 7);
  double* sw = dw + nlen + 3;  // s with three zeros before and four after (indices -3 .. nlen + 3)
  if (i < nw) {
    const double* d = Dn + (size_t)t * nstep + i;
    const double* s = Sn + (size_t)t * nstep + i;
    for (int j = lane; j < nlen; j += 64) { dw[j] = d[j] * wt[j]; sw[j] = s[j] * wt[j]; }
    if (lane < 3) sw[-1 - lane] = 0.0;
    if (lane < 4) sw[nlen + lane] = 0.0;
  }
  __syncthreads();
  if (i >= nw) return;
  double best = -INFINITY;
  int bm = 0x7fffffff;
  const int nlag = 2 * nlen - 1, nb = (nlag + 3) / 4, cb = (nlen - 1) / 4;
  for (int r = lane; r < nb; r += 64) {
    const int mb = 4 * ws_block_of_rank(r, cb, nb);
    const int lo = mb - (nlen - 1) > 0 ? mb - (nlen - 1) : 0, hi = mb + 3 < nlen - 1 ? mb + 3 : nlen - 1;
    const int off = nlen - 1 - mb;  // s index of lag mb + q at i1: i1 + off - q, in [-3, nlen + 3]
    double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
    double r0 = sw[lo + off], r1 = sw[lo + off - 1], r2 = sw[lo + off - 2], r3 = sw[lo + off - 3];
    int i1 = lo;
    for (; i1 + 3 <= hi; i1 += 4) {
      const double d0 = dw[i1], d1 = dw[i1 + 1], d2 = dw[i1 + 2], d3 = dw[i1 + 3];
      const double s0 = sw[i1 + 1 + off], s1 = sw[i1 + 2 + off], s2 = sw[i1 + 3 + off], s3 = sw[i1 + 4 + off];
      c0 = c0 + d0 * r0; c1 = c1 + d0 * r1; c2 = c2 + d0 * r2; c3 = c3 + d0 * r3;
      c0 = c0 + d1 * s0; c1 = c1 + d1 * r0; c2 = c2 + d1 * r1; c3 = c3 + d1 * r2;
      c0 = c0 + d2 * s1; c1 = c1 + d2 * s0; c2 = c2 + d2 * r0; c3 = c3 + d2 * r1;
      c0 = c0 + d3 * s2; c1 = c1 + d3 * s1; c2 = c2 + d3 * s0; c3 = c3 + d3 * r0;
      r0 = s3; r1 = s2; r2 = s1; r3 = s0;
    }
    for (; i1 <= hi; ++i1) {
      const double dv = dw[i1];
      c0 = c0 + dv * r0;
      c1 = c1 + dv * r1;
      c2 = c2 + dv * r2;
      c3 = c3 + dv * r3;
      r3 = r2;
      r2 = r1;
      r1 = r0;
      r0 = sw[i1 + 1 + off];
    }
    const double cv[4] = {c0, c1, c2, c3};
    for (int q = 0; q < 4; ++q)  // blocks arrive out of lag order: ties go to the smaller lag
      if (mb + q < nlag && (cv[q] > best || (cv[q] == best && mb + q < bm))) { best = cv[q]; bm = mb + q; }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double v2 = __shfl_xor(best, o);
    const int m2 = __shfl_xor(bm, o);
    if (v2 > best || (v2 == best && m2 < bm)) { best = v2; bm = m2; }
  }
  if (lane == 0) ish_out[(size_t)t * nw + i] = bm + 1 - nlen;
}

// xcorr_shift_coe's coefficient (win_sel.f90:439-469) of every window start,
// one window start per lane: a = sum(d s_shifted), b = sum(d^2), c =
// sum(s_shifted^2) over the tapered window with d's emptied edge zeroed, each
// sum in the reference's order; neighbouring lanes read neighbouring samples
__global__ void __launch_bounds__(256) k_ws_coef(int nstep, int nlen, int nw, double dt,
                                                 const double* __restrict__ wt, const double* __restrict__ Dn,
                                                 const double* __restrict__ Sn, const int* __restrict__ ish_in,
                                                 double* __restrict__ ts, double* __restrict__ cc) {
  const int t = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nw) return;
  const int ish = ish_in[(size_t)t * nw + i];
  const double* d = Dn + (size_t)t * nstep + i;
  const double* s = Sn + (size_t)t * nstep + i;
  // d's kept range [jd0, jd1) and s's shifted range [js0, js1) within 0 .. nlen
  const int jd0 = ish > 0 ? ish : 0, jd1 = ish < 0 ? nlen + ish : nlen;
  const int js0 = ish > 0 ? ish : 0, js1 = ish < 0 ? nlen + ish : nlen;
  double a = 0.0, b = 0.0, c = 0.0;
  for (int j = 0; j < nlen; ++j) {
    const double dv = (j >= jd0 && j < jd1) ? d[j] * wt[j] : 0.0;
    const double sv = (j >= js0 && j < js1) ? s[j - ish] * wt[j - ish] : 0.0;
    a = a + dv * sv;
    b = b + dv * dv;
    c = c + sv * sv;
  }
  ts[(size_t)t * nw + i] = (double)ish * dt;
  cc[(size_t)t * nw + i] = a / sqrt(b * c);
}

// envelope |hilbert(syn)| of every trace's normalised synthetic for the phase
// splitting (win_sel.f90:121-133): one workgroup per trace, the analytic signal
// on nf = 2 ** exponent(real(nstep)) points through the LDS real transform (as
// k_ep_window), Re and Im as real-output inverses, env = |xa|.  G0, G1: M + 1
// double2 per trace
__global__ void __launch_bounds__(MT_THREADS) k_ws_env(int nstep, int nf, int logM, const double2* __restrict__ tw,
                                                       const double* __restrict__ Sn, double2* __restrict__ G0,
                                                       double2* __restrict__ G1, double* __restrict__ env) {
  __shared__ double2 buf[MT_MAXM];
  const int t = blockIdx.x, M = nf / 2, lane = threadIdx.x;
  const unsigned shb = 32 - logM;
  const double* x = Sn + (size_t)t * nstep;
  double* e = env + (size_t)t * nstep;
  double2* g0 = G0 + (size_t)t * (M + 1);
  double2* g1 = G1 + (size_t)t * (M + 1);
  const double hnyq = (nstep % 2 == 0) ? 1.0 : 0.0;
  mt_load_real<MT_THREADS>(buf, M, logM, nstep, [&](int n) { return x[n]; });
  mt_fft<MT_THREADS>(buf, M, logM, nf, tw, false);
  for (int k = lane; k <= M; k += MT_THREADS) {
    double2 X;
    if (k == M) { const double2 z = buf[0]; X = make_double2(z.x - z.y, 0.0); }
    else if (k == 0) { const double2 z = buf[0]; X = make_double2(z.x + z.y, 0.0); }
    else X = mt_half(buf[k], buf[M - k], tw[k]);
    g0[k] = k == 0 ? make_double2(X.x, 0.0) : (k == M ? make_double2(hnyq * X.x, 0.0) : X);
    g1[k] = (k == 0 || k == M) ? make_double2(0.0, 0.0) : make_double2(X.y, -X.x);
  }
  __syncthreads();
  for (int pass = 0; pass < 2; ++pass) {
    const double2* g = pass == 0 ? g0 : g1;
    for (int k = lane; k < M; k += MT_THREADS)
      buf[__brev((unsigned)k) >> shb] = mt_repack(g[k], g[M - k], tw[k]);
    __syncthreads();
    mt_fft<MT_THREADS>(buf, M, logM, nf, tw, true);
    for (int n = lane; n < nstep; n += MT_THREADS) {
      const double2 z = buf[n >> 1];
      const double y = ((n & 1) ? z.y : z.x) / M;
      e[n] = pass == 0 ? y : hypot(e[n], y);
    }
    __syncthreads();
  }
}

// select_windows (win_sel.f90:102-277) without phase splitting, one workgroup
// per trace: the good-sample mask (cc, shift, time range, and no time-shift
// jump within nint(jump / dt)) over the lanes into G (one byte per sliding
// window); lane 0 collects the contiguous groups (at most WS_MAXG, the
// reference's max_good_win); one lane per group runs the length, peak-count,
// SNR and energy-ratio checks (sums in the reference's order); lane 0 writes
// the accepted windows in group order (seconds, win[t][k][2], k < maxwin) and
// their count to nwin[t] (-1 when the groups overflow WS_MAXG).
constexpr int WS_MAXG = 1000;
__global__ void __launch_bounds__(256) k_ws_select(int nstep, int nlen, int nw, double dt, double t0,
                                                   double min_period, WsCfg cfg, const double* __restrict__ tp,
                                                   const double* __restrict__ dist, const double* __restrict__ noise,
                                                   const double* __restrict__ Dn, const double* __restrict__ Sn,
                                                   const double* __restrict__ ts, const double* __restrict__ cc,
                                                   unsigned char* __restrict__ G, const double* __restrict__ env,
                                                   double* __restrict__ stl, int maxwin, int* __restrict__ nwin,
                                                   double* __restrict__ win) {
  __shared__ int ga[WS_MAXG], gb[WS_MAXG];
  __shared__ unsigned char ok[WS_MAXG];
  __shared__ int ng;
  const int t = blockIdx.x;
  const double tstart = tp[t] - min_period * 1.5 + t0;
  const double tend = dist[t] / cfg.min_velocity + min_period * 1.5 + t0;
  const double jump = cfg.jump_fac * min_period;
  const int nj = (int)round(jump / dt);
  const double* tsr = ts + (size_t)t * nw;
  const double* ccr = cc + (size_t)t * nw;
  unsigned char* g = G + (size_t)t * nw;
  auto tcc = [&](int i) { return ((double)i + (double)nlen / 2.0) * dt; };
  auto is_jump = [&](int i) { return i > 0 && i < nw && fabs(tsr[i] - tsr[i - 1]) > jump; };
  for (int i = threadIdx.x; i < nw; i += 256) {
    const double tc = tcc(i);
    bool good = ccr[i] > cfg.thr_corr && fabs(tsr[i]) < cfg.thr_shift_fac * min_period && tc >= tstart && tc <= tend;
    for (int k = max(1, i - nj); good && k <= min(nw - 1, i + nj); ++k)
      if (is_jump(k)) good = false;
    g[i] = good ? 1 : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int n = 0, start = -1;
    for (int i = 0; i <= nw; ++i) {
      const bool gd = i < nw && g[i];
      if (gd && start < 0) start = i;
      else if (!gd && start >= 0) {
        if (n < WS_MAXG) { ga[n] = start; gb[n] = i - 1; }
        ++n;
        start = -1;
      }
    }
    ng = n;
  }
  __syncthreads();
  const int n = min(ng, WS_MAXG);
  const double* dat = Dn + (size_t)t * nstep;
  const double* syn = Sn + (size_t)t * nstep;
  auto check = [&](int a, int b) {
    bool pass = true;
    if ((b - a) * dt < cfg.min_win_len_fac * min_period) pass = false;
    int ib = (int)(tcc(a) / dt) + 1, ie = (int)(tcc(b) / dt) + 1;
    ib = max(1, min(ib, nstep));
    ie = max(1, min(ie, nstep));
    if (ie <= ib) pass = false;
    if (pass) {
      const double* dp = dat + ib - 1;
      const double* sp = syn + ib - 1;
      const int m = ie - ib + 1;
      int pd = 0, td = 0, ps = 0, tq = 0;
      double amax = 0.0, ed = 0.0, es = 0.0;
      for (int k = 0; k < m; ++k) {
        if (k >= 1 && k < m - 1) {
          if (dp[k] - dp[k - 1] > 0.0 && dp[k + 1] - dp[k] < 0.0) ++pd;
          if (-dp[k] - -dp[k - 1] > 0.0 && -dp[k + 1] - -dp[k] < 0.0) ++td;
          if (sp[k] - sp[k - 1] > 0.0 && sp[k + 1] - sp[k] < 0.0) ++ps;
          if (-sp[k] - -sp[k - 1] > 0.0 && -sp[k + 1] - -sp[k] < 0.0) ++tq;
        }
        amax = fmax(amax, fabs(dp[k]));
        ed = ed + dp[k] * dp[k];
        es = es + sp[k] * sp[k];
      }
      if (min(pd + td, ps + tq) < cfg.min_peaks) pass = false;
      else if (amax / noise[t] < cfg.min_snr) pass = false;
      else if (ed / es > cfg.min_energy_ratio || ed / es < 1 / cfg.min_energy_ratio) pass = false;
    }
    return pass;
  };
  if (cfg.split) {
    // phase splitting: lane 0 runs STA_LTA (sta_lta.f90:6-104) over the envelope,
    // then split_window_by_phases (win_sel.f90:279-437) and the checks group by group
    __shared__ double rm[256];
    double m = 0.0;
    for (int i = threadIdx.x; i < nstep; i += 256) m = fmax(m, env[(size_t)t * nstep + i]);
    rm[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) rm[threadIdx.x] = fmax(rm[threadIdx.x], rm[threadIdx.x + o]);
      __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double emax = rm[0];
    const double* e = env + (size_t)t * nstep;
    double* st = stl + (size_t)t * nstep;
    const double Cs = pow(10.0, -dt / min_period), Cl = pow(10.0, -dt / (12.0 * min_period));
    const double nz = emax / 1e5;
    double sta = 0.0, lta = 0.0;
    for (int i = 0; i < 1000; ++i) { sta = Cs * sta + nz; lta = Cl * lta + nz; }
    for (int i = 0; i < nstep; ++i) {
      const double x = e[i] + nz;
      sta = Cs * sta + x;
      lta = Cl * lta + x;
      st[i] = lta > 1.0e-9 ? sta / lta : nz;
    }
    auto peak = [&](int i, double sg) {  // 1-based sample i is a maximum of sg * stalta
      return i >= 2 && i <= nstep - 1 && sg * st[i - 1] - sg * st[i - 2] > 0.0 && sg * st[i] - sg * st[i - 1] < 0.0;
    };
    const int minlen = (int)(cfg.min_win_len_fac * min_period / dt);
    int count = 0;
    auto emit = [&](int a, int b) {
      if (!check(a, b)) return;
      if (count < maxwin) {
        win[((size_t)t * maxwin + count) * 2] = tcc(a);
        win[((size_t)t * maxwin + count) * 2 + 1] = tcc(b);
      }
      ++count;
    };
    for (int q = 0; q < n; ++q) {
      const int a = ga[q], b = gb[q];
      int s0 = (int)(tcc(a) / dt) + 1, s1 = (int)(tcc(b) / dt) + 1;
      s0 = max(1, min(s0, nstep));
      s1 = max(1, min(s1, nstep));
      int npk = 0;
      if (s0 < s1)
        for (int i = s0; i <= s1; ++i) npk += peak(i, 1.0);
      if (s0 >= s1 || npk <= 1) { emit(a, b); continue; }
      // split points: s0, the lowest trough between consecutive peaks (midpoint if none), s1
      int prev = s0, p1 = -1, nsub = 0;
      auto sub = [&](int x0, int x1, bool go) {
        if (x1 - x0 < minlen) return;
        ++nsub;
        if (!go) return;
        int j0 = a, j1 = b;
        for (int r = a; r <= b; ++r) {
          if (fabs(tcc(r) - x0 * dt) < fabs(tcc(j0) - x0 * dt)) j0 = r;
          if (fabs(tcc(r) - x1 * dt) < fabs(tcc(j1) - x1 * dt)) j1 = r;
        }
        emit(max(a, min(j0, b)), max(a, min(j1, b)));
      };
      // two passes: count the valid sub-windows, then emit them (or the group when none)
      for (int go = 0; go < 2; ++go) {
        prev = s0;
        p1 = -1;
        if (go == 1 && nsub == 0) { emit(a, b); break; }
        for (int i = s0; i <= s1; ++i) {
          if (!peak(i, 1.0)) continue;
          if (p1 >= 0) {
            int mi = -1;
            double mv = 1.79769313486231570815e308;
            for (int k = p1 + 1; k < i; ++k)
              if (peak(k, -1.0)) {
                if (mi < 0) mi = k;
                if (st[k - 1] < mv) { mv = st[k - 1]; mi = k; }
              }
            const int spt = mi >= 0 ? mi : (p1 + i) / 2;
            sub(prev, spt, go == 1);
            prev = spt;
          }
          p1 = i;
        }
        sub(prev, s1, go == 1);
      }
    }
    nwin[t] = ng > WS_MAXG ? -1 : count;
    return;
  }
  for (int q = threadIdx.x; q < n; q += 256) ok[q] = check(ga[q], gb[q]) ? 1 : 0;
  __syncthreads();
  if (threadIdx.x != 0) return;
  int count = 0;
  for (int q = 0; q < n; ++q) {
    if (!ok[q]) continue;
    if (count < maxwin) {
      win[((size_t)t * maxwin + count) * 2] = tcc(ga[q]);
      win[((size_t)t * maxwin + count) * 2 + 1] = tcc(gb[q]);
    }
    ++count;
  }
  nwin[t] = ng > WS_MAXG ? -1 : count;
}

// selected traces' adjoint-source rows added into the sample-major traces
// (the multitaper measurement accumulates its windows)
__global__ void __launch_bounds__(256) k_mt_put(int ntrace, int nb, int nl, const double* __restrict__ A,
                                                const double* __restrict__ out, int ostride,
                                                float* __restrict__ adj) {
  __shared__ double tile[64][65];
  const int t0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
  for (int r = threadIdx.x / 64; r < 64; r += 4) {
    const int t = t0 + r, k = k0 + (threadIdx.x & 63);
    if (t < ntrace && k < nl) tile[r][threadIdx.x & 63] = A[(size_t)t * nl + k];
  }
  __syncthreads();
  for (int r = threadIdx.x / 64; r < 64; r += 4) {
    const int k = k0 + r, t = t0 + (threadIdx.x & 63);
    if (t < ntrace && k < nl && out[ostride * (size_t)t + 6] != 0.0) {
      float* a = adj + (size_t)(nb + k) * ntrace + t;
      *a = (float)((double)*a + tile[threadIdx.x & 63][r]);
    }
  }
}

__global__ void k_fill_f32m(size_t n, float* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0f;
}
#pragma clang fp contract(fast)

double g_meas_ms = 0.0;

// window_taper's weights (signal.f90:419-491) for a window of npts samples
void taper_weights_into(int npts, double pct, int type, double* w);

void taper_weights(int npts, double pct, int type, std::vector<double>& w) {
  w.assign(npts, 1.0);
  taper_weights_into(npts, pct, type, w.data());
}

// window_taper's weights (signal.f90:419-491) written into w[0..npts)
void taper_weights_into(int npts, double pct, int type, double* w) {
  for (int i = 0; i < npts; ++i) w[i] = 1.0;
  const int frac = (pct == 0.0 || pct == 1.0) ? (int)(npts * pct / 2.0) : (int)(npts * pct / 2.0 + 0.5);
  if (frac <= 0) return;
  const double pi = 3.14159265358979323846;
  const int den = 2 * frac - 1;
  auto f = [&](int r) {
    switch (type) {
      case 1: return 0.5 - 0.5 * std::cos(2.0 * pi * r / den);
      case 2: return 0.54 - 0.46 * std::cos(2.0 * pi * r / den);
      case 3: return std::cos(pi * r / den - pi / 2.0);
      case 4: return 1.0 - std::pow(std::cos(pi * r / den), 10);
    }
    SEM_FAIL("window taper type %d not supported", type);
  };
  for (int i = 0; i < frac; ++i) w[i] = f(i);
  for (int i = 0; i < frac; ++i) w[npts - frac + i] = f(frac + i);
}

void window_range(double w0, double w1, double dt, int nstep, int& nb, int& nl) {
  const int b = (int)(w0 / dt) + 1, e = (int)(w1 / dt);  // get_window_info, 1-based
  if (e - b + 1 <= 0) SEM_FAIL("invalid measurement window [%g, %g]", w0, w1);
  if (b < 1 || e > nstep) SEM_FAIL("measurement window [%g, %g] outside the %d-sample trace", w0, w1, nstep);
  nb = b - 1;
  nl = e - b + 1;
}


// dpss_windows (utils/dpss.f90:137-262) for the multitaper measurement: the
// kmax largest eigenvalues of the DPSS tridiagonal matrix by bisection on
// Sturm counts (the reference calls LAPACK dsterf), eigenvectors by the
// reference's inverse iteration from sin(k pi i / (n - 1)) with its
// Golub-Van Loan solve, unit norm, the reference's sign convention; each
// taper is returned scaled by sqrt(n) as calc_adjoint_source uses it
void dpss_tapers(int n, double half_nbw, int kmax, std::vector<double>& out) {
  const double pi = 3.141592653589793;
  const double wbin = half_nbw / (double)n;
  std::vector<double> dg(n), od(n > 1 ? n - 1 : 0);
  for (int i = 0; i < n; ++i) {
    const double q = ((double)(n - 1) - 2.0 * i) / 2.0;
    dg[i] = q * q * std::cos(2.0 * pi * wbin);
  }
  for (int i = 0; i + 1 < n; ++i) od[i] = (double)(i + 1) * ((double)n - (double)(i + 1)) / 2.0;
  auto below = [&](double x) {
    int c = 0;
    double q = 1.0;
    for (int i = 0; i < n; ++i) {
      q = (dg[i] - x) - (i ? od[i - 1] * od[i - 1] / q : 0.0);
      if (q == 0.0) q = -1e-300;
      if (q < 0.0) ++c;
    }
    return c;
  };
  double lo = dg[0], hi = dg[0];
  for (int i = 0; i < n; ++i) {
    const double r = (i ? std::fabs(od[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(od[i]) : 0.0);
    lo = std::min(lo, dg[i] - r);
    hi = std::max(hi, dg[i] + r);
  }
  out.assign((size_t)kmax * n, 0.0);
  std::vector<double> ed(n), x(n), xp(n), dw(n), ew(od.size());
  auto norm = [&](std::vector<double>& v) {
    double ss = 0.0;
    for (double a : v) ss = ss + a * a;
    const double nv = std::sqrt(ss);
    for (double& a : v) a = a / nv;
  };
  for (int k = 0; k < kmax; ++k) {
    double a = lo, b = hi;
    for (int it = 0; it < 200; ++it) {
      const double m = 0.5 * (a + b);
      if (m <= a || m >= b) break;
      if (below(m) <= n - 1 - k) a = m; else b = m;
    }
    const double lam = 0.5 * (a + b);
    for (int i = 0; i < n; ++i) ed[i] = dg[i] - lam;
    for (int i = 0; i < n; ++i) x[i] = std::sin((double)(k + 1) * (pi * (double)i / (double)(n - 1)));
    norm(x);
    for (int it = 1;; ++it) {
      xp = x;
      dw = ed;
      ew = od;
      for (int i = 1; i < n; ++i) {
        const double tq = ew[i - 1];
        ew[i - 1] = tq / dw[i - 1];
        dw[i] = dw[i] - tq * ew[i - 1];
      }
      for (int i = 1; i < n; ++i) x[i] = x[i] - ew[i - 1] * x[i - 1];
      x[n - 1] = x[n - 1] / dw[n - 1];
      for (int i = n - 2; i >= 0; --i) x[i] = x[i] / dw[i] - ew[i] * x[i + 1];
      norm(x);
      double dn = 0.0;
      for (int i = 0; i < n; ++i) dn = dn + (std::fabs(x[i]) - std::fabs(xp[i])) * (std::fabs(x[i]) - std::fabs(xp[i]));
      if (std::sqrt(dn) <= 1e-8 || it > 1000) break;
    }
    double sgn = 1.0;
    if (k % 2 == 0) {
      double sm = 0.0;
      for (double v : x) sm = sm + v;
      if (sm < 0.0) sgn = -1.0;
    } else if (n > 1 && x[1] < 0.0) {
      sgn = -1.0;
    }
    const double sc = std::sqrt((double)n);
    for (int i = 0; i < n; ++i) {
      if (!std::isfinite(x[i])) SEM_FAIL("dpss taper %d of length %d did not converge", k + 1, n);
      out[(size_t)k * n + i] = sgn * x[i] * sc;
    }
  }
}
// per trace: its window misfits summed in window order (the host loop of
// fwat_waveform_adjoint_batch_), residuals (ntrace, nwin) when asked for
__global__ void k_trace_misfits(int ntrace, int nwin, const double* __restrict__ mr, double* __restrict__ misfits,
                                double* __restrict__ residuals) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntrace) return;
  const size_t T = (size_t)ntrace;
  double m = 0.0;
  for (int w = 0; w < nwin; ++w) {
    m += mr[(size_t)w * T + t];
    if (residuals) residuals[(size_t)t * nwin + w] = mr[(size_t)nwin * T + (size_t)w * T + t];
  }
  misfits[t] = m;
}

// ---- scratch of the allocation-free waveform entry point
// Limits of fwat_waveform_adjoint_batch_noalloc_ and the layout of its
// device scratch (doubles): taper weights of the windows [NA_MAX_WEIGHTS],
// taper weights of the filter window [NA_MAX_WEIGHTS], window misfits and
// residuals [2][NA_MAX_WIN][NA_MAX_TRACES], filter sections (6 floats each).
constexpr int NA_MAX_TRACES = 16384;
constexpr int NA_MAX_WIN = 8;
constexpr int NA_MAX_WEIGHTS = 65536;
constexpr int NA_MAX_SECT = 8;
constexpr size_t NA_OFF_FW = NA_MAX_WEIGHTS;
constexpr size_t NA_OFF_MR = 2 * (size_t)NA_MAX_WEIGHTS;
constexpr size_t NA_OFF_SOS = NA_OFF_MR + 2 * (size_t)NA_MAX_WIN * NA_MAX_TRACES;
constexpr size_t NA_SCRATCH_DOUBLES = NA_OFF_SOS + 3 * (size_t)NA_MAX_SECT;  // 3.1 MB

// One copy per GPU the library runs on, held until the process ends.  In host
// code `scratch` names a host-side stand-in, never the device array: the
// device address comes from hipGetSymbolAddress (na_scratch_device).
static __device__ __align__(256) double scratch[NA_SCRATCH_DOUBLES];

constexpr int NA_MAX_DEVICES = 64;
double* g_na_scratch[NA_MAX_DEVICES] = {};
std::mutex g_na_mutex;
// host staging of the weights and filter sections copied into the scratch
double g_na_stage[NA_OFF_MR + 3 * NA_MAX_SECT];

// Device address of this GPU's scratch, looked up once per device with that
// device current (the first call on the device).
double* na_scratch_device() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NA_MAX_DEVICES) SEM_FAIL("device %d outside the %d the scratch table holds", dev, NA_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_na_mutex);
  if (!g_na_scratch[dev]) {
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_na_scratch[dev] = static_cast<double*>(p);
  }
  return g_na_scratch[dev];
}

}  // namespace

extern "C" {

// obs, syn, adj: (NDIM, nrec, NSTEP) single precision = [nstep][ntrace];
// windows(2, nwin) start/end times (s); sos(6, nsect) b0 b1 b2 1 a1 a2 per
// section of the adjoint-source bandpass (nsect = 0: no filter);
// misfits(ntrace) out (window misfits summed in window order), residuals(nwin, ntrace) out or NULL
void fwat_waveform_adjoint_batch_(int* nstep, int* ntrace, double* dt, realw* obs, realw* syn, int* nwin,
                                  double* windows, double* taper_percentage, int* itaper_type, int* nsect, realw* sos,
                                  double* misfits, double* residuals, realw* adj) {
  const int N = *nstep, T = *ntrace, nw = *nwin;
  if (N <= 0 || T <= 0) SEM_FAIL("fwat_waveform_adjoint_batch_: sizes must be positive");
  if (nw < 1 || nw > MAX_WIN) SEM_FAIL("fwat_waveform_adjoint_batch_: 1..%d windows", MAX_WIN);
  if (*nsect < 0 || *nsect > 8) SEM_FAIL("fwat_waveform_adjoint_batch_: 0..8 filter sections");
  if (*taper_percentage < 0.0 || *taper_percentage > 1.0) SEM_FAIL("taper_percentage must be 0 < %% < 1");
  WinInfo W{};
  std::vector<double> wts, w1;
  double wmin = windows[0], wmax = windows[0];
  for (int w = 0; w < nw; ++w) {
    window_range(windows[2 * w], windows[2 * w + 1], *dt, N, W.nb[w], W.nl[w]);
    taper_weights(W.nl[w], *taper_percentage, *itaper_type, w1);
    W.woff[w] = (int)wts.size();
    wts.insert(wts.end(), w1.begin(), w1.end());
    wmin = std::min(wmin, std::min(windows[2 * w], windows[2 * w + 1]));
    wmax = std::max(wmax, std::max(windows[2 * w], windows[2 * w + 1]));
  }
  int fnb = 0, fnl = 0;
  if (*nsect > 0) {
    window_range(wmin, wmax, *dt, N, fnb, fnl);
    taper_weights(fnl, *taper_percentage, *itaper_type, w1);
  }
  const size_t n = (size_t)N * T;
  hipStream_t s;
  HIP_CHECK(hipStreamCreate(&s));
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  float* d_obs = dev_upload(obs, n, s);
  float* d_syn = dev_upload(syn, n, s);
  float* d_adj = dev_alloc<float>(n);
  double* d_wts = dev_upload(wts.data(), wts.size(), s);
  double* d_mr = dev_alloc<double>(2 * (size_t)nw * T);
  double* d_fw = *nsect > 0 ? dev_upload(w1.data(), w1.size(), s) : nullptr;
  float* d_sos = *nsect > 0 ? dev_upload(sos, 6 * (size_t)*nsect, s) : nullptr;
  HIP_CHECK(hipEventRecord(e0, s));
  const unsigned nb = (unsigned)((T + MTPB - 1) / MTPB);
  hipLaunchKernelGGL(k_waveform_adj, dim3(nb), dim3(MTPB), 0, s, N, T, *dt, nw, W, d_wts, d_obs, d_syn, d_adj, d_mr,
                     d_mr + (size_t)nw * T);
  if (*nsect > 0) {
    launch_sos(s, N, T, *nsect, d_sos, d_adj);
    hipLaunchKernelGGL(k_window_final, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, N, T, fnb, fnl, d_fw,
                       d_adj);
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipEventRecord(e1, s));
  std::vector<double> mr(2 * (size_t)nw * T);
  HIP_CHECK(hipMemcpyAsync(adj, d_adj, n * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(mr.data(), d_mr, mr.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  float ms = 0.0f;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  g_meas_ms = ms;
  // per trace: sum of its window misfits in window order; residuals (ntrace, nwin)
  for (int t = 0; t < T; ++t) {
    double m = 0.0;
    for (int w = 0; w < nw; ++w) {
      m += mr[(size_t)w * T + t];
      if (residuals) residuals[(size_t)t * nw + w] = mr[(size_t)nw * T + (size_t)w * T + t];
    }
    misfits[t] = m;
  }
  dev_free(d_obs); dev_free(d_syn); dev_free(d_adj); dev_free(d_wts); dev_free(d_mr); dev_free(d_fw); dev_free(d_sos);
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  HIP_CHECK(hipStreamDestroy(s));
}

void fwat_waveform_adjoint_last_time_(double* ms) { *ms = g_meas_ms; }

// fwat_waveform_adjoint_batch_ for callers that cannot allocate device memory
// (code under HIP graph capture, serving loops): the same measurement with
// obs, syn (inputs), adj, misfits and residuals (outputs) in the caller's
// device memory, on the caller's stream (*stream: a hipStream_t, 0 = the null
// stream), and the intermediate arrays in this library's per-GPU static
// scratch -- no device allocation or release.  windows and sos are host arrays
// as in fwat_waveform_adjoint_batch_; misfits (ntrace) and residuals
// (nwin, ntrace, or NULL) are device doubles.  Limits: ntrace <= 16384, nwin
// <= 8, nsect <= 8, window lengths summed <= 65536 samples and the filter
// window <= 65536 samples.  Returns when the outputs are written.  Calls that
// overlap on one GPU share its scratch and race.
void fwat_waveform_adjoint_batch_noalloc_(int* nstep, int* ntrace, double* dt, realw* d_obs, realw* d_syn, int* nwin,
                                          double* windows, double* taper_percentage, int* itaper_type, int* nsect,
                                          realw* sos, double* d_misfits, double* d_residuals, realw* d_adj,
                                          long* stream) {
  const int N = *nstep, T = *ntrace, nw = *nwin;
  if (N <= 0 || T <= 0) SEM_FAIL("fwat_waveform_adjoint_batch_noalloc_: sizes must be positive");
  if (T > NA_MAX_TRACES) SEM_FAIL("fwat_waveform_adjoint_batch_noalloc_: %d traces, at most %d", T, NA_MAX_TRACES);
  if (nw < 1 || nw > NA_MAX_WIN) SEM_FAIL("fwat_waveform_adjoint_batch_noalloc_: 1..%d windows", NA_MAX_WIN);
  if (*nsect < 0 || *nsect > NA_MAX_SECT) SEM_FAIL("fwat_waveform_adjoint_batch_noalloc_: 0..%d filter sections", NA_MAX_SECT);
  if (*taper_percentage < 0.0 || *taper_percentage > 1.0) SEM_FAIL("taper_percentage must be 0 < %% < 1");
  if (!d_obs || !d_syn || !d_adj || !d_misfits) SEM_FAIL("fwat_waveform_adjoint_batch_noalloc_: null device buffer");
  double* sc = na_scratch_device();
  hipStream_t s = reinterpret_cast<hipStream_t>(*stream);
  WinInfo W{};
  int nwts = 0;
  double wmin = windows[0], wmax = windows[0];
  for (int w = 0; w < nw; ++w) {
    window_range(windows[2 * w], windows[2 * w + 1], *dt, N, W.nb[w], W.nl[w]);
    if (nwts + W.nl[w] > NA_MAX_WEIGHTS)
      SEM_FAIL("fwat_waveform_adjoint_batch_noalloc_: windows hold more than %d samples", NA_MAX_WEIGHTS);
    W.woff[w] = nwts;
    taper_weights_into(W.nl[w], *taper_percentage, *itaper_type, g_na_stage + nwts);
    nwts += W.nl[w];
    wmin = std::min(wmin, std::min(windows[2 * w], windows[2 * w + 1]));
    wmax = std::max(wmax, std::max(windows[2 * w], windows[2 * w + 1]));
  }
  int fnb = 0, fnl = 0;
  if (*nsect > 0) {
    window_range(wmin, wmax, *dt, N, fnb, fnl);
    if (fnl > NA_MAX_WEIGHTS)
      SEM_FAIL("fwat_waveform_adjoint_batch_noalloc_: filter window of %d samples, at most %d", fnl, NA_MAX_WEIGHTS);
    taper_weights_into(fnl, *taper_percentage, *itaper_type, g_na_stage + NA_OFF_FW);
    std::memcpy(g_na_stage + NA_OFF_MR, sos, 6 * sizeof(float) * (size_t)*nsect);
  }
  double* d_wts = sc;
  double* d_fw = sc + NA_OFF_FW;
  double* d_mr = sc + NA_OFF_MR;
  float* d_sos = reinterpret_cast<float*>(sc + NA_OFF_SOS);
  HIP_CHECK(hipMemcpyAsync(d_wts, g_na_stage, (size_t)nwts * sizeof(double), hipMemcpyHostToDevice, s));
  if (*nsect > 0) {
    HIP_CHECK(hipMemcpyAsync(d_fw, g_na_stage + NA_OFF_FW, (size_t)fnl * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(d_sos, g_na_stage + NA_OFF_MR, 6 * sizeof(float) * (size_t)*nsect, hipMemcpyHostToDevice, s));
  }
  const unsigned nb = (unsigned)((T + MTPB - 1) / MTPB);
  hipLaunchKernelGGL(k_waveform_adj, dim3(nb), dim3(MTPB), 0, s, N, T, *dt, nw, W, d_wts, d_obs, d_syn, d_adj, d_mr,
                     d_mr + (size_t)nw * T);
  if (*nsect > 0) {
    const size_t n = (size_t)N * T;
    launch_sos(s, N, T, *nsect, d_sos, d_adj);
    hipLaunchKernelGGL(k_window_final, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, N, T, fnb, fnl, d_fw,
                       d_adj);
  }
  hipLaunchKernelGGL(k_trace_misfits, dim3(nb), dim3(MTPB), 0, s, T, nw, d_mr, d_misfits, d_residuals);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(s));
}

// cc_tt_misfit.f90 CCTTMisfit%calc_adjoint_source for every trace: imeas 11
// (traveltime) or 12 (amplitude), or ccc_misfit.f90 CCCMisfit (imeas 5,
// zero-lag correlation coefficient; uses CC_MIN only); cfg(7) = dt_sigma_min, dlna_sigma_min,
// TSHIFT_MIN, TSHIFT_MAX, DLNA_MIN, DLNA_MAX, CC_MIN; meas(8, nwin, ntrace) =
// misfit, residual, error, cc_max, tshift, dlna, selected, ishift; misfits
// (ntrace) summed over windows; sos / nsect as for the waveform measurement
void fwat_cc_adjoint_batch_(int* nstep, int* ntrace, double* dt, realw* obs, realw* syn, int* nwin, double* windows,
                            double* taper_percentage, int* itaper_type, int* imeas, double* cfg, int* nsect,
                            realw* sos, double* misfits, double* meas, realw* adj) {
  const int N = *nstep, T = *ntrace, nw = *nwin;
  if (N <= 0 || T <= 0) SEM_FAIL("fwat_cc_adjoint_batch_: sizes must be positive");
  if (nw < 1 || nw > MAX_WIN) SEM_FAIL("fwat_cc_adjoint_batch_: 1..%d windows", MAX_WIN);
  if (*imeas != 11 && *imeas != 12 && *imeas != 5)
    SEM_FAIL("fwat_cc_adjoint_batch_: imeas must be 11 (CC-TT), 12 (CC-DLNA) or 5 (CCC)");
  if (T > 65535) SEM_FAIL("fwat_cc_adjoint_batch_: at most 65535 traces per call");
  if (*nsect < 0 || *nsect > 8) SEM_FAIL("fwat_cc_adjoint_batch_: 0..8 filter sections");
  if (*taper_percentage < 0.0 || *taper_percentage > 1.0) SEM_FAIL("taper_percentage must be 0 < %% < 1");
  const CcCfg C{cfg[0], cfg[1], cfg[2], cfg[3], cfg[4], cfg[5], cfg[6]};
  const size_t n = (size_t)N * T;
  hipStream_t s;
  HIP_CHECK(hipStreamCreate(&s));
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  float* d_obs = dev_upload(obs, n, s);
  float* d_syn = dev_upload(syn, n, s);
  float* d_adj = dev_alloc<float>(n);
  int* d_ish = dev_alloc<int>(T);
  double* d_out = dev_alloc<double>(8 * (size_t)nw * T);
  int maxnl = 0;
  std::vector<int> nbs(nw), nls(nw);
  std::vector<std::vector<double>> tws(nw);
  double wmin = windows[0], wmax = windows[0];
  for (int w = 0; w < nw; ++w) {
    window_range(windows[2 * w], windows[2 * w + 1], *dt, N, nbs[w], nls[w]);
    taper_weights(nls[w], *taper_percentage, *itaper_type, tws[w]);
    maxnl = std::max(maxnl, nls[w]);
    wmin = std::min(wmin, std::min(windows[2 * w], windows[2 * w + 1]));
    wmax = std::max(wmax, std::max(windows[2 * w], windows[2 * w + 1]));
  }
  double* d_D = dev_alloc<double>((size_t)T * maxnl);
  double* d_S = dev_alloc<double>((size_t)T * maxnl);
  double* d_tw = dev_alloc<double>(maxnl);
  double* d_A = dev_alloc<double>((size_t)T * maxnl);
  const int maxblk = (2 * maxnl - 1 + CC_LAGS - 1) / CC_LAGS;
  double* d_bv = dev_alloc<double>((size_t)T * maxblk);
  int* d_bi = dev_alloc<int>((size_t)T * maxblk);
  HIP_CHECK(hipEventRecord(e0, s));
  hipLaunchKernelGGL(k_fill_f32m, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, d_adj);
  const unsigned nbk = (unsigned)((T + MTPB - 1) / MTPB);
  for (int w = 0; w < nw; ++w) {
    HIP_CHECK(hipMemcpyAsync(d_tw, tws[w].data(), nls[w] * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_cc_rows, dim3((unsigned)((T + 63) / 64), (unsigned)((nls[w] + 63) / 64)), dim3(256), 0, s, T,
                       nbs[w], nls[w], d_tw, d_obs, d_syn, d_D, d_S);
    if (*imeas == 5) {
      hipLaunchKernelGGL(k_ccc_finish, dim3((unsigned)T), dim3(256), 0, s, nls[w], *dt, C.cc_min, d_tw, d_D, d_S, d_A,
                         d_out + 8 * (size_t)w * T);
    } else {
      const int nblk = (2 * nls[w] - 1 + CC_LAGS - 1) / CC_LAGS;
      hipLaunchKernelGGL(k_cc_corr, dim3((unsigned)nblk, (unsigned)T), dim3(256), 0, s, nls[w], d_D, d_S, d_bv,
                         d_bi);
      hipLaunchKernelGGL(k_cc_pick, dim3(nbk), dim3(MTPB), 0, s, T, nblk, nls[w], d_bv, d_bi, d_ish);
      hipLaunchKernelGGL(k_cc_finish, dim3((unsigned)T), dim3(256), 0, s, nls[w], *dt, *imeas, C, d_tw, d_D, d_S,
                         d_ish, d_A, d_out + 8 * (size_t)w * T, 8);
    }
    hipLaunchKernelGGL(k_cc_put, dim3((unsigned)((T + 63) / 64), (unsigned)((nls[w] + 63) / 64)), dim3(256), 0, s, T,
                       nbs[w], nls[w], d_A, d_out + 8 * (size_t)w * T, d_adj);
  }
  std::vector<double> fw;
  double* d_fw = nullptr;
  float* d_sos = nullptr;
  if (*nsect > 0) {
    int fnb = 0, fnl = 0;
    window_range(wmin, wmax, *dt, N, fnb, fnl);
    taper_weights(fnl, *taper_percentage, *itaper_type, fw);
    d_fw = dev_upload(fw.data(), fw.size(), s);
    d_sos = dev_upload(sos, 6 * (size_t)*nsect, s);
    launch_sos(s, N, T, *nsect, d_sos, d_adj);
    hipLaunchKernelGGL(k_window_final, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, N, T, fnb, fnl, d_fw,
                       d_adj);
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipEventRecord(e1, s));
  std::vector<double> out(8 * (size_t)nw * T);
  HIP_CHECK(hipMemcpyAsync(adj, d_adj, n * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(out.data(), d_out, out.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  float ms = 0.0f;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  g_meas_ms = ms;
  for (int t = 0; t < T; ++t) {
    double m = 0.0;
    for (int w = 0; w < nw; ++w) {
      const double* o = &out[8 * ((size_t)w * T + t)];
      m += o[0];
      if (meas) for (int q = 0; q < 8; ++q) meas[8 * ((size_t)t * nw + w) + q] = o[q];
    }
    misfits[t] = m;
  }
  dev_free(d_obs); dev_free(d_syn); dev_free(d_adj); dev_free(d_ish); dev_free(d_out); dev_free(d_D); dev_free(d_S);
  dev_free(d_tw); dev_free(d_fw); dev_free(d_sos); dev_free(d_bv); dev_free(d_bi); dev_free(d_A);
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  HIP_CHECK(hipStreamDestroy(s));
}

// mt_tt_misfit.f90 MTTTMisfit%calc_adjoint_source for every trace: imeas 13
// (multitaper traveltime) or 14 (amplitude), with the cross-correlation
// fallback of the reference.  cfg(19) = the 7 cross-correlation values of
// fwat_cc_adjoint_batch_, then transfunc_waterlevel, water_threshold, mt_nw,
// num_taper, dt_fac, err_fac, dt_max_scale, min_cycle_in_window,
// use_cc_error (0/1), min_period, max_period, use_mt_error (0/1);
// meas(10, nwin, ntrace) = misfit, residual, error, cc_max, tshift, dlna,
// selected, ishift, measurement used (13/14, 11/12 after the fallback, 0 not
// selected), the other misfit; window adjoint sources are summed as the
// reference does, then band-passed (sos, nsect) and tapered as filter_adj.
// Limits: 32 <= nstep <= 16383 (transforms of up to 16384 samples in LDS).
void fwat_mt_adjoint_batch_(int* nstep, int* ntrace, double* dt, realw* obs, realw* syn, int* nwin, double* windows,
                            double* taper_percentage, int* itaper_type, int* imeas, double* cfg, int* nsect,
                            realw* sos, double* misfits, double* meas, realw* adj) {
  const int N = *nstep, T = *ntrace, nw = *nwin;
  if (N < 32 || N > 16383) SEM_FAIL("fwat_mt_adjoint_batch_: nstep must be in [32, 16383] (got %d)", N);
  if (T <= 0) SEM_FAIL("fwat_mt_adjoint_batch_: ntrace must be positive");
  if (nw < 1 || nw > MAX_WIN) SEM_FAIL("fwat_mt_adjoint_batch_: 1..%d windows", MAX_WIN);
  if (*imeas != 13 && *imeas != 14) SEM_FAIL("fwat_mt_adjoint_batch_: imeas must be 13 (MT-TT) or 14 (MT-DLNA)");
  if (T > 65535) SEM_FAIL("fwat_mt_adjoint_batch_: at most 65535 traces per call");
  if (*nsect < 0 || *nsect > 8) SEM_FAIL("fwat_mt_adjoint_batch_: 0..8 filter sections");
  if (*taper_percentage < 0.0 || *taper_percentage > 1.0) SEM_FAIL("taper_percentage must be 0 < %% < 1");
  const int ntaper = (int)cfg[10];
  if (ntaper < 2 || ntaper > 16) SEM_FAIL("fwat_mt_adjoint_batch_: num_taper must be in [2, 16]");
  if (!(cfg[16] > 0.0) || !(cfg[17] > 0.0)) SEM_FAIL("fwat_mt_adjoint_batch_: min/max period must be positive");
  const CcCfg C{cfg[0], cfg[1], cfg[2], cfg[3], cfg[4], cfg[5], cfg[6]};
  const MtCfg MC{cfg[7],   cfg[11], cfg[12], cfg[13], cfg[14], cfg[16], cfg[8], ntaper, cfg[15] != 0.0 ? 1 : 0,
                 cfg[18] != 0.0 ? 1 : 0};
  int e2 = 0;
  std::frexp((float)N, &e2);
  const int nf = 1 << e2, M = nf / 2;  // 2 ** exponent(real(nlen))
  int logM = 0;
  while ((1 << logM) < M) ++logM;
  // twiddles exp(-2 pi i j / nf), j < nf / 2
  std::vector<double> twh(2 * (size_t)M);
  for (int j = 0; j < M; ++j) {
    const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)j / (long double)nf;
    twh[2 * j] = (double)std::cos(a);
    twh[2 * j + 1] = (double)std::sin(a);
  }
  const size_t n = (size_t)N * T;
  hipStream_t s;
  HIP_CHECK(hipStreamCreate(&s));
  hipEvent_t ev0, ev1;
  HIP_CHECK(hipEventCreate(&ev0));
  HIP_CHECK(hipEventCreate(&ev1));
  float* d_obs = dev_upload(obs, n, s);
  float* d_syn = dev_upload(syn, n, s);
  float* d_adj = dev_alloc<float>(n);
  double2* d_twd = reinterpret_cast<double2*>(dev_upload(twh.data(), twh.size(), s));
  int* d_ish = dev_alloc<int>(T);
  int* d_flim = dev_alloc<int>(3 * (size_t)T);
  double* d_out = dev_alloc<double>(10 * (size_t)nw * T);
  int maxnl = 0;
  std::vector<int> nbs(nw), nls(nw);
  std::vector<std::vector<double>> tws(nw), taps(nw);
  double wmin = windows[0], wmax = windows[0];
  for (int w = 0; w < nw; ++w) {
    window_range(windows[2 * w], windows[2 * w + 1], *dt, N, nbs[w], nls[w]);
    taper_weights(nls[w], *taper_percentage, *itaper_type, tws[w]);
    bool same = false;
    for (int v = 0; v < w && !same; ++v)
      if (nls[v] == nls[w]) { taps[w] = taps[v]; same = true; }
    if (!same) dpss_tapers(nls[w], cfg[9], ntaper, taps[w]);
    maxnl = std::max(maxnl, nls[w]);
    wmin = std::min(wmin, std::min(windows[2 * w], windows[2 * w + 1]));
    wmax = std::max(wmax, std::max(windows[2 * w], windows[2 * w + 1]));
  }
  double* d_D = dev_alloc<double>((size_t)T * maxnl);
  double* d_S = dev_alloc<double>((size_t)T * maxnl);
  double* d_A = dev_alloc<double>((size_t)T * maxnl);
  double2* d_g01 = dev_alloc<double2>(2 * (size_t)T * (M + 1));
  double* d_g2 = dev_alloc<double>((size_t)T * (M + 1));
  MtJack jack{nullptr, nullptr, nullptr, nullptr};
  if (MC.use_mt_error) {
    const size_t rows = (size_t)T * ntaper * (M + 1);
    jack.pd = dev_alloc<double2>(rows);
    jack.pb = dev_alloc<double>(rows);
    jack.xj = dev_alloc<double2>(rows);
    jack.er = dev_alloc<double2>((size_t)T * (M + 1));
  }
  double* d_tw = dev_alloc<double>(maxnl);
  double* d_tap = dev_alloc<double>((size_t)ntaper * maxnl);
  const int maxblk = (2 * maxnl - 1 + CC_LAGS - 1) / CC_LAGS;
  double* d_bv = dev_alloc<double>((size_t)T * maxblk);
  int* d_bi = dev_alloc<int>((size_t)T * maxblk);
  const int imcc = *imeas == 13 ? 11 : 12;
  float* d_obsT = dev_alloc<float>(n);
  float* d_synT = dev_alloc<float>(n);
  HIP_CHECK(hipEventRecord(ev0, s));
  hipLaunchKernelGGL(k_fill_f32m, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, d_adj);
  const dim3 ttiles((unsigned)((T + 63) / 64), (unsigned)((N + 63) / 64));
  hipLaunchKernelGGL(k_rows_f32, ttiles, dim3(256), 0, s, N, T, d_obs, d_obsT);
  hipLaunchKernelGGL(k_rows_f32, ttiles, dim3(256), 0, s, N, T, d_syn, d_synT);
  hipLaunchKernelGGL(k_mt_flim, dim3((unsigned)T), dim3(MT_THREADS), 0, s, N, T, *dt, nf, logM, d_twd, d_synT,
                     cfg[8], cfg[16], cfg[17], cfg[9], cfg[14], d_flim);
  const unsigned nbk = (unsigned)((T + MTPB - 1) / MTPB);
  for (int w = 0; w < nw; ++w) {
    const int nl = nls[w];
    double* ow = d_out + 10 * (size_t)w * T;
    HIP_CHECK(hipMemcpyAsync(d_tw, tws[w].data(), nl * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(d_tap, taps[w].data(), (size_t)ntaper * nl * sizeof(double), hipMemcpyHostToDevice, s));
    const dim3 tiles((unsigned)((T + 63) / 64), (unsigned)((nl + 63) / 64));
    hipLaunchKernelGGL(k_cc_rows, tiles, dim3(256), 0, s, T, nbs[w], nl, d_tw, d_obs, d_syn, d_D, d_S);
    const int nblk = (2 * nl - 1 + CC_LAGS - 1) / CC_LAGS;
    hipLaunchKernelGGL(k_cc_corr, dim3((unsigned)nblk, (unsigned)T), dim3(256), 0, s, nl, d_D, d_S, d_bv, d_bi);
    hipLaunchKernelGGL(k_cc_pick, dim3(nbk), dim3(MTPB), 0, s, T, nblk, nl, d_bv, d_bi, d_ish);
    hipLaunchKernelGGL(k_cc_finish, dim3((unsigned)T), dim3(256), 0, s, nl, *dt, imcc, C, d_tw, d_D, d_S, d_ish, d_A,
                       ow, 10);
    hipLaunchKernelGGL(k_mt_window, dim3((unsigned)T), dim3(MT_THREADS), 0, s, N, T, *dt, nf, logM, d_twd, nbs[w], nl,
                       *imeas, MC, d_tw, d_tap, d_obsT, d_S, d_flim, d_g01, d_g01 + (size_t)T * (M + 1), d_g2, d_A,
                       ow, jack);
    hipLaunchKernelGGL(k_mt_put, tiles, dim3(256), 0, s, T, nbs[w], nl, d_A, ow, 10, d_adj);
  }
  std::vector<double> fw;
  double* d_fw = nullptr;
  float* d_sos = nullptr;
  if (*nsect > 0) {
    int fnb = 0, fnl = 0;
    window_range(wmin, wmax, *dt, N, fnb, fnl);
    taper_weights(fnl, *taper_percentage, *itaper_type, fw);
    d_fw = dev_upload(fw.data(), fw.size(), s);
    d_sos = dev_upload(sos, 6 * (size_t)*nsect, s);
    launch_sos(s, N, T, *nsect, d_sos, d_adj);
    hipLaunchKernelGGL(k_window_final, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, N, T, fnb, fnl, d_fw,
                       d_adj);
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipEventRecord(ev1, s));
  std::vector<double> out(10 * (size_t)nw * T);
  HIP_CHECK(hipMemcpyAsync(adj, d_adj, n * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(out.data(), d_out, out.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  float ms = 0.0f;
  HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
  g_meas_ms = ms;
  for (int t = 0; t < T; ++t) {
    double m = 0.0;
    for (int w = 0; w < nw; ++w) {
      const double* o = &out[10 * ((size_t)w * T + t)];
      m += o[0];
      if (meas) for (int q = 0; q < 10; ++q) meas[10 * ((size_t)t * nw + w) + q] = o[q];
    }
    misfits[t] = m;
  }
  dev_free(d_obs); dev_free(d_syn); dev_free(d_adj); dev_free(d_twd); dev_free(d_ish); dev_free(d_flim);
  dev_free(d_out); dev_free(d_D); dev_free(d_S); dev_free(d_A); dev_free(d_tw); dev_free(d_tap); dev_free(d_bv);
  dev_free(d_bi); dev_free(d_fw); dev_free(d_sos); dev_free(d_g01); dev_free(d_g2); dev_free(d_obsT); dev_free(d_synT);
  dev_free(jack.pd); dev_free(jack.pb); dev_free(jack.xj); dev_free(jack.er);
  HIP_CHECK(hipEventDestroy(ev0));
  HIP_CHECK(hipEventDestroy(ev1));
  HIP_CHECK(hipStreamDestroy(s));
}

// exponentiated_phase_misfit.f90 ExponentiatedPhaseMisfit%calc_adjoint_source
// (imeas 4) for every trace: wtr_env the envelope water level (adj_config);
// meas(8, nwin, ntrace) = misfit, residual, misfit_real, misfit_imag,
// residual_real, residual_imag, selected (1), 0; misfits(ntrace) summed over
// windows; window adjoint sources summed, then band-passed (sos, nsect) and
// tapered as filter_adj.  Windows of 32 to 16383 samples.
void fwat_ep_adjoint_batch_(int* nstep, int* ntrace, double* dt, realw* obs, realw* syn, int* nwin, double* windows,
                            double* taper_percentage, int* itaper_type, double* wtr_env, int* nsect, realw* sos,
                            double* misfits, double* meas, realw* adj) {
  const int N = *nstep, T = *ntrace, nw = *nwin;
  if (N <= 0 || T <= 0) SEM_FAIL("fwat_ep_adjoint_batch_: sizes must be positive");
  if (nw < 1 || nw > MAX_WIN) SEM_FAIL("fwat_ep_adjoint_batch_: 1..%d windows", MAX_WIN);
  if (T > 65535) SEM_FAIL("fwat_ep_adjoint_batch_: at most 65535 traces per call");
  if (*nsect < 0 || *nsect > 8) SEM_FAIL("fwat_ep_adjoint_batch_: 0..8 filter sections");
  if (*taper_percentage < 0.0 || *taper_percentage > 1.0) SEM_FAIL("taper_percentage must be 0 < %% < 1");
  const size_t n = (size_t)N * T;
  int maxnl = 0, maxM = 0;
  std::vector<int> nbs(nw), nls(nw), nfes(nw);
  std::vector<std::vector<double>> tws(nw);
  double wmin = windows[0], wmax = windows[0];
  for (int w = 0; w < nw; ++w) {
    window_range(windows[2 * w], windows[2 * w + 1], *dt, N, nbs[w], nls[w]);
    if (nls[w] < 32 || nls[w] > 16383) SEM_FAIL("fwat_ep_adjoint_batch_: windows of 32 to 16383 samples (got %d)", nls[w]);
    taper_weights(nls[w], *taper_percentage, *itaper_type, tws[w]);
    int e2 = 0;
    std::frexp((float)nls[w], &e2);
    nfes[w] = 1 << e2;  // 2 ** exponent(real(n)) of hilbert
    maxnl = std::max(maxnl, nls[w]);
    maxM = std::max(maxM, nfes[w] / 2);
    wmin = std::min(wmin, std::min(windows[2 * w], windows[2 * w + 1]));
    wmax = std::max(wmax, std::max(windows[2 * w], windows[2 * w + 1]));
  }
  hipStream_t s;
  HIP_CHECK(hipStreamCreate(&s));
  hipEvent_t ev0, ev1;
  HIP_CHECK(hipEventCreate(&ev0));
  HIP_CHECK(hipEventCreate(&ev1));
  float* d_obs = dev_upload(obs, n, s);
  float* d_syn = dev_upload(syn, n, s);
  float* d_adj = dev_alloc<float>(n);
  double* d_out = dev_alloc<double>(8 * (size_t)nw * T);
  double* d_D = dev_alloc<double>((size_t)T * maxnl);
  double* d_S = dev_alloc<double>((size_t)T * maxnl);
  double* d_A = dev_alloc<double>((size_t)T * maxnl);
  double* d_R = dev_alloc<double>(4 * (size_t)T * maxnl);
  double2* d_g01 = dev_alloc<double2>(2 * (size_t)T * (maxM + 1));
  // twiddles exp(-2 pi i j / nfe), j < nfe / 2, and the window tapers of every window, uploaded once
  std::vector<double> twh, twin;
  std::vector<size_t> toff(nw), woff(nw);
  for (int w = 0; w < nw; ++w) {
    const int nfe = nfes[w], M = nfe / 2;
    toff[w] = twh.size() / 2;
    for (int j = 0; j < M; ++j) {
      const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)j / (long double)nfe;
      twh.push_back((double)std::cos(a));
      twh.push_back((double)std::sin(a));
    }
    woff[w] = twin.size();
    twin.insert(twin.end(), tws[w].begin(), tws[w].end());
  }
  double* d_tw = dev_upload(twin.data(), twin.size(), s);
  double2* d_twd = reinterpret_cast<double2*>(dev_upload(twh.data(), twh.size(), s));
  HIP_CHECK(hipEventRecord(ev0, s));
  hipLaunchKernelGGL(k_fill_f32m, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, d_adj);
  for (int w = 0; w < nw; ++w) {
    const int nl = nls[w], nfe = nfes[w], M = nfe / 2;
    int logM = 0;
    while ((1 << logM) < M) ++logM;
    double* ow = d_out + 8 * (size_t)w * T;
    const dim3 tiles((unsigned)((T + 63) / 64), (unsigned)((nl + 63) / 64));
    hipLaunchKernelGGL(k_cc_rows, tiles, dim3(256), 0, s, T, nbs[w], nl, d_tw + woff[w], d_obs, d_syn, d_D, d_S);
    hipLaunchKernelGGL(k_ep_window, dim3((unsigned)T), dim3(MT_THREADS), 0, s, nl, *dt, nfe, logM, *wtr_env,
                       d_twd + toff[w], d_tw + woff[w], d_D, d_S, d_g01, d_g01 + (size_t)T * (M + 1), d_R, d_A, ow);
    hipLaunchKernelGGL(k_mt_put, tiles, dim3(256), 0, s, T, nbs[w], nl, d_A, ow, 8, d_adj);
  }
  const unsigned nbk = (unsigned)((T + MTPB - 1) / MTPB);
  std::vector<double> fw;
  double* d_fw = nullptr;
  float* d_sos = nullptr;
  if (*nsect > 0) {
    int fnb = 0, fnl = 0;
    window_range(wmin, wmax, *dt, N, fnb, fnl);
    taper_weights(fnl, *taper_percentage, *itaper_type, fw);
    d_fw = dev_upload(fw.data(), fw.size(), s);
    d_sos = dev_upload(sos, 6 * (size_t)*nsect, s);
    launch_sos(s, N, T, *nsect, d_sos, d_adj);
    hipLaunchKernelGGL(k_window_final, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, N, T, fnb, fnl, d_fw,
                       d_adj);
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipEventRecord(ev1, s));
  std::vector<double> out(8 * (size_t)nw * T);
  HIP_CHECK(hipMemcpyAsync(adj, d_adj, n * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(out.data(), d_out, out.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  float ms = 0.0f;
  HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
  g_meas_ms = ms;
  for (int t = 0; t < T; ++t) {
    double m = 0.0;
    for (int w = 0; w < nw; ++w) {
      const double* o = &out[8 * ((size_t)w * T + t)];
      m += o[0];
      if (meas) for (int q = 0; q < 8; ++q) meas[8 * ((size_t)t * nw + w) + q] = o[q];
    }
    misfits[t] = m;
  }
  dev_free(d_obs); dev_free(d_syn); dev_free(d_adj); dev_free(d_out); dev_free(d_D); dev_free(d_S); dev_free(d_A);
  dev_free(d_R); dev_free(d_g01); dev_free(d_tw); dev_free(d_twd); dev_free(d_fw); dev_free(d_sos);
  HIP_CHECK(hipEventDestroy(ev0));
  HIP_CHECK(hipEventDestroy(ev1));
  HIP_CHECK(hipStreamDestroy(s));
}

// win_sel_type%select_windows (window_selector/win_sel.f90:35-277) for every
// trace of an event: obs, syn [nstep][ntrace] (band-passed as the caller does
// before selection), tp(ntrace) first-arrival times, dist(ntrace) epicentral
// distances, t0 the simulation start offset; cfg(10) = sliding_win_len_fac,
// min_velocity, threshold_corr, threshold_shift_fac, jump_fac, min_win_len_fac,
// min_snr_window, min_energy_ratio, min_peaks_troughs, is_split_phases (groups
// split at the STA/LTA troughs between phases; traces of at most 16383 samples);
// nwin(ntrace) out (windows found, may exceed maxwin), windows(2, maxwin,
// ntrace) out in seconds
void fwat_win_sel_batch_(int* nstep, int* ntrace, double* dt, realw* obs, realw* syn, double* t0, double* tp,
                         double* dist, double* min_period, double* cfg, int* maxwin, int* nwin, double* windows) {
  const int N = *nstep, T = *ntrace;
  if (N <= 0 || T <= 0) SEM_FAIL("fwat_win_sel_batch_: sizes must be positive");
  if (T > 65535) SEM_FAIL("fwat_win_sel_batch_: at most 65535 traces per call");
  const bool split = cfg[9] != 0.0;
  if (split && N > 16383) SEM_FAIL("fwat_win_sel_batch_: phase splitting takes traces of at most 16383 samples");
  if (*maxwin < 1) SEM_FAIL("fwat_win_sel_batch_: maxwin must be positive");
  const int nlen = (int)(cfg[0] * *min_period / *dt) + 1;
  if (nlen < 2 || nlen > 4090) SEM_FAIL("fwat_win_sel_batch_: sliding window of %d samples (2..4090)", nlen);
  const int nw = N - nlen;
  const WsCfg C{cfg[0], cfg[1], cfg[2], cfg[3], cfg[4], cfg[5], cfg[6], cfg[7], (int)cfg[8], split ? 1 : 0};
  if (nw <= 0) {
    for (int t = 0; t < T; ++t) nwin[t] = 0;
    return;
  }
  std::vector<double> wt;
  taper_weights(nlen, 1.0, 1, wt);
  const size_t n = (size_t)N * T;
  hipStream_t s;
  HIP_CHECK(hipStreamCreate(&s));
  hipEvent_t ev0, ev1;
  HIP_CHECK(hipEventCreate(&ev0));
  HIP_CHECK(hipEventCreate(&ev1));
  float* d_obs = dev_upload(obs, n, s);
  float* d_syn = dev_upload(syn, n, s);
  double* d_tp = dev_upload(tp, (size_t)T, s);
  double* d_dist = dev_upload(dist, (size_t)T, s);
  double* d_wt = dev_upload(wt.data(), wt.size(), s);
  float* d_obsT = dev_alloc<float>(n);
  float* d_synT = dev_alloc<float>(n);
  double* d_Dn = dev_alloc<double>(n);
  double* d_Sn = dev_alloc<double>(n);
  double* d_noise = dev_alloc<double>(T);
  double* d_ts = dev_alloc<double>((size_t)nw * T);
  double* d_cc = dev_alloc<double>((size_t)nw * T);
  int* d_ish = dev_alloc<int>((size_t)nw * T);
  int* d_nwin = dev_alloc<int>(T);
  double* d_win = dev_alloc<double>(2 * (size_t)*maxwin * T);
  unsigned char* d_good = dev_alloc<unsigned char>((size_t)nw * T);
  // phase splitting: envelope transform on nf = 2 ** exponent(real(nstep)) points
  int nfe = 0, Me = 0, logMe = 0;
  double *d_env = nullptr, *d_stl = nullptr;
  double2 *d_g01 = nullptr, *d_twe = nullptr;
  if (split) {
    int e2 = 0;
    std::frexp((float)N, &e2);
    nfe = 1 << e2;
    Me = nfe / 2;
    while ((1 << logMe) < Me) ++logMe;
    std::vector<double> twh(2 * (size_t)Me);
    for (int j = 0; j < Me; ++j) {
      const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)j / (long double)nfe;
      twh[2 * j] = (double)std::cos(a);
      twh[2 * j + 1] = (double)std::sin(a);
    }
    d_twe = reinterpret_cast<double2*>(dev_upload(twh.data(), twh.size(), s));
    d_env = dev_alloc<double>(n);
    d_stl = dev_alloc<double>(n);
    d_g01 = dev_alloc<double2>(2 * (size_t)T * (Me + 1));
  }
  HIP_CHECK(hipEventRecord(ev0, s));
  const dim3 ttiles((unsigned)((T + 63) / 64), (unsigned)((N + 63) / 64));
  hipLaunchKernelGGL(k_rows_f32, ttiles, dim3(256), 0, s, N, T, d_obs, d_obsT);
  hipLaunchKernelGGL(k_rows_f32, ttiles, dim3(256), 0, s, N, T, d_syn, d_synT);
  hipLaunchKernelGGL(k_ws_prep, dim3((unsigned)T), dim3(256), 0, s, N, *dt, *t0, *min_period, d_tp, d_obsT, d_synT,
                     d_Dn, d_Sn, d_noise);
  const int nwv = nlen <= 1020 ? 4 : (nlen <= 2044 ? 2 : 1);  // at most 64 KB of LDS per workgroup
  hipLaunchKernelGGL(k_ws_sliding, dim3((unsigned)((nw + nwv - 1) / nwv), (unsigned)T), dim3(64 * nwv),
                     (size_t)nwv * (2 * nlen + 7) * sizeof(double), s, N, nlen, nw, d_wt, d_Dn, d_Sn, d_ish);
  hipLaunchKernelGGL(k_ws_coef, dim3((unsigned)((nw + 255) / 256), (unsigned)T), dim3(256), 0, s, N, nlen, nw, *dt,
                     d_wt, d_Dn, d_Sn, d_ish, d_ts, d_cc);
  if (split)
    hipLaunchKernelGGL(k_ws_env, dim3((unsigned)T), dim3(MT_THREADS), 0, s, N, nfe, logMe, d_twe, d_Sn, d_g01,
                       d_g01 + (size_t)T * (Me + 1), d_env);
  hipLaunchKernelGGL(k_ws_select, dim3((unsigned)T), dim3(256), 0, s, N, nlen, nw, *dt, *t0, *min_period, C, d_tp,
                     d_dist, d_noise, d_Dn, d_Sn, d_ts, d_cc, d_good, d_env, d_stl, *maxwin, d_nwin, d_win);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipEventRecord(ev1, s));
  HIP_CHECK(hipMemcpyAsync(nwin, d_nwin, T * sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(windows, d_win, 2 * (size_t)*maxwin * T * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  float ms = 0.0f;
  HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
  g_meas_ms = ms;
  dev_free(d_obs); dev_free(d_syn); dev_free(d_tp); dev_free(d_dist); dev_free(d_wt); dev_free(d_obsT);
  dev_free(d_synT); dev_free(d_Dn); dev_free(d_Sn); dev_free(d_noise); dev_free(d_ts); dev_free(d_cc); dev_free(d_ish);
  dev_free(d_nwin); dev_free(d_win); dev_free(d_good); dev_free(d_env); dev_free(d_stl); dev_free(d_g01);
  dev_free(d_twe);
  for (int t = 0; t < T; ++t)
    if (nwin[t] < 0) SEM_FAIL("fwat_win_sel_batch_: trace %d has more than %d good groups", t, WS_MAXG);
  HIP_CHECK(hipEventDestroy(ev0));
  HIP_CHECK(hipEventDestroy(ev1));
  HIP_CHECK(hipStreamDestroy(s));
}

// waveform_conv_misfit.f90 WaveformConvMisfit%calc_adjoint_source for every
// receiver (telecc_data.f90:175-178): dat_r, dat_z, syn_r, syn_z [nstep][nrec]
// single precision; windows(2, nrec) in seconds; misfits(nrec), residuals(nrec)
// out; adj_r, adj_z [nstep][nrec] out (zero outside each window; the caller
// rotates R to N/E and weights as telecc_data does)
void fwat_conv_adjoint_batch_(int* nstep, int* nrec, double* dt, realw* dat_r, realw* dat_z, realw* syn_r,
                              realw* syn_z, double* windows, double* taper_percentage, int* itaper_type,
                              double* misfits, double* residuals, realw* adj_r, realw* adj_z) {
  const int N = *nstep, R = *nrec;
  if (N < 2 || R <= 0) SEM_FAIL("fwat_conv_adjoint_batch_: sizes must be positive");
  if (R > 65535) SEM_FAIL("fwat_conv_adjoint_batch_: at most 65535 receivers per call");
  if (*taper_percentage < 0.0 || *taper_percentage > 1.0) SEM_FAIL("taper_percentage must be 0 < %% < 1");
  const int ind0 = (N % 2 == 0) ? N / 2 : N / 2 - 1;  // 0-based first sample of myconvolution_dp's part 0
  std::vector<int> nbw(R), nlw(R), kb0(R, ind0), nall(R, N), kbw(R);
  int maxnl = 0;
  for (int r = 0; r < R; ++r) {
    window_range(windows[2 * r], windows[2 * r + 1], *dt, N, nbw[r], nlw[r]);
    kbw[r] = ind0 + nbw[r];
    maxnl = std::max(maxnl, nlw[r]);
  }
  const size_t n = (size_t)N * R;
  hipStream_t s;
  HIP_CHECK(hipStreamCreate(&s));
  hipEvent_t ev0, ev1;
  HIP_CHECK(hipEventCreate(&ev0));
  HIP_CHECK(hipEventCreate(&ev1));
  float* d_in[4];
  realw* h_in[4] = {dat_r, dat_z, syn_r, syn_z};
  for (int q = 0; q < 4; ++q) d_in[q] = dev_upload(h_in[q], n, s);
  double* d_rows = dev_alloc<double>(4 * n);  // dat_r, dat_z, syn_r, syn_z rows
  double* d_cd = dev_alloc<double>(n);
  double* d_ar = dev_alloc<double>(n);
  double* d_az = dev_alloc<double>(n);
  int* d_nb = dev_upload(nbw.data(), (size_t)R, s);
  int* d_nl = dev_upload(nlw.data(), (size_t)R, s);
  int* d_kb0 = dev_upload(kb0.data(), (size_t)R, s);
  int* d_nall = dev_upload(nall.data(), (size_t)R, s);
  int* d_kbw = dev_upload(kbw.data(), (size_t)R, s);
  double* d_out = dev_alloc<double>(2 * (size_t)R);
  float* d_ar_out = dev_alloc<float>(n);
  float* d_az_out = dev_alloc<float>(n);
  HIP_CHECK(hipEventRecord(ev0, s));
  hipLaunchKernelGGL(k_fill_f32m, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, d_ar_out);
  hipLaunchKernelGGL(k_fill_f32m, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, d_az_out);
  const dim3 tt((unsigned)((R + 63) / 64), (unsigned)((N + 63) / 64));
  for (int q = 0; q < 4; ++q) hipLaunchKernelGGL(k_rows_f64, tt, dim3(256), 0, s, N, R, d_in[q], d_rows + q * n);
  const double* dr = d_rows;
  const double* dz = d_rows + n;
  const double* sr = d_rows + 2 * n;
  const double* sz = d_rows + 3 * n;
  const dim3 gall((unsigned)((N + CV_OUT - 1) / CV_OUT), (unsigned)R);
  const dim3 gwin((unsigned)((maxnl + CV_OUT - 1) / CV_OUT), (unsigned)R);
  // diff = (syn_r * dat_z - dat_r * syn_z) dt
  hipLaunchKernelGGL(k_conv_mid, gall, dim3(256), 0, s, N, N, sr, 0, 1.0, dz, d_kb0, d_nall, N, 0, 1.0, d_cd);
  hipLaunchKernelGGL(k_conv_mid, gall, dim3(256), 0, s, N, N, dr, 0, 1.0, sz, d_kb0, d_nall, N, 1, *dt, d_cd);
  // adjoint sources on the window: conv(rev(dat_z), diff) dt, conv(-rev(dat_r), diff) dt
  hipLaunchKernelGGL(k_conv_mid, gwin, dim3(256), 0, s, N, N, dz, 1, 1.0, d_cd, d_kbw, d_nl, N, 0, *dt, d_ar);
  hipLaunchKernelGGL(k_conv_mid, gwin, dim3(256), 0, s, N, N, dr, 1, -1.0, d_cd, d_kbw, d_nl, N, 0, *dt, d_az);
  hipLaunchKernelGGL(k_conv_finish, dim3((unsigned)R), dim3(256), 0, s, N, R, *dt, d_nb, d_nl, *taper_percentage,
                     *itaper_type, d_cd, d_ar, d_az, d_out, d_ar_out, d_az_out);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipEventRecord(ev1, s));
  std::vector<double> out(2 * (size_t)R);
  HIP_CHECK(hipMemcpyAsync(adj_r, d_ar_out, n * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(adj_z, d_az_out, n * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(out.data(), d_out, out.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  float ms = 0.0f;
  HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
  g_meas_ms = ms;
  for (int r = 0; r < R; ++r) {
    misfits[r] = out[2 * r];
    if (residuals) residuals[r] = out[2 * r + 1];
  }
  for (int q = 0; q < 4; ++q) dev_free(d_in[q]);
  dev_free(d_rows); dev_free(d_cd); dev_free(d_ar); dev_free(d_az); dev_free(d_nb); dev_free(d_nl); dev_free(d_kb0);
  dev_free(d_nall); dev_free(d_kbw); dev_free(d_out); dev_free(d_ar_out); dev_free(d_az_out);
  HIP_CHECK(hipEventDestroy(ev0));
  HIP_CHECK(hipEventDestroy(ev1));
  HIP_CHECK(hipStreamDestroy(s));
}

}  // extern "C"

// ---------------------------------------------------------------------------
// FLEXWIN window selection of every trace (flexwin_type%select_windows,
// src/preproc/window_selector/flexwin.f90:75-734, dispatched for WIN_TYPE 2 at
// leq_data.f90:116-121).  Stages, each a launch over all traces:
//   k_fw_prep    normalisation by max|syn| and the noise level (one workgroup
//                per trace; the noise sum on one lane in the reference's order)
//   k_ws_env     envelope |hilbert(syn)| through the LDS real transform
//   k_fw_stalta  STA/LTA on one lane (a recursion), then the STA/LTA peaks and
//                troughs as ordered index lists (workgroup scans)
//   k_fw_cand    every candidate (trough, peak, trough) in the reference's
//                enumeration order, one lane per candidate: clipping to the
//                time range and the per-candidate passes 4.0-4.5 (travel
//                times, minimum length, minima water level, prominence, phase
//                separation, curtailing); survivors compacted in order, then
//                duplicates (the first of each (left, right) pair stays, peak
//                recentred) and the second length pass
//   k_fw_crit    one workgroup per surviving window: the full cross-correlation
//                (one lag per lane, each lag's sum in the reference's order,
//                first maximum), dlnA, cc_max_coef, SNR, peak/trough count
//   k_fw_finish  one lane per trace: the data-fit rejections, then merging
//                (windows sorted by left end, touching ones joined; criteria
//                recomputed by k_fw_crit, written by k_fw_out) or weighted
//                interval scheduling (quicksort by right end with the
//                reference's pivot and 1e-6 ties, p(j) by bisection, optimum,
//                backtrack)
#pragma clang fp contract(off)
namespace {

constexpr int FW_THREADS = 256;
constexpr int FW_CAP = 16384;       // candidates kept through the per-candidate passes, per trace
constexpr int FW_MAXCAND = 500000;  // max_good_win (flexwin.f90:11)

struct FwCfg {
  double min_len;    // min_win_len_fac * min_period
  double thr_corr;   // threshold_corr
  double thr_shift;  // threshold_shift_fac * min_period
  double thr_dlna, min_snr, wl, wmid, c2, c3a, c3b, tdl, tdr, min_period, dt;
  int min_pt, split;
};

// per-trace time range: tstart, tend, min_time, max_time (s) and nstart, nend,
// the number of samples of the noise level (0: fixed 0.001)
struct FwTrace {
  double tstart, tend, min_time, max_time;
  int nstart, nend, nnoise, pad;
};

// exclusive prefix of v over the workgroup; *total gets the sum
__device__ int fw_scan(int v, int* sh, int* total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = 1; o < FW_THREADS; o <<= 1) {
    const int x = t >= o ? sh[t - o] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  *total = sh[FW_THREADS - 1];
  const int r = sh[t] - v;
  __syncthreads();
  return r;
}

// first index in a[0..n) with a[i] >= v (a ascending)
__device__ int fw_lower(const int* a, int n, int v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (a[m] < v) lo = m + 1;
    else hi = m;
  }
  return lo;
}

__global__ void __launch_bounds__(FW_THREADS) k_fw_prep(int nstep, const FwTrace* __restrict__ tr,
                                                       const float* __restrict__ obsT,
                                                       const float* __restrict__ synT, double* __restrict__ Dn,
                                                       double* __restrict__ Sn, double* __restrict__ noise) {
  __shared__ double sh[FW_THREADS];
  const int t = blockIdx.x;
  const float* d = obsT + (size_t)t * nstep;
  const float* s = synT + (size_t)t * nstep;
  double m = 0.0;
  for (int i = threadIdx.x; i < nstep; i += FW_THREADS) m = fmax(m, fabs((double)s[i]));
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int o = FW_THREADS / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  const double smax = sh[0];
  for (int i = threadIdx.x; i < nstep; i += FW_THREADS) {
    Dn[(size_t)t * nstep + i] = (double)d[i] / smax;
    Sn[(size_t)t * nstep + i] = (double)s[i] / smax;
  }
  if (threadIdx.x == 0) {  // initialize (flexwin.f90:92-106): the un-normalised data, summed in order
    const int m2 = tr[t].nnoise;
    double e = 0.0;
    for (int i = 0; i < m2; ++i) e = e + (double)d[i] * (double)d[i];
    noise[t] = m2 > 0 ? sqrt(e / (double)m2) : 0.001;
  }
}

// STA_LTA (sta_lta.f90:6-104, TWEAK off) of the envelope on lane 0, then
// find_maxima of STA/LTA and of its negative as 1-based index lists
__global__ void __launch_bounds__(FW_THREADS) k_fw_stalta(int nstep, double Cs, double Cl,
                                                         const double* __restrict__ env, double* __restrict__ SL,
                                                         int* __restrict__ PK, int* __restrict__ TR,
                                                         int* __restrict__ NPT) {
  __shared__ double rm[FW_THREADS];
  __shared__ int sh[FW_THREADS];
  const int t = blockIdx.x, lane = threadIdx.x, cap = nstep / 2 + 1;
  const double* e = env + (size_t)t * nstep;
  double* sl = SL + (size_t)t * nstep;
  double m = 0.0;
  for (int i = lane; i < nstep; i += FW_THREADS) m = fmax(m, e[i]);
  rm[lane] = m;
  __syncthreads();
  for (int o = FW_THREADS / 2; o > 0; o >>= 1) {
    if (lane < o) rm[lane] = fmax(rm[lane], rm[lane + o]);
    __syncthreads();
  }
  if (lane == 0) {
    const double noise = rm[0] / 1e5;
    double sta = 0.0, lta = 0.0;
    for (int i = 0; i < 1000; ++i) {
      sta = Cs * sta + noise;
      lta = Cl * lta + noise;
    }
    for (int i = 0; i < nstep; ++i) {
      const double x = e[i] + noise;
      sta = Cs * sta + x;
      lta = Cl * lta + x;
      sl[i] = lta > 1.0e-9 ? sta / lta : noise;
    }
  }
  __syncthreads();
  int np = 0, nt = 0;
  for (int base = 0; base < nstep; base += FW_THREADS) {
    const int i = base + lane;
    bool pk = false, tr = false;
    if (i >= 1 && i < nstep - 1) {
      const double a = sl[i - 1], b = sl[i], c = sl[i + 1];
      pk = b - a > 0.0 && c - b < 0.0;
      tr = (-b) - (-a) > 0.0 && (-c) - (-b) < 0.0;
    }
    int tot = 0;
    const int op = fw_scan(pk ? 1 : 0, sh, &tot);
    if (pk) PK[(size_t)t * cap + np + op] = i + 1;
    np += tot;
    const int ot = fw_scan(tr ? 1 : 0, sh, &tot);
    if (tr) TR[(size_t)t * cap + nt + ot] = i + 1;
    nt += tot;
  }
  if (lane == 0) {
    NPT[2 * t] = np;
    NPT[2 * t + 1] = nt;
  }
}

// candidates and the passes before the data-fit criteria (flexwin.f90:205-294,
// 393-625).  SV[t][k] = (left, right, peak) 1-based; NSV[t] = survivors, -1 when
// the trace has no peak or fewer than two troughs, -2 when more than FW_CAP
// candidates survive the per-candidate passes.
__global__ void __launch_bounds__(FW_THREADS) k_fw_cand(int nstep, FwCfg c, const FwTrace* __restrict__ trc,
                                                       const double* __restrict__ SL, const int* __restrict__ PK,
                                                       const int* __restrict__ TR, const int* __restrict__ NPT,
                                                       long long* __restrict__ PRE, int4* __restrict__ SV,
                                                       int* __restrict__ NSV) {
  __shared__ int sh[FW_THREADS];
  __shared__ unsigned char keep[FW_CAP];
  __shared__ long long total_s;
  const int t = blockIdx.x, lane = threadIdx.x, cap = nstep / 2 + 1;
  const int np = NPT[2 * t], nt = NPT[2 * t + 1];
  if (np == 0 || nt < 2) {
    if (lane == 0) NSV[t] = -1;
    return;
  }
  const double* sl = SL + (size_t)t * nstep;
  const int* pk = PK + (size_t)t * cap;
  const int* tr = TR + (size_t)t * cap;
  long long* pre = PRE + (size_t)t * (cap + 1);
  int4* sv = SV + (size_t)t * FW_CAP;
  const FwTrace T = trc[t];
  if (lane == 0) {  // candidates per peak in the enumeration order, capped at max_good_win
    long long acc = 0;
    pre[0] = 0;
    for (int i = 0; i < np; ++i) {
      if (acc < FW_MAXCAND && !(sl[pk[i] - 1] < c.wl)) {
        const int L = fw_lower(tr, nt, pk[i]);
        acc += (long long)L * (nt - L);
      }
      pre[i + 1] = acc;
    }
    total_s = acc < FW_MAXCAND ? acc : FW_MAXCAND;
  }
  __syncthreads();
  const long long total = total_s;
  int ns = 0;
  bool over = false;
  for (long long base = 0; base < total; base += FW_THREADS) {
    const long long cidx = base + lane;
    bool ok = false;
    int left = 0, right = 0, peak = 0;
    if (cidx < total) {
      int lo = 0, hi = np - 1;  // last peak with pre[i] <= cidx
      while (lo < hi) {
        const int m = (lo + hi + 1) >> 1;
        if (pre[m] <= cidx) lo = m;
        else hi = m - 1;
      }
      const int i = lo;
      peak = pk[i];
      const int L = fw_lower(tr, nt, peak), R = nt - L;
      const long long o = cidx - pre[i];
      left = tr[(int)(o / R)];
      right = tr[L + (int)(o % R)];
      if ((double)(left - 1) * c.dt < T.tstart) left = T.nstart;
      if ((double)(right - 1) * c.dt > T.tend) right = T.nend;
      ok = true;
      // 4.0 travel times
      const double ws = (double)(left - 1) * c.dt, we = (double)(right - 1) * c.dt;
      if (!(we >= T.min_time && ws <= T.max_time)) ok = false;
      // 4.1 minimum length
      if (ok && (double)(right - left) * c.dt < c.min_len) ok = false;
      // 4.2 minima water level: troughs strictly inside
      if (ok)
        for (int j = fw_lower(tr, nt, left + 1); j < nt && tr[j] < right; ++j)
          if (sl[tr[j] - 1] <= c.wmid) { ok = false; break; }
      // 4.3 prominence of the central peak (skipped when c_2 = 0)
      if (ok && c.c2 != 0.0) {
        if (L == 0 || L == nt) ok = false;
        else {
          const double cv = sl[peak - 1], dl = cv - sl[tr[L - 1] - 1], dr = cv - sl[tr[L] - 1];
          if (dl < c.c2 * cv || dr < c.c2 * cv) ok = false;
        }
      }
      // 4.4 phase separation
      if (ok) {
        double smin = 1.7976931348623157e308;
        for (int j = fw_lower(tr, nt, left); j < nt && tr[j] <= right; ++j)
          if (sl[tr[j] - 1] < smin) smin = sl[tr[j] - 1];
        const double dcen = sl[peak - 1] - smin;
        for (int j = fw_lower(pk, np, left); j < np && pk[j] <= right; ++j) {
          if (pk[j] == peak) continue;
          const double ds = sl[pk[j] - 1] - smin;
          const double dtm = (double)abs(peak - pk[j]) * c.dt / c.min_period;
          double ft = 1.0;
          if (dtm >= c.c3b) {
            const double a = (dtm - c.c3b) / c.c3b;
            ft = exp(-a * a);
          }
          if (ds > c.c3a * dcen * ft) { ok = false; break; }
        }
      }
      // 4.5 curtail to the internal maxima
      if (ok) {
        int nin = 0, il = 0, ir = 0;
        for (int j = fw_lower(pk, np, left); j < np && pk[j] <= right; ++j) {
          if (pk[j] == peak) continue;
          if (nin == 0) il = pk[j];
          ir = pk[j];
          ++nin;
        }
        if (nin >= 2) {
          const double dl = (double)(il - left), dr = (double)(right - ir);
          if (dl > c.tdl) left = (int)((double)il - c.tdl);
          if (dr > c.tdr) right = (int)((double)ir + c.tdr);
        }
      }
    }
    int tot = 0;
    const int off = fw_scan(ok ? 1 : 0, sh, &tot);
    if (ok) {
      if (ns + off < FW_CAP) sv[ns + off] = make_int4(left, right, peak, 0);
    }
    if (ns + tot > FW_CAP) over = true;
    ns = min(ns + tot, FW_CAP);
  }
  if (over) {
    if (lane == 0) NSV[t] = -2;
    return;
  }
  // 4.6 duplicates: a pair stays with its first candidate; 4.7 minimum length
  for (int k = lane; k < ns; k += FW_THREADS) {
    const int4 a = sv[k];
    bool dup = false;
    for (int q = 0; q < k && !dup; ++q) {
      const int4 b = sv[q];
      dup = b.x == a.x && b.y == a.y;
    }
    keep[k] = (!dup && !((double)(a.y - a.x) * c.dt < c.min_len)) ? 1 : 0;
  }
  __syncthreads();
  int nk = 0;
  for (int base = 0; base < ns; base += FW_THREADS) {
    const int k = base + lane;
    int4 a = make_int4(0, 0, 0, 0);
    const bool kp = k < ns && keep[k];
    if (kp) a = sv[k];
    int tot = 0;
    const int off = fw_scan(kp ? 1 : 0, sh, &tot);
    if (kp) sv[nk + off] = make_int4(a.x, a.y, (a.x + a.y) / 2, 0);
    nk += tot;
    __syncthreads();
  }
  if (lane == 0) NSV[t] = nk;
}

// calculate_criteria (flexwin.f90:110-181) of window e of trace blockIdx.y:
// CR[t][e] = cc, dlnA, time shift, weight, SNR (amplitude), rejected by the
// peak/trough count (0/1)
__global__ void __launch_bounds__(FW_THREADS) k_fw_crit(int nstep, int stride, FwCfg c,
                                                       const double* __restrict__ Dn,
                                                       const double* __restrict__ Sn,
                                                       const double* __restrict__ noise,
                                                       const int4* __restrict__ W, const int* __restrict__ NW,
                                                       double* __restrict__ CR) {
  __shared__ double bv[FW_THREADS];
  __shared__ int bi[FW_THREADS];
  const int t = blockIdx.y, e = blockIdx.x, lane = threadIdx.x;
  if (e >= NW[t]) return;
  const int4 w = W[(size_t)t * stride + e];
  const int nl = w.y - w.x + 1;
  const double* d = Dn + (size_t)t * nstep + (w.x - 1);
  const double* s = Sn + (size_t)t * nstep + (w.x - 1);
  // mycorrelation_dp(d, s, cc, 1): lag m sums d(i1) s(nl - (m - i1)) over increasing i1
  double best = -INFINITY;
  int bm = 0x7fffffff;
  for (int m = lane; m < 2 * nl - 1; m += FW_THREADS) {
    const int lo = m - (nl - 1) > 0 ? m - (nl - 1) : 0, hi = m < nl - 1 ? m : nl - 1;
    double acc = 0.0;
    for (int i1 = lo; i1 <= hi; ++i1) acc = acc + d[i1] * s[nl - 1 - (m - i1)];
    if (acc > best) { best = acc; bm = m; }
  }
  bv[lane] = best;
  bi[lane] = bm;
  __syncthreads();
  for (int o = FW_THREADS / 2; o > 0; o >>= 1) {
    if (lane < o) {
      const double v2 = bv[lane + o];
      const int i2 = bi[lane + o];
      if (v2 > bv[lane] || (v2 == bv[lane] && i2 < bi[lane])) { bv[lane] = v2; bi[lane] = i2; }
    }
    __syncthreads();
  }
  if (lane != 0) return;
  const int ish = bi[0] + 1 - nl;
  double sd2 = 0.0, ss2 = 0.0;
  for (int i = 0; i < nl; ++i) sd2 = sd2 + d[i] * d[i];
  for (int i = 0; i < nl; ++i) ss2 = ss2 + s[i] * s[i];
  const double dlna = 0.5 * log(sd2 / ss2);
  // cc_max_coef (utils/cc.f90:132-160): s shifted by ish, d's emptied edge zeroed
  double a = 0.0, b = 0.0, q = 0.0, amax = 0.0;
  for (int i = 0; i < nl; ++i) {
    const int j = i - ish;
    const double sv = (j >= 0 && j < nl) ? s[j] : 0.0;
    const bool z = (ish > 0 && i < ish) || (ish < 0 && i >= nl + ish);
    const double dv = z ? 0.0 : d[i];
    a = a + dv * sv;
    b = b + dv * dv;
    q = q + sv * sv;
    amax = fmax(amax, fabs(d[i]));
  }
  const double cc = a / sqrt(b * q);
  int npk = 0;
  for (int i = 1; i < nl - 1; ++i) {
    if (s[i] - s[i - 1] > 0.0 && s[i + 1] - s[i] < 0.0) ++npk;
    if ((-s[i]) - (-s[i - 1]) > 0.0 && (-s[i + 1]) - (-s[i]) < 0.0) ++npk;
  }
  double* o = CR + 6 * ((size_t)t * stride + e);
  o[0] = cc;
  o[1] = dlna;
  o[2] = ish * c.dt;
  o[3] = (double)(w.y - w.x) * c.dt / c.min_period * cc;
  o[4] = noise[t] > 0.0 ? amax / noise[t] : 1.0e5;
  o[5] = (c.min_pt > 0 && npk < c.min_pt) ? 1.0 : 0.0;
}

// data-fit rejections (flexwin.f90:270-293), then merge_windows (:627-700) into
// MW/NM (merging mode) or schedule_weighted_intervals (interval_scheduling.f90)
// straight into the outputs; one lane per trace.  Scratch rows per trace of
// `stride` entries: IX (int), IV (left, right, weight as double), PJ (int),
// OPT (stride + 1 doubles); STK holds 2 stride + 2 ranges (int2) per trace.
__global__ void k_fw_finish(int ntrace, int stride, FwCfg c, double t0, const int4* __restrict__ SV,
                            const int* __restrict__ NSV, const double* __restrict__ CR, int* __restrict__ IX,
                            double* __restrict__ IV, int* __restrict__ PJ, double* __restrict__ OPT,
                            int2* __restrict__ STK, int4* __restrict__ MW, int* __restrict__ NM, int maxwin,
                            int* __restrict__ nwin, double* __restrict__ win, double* __restrict__ info) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntrace) return;
  const int n0 = NSV[t];
  nwin[t] = 0;
  NM[t] = 0;
  if (n0 <= 0) return;
  const int4* sv = SV + (size_t)t * stride;
  const double* cr = CR + 6 * (size_t)t * stride;
  int* ix = IX + (size_t)t * stride;
  int nv = 0;
  for (int e = 0; e < n0; ++e) {
    const double* o = cr + 6 * e;
    bool rej = o[5] != 0.0;
    if (o[0] < c.thr_corr) rej = true;
    if (fabs(o[2]) > c.thr_shift) rej = true;
    if (fabs(o[1]) > c.thr_dlna) rej = true;
    if (o[4] < c.min_snr) rej = true;
    if (!rej) ix[nv++] = e;
  }
  if (!c.split) {
    // sort_candidates_by_left, then merge
    for (int i = 0; i < nv - 1; ++i)
      for (int j = i + 1; j < nv; ++j)
        if (sv[ix[j]].x < sv[ix[i]].x) { const int tmp = ix[i]; ix[i] = ix[j]; ix[j] = tmp; }
    int4* mw = MW + (size_t)t * stride;
    int nm = 0;
    for (int i = 0; i < nv; ++i) {
      const int4 w = sv[ix[i]];
      if (nm == 0 || mw[nm - 1].y + 1 < w.x) mw[nm++] = w;
      else mw[nm - 1].y = max(mw[nm - 1].y, w.y);
    }
    for (int i = 0; i < nm; ++i) mw[i].z = (mw[i].x + mw[i].y) / 2;
    NM[t] = nm;
    return;
  }
  NM[t] = 0;
  if (nv == 0) {
    nwin[t] = 0;
    return;
  }
  // intervals (left, right in s as left_idx dt, right_idx dt; weight)
  double* iv = IV + 3 * (size_t)t * stride;
  int* id = PJ + (size_t)t * stride;  // original candidate of each interval, then p(j)
  for (int k = 0; k < nv; ++k) {
    const int4 w = sv[ix[k]];
    iv[3 * k] = (double)w.x * c.dt;
    iv[3 * k + 1] = (double)w.y * c.dt;
    iv[3 * k + 2] = cr[6 * ix[k] + 3];
  }
  for (int k = 0; k < nv; ++k) id[k] = ix[k];
  auto swp = [&](int a, int b) {
    for (int q = 0; q < 3; ++q) { const double x = iv[3 * a + q]; iv[3 * a + q] = iv[3 * b + q]; iv[3 * b + q] = x; }
    const int y = id[a]; id[a] = id[b]; id[b] = y;
  };
  // quicksort (interval_scheduling.f90:130-163) with an explicit stack of ranges
  const int sstride = 2 * stride + 2;
  int2* stk = STK + (size_t)t * sstride;
  int sp = 0;
  stk[sp++] = make_int2(0, nv - 1);
  while (sp > 0) {
    const int2 r = stk[--sp];
    const int first = r.x, last = r.y;
    if (first >= last) continue;
    const int pi = (first + last) / 2;
    const double pl = iv[3 * pi], pr = iv[3 * pi + 1];
    swp(last, pi);
    int i = first - 1;
    for (int j = first; j <= last - 1; ++j) {
      const double l = iv[3 * j], rr = iv[3 * j + 1];
      if (rr < pr - 1.0e-6 || (fabs(rr - pr) <= 1.0e-6 && l <= pl)) {
        ++i;
        swp(i, j);
      }
    }
    swp(i + 1, last);
    if (sp + 2 > sstride) break;  // cannot happen: pending ranges are disjoint, plus one empty one
    stk[sp++] = make_int2(i + 2, last);
    stk[sp++] = make_int2(first, i);
  }
  // p(j) (1-based) by bisection; orig index kept in ix
  for (int k = 0; k < nv; ++k) ix[k] = id[k];
  double* opt = OPT + (size_t)t * (stride + 1);
  opt[0] = 0.0;
  for (int j = 1; j <= nv; ++j) {
    int lo = 1, hi = j - 1, p = 0;
    const double v = iv[3 * (j - 1)];
    while (lo <= hi) {
      const int mid = (lo + hi) / 2;
      if (iv[3 * (mid - 1) + 1] <= v) { p = mid; lo = mid + 1; }
      else hi = mid - 1;
    }
    id[j - 1] = p;
    const double a = iv[3 * (j - 1) + 2] + opt[p], b = opt[j - 1];
    opt[j] = a > b ? a : b;
  }
  // backtrack from the last interval, then reverse into time order
  int ns = 0;
  int* selr = reinterpret_cast<int*>(stk);  // the stack is free again
  for (int j = nv; j > 0;) {
    if (iv[3 * (j - 1) + 2] + opt[id[j - 1]] > opt[j - 1]) { selr[ns++] = ix[j - 1]; j = id[j - 1]; }
    else j = j - 1;
  }
  nwin[t] = ns;
  for (int k = 0; k < ns && k < maxwin; ++k) {
    const int e = selr[ns - 1 - k];
    const int4 w = sv[e];
    double* wo = win + 2 * ((size_t)t * maxwin + k);
    wo[0] = (double)(w.x - 1) * c.dt - t0;
    wo[1] = (double)(w.y - 1) * c.dt - t0;
    double* io = info + 4 * ((size_t)t * maxwin + k);
    io[0] = cr[6 * e + 3];
    io[1] = cr[6 * e];
    io[2] = cr[6 * e + 2];
    io[3] = cr[6 * e + 1];
  }
}

// outputs of the merging mode: merged windows with their recomputed criteria
__global__ void k_fw_out(int ntrace, int stride, FwCfg c, double t0, const int4* __restrict__ MW,
                         const int* __restrict__ NM, const double* __restrict__ CRM, int maxwin,
                         int* __restrict__ nwin, double* __restrict__ win, double* __restrict__ info) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntrace || c.split) return;
  const int nm = NM[t];
  nwin[t] = nm;
  for (int k = 0; k < nm && k < maxwin; ++k) {
    const int4 w = MW[(size_t)t * stride + k];
    const double* o = CRM + 6 * ((size_t)t * stride + k);
    double* wo = win + 2 * ((size_t)t * maxwin + k);
    wo[0] = (double)(w.x - 1) * c.dt - t0;
    wo[1] = (double)(w.y - 1) * c.dt - t0;
    double* io = info + 4 * ((size_t)t * maxwin + k);
    io[0] = o[3];
    io[1] = o[0];
    io[2] = o[2];
    io[3] = o[1];
  }
}

}  // namespace

extern "C" {

// FLEXWIN window selection (flexwin_type, src/preproc/window_selector/flexwin.f90)
// of every trace: obs, syn [nstep][ntrace] (filtered as leq_data.f90:93-105 does),
// tp, dist (ntrace) first-arrival time and distance, t0, min/max period of the
// band; cfg(16) = min_velocity, min_win_len_fac, threshold_corr,
// threshold_shift_fac, min_peaks_troughs, threshold_dlna, min_snr_window,
// is_split_phases (0/1), max_time_before_first_arrival, stalta_water_level, c_0,
// c_2, c_3a, c_3b, c_4a, c_4b (input_params.f90:400-418).  nwin(ntrace),
// windows(2, maxwin, ntrace) in s relative to t0 (the reference's twin),
// info(4, maxwin, ntrace) = weight, cc, time shift, dlnA.  A trace keeps at most
// 16384 candidates through the per-candidate passes (refused beyond).
void fwat_flexwin_batch_(int* nstep, int* ntrace, double* dt, realw* obs, realw* syn, double* t0, double* tp,
                         double* dist, double* min_period, double* max_period, double* cfg, int* maxwin, int* nwin,
                         double* windows, double* info) {
  const int N = *nstep, T = *ntrace;
  if (N < 3 || T <= 0) SEM_FAIL("fwat_flexwin_batch_: nstep >= 3 and ntrace > 0 needed");
  if (N > 16383) SEM_FAIL("fwat_flexwin_batch_: traces of at most 16383 samples (envelope transform in LDS)");
  if (T > 65535) SEM_FAIL("fwat_flexwin_batch_: at most 65535 traces per call");
  if (*maxwin < 1) SEM_FAIL("fwat_flexwin_batch_: maxwin must be positive");
  if (!(*min_period > 0.0) || !(*dt > 0.0)) SEM_FAIL("fwat_flexwin_batch_: dt and min_period must be positive");
  if (!(cfg[1] >= 0.0) || !(cfg[0] > 0.0)) SEM_FAIL("fwat_flexwin_batch_: min_win_len_fac >= 0, min_velocity > 0");
  const double mp = *min_period, Mp = *max_period, h = *dt;
  FwCfg C;
  C.min_len = cfg[1] * mp;
  C.thr_corr = cfg[2];
  C.thr_shift = cfg[3] * mp;
  C.min_pt = (int)cfg[4];
  C.thr_dlna = cfg[5];
  C.min_snr = cfg[6];
  C.split = cfg[7] != 0.0 ? 1 : 0;
  C.wl = cfg[9];
  C.wmid = cfg[10] * cfg[9];
  C.c2 = cfg[11];
  C.c3a = cfg[12];
  C.c3b = cfg[13];
  C.tdl = mp * cfg[14] / h;
  C.tdr = mp * cfg[15] / h;
  C.min_period = mp;
  C.dt = h;
  // per-trace time range (initialize, flexwin.f90:86-106; reject_on_traveltimes :402-403)
  std::vector<FwTrace> trc(T);
  for (int t = 0; t < T; ++t) {
    FwTrace& r = trc[t];
    r.tstart = tp[t] - cfg[8] + *t0;
    r.tend = dist[t] / cfg[0] + Mp + *t0;
    r.nstart = (int)(r.tstart / h) + 1;
    r.nend = std::min(N, (int)(r.tend / h) + 1);
    r.min_time = r.tstart + cfg[8] - mp;
    r.max_time = r.tend - Mp;
    if (r.nstart > 0) {
      r.nnoise = std::min(r.nstart, N);
    } else {
      const int ntp = (int)((tp[t] + *t0) / h) + 1;
      r.nnoise = ntp > 1 ? std::min(ntp, N) : 0;
    }
    r.pad = 0;
  }
  const double Cs = std::pow(10.0, -h / mp), Cl = std::pow(10.0, -h / (12.0 * mp));
  int e2 = 0;
  std::frexp((float)N, &e2);
  const int nfe = 1 << e2, Me = nfe / 2;
  int logMe = 0;
  while ((1 << logMe) < Me) ++logMe;
  std::vector<double> twh(2 * (size_t)Me);
  for (int j = 0; j < Me; ++j) {
    const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)j / (long double)nfe;
    twh[2 * j] = (double)std::cos(a);
    twh[2 * j + 1] = (double)std::sin(a);
  }
  const size_t n = (size_t)N * T;
  const int cap = N / 2 + 1;
  hipStream_t s;
  HIP_CHECK(hipStreamCreate(&s));
  hipEvent_t ev0, ev1;
  HIP_CHECK(hipEventCreate(&ev0));
  HIP_CHECK(hipEventCreate(&ev1));
  float* d_obs = dev_upload(obs, n, s);
  float* d_syn = dev_upload(syn, n, s);
  FwTrace* d_trc = reinterpret_cast<FwTrace*>(dev_upload(reinterpret_cast<double*>(trc.data()),
                                                         trc.size() * sizeof(FwTrace) / sizeof(double), s));
  double2* d_twe = reinterpret_cast<double2*>(dev_upload(twh.data(), twh.size(), s));
  float* d_obsT = dev_alloc<float>(n);
  float* d_synT = dev_alloc<float>(n);
  double* d_Dn = dev_alloc<double>(n);
  double* d_Sn = dev_alloc<double>(n);
  double* d_env = dev_alloc<double>(n);
  double* d_sl = dev_alloc<double>(n);
  double* d_noise = dev_alloc<double>(T);
  double2* d_g01 = dev_alloc<double2>(2 * (size_t)T * (Me + 1));
  int* d_pk = dev_alloc<int>((size_t)T * cap);
  int* d_tr = dev_alloc<int>((size_t)T * cap);
  int* d_npt = dev_alloc<int>(2 * (size_t)T);
  long long* d_pre = dev_alloc<long long>((size_t)T * (cap + 1));
  int4* d_sv = dev_alloc<int4>((size_t)T * FW_CAP);
  int* d_nsv = dev_alloc<int>(T);
  int* d_nwin = dev_alloc<int>(T);
  int* d_nm = dev_alloc<int>(T);
  double* d_win = dev_alloc<double>(2 * (size_t)*maxwin * T);
  double* d_info = dev_alloc<double>(4 * (size_t)*maxwin * T);
  HIP_CHECK(hipEventRecord(ev0, s));
  const dim3 ttiles((unsigned)((T + 63) / 64), (unsigned)((N + 63) / 64));
  hipLaunchKernelGGL(k_rows_f32, ttiles, dim3(256), 0, s, N, T, d_obs, d_obsT);
  hipLaunchKernelGGL(k_rows_f32, ttiles, dim3(256), 0, s, N, T, d_syn, d_synT);
  hipLaunchKernelGGL(k_fw_prep, dim3((unsigned)T), dim3(FW_THREADS), 0, s, N, d_trc, d_obsT, d_synT, d_Dn, d_Sn,
                     d_noise);
  hipLaunchKernelGGL(k_ws_env, dim3((unsigned)T), dim3(MT_THREADS), 0, s, N, nfe, logMe, d_twe, d_Sn, d_g01,
                     d_g01 + (size_t)T * (Me + 1), d_env);
  hipLaunchKernelGGL(k_fw_stalta, dim3((unsigned)T), dim3(FW_THREADS), 0, s, N, Cs, Cl, d_env, d_sl, d_pk, d_tr,
                     d_npt);
  hipLaunchKernelGGL(k_fw_cand, dim3((unsigned)T), dim3(FW_THREADS), 0, s, N, C, d_trc, d_sl, d_pk, d_tr, d_npt,
                     d_pre, d_sv, d_nsv);
  HIP_CHECK(hipGetLastError());
  std::vector<int> nsv(T);
  HIP_CHECK(hipMemcpyAsync(nsv.data(), d_nsv, T * sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  int maxs = 1;
  for (int t = 0; t < T; ++t) {
    if (nsv[t] == -2) SEM_FAIL("fwat_flexwin_batch_: trace %d keeps more than %d candidates", t, FW_CAP);
    maxs = std::max(maxs, nsv[t]);
  }
  // criteria of the survivors; then rejections and merging or scheduling
  double* d_cr = dev_alloc<double>(6 * (size_t)T * maxs);
  int* d_ix = dev_alloc<int>((size_t)T * maxs);
  double* d_iv = dev_alloc<double>(3 * (size_t)T * maxs);
  int* d_pj = dev_alloc<int>((size_t)T * maxs);
  double* d_opt = dev_alloc<double>((size_t)T * (maxs + 1));
  int2* d_stk = dev_alloc<int2>((size_t)T * (2 * (size_t)maxs + 2));
  int4* d_mw = dev_alloc<int4>((size_t)T * maxs);
  double* d_crm = dev_alloc<double>(6 * (size_t)T * maxs);
  // the survivors' rows were written with stride FW_CAP; the criteria rows use maxs
  int4* d_svc = dev_alloc<int4>((size_t)T * maxs);
  HIP_CHECK(hipMemcpy2DAsync(d_svc, (size_t)maxs * sizeof(int4), d_sv, (size_t)FW_CAP * sizeof(int4),
                             (size_t)maxs * sizeof(int4), (size_t)T, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_fw_crit, dim3((unsigned)maxs, (unsigned)T), dim3(FW_THREADS), 0, s, N, maxs, C, d_Dn, d_Sn,
                     d_noise, d_svc, d_nsv, d_cr);
  const unsigned nb = (unsigned)((T + 63) / 64);
  hipLaunchKernelGGL(k_fw_finish, dim3(nb), dim3(64), 0, s, T, maxs, C, *t0, d_svc, d_nsv, d_cr, d_ix, d_iv, d_pj,
                     d_opt, d_stk, d_mw, d_nm, *maxwin, d_nwin, d_win, d_info);
  if (!C.split) {
    hipLaunchKernelGGL(k_fw_crit, dim3((unsigned)maxs, (unsigned)T), dim3(FW_THREADS), 0, s, N, maxs, C, d_Dn, d_Sn,
                       d_noise, d_mw, d_nm, d_crm);
    hipLaunchKernelGGL(k_fw_out, dim3(nb), dim3(64), 0, s, T, maxs, C, *t0, d_mw, d_nm, d_crm, *maxwin, d_nwin, d_win,
                       d_info);
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipEventRecord(ev1, s));
  HIP_CHECK(hipMemcpyAsync(nwin, d_nwin, T * sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(windows, d_win, 2 * (size_t)*maxwin * T * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(info, d_info, 4 * (size_t)*maxwin * T * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  float ms = 0.0f;
  HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
  g_meas_ms = ms;
  dev_free(d_obs); dev_free(d_syn); dev_free(d_trc); dev_free(d_twe); dev_free(d_obsT); dev_free(d_synT);
  dev_free(d_Dn); dev_free(d_Sn); dev_free(d_env); dev_free(d_sl); dev_free(d_noise); dev_free(d_g01);
  dev_free(d_pk); dev_free(d_tr); dev_free(d_npt); dev_free(d_pre); dev_free(d_sv); dev_free(d_nsv);
  dev_free(d_nwin); dev_free(d_nm); dev_free(d_win); dev_free(d_info); dev_free(d_cr); dev_free(d_ix);
  dev_free(d_iv); dev_free(d_pj); dev_free(d_opt); dev_free(d_stk); dev_free(d_mw); dev_free(d_crm);
  dev_free(d_svc);
  HIP_CHECK(hipEventDestroy(ev0));
  HIP_CHECK(hipEventDestroy(ev1));
  HIP_CHECK(hipStreamDestroy(s));
}

}  // extern "C"
