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
#include <vector>

#include "sem_device.h"

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

// xcorr_shift, first stage: workgroup (trace, block of 256 lags) of the full
// correlation; d and the matching stretch of s come through LDS in chunks of
// 1024 samples taken in increasing i1, so each lag's sum runs over i1 in
// increasing order as in myconvolution_dp.  Writes the block's first maximum.
constexpr int CC_LAGS = 256, CC_CHUNK = 1024;
__global__ void __launch_bounds__(CC_LAGS) k_cc_corr(int nl, const double* __restrict__ D,
                                                     const double* __restrict__ S, double* __restrict__ bval,
                                                     int* __restrict__ bidx) {
  __shared__ double sd[CC_CHUNK], ss[CC_CHUNK + CC_LAGS];
  const int t = blockIdx.y, nblk = gridDim.x;
  const double* d = D + (size_t)t * nl;
  const double* s = S + (size_t)t * nl;
  const int m0 = blockIdx.x * CC_LAGS, m = m0 + (int)threadIdx.x;
  const int nlag = 2 * nl - 1;
  // lag m reads s[i1 + nl - 1 - m] for i1 in [max(0, m - nl + 1), min(nl - 1, m)]
  const int lo = m - (nl - 1) > 0 ? m - (nl - 1) : 0, hi = m < nl - 1 ? m : nl - 1;
  const int blo = m0 - (nl - 1) > 0 ? m0 - (nl - 1) : 0;
  const int mlast = min(m0 + CC_LAGS - 1, nlag - 1);
  const int bhi = mlast < nl - 1 ? mlast : nl - 1;
  double c = 0.0;
  for (int i0 = blo; i0 <= bhi; i0 += CC_CHUNK) {
    const int i1end = min(i0 + CC_CHUNK - 1, bhi);
    // s indices of this chunk over the block's lags: [i0 + nl - 1 - (m0 + 255), i1end + nl - 1 - m0]
    const int sbase = i0 + nl - 1 - (m0 + CC_LAGS - 1);
    __syncthreads();
    for (int k = threadIdx.x; k < CC_CHUNK; k += CC_LAGS) sd[k] = (i0 + k <= i1end) ? d[i0 + k] : 0.0;
    for (int k = threadIdx.x; k < CC_CHUNK + CC_LAGS; k += CC_LAGS) {
      const int j = sbase + k;
      ss[k] = (j >= 0 && j < nl) ? s[j] : 0.0;
    }
    __syncthreads();
    if (m < nlag) {
      const int a = max(lo, i0), b = min(hi, i1end);
      // s[i1 + nl - 1 - m] sits at ss[i1 - i0 + (m0 + 255 - m)]
      const int off = m0 + CC_LAGS - 1 - m - i0;
      for (int i1 = a; i1 <= b; ++i1) c = c + sd[i1 - i0] * ss[i1 + off];
    }
  }
  __shared__ double rv[CC_LAGS];
  __shared__ int ri[CC_LAGS];
  rv[threadIdx.x] = m < nlag ? c : -INFINITY;
  ri[threadIdx.x] = m < nlag ? m : 0x7fffffff;
  __syncthreads();
  for (int o = CC_LAGS / 2; o > 0; o >>= 1) {
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
                                                   double* __restrict__ A, double* __restrict__ out) {
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
    double* o = out + 8 * (size_t)t;
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

__global__ void k_fill_f32m(size_t n, float* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0f;
}
#pragma clang fp contract(fast)

double g_meas_ms = 0.0;

// window_taper's weights (signal.f90:419-491) for a window of npts samples
void taper_weights(int npts, double pct, int type, std::vector<double>& w) {
  w.assign(npts, 1.0);
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
    hipLaunchKernelGGL(k_sos_filter, dim3(nb), dim3(MTPB), 0, s, N, T, *nsect, d_sos, d_adj);
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

// cc_tt_misfit.f90 CCTTMisfit%calc_adjoint_source for every trace: imeas 11
// (traveltime) or 12 (amplitude); cfg(7) = dt_sigma_min, dlna_sigma_min,
// TSHIFT_MIN, TSHIFT_MAX, DLNA_MIN, DLNA_MAX, CC_MIN; meas(8, nwin, ntrace) =
// misfit, residual, error, cc_max, tshift, dlna, selected, ishift; misfits
// (ntrace) summed over windows; sos / nsect as for the waveform measurement
void fwat_cc_adjoint_batch_(int* nstep, int* ntrace, double* dt, realw* obs, realw* syn, int* nwin, double* windows,
                            double* taper_percentage, int* itaper_type, int* imeas, double* cfg, int* nsect,
                            realw* sos, double* misfits, double* meas, realw* adj) {
  const int N = *nstep, T = *ntrace, nw = *nwin;
  if (N <= 0 || T <= 0) SEM_FAIL("fwat_cc_adjoint_batch_: sizes must be positive");
  if (nw < 1 || nw > MAX_WIN) SEM_FAIL("fwat_cc_adjoint_batch_: 1..%d windows", MAX_WIN);
  if (*imeas != 11 && *imeas != 12) SEM_FAIL("fwat_cc_adjoint_batch_: imeas must be 11 (CC-TT) or 12 (CC-DLNA)");
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
    const int nblk = (2 * nls[w] - 1 + CC_LAGS - 1) / CC_LAGS;
    hipLaunchKernelGGL(k_cc_corr, dim3((unsigned)nblk, (unsigned)T), dim3(CC_LAGS), 0, s, nls[w], d_D, d_S, d_bv,
                       d_bi);
    hipLaunchKernelGGL(k_cc_pick, dim3(nbk), dim3(MTPB), 0, s, T, nblk, nls[w], d_bv, d_bi, d_ish);
    hipLaunchKernelGGL(k_cc_finish, dim3((unsigned)T), dim3(256), 0, s, nls[w], *dt, *imeas, C, d_tw, d_D, d_S,
                       d_ish, d_A, d_out + 8 * (size_t)w * T);
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
    hipLaunchKernelGGL(k_sos_filter, dim3(nbk), dim3(MTPB), 0, s, N, T, *nsect, d_sos, d_adj);
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

}  // extern "C"
