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

// zero-phase cascade of second-order sections over each trace, in place
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
      for (int k = 0; k < nstep; ++k) {
        const size_t a = (size_t)(pass == 0 ? k : nstep - 1 - k) * T + t;
        const float xi = x[a];
        float out = b0 * xi + b1 * x1 + b2 * x2;
        out = out - (a1 * y1 + a2 * y2);
        y2 = y1;
        y1 = out;
        x2 = x1;
        x1 = xi;
        x[a] = out;
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

}  // extern "C"
