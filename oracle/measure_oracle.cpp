// CPU restatement of the adjoint-source filter of SpecFWAT's measurements.
//
// TEST INFRASTRUCTURE ONLY: compiled into liboracle_sem.so and used by tests/
// and bench.py as the checker of the product's measurement code; the product
// never links it.
//
//   orc_bandpass_design   the filter bandpass_dp asks of SAC's xapiir
//                         (src/shared/signal.f90:7-36; xapiir_sub.f90 DESIGN with
//                         APROTO 'BU', TYPE 'BP', IORD = 2, config.f90:74-77):
//                         Butterworth lowpass prototype, lowpass-to-bandpass
//                         mapping s -> (s^2 + w0^2) / (B s) on tangent-prewarped
//                         band edges w = tan(pi f dt), bilinear mapping
//                         s = (1 - 1/z) / (1 + 1/z), one second-order section per
//                         bandpass pole pair.  Restated from the mathematics
//                         (coefficients in double, stored single); pinned by the
//                         frequency response of an independent implementation.
//   orc_bandpass_apply    xapiir_sub.f90 APPLY: direct form I per section over the
//                         whole trace, then (zero phase, PASSES = 2) the same
//                         sections over the reversed trace; single precision
//   orc_filter_adj        filter_adj, src/preproc/adjoint_sources/adjoint_source.f90:
//                         92-114: bandpass 1/max_period .. 1/min_period, then taper
//                         over [min window start, max window end], zero outside
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

extern "C" void orc_window_taper(double* sig, int npts, double pct, int itype);

extern "C" {

// sn, sd: [nsect][3] in double; returns nsect (2 for the 2-pole prototype)
int orc_bandpass_design_d(double dt, double f1, double f2, double* sn, double* sd) {
  using cd = std::complex<double>;
  const double pi = 3.14159265358979323846;
  const double wl = std::tan(pi * f1 * dt), wh = std::tan(pi * f2 * dt);
  const double w0sq = wl * wh, B = wh - wl;
  const cd p(-std::sqrt(0.5), std::sqrt(0.5));  // upper pole of the 2-pole Butterworth prototype
  const cd disc = std::sqrt(p * p * B * B - 4.0 * w0sq);
  const cd q[2] = {(p * B + disc) / 2.0, (p * B - disc) / 2.0};
  for (int k = 0; k < 2; ++k) {
    const double a1 = -2.0 * q[k].real(), a0 = std::norm(q[k]);
    const double D = 1.0 + a1 + a0;
    sn[3 * k] = B / D;
    sn[3 * k + 1] = 0.0;
    sn[3 * k + 2] = -B / D;
    sd[3 * k] = 1.0;
    sd[3 * k + 1] = (2.0 * a0 - 2.0) / D;
    sd[3 * k + 2] = (1.0 - a1 + a0) / D;
  }
  return 2;
}

// the same stored in single precision, as the filter runs
int orc_bandpass_design(double dt, float f1, float f2, float* sn, float* sd) {
  double a[6], b[6];
  const int ns = orc_bandpass_design_d(dt, (double)f1, (double)f2, a, b);
  for (int i = 0; i < 3 * ns; ++i) { sn[i] = (float)a[i]; sd[i] = (float)b[i]; }
  return ns;
}

void orc_bandpass_apply(float* x, long n, const float* sn, const float* sd, int nsect, int zero_phase) {
  for (int pass = 0; pass < (zero_phase ? 2 : 1); ++pass)
    for (int j = 0; j < nsect; ++j) {
      const float b0 = sn[3 * j], b1 = sn[3 * j + 1], b2 = sn[3 * j + 2];
      const float a1 = sd[3 * j + 1], a2 = sd[3 * j + 2];
      float x1 = 0.0f, x2 = 0.0f, y1 = 0.0f, y2 = 0.0f;
      for (long k = 0; k < n; ++k) {
        const long i = pass == 0 ? k : n - 1 - k;
        float out = b0 * x[i] + b1 * x1 + b2 * x2;
        out = out - (a1 * y1 + a2 * y2);
        y2 = y1;
        y1 = out;
        x2 = x1;
        x1 = x[i];
        x[i] = out;
      }
    }
}

void orc_filter_adj(double* adj, int n, double dt, float f1, float f2, const double* windows, int nwin,
                    double taper_pct, int itaper) {
  float sn[6], sd[6];
  const int ns = orc_bandpass_design(dt, f1, f2, sn, sd);
  std::vector<float> xs(n);
  for (int i = 0; i < n; ++i) xs[i] = (float)adj[i];
  orc_bandpass_apply(xs.data(), n, sn, sd, ns, 1);
  for (int i = 0; i < n; ++i) adj[i] = (double)xs[i];
  double w0 = windows[0], w1 = windows[0];
  for (int i = 0; i < 2 * nwin; ++i) { w0 = std::min(w0, windows[i]); w1 = std::max(w1, windows[i]); }
  // get_window_info; a window reaching past the trace is cut to it
  const int nb = std::max((int)(w0 / dt) + 1, 1), ne = std::min((int)(w1 / dt), n);
  const int nl = ne - nb + 1;
  std::vector<double> tw(adj + nb - 1, adj + ne);
  orc_window_taper(tw.data(), nl, taper_pct, itaper);
  for (int i = 0; i < n; ++i) adj[i] = 0.0;
  for (int i = 0; i < nl; ++i) adj[nb - 1 + i] = tw[i];
}

}  // extern "C"
