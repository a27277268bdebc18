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
//   orc_cc_adjoint        CCTTMisfit%calc_adjoint_source, src/preproc/adjoint_sources/
//                         cc_tt_misfit.f90:17-85 with cross_correlate (utils/cc.f90:
//                         10-160: calc_cc_shift, calc_cc_error, cc_correction,
//                         xcorr_shift over mycorrelation_dp / myconvolution_dp,
//                         signal.f90:146-220, cc_max_coef, calc_cc_adjsrc), gradient
//                         (utils.f90:1613-1628) and cc_measure_select (cc_tt_misfit.f90:
//                         126-146); double precision, loops in the reference's order
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

extern "C" void orc_window_taper(double* sig, int npts, double pct, int itype);
extern "C" double orc_simpson(const double* y, int n, double dx);

namespace {

// gradient (utils.f90:1613-1628): central differences, one-sided at the ends
std::vector<double> gradient(const std::vector<double>& y, double dx) {
  const int n = (int)y.size();
  std::vector<double> g(n, 0.0);
  for (int i = 1; i < n - 1; ++i) g[i] = (y[i + 1] - y[i - 1]) / (2.0 * dx);
  if (n >= 2) { g[0] = (y[1] - y[0]) / dx; g[n - 1] = (y[n - 1] - y[n - 2]) / dx; }
  return g;
}

// xcorr_shift: full correlation of d with s, first maximum, lag = index - n
int xcorr_shift(const std::vector<double>& d, const std::vector<double>& s) {
  const int n = (int)d.size();
  std::vector<double> cc(2 * n - 1, 0.0);
  // myconvolution_dp(d, flip(s)): conv(i1 + i2 - 1) += d(i1) * s(n - i2 + 1), i1 outer
  for (int i1 = 0; i1 < n; ++i1)
    for (int i2 = 0; i2 < n; ++i2) cc[i1 + i2] = cc[i1 + i2] + d[i1] * s[n - 1 - i2];
  int best = 0;
  for (int m = 1; m < 2 * n - 1; ++m)
    if (cc[m] > cc[best]) best = m;
  return best + 1 - n;
}

}  // namespace

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

// Cross-correlation traveltime (imeas 11) or amplitude (12) measurement of one
// trace: cfg = {dt_sigma_min, dlna_sigma_min, tshift_min, tshift_max, dlna_min,
// dlna_max, cc_min}; per window out: misfit, residual (tshift or dlna), error,
// cc_max, tshift, dlna, selected.  Returns the total misfit.
double orc_cc_adjoint(const double* dat, const double* syn, int n, double dt, const double* windows, int nwin,
                      double taper_pct, int itaper, int imeas, const double* cfg, double* adj, double* misfits,
                      double* residuals, double* errors, double* ccmax, double* tshift_o, double* dlna_o,
                      int* selected) {
  double total = 0.0;
  for (int i = 0; i < n; ++i) adj[i] = 0.0;
  for (int w = 0; w < nwin; ++w) {
    const int nb = (int)(windows[2 * w] / dt) + 1, ne = (int)(windows[2 * w + 1] / dt);
    const int nl = ne - nb + 1;
    std::vector<double> sv(syn + nb - 1, syn + ne), dv(dat + nb - 1, dat + ne);
    orc_window_taper(sv.data(), nl, taper_pct, itaper);
    orc_window_taper(dv.data(), nl, taper_pct, itaper);
    // calc_cc_shift
    const int ishift = xcorr_shift(dv, sv);
    const double tshift = ishift * dt;
    double sd2 = 0.0, ss2 = 0.0;
    for (int i = 0; i < nl; ++i) sd2 = sd2 + dv[i] * dv[i];
    for (int i = 0; i < nl; ++i) ss2 = ss2 + sv[i] * sv[i];
    const double dlna = 0.5 * std::log(sd2 / ss2);
    // calc_cc_error with cc_correction
    std::vector<double> scc(nl, 0.0), sccd(nl, 0.0);
    for (int i = 0; i < nl; ++i) {
      const int j = i - ishift;
      if (j >= 0 && j < nl) { scc[i] = sv[j]; sccd[i] = std::exp(dlna) * sv[j]; }
    }
    const std::vector<double> vel = gradient(sccd, dt);
    double num = 0.0, den = 0.0, den2 = 0.0;
    for (int i = 0; i < nl; ++i) num = num + (dv[i] - sccd[i]) * (dv[i] - sccd[i]);
    for (int i = 0; i < nl; ++i) den = den + vel[i] * vel[i];
    for (int i = 0; i < nl; ++i) den2 = den2 + scc[i] * scc[i];
    double sig_dt = std::sqrt(num / den), sig_dlna = std::sqrt(num / den2);
    if (sig_dt < cfg[0] || std::isnan(sig_dt)) sig_dt = cfg[0];
    if (sig_dlna < cfg[1] || std::isnan(sig_dlna)) sig_dlna = cfg[1];
    // cc_max_coef (same shift; d's edges zeroed where the shifted s is empty)
    std::vector<double> dsh(dv);
    if (ishift > 0) for (int i = 0; i < ishift && i < nl; ++i) dsh[i] = 0.0;
    else if (ishift < 0) for (int i = nl + ishift; i < nl; ++i) if (i >= 0) dsh[i] = 0.0;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = 0; i < nl; ++i) a = a + dsh[i] * scc[i];
    for (int i = 0; i < nl; ++i) b = b + dsh[i] * dsh[i];
    for (int i = 0; i < nl; ++i) c = c + scc[i] * scc[i];
    const double cmax = a / std::sqrt(b * c);
    ccmax[w] = cmax;
    tshift_o[w] = tshift;
    dlna_o[w] = dlna;
    const bool sel = !(cmax < cfg[6] || tshift < cfg[2] || tshift > cfg[3] || dlna < cfg[4] || dlna > cfg[5]);
    selected[w] = sel;
    residuals[w] = imeas == 11 ? tshift : dlna;
    errors[w] = imeas == 11 ? sig_dt : sig_dlna;
    misfits[w] = 0.0;
    if (!sel) continue;
    // calc_cc_adjsrc
    const double mp = 0.5 * (tshift / sig_dt) * (tshift / sig_dt);
    const double mq = 0.5 * (dlna / sig_dlna) * (dlna / sig_dlna);
    std::vector<double> out(nl);
    if (imeas == 11) {
      const std::vector<double> dsdt = gradient(sv, dt);
      std::vector<double> sq(nl);
      for (int i = 0; i < nl; ++i) sq[i] = dsdt[i] * dsdt[i];
      const double nn = orc_simpson(sq.data(), nl, dt);
      for (int i = 0; i < nl; ++i) out[i] = dsdt[i] * tshift / nn / (sig_dt * sig_dt);
    } else {
      std::vector<double> sq(nl);
      for (int i = 0; i < nl; ++i) sq[i] = sv[i] * sv[i];
      const double nn = orc_simpson(sq.data(), nl, dt);
      for (int i = 0; i < nl; ++i) out[i] = -1.0 * sv[i] * dlna / nn / (sig_dlna * sig_dlna);
    }
    orc_window_taper(out.data(), nl, taper_pct, itaper);
    const double mf = imeas == 11 ? mp : mq;
    misfits[w] = mf;
    total = total + mf;
    for (int i = 0; i < nl; ++i) adj[nb - 1 + i] = out[i];
  }
  return total;
}

}  // extern "C"
