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
//                         imeas 5: CCCMisfit%calc_adjoint_source (ccc_misfit.f90:15-71)
//   orc_hilbert           hilbert (src/shared/fftpack.f90:135-201)
//   orc_ep_adjoint        ExponentiatedPhaseMisfit%calc_adjoint_source
//                         (exponentiated_phase_misfit.f90:21-104)
//   orc_win_sel           win_sel_type%select_windows (window_selector/win_sel.f90:56-469,
//                         sta_lta.f90:6-104)
//   orc_flexwin           flexwin_type%select_windows (window_selector/flexwin.f90:75-734)
//                         with schedule_weighted_intervals (interval_scheduling.f90:13-165)
//   orc_conv_adjoint      WaveformConvMisfit%calc_adjoint_source (waveform_conv_misfit.f90:17-64)
//                         with myconvolution_dp part 0 (signal.f90:146-189)
//   orc_fft               fft_raw_dp / fft_raw (src/shared/fftpack.f90:70-100, 202-241)
//   orc_dpss              dpss_windows (utils/dpss.f90:137-262); eigenvalues by bisection
//                         where the reference calls LAPACK dsterf
//   orc_mt_adjoint        MTTTMisfit%calc_adjoint_source (mt_tt_misfit.f90:29-791) with
//                         process_cycle_skipping (signal.f90:510-598) and, with use_mt_error,
//                         the jackknife errors of calculate_mt_error (:302-409)
#include <algorithm>
#include <cmath>
#include <complex>
#include <functional>
#include <limits>
#include <utility>
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


struct CcStats {
  int ishift;
  double tshift, dlna, sig_dt, sig_dlna, cmax;
  bool sel;
};

// calc_cc_shift, calc_cc_error with cc_correction, cc_max_coef and
// cc_measure_select on the tapered windows dv, sv
CcStats cc_stats(const std::vector<double>& dv, const std::vector<double>& sv, double dt, const double* cfg) {
  const int nl = (int)dv.size();
  CcStats c;
  c.ishift = xcorr_shift(dv, sv);
  c.tshift = c.ishift * dt;
  double sd2 = 0.0, ss2 = 0.0;
  for (int i = 0; i < nl; ++i) sd2 = sd2 + dv[i] * dv[i];
  for (int i = 0; i < nl; ++i) ss2 = ss2 + sv[i] * sv[i];
  c.dlna = 0.5 * std::log(sd2 / ss2);
  std::vector<double> scc(nl, 0.0), sccd(nl, 0.0);
  for (int i = 0; i < nl; ++i) {
    const int j = i - c.ishift;
    if (j >= 0 && j < nl) { scc[i] = sv[j]; sccd[i] = std::exp(c.dlna) * sv[j]; }
  }
  const std::vector<double> vel = gradient(sccd, dt);
  double num = 0.0, den = 0.0, den2 = 0.0;
  for (int i = 0; i < nl; ++i) num = num + (dv[i] - sccd[i]) * (dv[i] - sccd[i]);
  for (int i = 0; i < nl; ++i) den = den + vel[i] * vel[i];
  for (int i = 0; i < nl; ++i) den2 = den2 + scc[i] * scc[i];
  c.sig_dt = std::sqrt(num / den);
  c.sig_dlna = std::sqrt(num / den2);
  if (c.sig_dt < cfg[0] || std::isnan(c.sig_dt)) c.sig_dt = cfg[0];
  if (c.sig_dlna < cfg[1] || std::isnan(c.sig_dlna)) c.sig_dlna = cfg[1];
  // cc_max_coef (same shift; d's edges zeroed where the shifted s is empty)
  std::vector<double> dsh(dv);
  if (c.ishift > 0) for (int i = 0; i < c.ishift && i < nl; ++i) dsh[i] = 0.0;
  else if (c.ishift < 0) for (int i = nl + c.ishift; i < nl; ++i) if (i >= 0) dsh[i] = 0.0;
  double a = 0.0, b = 0.0, cc = 0.0;
  for (int i = 0; i < nl; ++i) a = a + dsh[i] * scc[i];
  for (int i = 0; i < nl; ++i) b = b + dsh[i] * dsh[i];
  for (int i = 0; i < nl; ++i) cc = cc + scc[i] * scc[i];
  c.cmax = a / std::sqrt(b * cc);
  c.sel = !(c.cmax < cfg[6] || c.tshift < cfg[2] || c.tshift > cfg[3] || c.dlna < cfg[4] || c.dlna > cfg[5]);
  return c;
}

// calc_cc_adjsrc (utils/cc.f90): both adjoint sources when imeas is 0, else
// the one imeas (11 or 12) asks for; misfits 0.5 (tshift/sigma)^2, 0.5 (dlna/sigma)^2
void cc_adjsrc(const std::vector<double>& sv, const CcStats& c, double dt, int imeas, std::vector<double>& p,
               std::vector<double>& q, double& mp, double& mq) {
  const int nl = (int)sv.size();
  mp = 0.5 * (c.tshift / c.sig_dt) * (c.tshift / c.sig_dt);
  mq = 0.5 * (c.dlna / c.sig_dlna) * (c.dlna / c.sig_dlna);
  p.assign(nl, 0.0);
  q.assign(nl, 0.0);
  if (imeas != 12) {
    const std::vector<double> dsdt = gradient(sv, dt);
    std::vector<double> sq(nl);
    for (int i = 0; i < nl; ++i) sq[i] = dsdt[i] * dsdt[i];
    const double nn = orc_simpson(sq.data(), nl, dt);
    for (int i = 0; i < nl; ++i) p[i] = dsdt[i] * c.tshift / nn / (c.sig_dt * c.sig_dt);
  }
  if (imeas != 11) {
    std::vector<double> sq(nl);
    for (int i = 0; i < nl; ++i) sq[i] = sv[i] * sv[i];
    const double nn = orc_simpson(sq.data(), nl, dt);
    for (int i = 0; i < nl; ++i) q[i] = -1.0 * sv[i] * c.dlna / nn / (c.sig_dlna * c.sig_dlna);
  }
}

// complex over a real-valued complex, as Smith's division gives it (a/c, b/c);
// the spectra sum(s conj(s)) have an exactly zero imaginary part
std::complex<double> cdivr(std::complex<double> a, double c) { return {a.real() / c, a.imag() / c}; }

// smallest power of two above n: 2 ** exponent(real(n)) of the reference
int next_pow2_above(int n) {
  int e = 0;
  std::frexp((float)n, &e);
  return 1 << e;
}

// fft_raw_dp (fftpack.f90:202-241): bit reversal as written, then radix-2
// butterflies with twiddle exp(i pi ind (k - 1) / kmax)
template <class T>
void fft_raw(std::vector<std::complex<T>>& x, int ind, T pi) {
  const int n = (int)x.size();
  int j = 1;
  for (int i = 1; i <= n; ++i) {
    if (i < j) std::swap(x[j - 1], x[i - 1]);
    int m = n / 2;
    while (true) {
      if (j > m) {
        j = j - m;
        m = m / 2;
        if (m < 2) break;
      } else {
        break;
      }
    }
    j = j + m;
  }
  for (int kmax = 1; kmax < n; kmax *= 2) {
    const int istep = kmax * 2;
    for (int k = 1; k <= kmax; ++k) {
      const T th = pi * (T)ind * (T)(k - 1) / (T)kmax;
      const std::complex<T> w(std::cos(th), std::sin(th));
      for (int i = k; i <= n; i += istep) {
        const int jj = i + kmax;
        const std::complex<T> t = x[jj - 1] * w;
        x[jj - 1] = x[i - 1] - t;
        x[i - 1] = x[i - 1] + t;
      }
    }
  }
}


// hilbert (fftpack.f90:135-201): analytic signal of x on nft = 2 ** exponent(real(n))
// points; the Nyquist weight follows the parity of the input length n
std::vector<std::complex<double>> hilbert(const std::vector<double>& x) {
  using cd = std::complex<double>;
  const int n = (int)x.size(), nft = next_pow2_above(n);
  std::vector<cd> xf(nft, cd(0.0, 0.0));
  for (int i = 0; i < n; ++i) xf[i] = cd(x[i], 0.0);
  fft_raw(xf, -1, 3.14159265358979323846);
  std::vector<double> h(nft, 0.0);
  h[0] = 1.0;
  if (n % 2 == 0) {
    h[nft / 2] = 1.0;
    for (int i = 1; i < nft / 2; ++i) h[i] = 2.0;
  } else {
    for (int i = 1; i < (nft + 1) / 2; ++i) h[i] = 2.0;
  }
  for (int i = 0; i < nft; ++i) xf[i] = cd(xf[i].real() * h[i], xf[i].imag() * h[i]);
  fft_raw(xf, 1, 3.14159265358979323846);
  std::vector<cd> xa(n);
  for (int i = 0; i < n; ++i) xa[i] = cd(xf[i].real() / (double)nft, xf[i].imag() / (double)nft);
  return xa;
}


// find_maxima (signal.f90:600-631): number of interior samples with a strict
// rise before and a strict fall after
int count_maxima(const double* x, int n, double sign) {
  if (n < 3) return 0;
  int c = 0;
  for (int i = 1; i < n - 1; ++i)
    if (sign * x[i] - sign * x[i - 1] > 0.0 && sign * x[i + 1] - sign * x[i] < 0.0) ++c;
  return c;
}

// xcorr_shift_coe (win_sel.f90:439-469): lag of the first correlation maximum and
// the correlation coefficient of d with s shifted by it (d's emptied edge zeroed)
void xcorr_shift_coe(const std::vector<double>& d, const std::vector<double>& s, double dt, double& tshift,
                     double& coef) {
  const int n = (int)s.size();
  const int ish = xcorr_shift(d, s);
  std::vector<double> ss(n, 0.0), ds(d);
  for (int i = 0; i < n; ++i) {
    const int j = i - ish;
    if (j >= 0 && j < n) ss[i] = s[j];
  }
  tshift = (double)ish * dt;
  if (ish > 0) for (int i = 0; i < ish && i < n; ++i) ds[i] = 0.0;
  else if (ish < 0) for (int i = n + ish; i < n; ++i) if (i >= 0) ds[i] = 0.0;
  double a = 0.0, b = 0.0, c = 0.0;
  for (int i = 0; i < n; ++i) a = a + ds[i] * ss[i];
  for (int i = 0; i < n; ++i) b = b + ds[i] * ds[i];
  for (int i = 0; i < n; ++i) c = c + ss[i] * ss[i];
  coef = a / std::sqrt(b * c);
}


// STA_LTA (window_selector/sta_lta.f90:6-104, TWEAK off): Bai & Kennett
// recursive averages of the envelope over a 1000-sample pre-extension at the
// noise level max(env) / 1e5
std::vector<double> sta_lta(const std::vector<double>& env, double dt, double min_period) {
  const int n = (int)env.size(), next = 1000;
  const double Cs = std::pow(10.0, -dt / min_period), Cl = std::pow(10.0, -dt / (12.0 * min_period));
  const double tol = 1.0e-9;
  double mx = env[0];
  for (double v : env) mx = std::max(mx, v);
  const double noise = mx / 1e5;
  double sta = 0.0, lta = 0.0;
  for (int i = 0; i < next; ++i) { sta = Cs * sta + noise; lta = Cl * lta + noise; }
  std::vector<double> r(n, 0.0);
  for (int i = 0; i < n; ++i) {
    const double x = env[i] + noise;
    sta = Cs * sta + x;
    lta = Cl * lta + x;
    r[i] = lta > tol ? sta / lta : noise;
  }
  return r;
}

// find_maxima as index lists (1-based, as the reference)
std::vector<int> maxima(const std::vector<double>& x, double sign) {
  std::vector<int> r;
  const int n = (int)x.size();
  for (int i = 1; i < n - 1; ++i)
    if (sign * x[i] - sign * x[i - 1] > 0.0 && sign * x[i + 1] - sign * x[i] < 0.0) r.push_back(i + 1);
  return r;
}


// myconvolution_dp (signal.f90:146-189) with part 0: the n2 samples of the full
// convolution from index n2/2 + 1 (even n2) or n2/2 (odd n2), 1-based, each
// output summed over sig1's index in increasing order
std::vector<double> conv_mid(const std::vector<double>& x, const std::vector<double>& y) {
  const int n1 = (int)x.size(), n2 = (int)y.size();
  std::vector<double> full(n1 + n2 - 1, 0.0);
  for (int i1 = 0; i1 < n1; ++i1)
    for (int i2 = 0; i2 < n2; ++i2) full[i1 + i2] = full[i1 + i2] + x[i1] * y[i2];
  const int ind = (n2 % 2 == 0) ? n2 / 2 + 1 : n2 / 2;
  return std::vector<double>(full.begin() + (ind - 1), full.begin() + (ind - 1) + n2);
}

// number of eigenvalues of the symmetric tridiagonal (d, e) below x (Sturm count)
int sturm_below(const std::vector<double>& d, const std::vector<double>& e, double x) {
  int cnt = 0;
  double q = 1.0;
  for (size_t i = 0; i < d.size(); ++i) {
    const double e2 = i ? e[i - 1] * e[i - 1] : 0.0;
    q = (d[i] - x) - (i ? e2 / q : 0.0);
    if (q == 0.0) q = -1e-300;
    if (q < 0.0) ++cnt;
  }
  return cnt;
}

// tridisolve (dpss.f90:14-45), Golub and Van Loan's symmetric tridiagonal solve
void tridisolve(const std::vector<double>& d, const std::vector<double>& e, std::vector<double>& x) {
  const int n = (int)d.size();
  std::vector<double> dw(d), ew(e);
  for (int k = 1; k < n; ++k) {
    const double t = ew[k - 1];
    ew[k - 1] = t / dw[k - 1];
    dw[k] = dw[k] - t * ew[k - 1];
  }
  for (int k = 1; k < n; ++k) x[k] = x[k] - ew[k - 1] * x[k - 1];
  x[n - 1] = x[n - 1] / dw[n - 1];
  for (int k = n - 2; k >= 0; --k) x[k] = x[k] / dw[k] - ew[k] * x[k + 1];
}

double l2norm(const std::vector<double>& x) {
  double s = 0.0;
  for (double v : x) s = s + v * v;
  return std::sqrt(s);
}

// dpss_windows: the tridiagonal matrix of the discrete prolate spheroidal
// sequences, its kmax largest eigenvalues, inverse iteration
// (tridi_inverse_iteration, dpss.f90:49-107) from sin(k pi i / (n - 1)), signs
void dpss(int n, double half_nbw, int kmax, std::vector<std::vector<double>>& v) {
  const double pi = 3.141592653589793;
  const double wbin = half_nbw / (double)n;
  std::vector<double> dg(n), od(std::max(n - 1, 0));
  for (int i = 0; i < n; ++i) {
    const double q = ((double)(n - 1) - 2.0 * i) / 2.0;
    dg[i] = q * q * std::cos(2.0 * pi * wbin);
  }
  for (int i = 0; i < n - 1; ++i) od[i] = (double)(i + 1) * ((double)n - (double)(i + 1)) / 2.0;
  double lo = dg[0], hi = dg[0];
  for (int i = 0; i < n; ++i) {
    const double r = (i ? std::fabs(od[i - 1]) : 0.0) + (i < n - 1 ? std::fabs(od[i]) : 0.0);
    lo = std::min(lo, dg[i] - r);
    hi = std::max(hi, dg[i] + r);
  }
  v.assign(kmax, std::vector<double>(n, 0.0));
  for (int k = 0; k < kmax; ++k) {
    // eigenvalue number n - 1 - k (0-based ascending) by bisection
    const int target = n - 1 - k;
    double a = lo, b = hi;
    for (int it = 0; it < 200; ++it) {
      const double m = 0.5 * (a + b);
      if (m <= a || m >= b) break;
      if (sturm_below(dg, od, m) <= target) a = m; else b = m;
    }
    const double lam = 0.5 * (a + b);
    std::vector<double> ed(n), x(n), xp(n);
    for (int i = 0; i < n; ++i) ed[i] = dg[i] - lam;
    for (int i = 0; i < n; ++i) x[i] = std::sin((double)(k + 1) * (pi * (double)i / (double)(n - 1)));
    double nx = l2norm(x);
    for (double& t : x) t = t / nx;
    for (int it = 1;; ++it) {
      xp = x;
      tridisolve(ed, od, x);
      nx = l2norm(x);
      for (double& t : x) t = t / nx;
      double dn = 0.0;
      for (int i = 0; i < n; ++i) dn = dn + (std::fabs(x[i]) - std::fabs(xp[i])) * (std::fabs(x[i]) - std::fabs(xp[i]));
      if (std::sqrt(dn) <= 1e-8 || it > 1000) break;
    }
    v[k] = x;
  }
  for (int k = 0; k < kmax; k += 2) {
    double s = 0.0;
    for (double t : v[k]) s = s + t;
    if (s < 0.0) for (double& t : v[k]) t = -t;
  }
  for (int k = 1; k < kmax; k += 2)
    if (n > 1 && v[k][1] < 0.0) for (double& t : v[k]) t = -t;
}

// process_cycle_skipping (signal.f90:510-598): 2 pi jumps of the phase between
// neighbouring frequencies, corrected from there on
void cycle_skipping(std::vector<double>& phi, int nfmin, int nfmax, double step) {
  const double pi = 3.141592653589793;
  for (int iw = nfmin + 2; iw <= nfmax - 1; ++iw) {
    const double p0 = phi[iw - 2], p1 = phi[iw - 1], p2 = phi[iw];
    const double s0 = std::fabs(p2 + p0 - 2.0 * p1);
    const double s1 = std::fabs((p2 + 2.0 * pi) + p0 - 2.0 * p1);
    const double s2 = std::fabs((p2 - 2.0 * pi) + p0 - 2.0 * p1);
    if (std::fabs(p1 - p2) > step) {
      if (s1 < s0 && s1 < s2) for (int i = iw + 1; i <= nfmax; ++i) phi[i - 1] = phi[i - 1] + 2.0 * pi;
      if (s2 < s0 && s2 < s1) for (int i = iw + 1; i <= nfmax; ++i) phi[i - 1] = phi[i - 1] - 2.0 * pi;
    }
  }
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
    if (imeas == 5) {
      // CCCMisfit%calc_adjoint_source (ccc_misfit.f90:15-71): zero-lag correlation coefficient
      double p = 0.0, de = 0.0, se = 0.0;
      for (int i = 0; i < nl; ++i) p = p + dv[i] * sv[i];
      for (int i = 0; i < nl; ++i) de = de + dv[i] * dv[i];
      for (int i = 0; i < nl; ++i) se = se + sv[i] * sv[i];
      p = p * dt;
      de = de * dt;
      se = se * dt;
      const double coef = p / std::sqrt(de * se);
      ccmax[w] = coef;
      tshift_o[w] = 0.0;
      dlna_o[w] = 0.0;
      residuals[w] = coef;
      errors[w] = 0.0;
      misfits[w] = 0.0;
      selected[w] = !(coef < cfg[6]);
      if (!selected[w]) continue;
      std::vector<double> a(nl);
      for (int i = 0; i < nl; ++i) a[i] = -(dv[i] - (p / se) * sv[i]) / std::sqrt(de * se);
      orc_window_taper(a.data(), nl, taper_pct, itaper);
      misfits[w] = 1 - coef;
      total = total + misfits[w];
      for (int i = 0; i < nl; ++i) adj[nb - 1 + i] = a[i];
      continue;
    }
    const CcStats c = cc_stats(dv, sv, dt, cfg);
    ccmax[w] = c.cmax;
    tshift_o[w] = c.tshift;
    dlna_o[w] = c.dlna;
    selected[w] = c.sel;
    residuals[w] = imeas == 11 ? c.tshift : c.dlna;
    errors[w] = imeas == 11 ? c.sig_dt : c.sig_dlna;
    misfits[w] = 0.0;
    if (!c.sel) continue;
    std::vector<double> p, q;
    double mp, mq;
    cc_adjsrc(sv, c, dt, imeas, p, q, mp, mq);
    std::vector<double>& out = imeas == 11 ? p : q;
    orc_window_taper(out.data(), nl, taper_pct, itaper);
    const double mf = imeas == 11 ? mp : mq;
    misfits[w] = mf;
    total = total + mf;
    for (int i = 0; i < nl; ++i) adj[nb - 1 + i] = out[i];
  }
  return total;
}

// fft_raw_dp / fft_raw (fftpack.f90:70-100, 202-241) in double or single
// precision: forward ind = -1, inverse +1 (unnormalised)
void orc_fft(double* re, double* im, int n, int ind, int single) {
  if (single) {
    std::vector<std::complex<float>> x(n);
    for (int i = 0; i < n; ++i) x[i] = {(float)re[i], (float)im[i]};
    fft_raw(x, ind, 3.14159265358979323846f);
    for (int i = 0; i < n; ++i) { re[i] = x[i].real(); im[i] = x[i].imag(); }
  } else {
    std::vector<std::complex<double>> x(n);
    for (int i = 0; i < n; ++i) x[i] = {re[i], im[i]};
    fft_raw(x, ind, 3.14159265358979323846);
    for (int i = 0; i < n; ++i) { re[i] = x[i].real(); im[i] = x[i].imag(); }
  }
}

// dpss_windows (utils/dpss.f90:137-262): out[k][n], unit norm, sign convention
// of the reference; the eigenvalues of the tridiagonal matrix come from
// bisection (the reference calls LAPACK dsterf: any method yields them to
// rounding, and inverse iteration converges to the same vector)
void orc_dpss(int n, double half_nbw, int kmax, double* out) {
  std::vector<std::vector<double>> v;
  dpss(n, half_nbw, kmax, v);
  for (int k = 0; k < kmax; ++k)
    for (int i = 0; i < n; ++i) out[(size_t)k * n + i] = v[k][i];
}

// Multitaper traveltime (imeas 13) or amplitude (14) measurement of one trace,
// MTTTMisfit%calc_adjoint_source (mt_tt_misfit.f90:29-209) with its helpers
// (:211-300, 411-518, 520-791), process_cycle_skipping (signal.f90:510-598) and
// dpss_windows; falls back to the cross-correlation measurement as the
// reference does.  cfg: [0..6] as orc_cc_adjoint, 7 transfunc_waterlevel,
// 8 water_threshold, 9 mt_nw, 10 num_taper, 11 dt_fac, 12 err_fac,
// 13 dt_max_scale, 14 min_cycle_in_window, 15 use_cc_error, 16 min_period,
// 17 max_period, 18 use_mt_error (0/1; the reference's default is .false.).
// Per window out[10]: misfit, residual, error, cc_max, tshift, dlna, selected,
// ishift, measurement used (13/14, or 11/12 after the fallback, 0 unselected),
// the other misfit (misfit_q for 13, misfit_p for 14).  With use_mt_error and
// err non-null, err[w][2][nf] receives the window's jackknife err_dtau and
// err_dlna per frequency (zero when the window never reaches them).  Returns
// the total misfit.
double orc_mt_adjoint(const double* dat, const double* syn, int n, double dt, const double* windows, int nwin,
                      double taper_pct, int itaper, int imeas, const double* cfg, double* adj, double* out,
                      double* err) {
  using cd = std::complex<double>;
  const double pi = 3.141592653589793;
  double total = 0.0;
  for (int i = 0; i < n; ++i) adj[i] = 0.0;
  const int nf = next_pow2_above(n);  // 2 ** exponent(real(nlen))
  const int fnum = nf / 2 + 1;
  const double tlen = n * dt;
  // fftfreq (fftpack.f90:114-133), then df = freq(2) - freq(1), wvec = 2 pi freq
  std::vector<double> wvec(nf);
  for (int i = 1; i <= nf; ++i) {
    const double f = i <= nf / 2 ? (double)(i - 1) / (nf * dt) : -(double)(nf - i + 1) / (nf * dt);
    wvec[i - 1] = 2.0 * pi * f;
  }
  const double df = 1.0 / (nf * dt) - 0.0;
  const int ntaper = (int)cfg[10];
  const bool use_mt_error = cfg[18] != 0.0;
  if (err) for (size_t i = 0; i < 2 * (size_t)nf * nwin; ++i) err[i] = 0.0;
  for (int w = 0; w < nwin; ++w) {
    double* o = out + 10 * (size_t)w;
    const int nb = (int)(windows[2 * w] / dt) + 1, ne = (int)(windows[2 * w + 1] / dt);
    const int nl = ne - nb + 1;
    std::vector<double> sv(syn + nb - 1, syn + ne), dv(dat + nb - 1, dat + ne);
    orc_window_taper(sv.data(), nl, taper_pct, itaper);
    orc_window_taper(dv.data(), nl, taper_pct, itaper);
    const CcStats c = cc_stats(dv, sv, dt, cfg);
    o[0] = 0.0;
    o[1] = imeas == 13 ? c.tshift : c.dlna;
    o[2] = imeas == 13 ? c.sig_dt : c.sig_dlna;
    o[3] = c.cmax;
    o[4] = c.tshift;
    o[5] = c.dlna;
    o[6] = c.sel;
    o[7] = c.ishift;
    o[8] = 0.0;
    o[9] = 0.0;
    if (!c.sel) continue;
    bool mtm = true;
    double mp = 0.0, mq = 0.0;
    std::vector<double> ap, aq;
    do {
      // check_time_series_acceptability (:571-584): compares samples with seconds as written
      if (std::fabs(c.tshift) <= dt || cfg[14] * cfg[16] > nl) { mtm = false; break; }
      // prepare_data_for_mtm (:586-613)
      const int ish = (int)(c.tshift / dt);
      const int nbd = std::max(1, nb + ish), ned = std::min(n, ne + ish);
      if (ned - nbd + 1 != nl) { mtm = false; break; }
      std::vector<double> dm(nl);
      for (int i = 0; i < nl; ++i) dm[i] = dat[nbd - 1 + i] * std::exp(-c.dlna);
      orc_window_taper(dm.data(), nl, taper_pct, itaper);
      // calculate_freq_limits (:615-684) on the whole synthetic trace
      std::vector<cd> sspec(nf, cd(0.0, 0.0));
      for (int i = 0; i < n; ++i) sspec[i] = cd(syn[i], 0.0);
      fft_raw(sspec, -1, pi);
      double ampmax = -1.0;
      int iamp = 1;
      std::vector<double> amp(fnum);
      for (int i = 1; i <= fnum; ++i) {
        const cd v = sspec[i - 1] * dt;
        amp[i - 1] = std::abs(v);
        if (amp[i - 1] > ampmax) { ampmax = amp[i - 1]; iamp = i; }
      }
      const double wl = cfg[8] * ampmax;
      const int ifmin = std::max(1, (int)(1.0 / (cfg[17] * df)));
      const int ifmax = std::min(fnum, (int)(1.0 / (cfg[16] * df)));
      int nfmax = fnum, nfmin = 1;
      bool srch = true;
      auto thermostat = [&](int iw, int& lim) {
        if (amp[iw - 1] < wl && srch) { srch = false; lim = iw; }
        if (amp[iw - 1] > 10 * wl && !srch) { srch = true; lim = iw; }
      };
      if (iamp < fnum) for (int iw = iamp + 1; iw <= fnum; ++iw) thermostat(iw, nfmax);
      nfmax = std::min(std::min(nfmax, ifmax), (int)(1.0 / (2.0 * dt) / df));
      srch = true;
      if (iamp > 1) for (int iw = iamp - 1; iw >= 1; --iw) thermostat(iw, nfmin);
      nfmin = std::max(std::max(nfmin, ifmin), std::max(1, (int)(cfg[14] / tlen / df)));
      if ((nfmax - nfmin) * df < cfg[9] / (4.0 * tlen)) { mtm = false; break; }
      // dpss tapers of the window length, scaled by sqrt(nlen_win)
      std::vector<std::vector<double>> tap;
      dpss(nl, cfg[9], ntaper, tap);
      for (auto& t : tap) for (double& x : t) x = x * std::sqrt((double)nl);
      // calculate_multitaper (:211-300)
      // the tapered spectra of every taper are kept: the jackknife below forms
      // its leave-one-out sums from the same products
      std::vector<std::vector<cd>> pdw(ntaper), psw(ntaper);
      for (int k = 0; k < ntaper; ++k) {
        std::vector<cd>& dw = pdw[k];
        std::vector<cd>& sw = psw[k];
        dw.assign(nf, cd(0.0, 0.0));
        sw.assign(nf, cd(0.0, 0.0));
        for (int i = 0; i < nl; ++i) { dw[i] = cd(dm[i] * tap[k][i], 0.0); sw[i] = cd(sv[i] * tap[k][i], 0.0); }
        fft_raw(dw, -1, pi);
        fft_raw(sw, -1, pi);
        for (int i = 0; i < nf; ++i) {
          dw[i] = dw[i] * dt;
          sw[i] = sw[i] * dt;
        }
      }
      // calculate_multitaper over the tapers with use[k] set: dtau and dlnA
      // per frequency (phase and amplitude only feed these)
      auto multitaper = [&](const std::vector<char>& use, std::vector<double>& dtau, std::vector<double>& dlnw) {
        std::vector<cd> top(nf, cd(0.0, 0.0)), bot(nf, cd(0.0, 0.0));
        for (int k = 0; k < ntaper; ++k) {
          if (!use[k]) continue;
          for (int i = 0; i < nf; ++i) {
            top[i] = top[i] + pdw[k][i] * std::conj(psw[k][i]);
            bot[i] = bot[i] + psw[k][i] * std::conj(psw[k][i]);
          }
        }
        double bmax = 0.0;
        for (int i = 0; i < fnum; ++i) bmax = std::max(bmax, std::abs(bot[i]));
        const double wlevel = bmax * (cfg[7] * cfg[7]);
        std::vector<double> phi(nf, 0.0), absw(nf, 0.0);
        dtau.assign(nf, 0.0);
        dlnw.assign(nf, 0.0);
        std::vector<cd> tf(nf, cd(0.0, 0.0));
        for (int i = nfmin; i <= nfmax; ++i)
          tf[i - 1] = std::abs(bot[i - 1]) >= wlevel ? cdivr(top[i - 1], bot[i - 1].real())
                                                     : cdivr(top[i - 1], bot[i - 1].real() + wlevel);
        for (int i = nfmin; i <= nfmax; ++i) phi[i - 1] = std::atan2(tf[i - 1].imag(), tf[i - 1].real());
        cycle_skipping(phi, nfmin, nfmax, 1.5);
        for (int i = nfmin; i <= nfmax; ++i) absw[i - 1] = std::abs(tf[i - 1]);
        dtau[0] = c.tshift;
        for (int i = std::max(2, nfmin); i <= nfmax; ++i) dtau[i - 1] = -(1.0 / wvec[i - 1]) * phi[i - 1] + c.tshift;
        for (int i = nfmin; i <= nfmax; ++i) dlnw[i - 1] = std::log(absw[i - 1]) + c.dlna;
      };
      std::vector<double> dtau, dlnw;
      multitaper(std::vector<char>(ntaper, 1), dtau, dlnw);
      // calculate_mt_error (:302-409): jackknife over the tapers left out one at a
      // time.  The reference copies rows 1:nlen_f of each kept taper into an
      // nlen_t-row array; the rows that land in the array are those of the
      // taper itself, so each estimate uses the other tapers unchanged.
      std::vector<double> err_dtau(nf, 0.0), err_dlna(nf, 0.0);
      if (use_mt_error) {
        std::vector<std::vector<double>> xt(ntaper), xl(ntaper);
        std::vector<double> et(nf, 0.0), el(nf, 0.0);
        const double fn = (double)ntaper, fn1 = (double)(ntaper - 1);
        for (int k = 0; k < ntaper; ++k) {
          std::vector<char> use(ntaper, 1);
          use[k] = 0;
          multitaper(use, xt[k], xl[k]);
          for (int i = nfmin; i <= nfmax; ++i) {
            et[i - 1] = et[i - 1] + fn * dtau[i - 1] - fn1 * xt[k][i - 1];
            el[i - 1] = el[i - 1] + fn * dlnw[i - 1] - fn1 * xl[k][i - 1];
          }
        }
        for (int i = 0; i < nf; ++i) { et[i] = et[i] / fn; el[i] = el[i] / fn; }
        for (int k = 0; k < ntaper; ++k)
          for (int i = nfmin; i <= nfmax; ++i) {
            err_dtau[i - 1] = err_dtau[i - 1] + (xt[k][i - 1] - et[i - 1]) * (xt[k][i - 1] - et[i - 1]);
            err_dlna[i - 1] = err_dlna[i - 1] + (xl[k][i - 1] - el[i - 1]) * (xl[k][i - 1] - el[i - 1]);
          }
        const double nn = (double)(ntaper * (ntaper - 1));
        for (int i = nfmin; i <= nfmax; ++i) {
          err_dtau[i - 1] = std::sqrt(err_dtau[i - 1] / nn);
          err_dlna[i - 1] = std::sqrt(err_dlna[i - 1] / nn);
        }
        if (err)
          for (int i = 0; i < nf; ++i) {
            err[(2 * (size_t)w) * nf + i] = err_dtau[i];
            err[(2 * (size_t)w + 1) * nf + i] = err_dlna[i];
          }
      }
      // check_mtm_time_shift_acceptability (:520-569)
      for (int j = nfmin; j <= nfmax && mtm; ++j) {
        if (j * df > 0.0) {
          const double per = 1.0 / (j * df);
          if (std::fabs(dtau[j - 1]) > per / cfg[11]) mtm = false;
          else if (err_dtau[j - 1] > per / cfg[12]) mtm = false;
          else if (std::fabs(dtau[j - 1]) > cfg[13] * std::fabs(c.tshift)) mtm = false;
        }
      }
      if (!mtm) break;
      // calculate_freq_domain_taper (:709-791)
      const int L = nfmax - nfmin;
      std::vector<double> wt(nf, 0.0), wtap(std::max(L, 0), 1.0);
      orc_window_taper(wtap.data(), L, 1.0, 3);
      for (int i = 0; i < L; ++i) wt[nfmin - 1 + i] = wtap[i];
      double ssum = 0.0;
      for (int i = nfmin; i <= nfmax - 1; ++i) ssum = ssum + wt[i - 1];
      const double ffac = 2.0 * df * ssum;
      std::vector<double> wp(nf), wq(nf);
      for (int i = 0; i < nf; ++i) { wp[i] = wt[i] / ffac; wq[i] = wt[i] / ffac; }
      if (cfg[15] != 0.0) {
        for (int i = 0; i < nf; ++i) { wp[i] = wp[i] / (c.sig_dt * c.sig_dt); wq[i] = wq[i] / (c.sig_dlna * c.sig_dlna); }
      } else if (use_mt_error) {
        // water level of the multitaper errors: a fraction of the mean |dtau|,
        // |dlnA| over [nfreq_min, nfreq_max - 1], added where an error is below it
        double st = 0.0, sl = 0.0;
        for (int i = nfmin; i <= nfmax - 1; ++i) { st = st + std::fabs(dtau[i - 1]); sl = sl + std::fabs(dlnw[i - 1]); }
        const double dtau_wtr = cfg[8] * st / (double)(nfmax - nfmin);
        const double dlna_wtr = cfg[8] * sl / (double)(nfmax - nfmin);
        for (int i = nfmin; i <= nfmax; ++i) {
          double et = err_dtau[i - 1], el = err_dlna[i - 1];
          if (et < dtau_wtr) et = et + dtau_wtr;
          if (el < dlna_wtr) el = el + dlna_wtr;
          wp[i - 1] = wp[i - 1] / (et * et);
          wq[i - 1] = wq[i - 1] / (el * el);
        }
      }
      std::vector<double> yp(nf), yq(nf);
      for (int i = 0; i < nf; ++i) { yp[i] = dtau[i] * dtau[i] * wp[i]; yq[i] = dlnw[i] * dlnw[i] * wq[i]; }
      mp = orc_simpson(yp.data(), nf, df);
      mq = orc_simpson(yq.data(), nf, df);
      // calculate_mt_adjsrc (:411-518); the inverse transform in single precision
      std::vector<std::vector<cd>> stw(ntaper), stvw(ntaper);
      std::vector<cd> bp(nf, cd(0.0, 0.0)), bq(nf, cd(0.0, 0.0));
      for (int k = 0; k < ntaper; ++k) {
        std::vector<double> st(nl);
        for (int i = 0; i < nl; ++i) st[i] = sv[i] * tap[k][i];
        const std::vector<double> stv = gradient(st, dt);
        stw[k].assign(nf, cd(0.0, 0.0));
        stvw[k].assign(nf, cd(0.0, 0.0));
        for (int i = 0; i < nl; ++i) { stw[k][i] = cd(st[i], 0.0); stvw[k][i] = cd(stv[i], 0.0); }
        fft_raw(stw[k], -1, pi);
        fft_raw(stvw[k], -1, pi);
        for (int i = 0; i < nf; ++i) {
          stw[k][i] = stw[k][i] * dt;
          stvw[k][i] = stvw[k][i] * dt;
          bp[i] = bp[i] + stvw[k][i] * std::conj(stvw[k][i]);
          bq[i] = bq[i] + stw[k][i] * std::conj(stw[k][i]);
        }
      }
      ap.assign(nl, 0.0);
      aq.assign(nl, 0.0);
      for (int k = 0; k < ntaper; ++k) {
        std::vector<std::complex<float>> pw(nf, 0.0f), qw(nf, 0.0f);
        for (int i = 1; i <= nf; ++i) {
          cd p(0.0, 0.0), q(0.0, 0.0);
          if (i >= nfmin && i <= nfmax) {
            if (std::abs(bp[i - 1]) > 0.0) p = cdivr(stvw[k][i - 1], bp[i - 1].real());
            if (std::abs(bq[i - 1]) > 0.0) q = -cdivr(stw[k][i - 1], bq[i - 1].real());
          }
          p = p * dtau[i - 1] * wp[i - 1];
          q = q * dlnw[i - 1] * wq[i - 1];
          pw[i - 1] = std::complex<float>((float)p.real(), (float)p.imag()) / (float)nf;
          qw[i - 1] = std::complex<float>((float)q.real(), (float)q.imag()) / (float)nf;
        }
        fft_raw(pw, 1, 3.14159265358979323846f);
        fft_raw(qw, 1, 3.14159265358979323846f);
        for (int i = 0; i < nl; ++i) {
          ap[i] = ap[i] + (double)pw[i].real() * 2.0 / dt * tap[k][i];
          aq[i] = aq[i] + (double)qw[i].real() * 2.0 / dt * tap[k][i];
        }
      }
    } while (false);
    if (!mtm) {
      cc_adjsrc(sv, c, dt, 0, ap, aq, mp, mq);
      o[8] = imeas == 13 ? 11 : 12;
    } else {
      o[8] = imeas;
    }
    orc_window_taper(ap.data(), nl, taper_pct, itaper);
    orc_window_taper(aq.data(), nl, taper_pct, itaper);
    const std::vector<double>& a = imeas == 13 ? ap : aq;
    const double mf = imeas == 13 ? mp : mq;
    o[0] = mf;
    o[9] = imeas == 13 ? mq : mp;
    total = total + mf;
    for (int i = 0; i < nl; ++i) adj[nb - 1 + i] = adj[nb - 1 + i] + a[i];
  }
  return total;
}

// analytic signal of x (hilbert) for tests: re/im of length n
void orc_hilbert(const double* x, int n, double* re, double* im) {
  const auto xa = hilbert(std::vector<double>(x, x + n));
  for (int i = 0; i < n; ++i) { re[i] = xa[i].real(); im[i] = xa[i].imag(); }
}

// Exponentiated phase measurement of one trace, ExponentiatedPhaseMisfit%
// calc_adjoint_source (exponentiated_phase_misfit.f90:21-104); per window out[8]:
// misfit, residual, misfit_real, misfit_imag, residual_real, residual_imag,
// selected (always 1), 0.  Returns the total misfit.
double orc_ep_adjoint(const double* dat, const double* syn, int n, double dt, const double* windows, int nwin,
                      double taper_pct, int itaper, double wtr_env, double* adj, double* out) {
  double total = 0.0;
  for (int i = 0; i < n; ++i) adj[i] = 0.0;
  for (int w = 0; w < nwin; ++w) {
    const int nb = (int)(windows[2 * w] / dt) + 1, ne = (int)(windows[2 * w + 1] / dt);
    const int nl = ne - nb + 1;
    std::vector<double> sv(syn + nb - 1, syn + ne), dv(dat + nb - 1, dat + ne);
    orc_window_taper(sv.data(), nl, taper_pct, itaper);
    orc_window_taper(dv.data(), nl, taper_pct, itaper);
    const auto hs = hilbert(sv), hd = hilbert(dv);
    std::vector<double> es(nl), ed(nl);
    double ms = 0.0, md = 0.0;
    for (int i = 0; i < nl; ++i) {
      es[i] = std::abs(hs[i]);
      ed[i] = std::abs(hd[i]);
      ms = std::max(ms, es[i]);
      md = std::max(md, ed[i]);
    }
    const double ts = wtr_env * ms, td = wtr_env * md;
    std::vector<double> dr(nl), di(nl), dr2(nl), di2(nl), cub(nl);
    double rr = 0.0, ri = 0.0;
    for (int i = 0; i < nl; ++i) {
      const double esw = es[i] + ts, edw = ed[i] + td;
      dr[i] = dv[i] / edw - sv[i] / esw;
      di[i] = hd[i].imag() / edw - hs[i].imag() / esw;
      dr2[i] = dr[i] * dr[i];
      di2[i] = di[i] * di[i];
      cub[i] = esw * esw * esw;
      rr = rr + dr[i];
      ri = ri + di[i];
    }
    rr = rr / nl;
    ri = ri / nl;
    const double mr = 0.5 * orc_simpson(dr2.data(), nl, dt), mi = 0.5 * orc_simpson(di2.data(), nl, dt);
    double* o = out + 8 * (size_t)w;
    o[0] = mr + mi;
    o[1] = rr + ri;
    o[2] = mr;
    o[3] = mi;
    o[4] = rr;
    o[5] = ri;
    o[6] = 1.0;
    o[7] = 0.0;
    total = total + o[0];
    std::vector<double> u1(nl), u2(nl);
    for (int i = 0; i < nl; ++i) {
      const double h = hs[i].imag();
      u1[i] = dr[i] * sv[i] * h / cub[i];
      u2[i] = di[i] * (sv[i] * sv[i]) / cub[i];
    }
    const auto h1 = hilbert(u1), h2 = hilbert(u2);
    std::vector<double> ar(nl), ai(nl);
    for (int i = 0; i < nl; ++i) {
      const double h = hs[i].imag();
      ar[i] = -1.0 * (dr[i] * (h * h) / cub[i]) - h1[i].imag();
      ai[i] = di[i] * sv[i] * h / cub[i] + h2[i].imag();
    }
    orc_window_taper(ar.data(), nl, taper_pct, itaper);
    orc_window_taper(ai.data(), nl, taper_pct, itaper);
    for (int i = 0; i < nl; ++i) adj[nb - 1 + i] = adj[nb - 1 + i] + (ar[i] + ai[i]);
  }
  return total;
}

// win_sel_type (src/preproc/window_selector/win_sel.f90:56-437): normalisation by max|syn|, sliding_cc (Hann-tapered windows of
// int(sliding_win_len_fac min_period / dt) + 1 samples at every start),
// good samples (cc, shift, time range), jump exclusion, contiguous groups and
// the length / peak count / SNR / energy-ratio checks, with cfg[9] the groups
// split at the STA/LTA troughs between phases first.  cfg: sliding_win_len_fac,
// min_velocity, threshold_corr, threshold_shift_fac, jump_fac, min_win_len_fac,
// min_snr_window, min_energy_ratio, min_peaks_troughs, is_split_phases.  twin[maxwin][2] (s);
// returns the window count.  cc_out / ts_out (npts - nlen) receive the sliding
// correlation when not null.
int orc_win_sel(const double* dat_in, const double* syn_in, int npts, double dt, double t0, double tp, double dis,
                double min_period, const double* cfg, int maxwin, double* twin, double* cc_out, double* ts_out) {
  double smax = 0.0;
  for (int i = 0; i < npts; ++i) smax = std::max(smax, std::fabs(syn_in[i]));
  std::vector<double> dat(npts), syn(npts);
  for (int i = 0; i < npts; ++i) { dat[i] = dat_in[i] / smax; syn[i] = syn_in[i] / smax; }
  const double tstart = tp - min_period * 1.5 + t0;
  const double tend = dis / cfg[1] + min_period * 1.5 + t0;
  const double jump = cfg[4] * min_period;
  const int nstart = std::min(std::max((int)(tstart / dt) + 1, 1), npts);  // clamped to the trace
  double e = 0.0;
  for (int i = 0; i < nstart; ++i) e = e + dat_in[i] * dat_in[i];  // the un-normalised data, as written
  const double noise = std::sqrt(e / (double)nstart);
  // sliding_cc
  const int nlen = (int)(cfg[0] * min_period / dt) + 1;
  const int nw = npts - nlen;
  if (nw <= 0) return 0;
  std::vector<double> ts(nw), cc(nw), tc(nw);
  std::vector<double> dw(nlen), sw(nlen);
  for (int i = 0; i < nw; ++i) {
    for (int j = 0; j < nlen; ++j) { dw[j] = dat[i + j]; sw[j] = syn[i + j]; }
    orc_window_taper(dw.data(), nlen, 1.0, 1);
    orc_window_taper(sw.data(), nlen, 1.0, 1);
    xcorr_shift_coe(dw, sw, dt, ts[i], cc[i]);
    tc[i] = ((double)i + (double)nlen / 2.0) * dt;
    if (cc_out) cc_out[i] = cc[i];
    if (ts_out) ts_out[i] = ts[i];
  }
  std::vector<char> good(nw);
  for (int i = 0; i < nw; ++i)
    good[i] = cc[i] > cfg[2] && std::fabs(ts[i]) < cfg[3] * min_period && tc[i] >= tstart && tc[i] <= tend;
  for (int i = 0; i < nw; ++i) {
    const double g = i == 0 ? 0.0 : ts[i] - ts[i - 1];
    if (std::fabs(g) > jump) {
      const int nj = (int)std::round(jump / dt);  // nint: half away from zero
      for (int k = std::max(0, i - nj); k <= std::min(nw - 1, i + nj); ++k) good[k] = 0;
    }
  }
  int count = 0;
  auto check = [&](int a, int b) {  // 0-based group [a, b] of the sliding windows
    if ((b - a) * dt < cfg[5] * min_period) return;
    int ib = (int)(tc[a] / dt) + 1, ie = (int)(tc[b] / dt) + 1;
    ib = std::max(1, std::min(ib, npts));
    ie = std::max(1, std::min(ie, npts));
    if (ie <= ib) return;
    const int m = ie - ib + 1;
    const double* dp = dat.data() + ib - 1;
    const double* sp = syn.data() + ib - 1;
    const int np = std::min(count_maxima(dp, m, 1.0) + count_maxima(dp, m, -1.0),
                            count_maxima(sp, m, 1.0) + count_maxima(sp, m, -1.0));
    if (np < (int)cfg[8]) return;
    double amax = 0.0;
    for (int i = 0; i < m; ++i) amax = std::max(amax, std::fabs(dp[i]));
    if (amax / noise < cfg[6]) return;
    double ed = 0.0, es = 0.0;
    for (int i = 0; i < m; ++i) ed = ed + dp[i] * dp[i];
    for (int i = 0; i < m; ++i) es = es + sp[i] * sp[i];
    if (ed / es > cfg[7] || ed / es < 1 / cfg[7]) return;
    if (count < maxwin) { twin[2 * count] = tc[a]; twin[2 * count + 1] = tc[b]; }
    ++count;
  };
  // phase splitting (cfg[9]): envelope of the normalised synthetic, STA/LTA,
  // its peaks and troughs; split_window_by_phases (win_sel.f90:279-437)
  std::vector<double> stl;
  std::vector<int> peaks, troughs;
  if (cfg[9] != 0.0) {
    const auto xa = hilbert(syn);
    std::vector<double> env(npts);
    for (int i = 0; i < npts; ++i) env[i] = std::abs(xa[i]);
    stl = sta_lta(env, dt, min_period);
    peaks = maxima(stl, 1.0);
    troughs = maxima(stl, -1.0);
  }
  const int minlen = (int)(cfg[5] * min_period / dt);
  auto split = [&](int a, int b) {  // 0-based group -> 0-based sub-windows
    std::vector<std::pair<int, int>> sub;
    const int ns = (int)stl.size();
    int s0 = (int)(tc[a] / dt) + 1, s1 = (int)(tc[b] / dt) + 1;
    s0 = std::max(1, std::min(s0, ns));
    s1 = std::max(1, std::min(s1, ns));
    if (s0 >= s1) return std::vector<std::pair<int, int>>{{a, b}};
    std::vector<int> wp;
    for (int p : peaks) if (p >= s0 && p <= s1) wp.push_back(p);
    if (wp.size() <= 1) return std::vector<std::pair<int, int>>{{a, b}};
    std::vector<int> sp{s0};
    for (size_t i = 0; i + 1 < wp.size(); ++i) {
      const int p1 = wp[i], p2 = wp[i + 1];
      double mv = std::numeric_limits<double>::max();
      int mi = -1;
      for (int tr : troughs)
        if (tr > p1 && tr < p2) {
          if (mi < 0) mi = tr;
          if (stl[tr - 1] < mv) { mv = stl[tr - 1]; mi = tr; }
        }
      sp.push_back(mi >= 0 ? mi : (p1 + p2) / 2);
    }
    sp.push_back(s1);
    for (size_t i = 0; i + 1 < sp.size(); ++i) {
      if (sp[i + 1] - sp[i] < minlen) continue;
      int j0 = a, j1 = b;
      for (int q = a; q <= b; ++q) {
        if (std::fabs(tc[q] - sp[i] * dt) < std::fabs(tc[j0] - sp[i] * dt)) j0 = q;
        if (std::fabs(tc[q] - sp[i + 1] * dt) < std::fabs(tc[j1] - sp[i + 1] * dt)) j1 = q;
      }
      sub.push_back({std::max(a, std::min(j0, b)), std::max(a, std::min(j1, b))});
    }
    if (sub.empty()) sub.push_back({a, b});
    return sub;
  };
  auto group = [&](int a, int b) {
    if (cfg[9] == 0.0) { check(a, b); return; }
    for (const auto& ab : split(a, b)) check(ab.first, ab.second);
  };
  int start = -1;
  for (int i = 0; i < nw; ++i) {
    if (good[i] && start < 0) start = i;
    else if (!good[i] && start >= 0) { group(start, i - 1); start = -1; }
  }
  if (start >= 0) group(start, nw - 1);
  return count;
}

// WaveformConvMisfit%calc_adjoint_source (waveform_conv_misfit.f90:17-64): the
// cross-convolution of radial and vertical components, c1 = syn_r * dat_z,
// c2 = dat_r * syn_z, diff = (c1 - c2) dt, adjoint sources
// conv(rev(dat_z), diff) dt and conv(-rev(dat_r), diff) dt on the window;
// out[0] misfit 0.5 simpson(tapered diff^2), out[1] residual sum(diff) / nlen_win
void orc_conv_adjoint(const double* dat_r, const double* dat_z, const double* syn_r, const double* syn_z, int n,
                      double dt, double w0, double w1, double taper_pct, int itaper, double* adj_r, double* adj_z,
                      double* out) {
  const int nb = (int)(w0 / dt) + 1, ne = (int)(w1 / dt), nl = ne - nb + 1;
  const std::vector<double> dr(dat_r, dat_r + n), dz(dat_z, dat_z + n), sr(syn_r, syn_r + n), sz(syn_z, syn_z + n);
  const std::vector<double> c1 = conv_mid(sr, dz), c2 = conv_mid(dr, sz);
  std::vector<double> cd(n);
  for (int i = 0; i < n; ++i) cd[i] = (c1[i] - c2[i]) * dt;
  std::vector<double> rz(n), rr(n);
  for (int i = 0; i < n; ++i) { rz[i] = dz[n - 1 - i]; rr[i] = -dr[n - 1 - i]; }
  std::vector<double> ar = conv_mid(rz, cd), az = conv_mid(rr, cd);
  std::vector<double> tr(nl), tz(nl), tcd(cd.begin() + nb - 1, cd.begin() + ne);
  for (int i = 0; i < nl; ++i) { tr[i] = ar[nb - 1 + i] * dt; tz[i] = az[nb - 1 + i] * dt; }
  orc_window_taper(tcd.data(), nl, taper_pct, itaper);
  orc_window_taper(tr.data(), nl, taper_pct, itaper);
  orc_window_taper(tz.data(), nl, taper_pct, itaper);
  std::vector<double> sq(nl);
  double sm = 0.0;
  for (int i = 0; i < nl; ++i) { sq[i] = tcd[i] * tcd[i]; sm = sm + tcd[i]; }
  out[0] = 0.5 * orc_simpson(sq.data(), nl, dt);
  out[1] = sm / nl;
  for (int i = 0; i < n; ++i) { adj_r[i] = 0.0; adj_z[i] = 0.0; }
  for (int i = 0; i < nl; ++i) { adj_r[nb - 1 + i] = tr[i]; adj_z[nb - 1 + i] = tz[i]; }
}


// schedule_weighted_intervals (interval_scheduling.f90:13-165): intervals
// (left, right, weight) sorted by right end with the reference's quicksort
// (middle pivot swapped to the end, ties within 1e-6 ordered by left end),
// p(j) the last earlier interval ending at or before j's start (bisection),
// opt(j) = max(w_j + opt(p(j)), opt(j - 1)), backtracking from the last one.
// sel receives the chosen intervals' input positions in time order; returns
// their count.
int orc_schedule_intervals(int n, const double* left, const double* right, const double* weight, int* sel) {
  struct Iv { int idx; double left, right, weight; };
  std::vector<Iv> iv(n);
  for (int i = 0; i < n; ++i) iv[i] = {i, left[i], right[i], weight[i]};
  if (n == 0) return 0;
  std::function<void(int, int)> qs = [&](int first, int last) {
    if (first >= last) return;
    const int pi = (first + last) / 2;
    const Iv pv = iv[pi];
    std::swap(iv[last], iv[pi]);
    int i = first - 1;
    for (int j = first; j <= last - 1; ++j) {
      if (iv[j].right < pv.right - 1.0e-6 || (std::fabs(iv[j].right - pv.right) <= 1.0e-6 && iv[j].left <= pv.left)) {
        ++i;
        std::swap(iv[i], iv[j]);
      }
    }
    std::swap(iv[i + 1], iv[last]);
    qs(first, i);
    qs(i + 2, last);
  };
  qs(0, n - 1);
  std::vector<int> pj(n + 1, 0);
  for (int j = 1; j <= n; ++j) {
    int lo = 1, hi = j - 1, idx = 0;
    while (lo <= hi) {
      const int mid = (lo + hi) / 2;
      if (iv[mid - 1].right <= iv[j - 1].left) { idx = mid; lo = mid + 1; }
      else hi = mid - 1;
    }
    pj[j] = idx;
  }
  std::vector<double> opt(n + 1, 0.0);
  for (int j = 1; j <= n; ++j) {
    const double a = iv[j - 1].weight + opt[pj[j]], b = opt[j - 1];
    opt[j] = a > b ? a : b;
  }
  std::vector<int> back;
  for (int j = n; j > 0;) {
    if (iv[j - 1].weight + opt[pj[j]] > opt[j - 1]) { back.push_back(iv[j - 1].idx); j = pj[j]; }
    else j = j - 1;
  }
  const int ns = (int)back.size();
  for (int k = 0; k < ns; ++k) sel[k] = back[ns - 1 - k];
  return ns;
}

// STA_LTA of an envelope (for tests): out[n]
void orc_sta_lta(const double* env, int n, double dt, double min_period, double* out) {
  const std::vector<double> r = sta_lta(std::vector<double>(env, env + n), dt, min_period);
  for (int i = 0; i < n; ++i) out[i] = r[i];
}

// FLEXWIN window selection of one trace (flexwin_type%select_windows,
// src/preproc/window_selector/flexwin.f90:75-734, with schedule_weighted_intervals of
// interval_scheduling.f90:13-165): STA/LTA of the synthetic's envelope, candidate
// windows from every trough pair around each STA/LTA peak above the water level,
// the rejection passes in the reference's order, the data-fit criteria of the
// survivors, then merging (or, with is_split_phases, weighted interval scheduling).
// cfg[16]: min_velocity, min_win_len_fac, threshold_corr, threshold_shift_fac,
// min_peaks_troughs, threshold_dlna, min_snr_window, is_split_phases (0/1),
// max_time_before_first_arrival, stalta_water_level, c_0, c_2, c_3a, c_3b, c_4a, c_4b.
// twin[maxwin][2] (s, relative to t0 as the reference's twin), info[maxwin][4]:
// weight, cc, time shift, dlnA.  Returns the window count (windows beyond maxwin
// are counted, not written).  The noise level reads at most the trace's samples.
struct FwCand {
  int left, right, peak;
  double cc, dlna, tshift, weight, snr_amp;
  bool rejected;
};

int orc_flexwin(const double* dat_in, const double* syn_in, int npts, double dt, double t0, double tp, double dis,
                double min_period, double max_period, const double* cfg, int maxwin, double* twin, double* info) {
  // initialize (:75-108)
  double smax = 0.0;
  for (int i = 0; i < npts; ++i) smax = std::max(smax, std::fabs(syn_in[i]));
  std::vector<double> dat(npts), syn(npts);
  for (int i = 0; i < npts; ++i) { dat[i] = dat_in[i] / smax; syn[i] = syn_in[i] / smax; }
  const double tstart = tp - cfg[8] + t0;
  const double tend = dis / cfg[0] + max_period + t0;
  const int nstart = (int)(tstart / dt) + 1;
  const int nend = std::min(npts, (int)(tend / dt) + 1);
  double noise;
  if (nstart > 0) {
    const int m = std::min(nstart, npts);
    double e = 0.0;
    for (int i = 0; i < m; ++i) e = e + dat_in[i] * dat_in[i];
    noise = std::sqrt(e / (double)m);
  } else {
    int ntp = (int)((tp + t0) / dt) + 1;
    if (ntp > 1) {
      ntp = std::min(ntp, npts);
      double e = 0.0;
      for (int i = 0; i < ntp; ++i) e = e + dat_in[i] * dat_in[i];
      noise = std::sqrt(e / (double)ntp);
    } else {
      noise = 0.001;
    }
  }
  // calculate_criteria (:110-181) on samples [left, right] (1-based)
  auto criteria = [&](FwCand& c) {
    const int nl = c.right - c.left + 1;
    std::vector<double> dw(dat.begin() + (c.left - 1), dat.begin() + c.right);
    std::vector<double> sw(syn.begin() + (c.left - 1), syn.begin() + c.right);
    const int ish = xcorr_shift(dw, sw);
    c.tshift = ish * dt;
    double sd2 = 0.0, ss2 = 0.0;
    for (int i = 0; i < nl; ++i) sd2 = sd2 + dw[i] * dw[i];
    for (int i = 0; i < nl; ++i) ss2 = ss2 + sw[i] * sw[i];
    c.dlna = 0.5 * std::log(sd2 / ss2);
    // cc_max_coef (utils/cc.f90:132-160)
    std::vector<double> ssh(nl, 0.0), dsh(dw);
    for (int i = 0; i < nl; ++i) {
      const int j = i - ish;
      if (j >= 0 && j < nl) ssh[i] = sw[j];
    }
    if (ish > 0) for (int i = 0; i < ish && i < nl; ++i) dsh[i] = 0.0;
    else if (ish < 0) for (int i = nl + ish; i < nl; ++i) if (i >= 0) dsh[i] = 0.0;
    double a = 0.0, b = 0.0, q = 0.0;
    for (int i = 0; i < nl; ++i) a = a + dsh[i] * ssh[i];
    for (int i = 0; i < nl; ++i) b = b + dsh[i] * dsh[i];
    for (int i = 0; i < nl; ++i) q = q + ssh[i] * ssh[i];
    c.cc = a / std::sqrt(b * q);
    double amax = 0.0;
    for (int i = 0; i < nl; ++i) amax = std::max(amax, std::fabs(dw[i]));
    c.snr_amp = noise > 0.0 ? amax / noise : 1.0e5;
    if (cfg[4] > 0.0 && count_maxima(sw.data(), nl, 1.0) + count_maxima(sw.data(), nl, -1.0) < (int)cfg[4])
      c.rejected = true;
    c.weight = (double)(c.right - c.left) * dt / min_period * c.cc;
  };
  // select_windows (:183-391)
  const std::vector<std::complex<double>> xa = hilbert(syn);
  std::vector<double> env(npts);
  for (int i = 0; i < npts; ++i) env[i] = std::abs(xa[i]);
  const std::vector<double> sl = sta_lta(env, dt, min_period);
  const std::vector<int> peaks = maxima(sl, 1.0), troughs = maxima(sl, -1.0);
  const int np = (int)peaks.size(), nt = (int)troughs.size();
  if (np == 0 || nt < 2) return 0;
  const int max_good_win = 500000;
  std::vector<FwCand> cand;
  for (int i = 0; i < np && (int)cand.size() < max_good_win; ++i) {
    if (sl[peaks[i] - 1] < cfg[9]) continue;
    for (int j = 0; j < nt && (int)cand.size() < max_good_win; ++j) {
      if (troughs[j] >= peaks[i]) break;
      for (int k = 0; k < nt; ++k) {
        if (troughs[k] <= peaks[i]) continue;
        if ((int)cand.size() >= max_good_win) break;
        FwCand c{troughs[j], troughs[k], peaks[i], 0.0, 0.0, 0.0, 0.0, 0.0, false};
        if ((double)(c.left - 1) * dt < tstart) c.left = nstart;
        if ((double)(c.right - 1) * dt > tend) c.right = nend;
        cand.push_back(c);
      }
    }
  }
  const int nc = (int)cand.size();
  // 4.0 reject_on_traveltimes (:393-416)
  const double min_time = tstart + cfg[8] - min_period, max_time = tend - max_period;
  for (auto& c : cand) {
    if (c.rejected) continue;
    const double ws = (double)(c.left - 1) * dt, we = (double)(c.right - 1) * dt;
    if (!(we >= min_time && ws <= max_time)) c.rejected = true;
  }
  // 4.1 minimum length
  const auto min_length = [&]() {
    for (auto& c : cand)
      if (!c.rejected && (double)(c.right - c.left) * dt < cfg[1] * min_period) c.rejected = true;
  };
  min_length();
  // 4.2 reject_on_minima_water_level (:418-446)
  const double wmid = cfg[10] * cfg[9];
  for (auto& c : cand) {
    if (c.rejected) continue;
    for (int j = 0; j < nt; ++j)
      if (troughs[j] > c.left && troughs[j] < c.right && sl[troughs[j] - 1] <= wmid) { c.rejected = true; break; }
  }
  // 4.3 reject_on_prominence_of_central_peak (:448-499)
  if (cfg[11] != 0.0) {
    for (auto& c : cand) {
      if (c.rejected) continue;
      int lt = -1, rt = -1;
      for (int j = nt - 1; j >= 0; --j)
        if (troughs[j] < c.peak) { lt = troughs[j]; break; }
      for (int j = 0; j < nt; ++j)
        if (troughs[j] > c.peak) { rt = troughs[j]; break; }
      if (lt == -1 || rt == -1) { c.rejected = true; continue; }
      const double cv = sl[c.peak - 1], dl = cv - sl[lt - 1], dr = cv - sl[rt - 1];
      if (dl < cfg[11] * cv || dr < cfg[11] * cv) c.rejected = true;
    }
  }
  // 4.4 reject_on_phase_separation (:501-547)
  for (auto& c : cand) {
    if (c.rejected) continue;
    double smin = std::numeric_limits<double>::max();
    for (int j = 0; j < nt; ++j)
      if (troughs[j] >= c.left && troughs[j] <= c.right && sl[troughs[j] - 1] < smin) smin = sl[troughs[j] - 1];
    const double dcen = sl[c.peak - 1] - smin;
    for (int j = 0; j < np; ++j) {
      if (peaks[j] >= c.left && peaks[j] <= c.right && peaks[j] != c.peak) {
        const double ds = sl[peaks[j] - 1] - smin;
        const double dtm = (double)std::abs(c.peak - peaks[j]) * dt / min_period;
        const double ft = dtm >= cfg[13] ? std::exp(-((dtm - cfg[13]) / cfg[13]) * ((dtm - cfg[13]) / cfg[13])) : 1.0;
        if (ds > cfg[12] * dcen * ft) { c.rejected = true; break; }
      }
    }
  }
  // 4.5 curtail_length_of_windows (:549-601)
  const double tdl = min_period * cfg[14] / dt, tdr = min_period * cfg[15] / dt;
  for (auto& c : cand) {
    if (c.rejected) continue;
    int nin = 0, il = 0, ir = 0;
    for (int j = 0; j < np; ++j)
      if (peaks[j] >= c.left && peaks[j] <= c.right && peaks[j] != c.peak) {
        if (nin == 0) il = peaks[j];
        ir = peaks[j];
        ++nin;
      }
    if (nin < 2) continue;
    const double dl = (double)(il - c.left), dr = (double)(c.right - ir);
    if (dl > tdl) c.left = (int)((double)il - tdl);
    if (dr > tdr) c.right = (int)((double)ir + tdr);
  }
  // 4.6 remove_duplicates (:603-625)
  for (int i = 0; i < nc; ++i) {
    if (cand[i].rejected) continue;
    cand[i].peak = (cand[i].left + cand[i].right) / 2;
    for (int j = i + 1; j < nc; ++j)
      if (!cand[j].rejected && cand[j].left == cand[i].left && cand[j].right == cand[i].right) cand[j].rejected = true;
  }
  // 4.7 minimum length again; 4.8 criteria and the data-fit rejections
  min_length();
  for (auto& c : cand) {
    if (c.rejected) continue;
    criteria(c);
    if (c.cc < cfg[2]) c.rejected = true;
    if (std::fabs(c.tshift) > cfg[3] * min_period) c.rejected = true;
    if (std::fabs(c.dlna) > cfg[5]) c.rejected = true;
    if (c.snr_amp < cfg[6]) c.rejected = true;
  }
  std::vector<FwCand> out;
  if (cfg[7] == 0.0) {
    // merge_windows (:627-700): valid candidates sorted by left (sort_candidates_by_left,
    // :702-718), overlapping or touching windows merged, criteria recomputed
    std::vector<FwCand> v;
    for (auto& c : cand)
      if (!c.rejected) v.push_back(c);
    const int n = (int)v.size();
    for (int i = 0; i < n - 1; ++i)
      for (int j = i + 1; j < n; ++j)
        if (v[j].left < v[i].left) std::swap(v[i], v[j]);
    for (int i = 0; i < n; ++i) {
      if (out.empty() || out.back().right + 1 < v[i].left) out.push_back(v[i]);
      else out.back().right = std::max(out.back().right, v[i].right);
    }
    for (auto& c : out) {
      c.peak = (c.left + c.right) / 2;
      criteria(c);
      c.rejected = false;
    }
  } else {
    // schedule_weighted_intervals (interval_scheduling.f90:13-165)
    std::vector<int> idx;
    std::vector<double> lf, rt, wt;
    for (int i = 0; i < nc; ++i)
      if (!cand[i].rejected) {
        idx.push_back(i);
        lf.push_back((double)cand[i].left * dt);
        rt.push_back((double)cand[i].right * dt);
        wt.push_back(cand[i].weight);
      }
    const int n = (int)idx.size();
    if (n == 0) return 0;
    std::vector<int> sel(n);
    const int ns = orc_schedule_intervals(n, lf.data(), rt.data(), wt.data(), sel.data());
    for (int k = 0; k < ns; ++k) out.push_back(cand[idx[sel[k]]]);
  }
  const int nw = (int)out.size();
  for (int i = 0; i < nw && i < maxwin; ++i) {
    twin[2 * i] = (double)(out[i].left - 1) * dt - t0;
    twin[2 * i + 1] = (double)(out[i].right - 1) * dt - t0;
    info[4 * i] = out[i].weight;
    info[4 * i + 1] = out[i].cc;
    info[4 * i + 2] = out[i].tshift;
    info[4 * i + 3] = out[i].dlna;
  }
  return nw;
}

}  // extern "C"
