// CPU restatement of SpecFWAT's grid optimiser directions and model update.
//
// TEST INFRASTRUCTURE ONLY: compiled into liboracle_sem.so and used by tests/
// and bench.py's cpu_baseline leg as the checker of the product's optimiser
// kernels; the product never links it.
//
//   orc_sd_direction<T>     get_SD_direction, src/optimization/optimize_grid.f90:229-241
//   orc_lbfgs_direction<T>  get_lbfgs_direction, optimize_grid.f90:243-339, with
//                           inner_product (src/shared/utils.f90:1527-1534) and the
//                           log transform of read_model_grid(.., .true.)
//                           (src/optimization/opt_io.f90:79-81)
//   orc_cg_direction<T>     get_CG_direction, optimize_grid.f90:341-389
//   orc_model_update        model_update + alpha_scaling, optimize_grid.f90:176-203, 391-404
//
// T = float restates the reference (single-precision arrays and sequential
// single-precision sums); T = double is the same algorithm with double vectors
// and sums, the precision the device code keeps.  Arrays are grids
// (nx,ny,nz,nkernel) flattened, npg = nx*ny*nz points per parameter.
#include <algorithm>
#include <cmath>
#include <vector>

namespace {

// inner_product: sum(a * b * dx * dy * dz), evaluated left to right per element
template <class T>
T dot(const std::vector<T>& a, const std::vector<T>& b, const T* h) {
  T s = 0;
  for (size_t i = 0; i < a.size(); ++i) s = s + a[i] * b[i] * h[0] * h[1] * h[2];
  return s;
}

template <class T>
T maxabs(const std::vector<T>& a) {
  T m = 0;
  for (T v : a) m = std::max(m, (T)std::fabs(v));
  return m;
}

// model difference in the inversion's variables: log for vp, vs, rho (the
// first three parameters of model types 1 and 2), linear for gcp, gsp
template <class T>
std::vector<T> model_diff(const float* m1, const float* m0, long n, long npg) {
  std::vector<T> d(n);
  for (long i = 0; i < n; ++i) {
    if (i / npg < 3) d[i] = (T)std::log((T)m1[i]) - (T)std::log((T)m0[i]);
    else d[i] = (T)m1[i] - (T)m0[i];
  }
  return d;
}

template <class T>
std::vector<T> grad_diff(const float* g1, const float* g0, long n) {
  std::vector<T> d(n);
  for (long i = 0; i < n; ++i) d[i] = (T)g1[i] - (T)g0[i];
  return d;
}

template <class T>
void lbfgs(long n, long npg, int nstore, const float* grads, const float* models, const float* hess, float dx,
           float dy, float dz, float* dir, double* angle) {
  const T vol[3] = {(T)dx, (T)dy, (T)dz};
  std::vector<T> q(n), gcur(n);
  const float* gk = grads + (size_t)nstore * n;
  for (long i = 0; i < n; ++i) q[i] = gcur[i] = (T)gk[i];
  std::vector<T> p(nstore), a(nstore);
  for (int k = nstore - 1; k >= 0; --k) {
    std::vector<T> y = grad_diff<T>(grads + (size_t)(k + 1) * n, grads + (size_t)k * n, n);
    std::vector<T> s = model_diff<T>(models + (size_t)(k + 1) * n, models + (size_t)k * n, n, npg);
    p[k] = (T)1 / dot(y, s, vol);
    a[k] = p[k] * dot(s, q, vol);
    for (long i = 0; i < n; ++i) q[i] = q[i] - a[k] * y[i];
  }
  std::vector<T> r(n);
  {
    std::vector<T> y = grad_diff<T>(grads + (size_t)nstore * n, grads + (size_t)(nstore - 1) * n, n);
    std::vector<T> s = model_diff<T>(models + (size_t)nstore * n, models + (size_t)(nstore - 1) * n, n, npg);
    const T pk = dot(y, s, vol) / dot(y, y, vol);
    for (long i = 0; i < n; ++i) r[i] = pk * (T)hess[i] * q[i];
  }
  for (int k = 0; k < nstore; ++k) {
    std::vector<T> y = grad_diff<T>(grads + (size_t)(k + 1) * n, grads + (size_t)k * n, n);
    std::vector<T> s = model_diff<T>(models + (size_t)(k + 1) * n, models + (size_t)k * n, n, npg);
    const T b = p[k] * dot(y, r, vol);
    for (long i = 0; i < n; ++i) r[i] = r[i] + (a[k] - b) * s[i];
  }
  std::vector<T> d(n);
  for (long i = 0; i < n; ++i) d[i] = -r[i];
  const T m = maxabs(d);
  for (long i = 0; i < n; ++i) d[i] = d[i] / m;
  T gs = 0, gn = 0, dn = 0;
  for (long i = 0; i < n; ++i) {
    gs = gs + (-gcur[i]) * d[i];
    gn = gn + gcur[i] * gcur[i];
    dn = dn + d[i] * d[i];
  }
  *angle = (double)std::acos(gs / (std::sqrt(gn) * std::sqrt(dn))) * (180.0 / M_PI);
  for (long i = 0; i < n; ++i) dir[i] = (float)d[i];
}

}  // namespace

extern "C" {

// direction = -gradient * hess / max|gradient * hess|
void orc_sd_direction(long n, const float* grad, const float* hess, float* dir, int dbl) {
  if (dbl) {
    std::vector<double> d(n);
    for (long i = 0; i < n; ++i) d[i] = -(double)grad[i] * (double)hess[i];
    const double m = maxabs(d);
    for (long i = 0; i < n; ++i) dir[i] = (float)(d[i] / m);
  } else {
    std::vector<float> d(n);
    for (long i = 0; i < n; ++i) d[i] = -grad[i] * hess[i];
    const float m = maxabs(d);
    for (long i = 0; i < n; ++i) dir[i] = d[i] / m;
  }
}

// grads, models: (nx,ny,nz,nkernel, nstore+1) for iterations iter_store .. iter_current
void orc_lbfgs_direction(long n, long npg, int nstore, const float* grads, const float* models, const float* hess,
                         float dx, float dy, float dz, float* dir, double* angle, int dbl) {
  if (dbl) lbfgs<double>(n, npg, nstore, grads, models, hess, dx, dy, dz, dir, angle);
  else lbfgs<float>(n, npg, nstore, grads, models, hess, dx, dy, dz, dir, angle);
}

// Polak-Ribiere as written in the reference: alpha = sum_k <g1 - g0, g1>_k /
// sum_k <g0, g0>_k, direction = hess * (-g1 - alpha g0), normalised
void orc_cg_direction(long n, long npg, const float* g1, const float* g0, const float* hess, float* dir,
                      double* alpha_out, int dbl) {
  const int nk = (int)(n / npg);
  double n0 = 0, n1 = 0;
  float f0 = 0, f1 = 0;
  for (int k = 0; k < nk; ++k) {
    double a = 0, b = 0;
    float af = 0, bf = 0;
    for (long i = k * npg; i < (k + 1) * npg; ++i) {
      a += (double)g0[i] * g0[i];
      b += ((double)g1[i] - g0[i]) * g1[i];
      af = af + g0[i] * g0[i];
      bf = bf + (g1[i] - g0[i]) * g1[i];
    }
    n0 += a; n1 += b; f0 = f0 + af; f1 = f1 + bf;
  }
  if (dbl) {
    const double alpha = n1 / n0;
    std::vector<double> d(n);
    for (long i = 0; i < n; ++i) d[i] = (double)hess[i] * (-(double)g1[i] - alpha * g0[i]);
    const double m = maxabs(d);
    for (long i = 0; i < n; ++i) dir[i] = (float)(d[i] / m);
    *alpha_out = alpha;
  } else {
    const float alpha = f1 / f0;
    std::vector<float> d(n);
    for (long i = 0; i < n; ++i) d[i] = hess[i] * (-g1[i] - alpha * g0[i]);
    const float m = maxabs(d);
    for (long i = 0; i < n; ++i) dir[i] = d[i] / m;
    *alpha_out = alpha;
  }
}

// model_update: model_type 1 multiplies all parameters by exp(step d) and
// clamps vp/vs into [r0, r1] by moving vp; model_type 2 does that for
// parameters 1-3 when update_iso and adds step d to gcp, gsp when update_ani
void orc_model_update(long npg, int nk, int model_type, int update_iso, int update_ani, float step,
                      float r0, float r1, float* model, const float* dir) {
  const int niso = model_type == 1 ? nk : 3;
  if (model_type == 1 || update_iso) {
    for (long i = 0; i < npg * niso; ++i) model[i] = model[i] * std::exp(step * dir[i]);
    for (long i = 0; i < npg; ++i) {
      const float ratio = model[i] / model[npg + i];
      if (ratio < r0) model[i] = r0 * model[npg + i];
      if (ratio > r1) model[i] = r1 * model[npg + i];
    }
  }
  if (model_type == 2 && update_ani)
    for (long i = 3 * npg; i < 5 * npg; ++i) model[i] = model[i] + step * dir[i];
}

}  // extern "C"
