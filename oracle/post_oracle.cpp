// CPU restatement of SpecFWAT's kernel post-processing chain (xfwat_post_proc).
//
// TEST INFRASTRUCTURE ONLY.  This file is compiled into liboracle_sem.so next
// to sem_oracle.cpp and is used by tests/ and bench.py's cpu_baseline leg as
// the checker of the product's post-processing kernels; the product never
// links it.
//
// Functions and the reference code they restate:
//   orc_cijkl2hti        kernel_cijkl2hti, src/postproc/kernel_io.f90:100-156
//   orc_invert_hess      invert_hess, src/postproc/post_processing.f90:344-377
//   orc_gll2grid_coeff   compute_interpolation_coeff_FD_SEM,
//                        src/shared/projection_on_FD_grid_mod_fwat.f90:186-506,
//                        with locate_point_in_element (:935-1125) and SPECFEM's
//                        NGNOD = 8 recompute_jacobian / lagrange_any (absent from
//                        the tree: restated from the trilinear hexahedron map)
//   orc_gll2grid_project Project_model_SEM2FD_grid, same file :512-596
//   orc_taper_grid       taper_kernel_grid (post_processing.f90:296-333) with
//                        taper_cls Create / Interp / create_taper / coswin,
//                        src/postproc/taper_3D_mod.f90:16-141
//
// Precision follows the reference: single precision (CUSTOM_REAL = 4) where the
// Fortran declares real(cr)/real, double precision in the point location and the
// projection sum.  Built with -ffp-contract=off, so every expression rounds as
// written (left to right, as Fortran evaluates them).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" float orc_interp3_pre(const double* dx, const float* vx, int nx, const double* dy, const float* vy,
                                 int ny, const double* dz, const float* vz, int nz, const float* valin, float x,
                                 float y, float z);

namespace {

constexpr int NG = 5;

void gll_points(double* xi) {
  const double s = std::sqrt(3.0 / 7.0);
  xi[0] = -1.0; xi[1] = -s; xi[2] = 0.0; xi[3] = s; xi[4] = 1.0;
}

// lagrange_any (SPECFEM): h(dgr) = prod1 / prod2
void lagrange_any(double x, const double* xg, double* h) {
  for (int d = 0; d < NG; ++d) {
    double p1 = 1.0, p2 = 1.0;
    for (int i = 0; i < NG; ++i) {
      if (i == d) continue;
      p1 = p1 * (x - xg[i]);
      p2 = p2 * (xg[d] - xg[i]);
    }
    h[d] = p1 / p2;
  }
}

// NGNOD = 8 anchors in SPECFEM's usual_hex_nodes order: GLL index 0 or 4 per axis
const int AX[8] = {0, 4, 4, 0, 0, 4, 4, 0};
const int AY[8] = {0, 0, 4, 4, 0, 0, 4, 4};
const int AZ[8] = {0, 0, 0, 0, 4, 4, 4, 4};

// recompute_jacobian for the 8-node hexahedron: position and inverse Jacobian
void recompute_jacobian(const double* xe, const double* ye, const double* ze, double xi, double eta,
                        double gam, double* xyz, double* inv /* xix xiy xiz etax etay etaz gamx gamy gamz */) {
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
  for (int a = 0; a < 8; ++a) {
    x = x + sh[a] * xe[a]; y = y + sh[a] * ye[a]; z = z + sh[a] * ze[a];
    xxi = xxi + dxi[a] * xe[a]; xeta = xeta + deta[a] * xe[a]; xgam = xgam + dgam[a] * xe[a];
    yxi = yxi + dxi[a] * ye[a]; yeta = yeta + deta[a] * ye[a]; ygam = ygam + dgam[a] * ye[a];
    zxi = zxi + dxi[a] * ze[a]; zeta = zeta + deta[a] * ze[a]; zgam = zgam + dgam[a] * ze[a];
  }
  const double jac = xxi * (yeta * zgam - ygam * zeta) - xeta * (yxi * zgam - ygam * zxi) +
                     xgam * (yxi * zeta - yeta * zxi);
  xyz[0] = x; xyz[1] = y; xyz[2] = z;
  inv[0] = (yeta * zgam - ygam * zeta) / jac;
  inv[1] = (xgam * zeta - xeta * zgam) / jac;
  inv[2] = (xeta * ygam - xgam * yeta) / jac;
  inv[3] = (ygam * zxi - yxi * zgam) / jac;
  inv[4] = (xxi * zgam - xgam * zxi) / jac;
  inv[5] = (xgam * yxi - xxi * ygam) / jac;
  inv[6] = (yxi * zeta - yeta * zxi) / jac;
  inv[7] = (xeta * zxi - xxi * zeta) / jac;
  inv[8] = (xxi * yeta - xeta * yxi) / jac;
}

// locate_point_in_element: closest interior GLL point as the first guess, then
// num_iter Newton steps on the 8-node map, each clamped to the reference cube
void locate_in_element(const int* ib /* 125, 1-based */, const float* xs, const float* ys, const float* zs,
                       const double* xg, double xt, double yt, double zt, int num_iter, double* ref) {
  double dmin = 1e30;
  int gi = 0, gj = 0, gk = 0;
  for (int k = 1; k < NG - 1; ++k)
    for (int j = 1; j < NG - 1; ++j)
      for (int i = 1; i < NG - 1; ++i) {
        const int g = ib[i + NG * j + NG * NG * k] - 1;
        const double dx = xt - (double)xs[g], dy = yt - (double)ys[g], dz = zt - (double)zs[g];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < dmin) { dmin = d2; gi = i; gj = j; gk = k; }
      }
  double xi = xg[gi], eta = xg[gj], gam = xg[gk];
  double xe[8], ye[8], ze[8];
  for (int a = 0; a < 8; ++a) {
    const int g = ib[AX[a] + NG * AY[a] + NG * NG * AZ[a]] - 1;
    xe[a] = (double)xs[g]; ye[a] = (double)ys[g]; ze[a] = (double)zs[g];
  }
  double p[3], J[9];
  for (int it = 0; it < num_iter; ++it) {
    recompute_jacobian(xe, ye, ze, xi, eta, gam, p, J);
    const double dx = -(p[0] - xt), dy = -(p[1] - yt), dz = -(p[2] - zt);
    xi = xi + (J[0] * dx + J[1] * dy + J[2] * dz);
    eta = eta + (J[3] * dx + J[4] * dy + J[5] * dz);
    gam = gam + (J[6] * dx + J[7] * dy + J[8] * dz);
    xi = std::min(std::max(xi, -1.0), 1.0);
    eta = std::min(std::max(eta, -1.0), 1.0);
    gam = std::min(std::max(gam, -1.0), 1.0);
  }
  ref[0] = xi; ref[1] = eta; ref[2] = gam;
}

}  // namespace

extern "C" {

// kernel_cijkl2hti for one event (single precision, Fortran evaluation order).
//   model  : c44store, c45store, c55store, rho_vp, rho_vs, rhostore, each [n]
//   k      : [11][n] = rho, c11, c12, c13, c22, c23, c33, c44, c45, c55, c66 kernels
//   ker    : [5][n]  = alpha, beta, rho, gcp, gsp (alpha/beta/rho relative)
void orc_cijkl2hti(long n, const float* c44s, const float* c45s, const float* c55s, const float* rvp,
                   const float* rvs, const float* rho, const float* k, float* ker) {
  const float* rk = k;
  const float* K11 = k + 1 * n; const float* K12 = k + 2 * n; const float* K13 = k + 3 * n;
  const float* K22 = k + 4 * n; const float* K23 = k + 5 * n; const float* K33 = k + 6 * n;
  const float* K44 = k + 7 * n; const float* K45 = k + 8 * n; const float* K55 = k + 9 * n;
  const float* K66 = k + 10 * n;
  for (long i = 0; i < n; ++i) {
    const float L = 0.5f * (c44s[i] + c55s[i]);
    const float gc = 0.5f * (c44s[i] - c55s[i]);
    const float gs = c45s[i];
    const float gcp = gc / L, gsp = gs / L;
    const float vp = rvp[i] / rho[i], vs = rvs[i] / rho[i];
    const float a2 = 2.0f * rvp[i], b2 = 2.0f * rvs[i], b4 = -(4.0f * rvs[i]);
    float a = K11[i] * a2 + K12[i] * a2 + K13[i] * a2 + K22[i] * a2 + K23[i] * a2 + K33[i] * a2;
    float b = K12[i] * b4 + K13[i] * b4 + K23[i] * b4 + K44[i] * (2.0f * gcp * rvs[i] + b2) +
              K45[i] * (2.0f * gsp * rvs[i]) + K55[i] * (-(2.0f * gcp * rvs[i]) + b2) + K66[i] * b2;
    const float vp2 = vp * vp, vs2 = vs * vs, d = vp2 - 2.0f * vs2;
    float r = rk[i] + K11[i] * vp2 + K12[i] * d + K13[i] * d + K22[i] * vp2 + K23[i] * d + K33[i] * vp2 +
              K44[i] * (gcp * vs2 + vs2) + K45[i] * (gsp * vs2) + K55[i] * (-(gcp * vs2) + vs2) + K66[i] * vs2;
    const float g1 = (K44[i] - K55[i]) * rvs[i] * vs;
    const float g2 = K45[i] * rvs[i] * vs;
    ker[i] = a * vp;
    ker[n + i] = b * vs;
    ker[2 * n + i] = r * rho[i];
    ker[3 * n + i] = g1;
    ker[4 * n + i] = g2;
  }
}

// invert_hess (single precision): normalise by max|H|, invert above the 1e-3
// threshold, normalise again
void orc_invert_hess(long n, float* h) {
  float m = 0.0f;
  for (long i = 0; i < n; ++i) m = std::max(m, std::fabs(h[i]));
  if (m < 1.e-18f) {
    for (long i = 0; i < n; ++i) h[i] = 1.0f;
  } else {
    for (long i = 0; i < n; ++i) h[i] = h[i] / m;
  }
  const float thr = 1.e-3f;
  for (long i = 0; i < n; ++i) h[i] = std::fabs(h[i]) > thr ? 1.0f / h[i] : 1.0f / thr;
  m = 0.0f;
  for (long i = 0; i < n; ++i) m = std::max(m, std::fabs(h[i]));
  for (long i = 0; i < n; ++i) h[i] = h[i] / m;
}

// compute_interpolation_coeff_FD_SEM.  Grid point (i, j, k) (0-based) sits at
// (ox + i hx, oy + j hy, oz + k hz).  owner[p] = 1-based element that claimed
// point p first in element order (0 = no element), ref[3 p ..] its reference
// coordinates.  Element index ranges outside the grid are skipped (the
// reference requires the grid to cover the mesh).  Returns nb_fd_point.
long orc_gll2grid_coeff(int nspec, const int* ibool, const float* xs, const float* ys, const float* zs,
                        float ox, float oy, float oz, float hx, float hy, float hz, int nx, int ny, int nz,
                        int num_iter, int* owner, double* ref) {
  double xg[NG];
  gll_points(xg);
  const size_t npt = (size_t)nx * ny * nz;
  std::fill(owner, owner + npt, 0);
  long found = 0;
  for (int e = 0; e < nspec; ++e) {
    const int* ib = ibool + (size_t)e * 125;
    float xmin = 1e30f, xmax = -1e30f, ymin = 1e30f, ymax = -1e30f, zmin = 1e30f, zmax = -1e30f;
    for (int L = 0; L < 125; ++L) {
      const int g = ib[L] - 1;
      xmin = std::min(xmin, xs[g]); xmax = std::max(xmax, xs[g]);
      ymin = std::min(ymin, ys[g]); ymax = std::max(ymax, ys[g]);
      zmin = std::min(zmin, zs[g]); zmax = std::max(zmax, zs[g]);
    }
    // 1-based Fortran integer conversion (truncation); clamped to [-1, n + 1]
    // first, which changes nothing since indices outside 1..n are skipped
    auto idx = [](float t, int n) { return (int)std::min(std::max(t, -1.0f), (float)(n + 1)); };
    const int kmin = idx(1.0f + (zmin - oz) / hz, nz), kmax = idx(1.0f + (zmax - oz) / hz, nz);
    const int jmin = idx(1.0f + (ymin - oy) / hy, ny), jmax = idx(1.0f + (ymax - oy) / hy, ny);
    const int imin = idx(1.0f + (xmin - ox) / hx, nx), imax = idx(1.0f + (xmax - ox) / hx, nx);
    for (int kf = kmin; kf <= kmax; ++kf) {
      const float zf = oz + (float)(kf - 1) * hz;
      for (int jf = jmin; jf <= jmax; ++jf) {
        const float yf = oy + (float)(jf - 1) * hy;
        for (int i_f = imin; i_f <= imax; ++i_f) {
          const float xf = ox + (float)(i_f - 1) * hx;
          if (i_f < 1 || i_f > nx || jf < 1 || jf > ny || kf < 1 || kf > nz) continue;
          const size_t p = (size_t)(i_f - 1) + (size_t)nx * ((size_t)(jf - 1) + (size_t)ny * (kf - 1));
          if (owner[p] != 0) continue;
          double r[3];
          locate_in_element(ib, xs, ys, zs, xg, (double)xf, (double)yf, (double)zf, num_iter, r);
          if (std::fabs(r[0]) < 1.05 && std::fabs(r[1]) < 1.05 && std::fabs(r[2]) < 1.05) {
            owner[p] = e + 1;
            ref[3 * p] = r[0]; ref[3 * p + 1] = r[1]; ref[3 * p + 2] = r[2];
            ++found;
          }
        }
      }
    }
  }
  return found;
}

// Project_model_SEM2FD_grid on one rank: value = sum over the owner element's
// 125 points of hxi(i) heta(j) hgamma(k) model (double), stored single; points
// no element claimed stay 0.  model is (NGLLX,NGLLY,NGLLZ,NSPEC), grid (nx,ny,nz).
void orc_gll2grid_project(long npt, const int* owner, const double* ref, const float* model, float* grid) {
  double xg[NG];
  gll_points(xg);
  for (long p = 0; p < npt; ++p) {
    if (owner[p] == 0) { grid[p] = 0.0f; continue; }
    double hx[NG], hy[NG], hz[NG];
    lagrange_any(ref[3 * p], xg, hx);
    lagrange_any(ref[3 * p + 1], xg, hy);
    lagrange_any(ref[3 * p + 2], xg, hz);
    const float* m = model + (size_t)(owner[p] - 1) * 125;
    double v = 0.0;
    for (int k = 0; k < NG; ++k)
      for (int j = 0; j < NG; ++j)
        for (int i = 0; i < NG; ++i) v = v + hx[i] * hy[j] * hz[k] * (double)m[i + NG * j + NG * NG * k];
    grid[p] = (float)v;
  }
}

// coswin: first (opt 1) or last (opt 0) n samples of a 2n+1 Hann window
static void coswin(int n, int opt, float* out) {
  const float pi = 3.1415926535f;
  const int nn = 2 * n + 1;
  std::vector<float> w(nn);
  for (int i = 1; i <= nn; ++i) w[i - 1] = 0.5f - 0.5f * std::cos(2.0f * pi * (float)(i - 1) / (float)(nn - 1));
  for (int i = 0; i < n; ++i) out[i] = opt == 1 ? w[i] : w[n + 1 + i];
}

// taper_kernel_grid's mask: taper_cls%Create over the mesh box with spacing dh,
// then taper_cls%Interp at every grid point (i, j, k) of the output grid.
// Returns 0, or -1 when the taper ranges do not fit the taper grid.
int orc_taper_grid(float xmin, float xmax, float ymin, float ymax, float zmin, float zmax, float dh,
                   float h_suppress, float h_buffer, float v_suppress, float v_buffer, const float* gx, int gnx,
                   const float* gy, int gny, const float* gz, int gnz, float* mask) {
  const int nx = (int)std::floor((xmax - xmin) / dh + 1.0f);
  const int ny = (int)std::floor((ymax - ymin) / dh + 1.0f);
  const int nz = (int)std::floor((zmax - zmin) / dh + 1.0f);
  std::vector<float> xc(nx), yc(ny), zc(nz), xt(nx, 0.0f), yt(ny, 0.0f), zt(nz, 0.0f);
  for (int i = 0; i < nx; ++i) xc[i] = xmin + (float)i * dh;
  for (int i = 0; i < ny; ++i) yc[i] = ymin + (float)i * dh;
  for (int i = 0; i < nz; ++i) zc[i] = zmin + (float)i * dh;
  // lateral windows (1-based positions as in create_taper)
  auto lateral = [&](std::vector<float>& t, int n) -> int {
    const int n1 = (int)std::floor(h_suppress / dh), n2 = (int)std::floor(h_buffer / dh);
    const int x1 = n1 + 1, x2 = n1 + n2 + 1, x3 = n - n1 - n2, x4 = n - n1;
    if (x1 < 1 || x3 < 0 || x4 > n || x2 - 1 > n || x3 + 1 < 1) return -1;
    coswin(n2, 1, t.data() + (x1 - 1));
    coswin(n2, 0, t.data() + x3);
    for (int i = x2; i <= x3; ++i) t[i - 1] = 1.0f;
    return 0;
  };
  if (lateral(xt, nx) || lateral(yt, ny)) return -1;
  {
    const int n1 = (int)std::floor(v_suppress / dh), n2 = (int)std::floor(v_buffer / dh);
    const int z1 = n1 + n2 + 1;
    if (z1 - 1 > nz) return -1;
    coswin(n2, 1, zt.data() + n1);
    for (int i = z1; i <= nz; ++i) zt[i - 1] = 1.0f;
  }
  std::vector<float> tap((size_t)nx * ny * nz);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) tap[(size_t)i + (size_t)nx * (j + (size_t)ny * k)] = xt[i] * yt[j] * zt[k];
  const std::vector<double> dxc(xc.begin(), xc.end()), dyc(yc.begin(), yc.end()), dzc(zc.begin(), zc.end());
  for (int k = 0; k < gnz; ++k)
    for (int j = 0; j < gny; ++j)
      for (int i = 0; i < gnx; ++i) {
        const float x = gx[i], y = gy[j], z = gz[k];
        float v = 0.0f;
        if (!(x < xc[0] || x > xc[nx - 1] || y < yc[0] || y > yc[ny - 1] || z < zc[0] || z > zc[nz - 1]))
          v = orc_interp3_pre(dxc.data(), xc.data(), nx, dyc.data(), yc.data(), ny, dzc.data(), zc.data(), nz,
                              tap.data(), x, y, z);
        mask[(size_t)i + (size_t)gnx * (j + (size_t)gny * k)] = v;
      }
  return 0;
}

}  // extern "C"
