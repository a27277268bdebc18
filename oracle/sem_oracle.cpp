// CPU oracle for the SpecFWAT spectral-element forward/adjoint/kernel path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports, links or
// runs this file; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it, always as the checker or as the timed CPU baseline.
//
// What it restates (file:line are in the SpecFWAT tree):
//  * SpecFWAT's own in-tree leaves, loop for loop:
//      - smooth_sem_pde CPU path            src/postproc/smooth.f90:7-437, 603-689
//      - interp3_0_cr / locate_bissection   src/shared/utils.f90:1162-1232, 1490-1520
//      - interp3_nearest_simple             src/shared/utils.f90:1276-1325
//      - HTI -> c_ij (init_iso, hti2aniso, to_cijkl, glob2cart)
//                                            src/patch/aniso.f90:55-162
//      - kappa/mu/rho_vp/rho_vs from rho,vp,vs
//                                            src/mesh_database_api/get_model_fwat.f90:305-317
//      - waveform misfit + window taper + simpson
//                                            src/preproc/adjoint_sources/waveform_misfit.f90:15-64,
//                                            src/shared/signal.f90:419-508, src/shared/utils.f90:1630-1663
//  * the SPECFEM3D elastic solver that SpecFWAT drives through iterate_time
//    (src/preproc/prepare_fwd.f90:328-368).  Its source is an un-vendored
//    submodule (github.com/xumi1993/specfem3d, branch devel, "Specfem3D 4.1.1"
//    per include/config.h:98), absent from the tree, so this is a restatement of
//    the published SPECFEM3D Cartesian algorithm: GLL weak form with
//    hprime/hprimewgll contractions, Newmark central difference, Stacey with the
//    dt/2 mass-matrix term, FK injection of an incident velocity and traction
//    through the Stacey faces (the Veloc_FK / Tract_FK of
//    src/shared/fk_coupling.f90:71-297; the closed-form plane wave of
//    tests/fk_plane_wave.py pins it), moment-tensor point source with a Ricker STF,
//    Lagrange-interpolated receivers, backward reconstruction from the last
//    state plus replayed Stacey terms, and rho/kappa/mu (Tromp-Tape-Liu 2005)
//    or 21 c_ij kernels.  Parity against SPECFEM itself is unpinned; this file
//    is pinned by closed forms and invariants in tests/test_oracle_*.py.
//
// Layout conventions (identical to the product's C-ABI):
//   local point index  L = i + 5 j + 25 k, element-major: [ispec][L]
//   wavefields         [nglob][3]
//   seismograms        [nstep][nrec][3]
//   cijkl kernels      [ispec][L][21], upper-triangle Voigt order 11,12,..,16,22,..,66
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>
#include <omp.h>

namespace {

constexpr int NG = 5, NG2 = 25, NG3 = 125;

// ------------------------------------------------------------------ GLL basis
struct Gll {
  double xi[NG], w[NG];
  double hp[NG][NG];   // hp[i][j]  = l_j'(xi_i)        (SPECFEM hprime_xx(i,j))
  double hpw[NG][NG];  // hpw[i][j] = w_i * l_j'(xi_i)  (SPECFEM hprimewgll_xx(i,j))
};

double legendre(int n, double x) {
  double p0 = 1.0, p1 = x;
  if (n == 0) return p0;
  for (int k = 2; k <= n; ++k) {
    double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

Gll make_gll() {
  Gll g;
  const double s = std::sqrt(3.0 / 7.0);
  const double xi[NG] = {-1.0, -s, 0.0, s, 1.0};
  for (int i = 0; i < NG; ++i) {
    g.xi[i] = xi[i];
    double p = legendre(NG - 1, xi[i]);
    g.w[i] = 2.0 / ((NG - 1) * NG * p * p);
  }
  double c[NG];
  for (int i = 0; i < NG; ++i) {
    c[i] = 1.0;
    for (int k = 0; k < NG; ++k)
      if (k != i) c[i] *= (xi[i] - xi[k]);
  }
  for (int i = 0; i < NG; ++i)
    for (int j = 0; j < NG; ++j) {
      double d;
      if (i != j) {
        d = c[i] / (c[j] * (xi[i] - xi[j]));
      } else {
        d = 0.0;
        for (int k = 0; k < NG; ++k)
          if (k != i) d += 1.0 / (xi[i] - xi[k]);
      }
      g.hp[i][j] = d;
      g.hpw[i][j] = g.w[i] * d;
    }
  return g;
}

// Lagrange interpolants and derivatives at an arbitrary reference coordinate.
void lagrange(const Gll& g, double x, double* h, double* hd) {
  for (int j = 0; j < NG; ++j) {
    double p = 1.0;
    for (int k = 0; k < NG; ++k)
      if (k != j) p *= (x - g.xi[k]) / (g.xi[j] - g.xi[k]);
    h[j] = p;
    double s = 0.0;
    for (int m = 0; m < NG; ++m) {
      if (m == j) continue;
      double q = 1.0 / (g.xi[j] - g.xi[m]);
      for (int k = 0; k < NG; ++k)
        if (k != j && k != m) q *= (x - g.xi[k]) / (g.xi[j] - g.xi[k]);
      s += q;
    }
    hd[j] = s;
  }
}

// ------------------------------------------------------------------ mesh
struct Mesh {
  int nex = 0, ney = 0, nez = 0, nspec = 0, nglob = 0;
  double lx = 0, ly = 0, lz = 0;
  std::vector<double> anchor;           // [(nex+1)(ney+1)(nez+1)][3]
  std::vector<int> ibool;               // 0-based [nspec*125]
  std::vector<double> gx, gy, gz;       // [nglob]
  std::vector<double> jinv, jac;        // [nspec*125][9], [nspec*125]
  std::vector<int> irreg;               // [nspec], 0 = regular, else 1-based
  int nspec_irreg = 0;
  double xix_reg = 0, jac_reg = 0;
  int nfaces = 0;
  std::vector<int> f_ispec, f_ijk;      // [nfaces], [nfaces][25][3] 0-based
  std::vector<double> f_n, f_jw;        // [nfaces][25][3], [nfaces][25]
  std::vector<double> bbox;             // [nspec][6] element bounding boxes
  std::vector<int> color;               // greedy first-fit colour per element
  std::vector<std::vector<int>> color_sets;
};

// Deterministic interior-node perturbation used for the general-hexahedron
// (irregular) variant; boundary nodes stay put so the box faces stay planar.
void anchor_offset(int ix, int iy, int iz, double hx, double hy, double hz, double amp,
                   double* d) {
  d[0] = amp * hx * std::sin(1.3 * ix + 2.1 * iy + 0.7 * iz + 0.4);
  d[1] = amp * hy * std::sin(0.9 * ix - 1.7 * iy + 2.3 * iz + 1.1);
  d[2] = amp * hz * std::sin(-1.1 * ix + 0.6 * iy + 1.9 * iz + 2.2);
}

void trilinear(const double* X /*8x3*/, double xi, double eta, double gam, double* x,
               double J[3][3]) {
  for (int c = 0; c < 3; ++c) {
    x[c] = 0;
    J[c][0] = J[c][1] = J[c][2] = 0;
  }
  for (int n = 0; n < 8; ++n) {
    double sa = (n & 1) ? 1.0 : -1.0, sb = (n & 2) ? 1.0 : -1.0, sd = (n & 4) ? 1.0 : -1.0;
    double fa = 1 + sa * xi, fb = 1 + sb * eta, fd = 1 + sd * gam;
    double N = fa * fb * fd / 8.0;
    double dNa = sa * fb * fd / 8.0, dNb = fa * sb * fd / 8.0, dNd = fa * fb * sd / 8.0;
    for (int c = 0; c < 3; ++c) {
      x[c] += N * X[3 * n + c];
      J[c][0] += dNa * X[3 * n + c];
      J[c][1] += dNb * X[3 * n + c];
      J[c][2] += dNd * X[3 * n + c];
    }
  }
}

double inv3(const double J[3][3], double Ji[3][3]) {
  double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
               J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  double id = 1.0 / det;
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  return det;
}

void element_anchors(const Mesh& m, int ispec, double* X) {
  int ix = ispec % m.nex, iy = (ispec / m.nex) % m.ney, iz = ispec / (m.nex * m.ney);
  for (int n = 0; n < 8; ++n) {
    int a = ix + (n & 1), b = iy + ((n >> 1) & 1), d = iz + ((n >> 2) & 1);
    size_t id = (size_t)a + (size_t)(m.nex + 1) * (b + (size_t)(m.ney + 1) * d);
    for (int c = 0; c < 3; ++c) X[3 * n + c] = m.anchor[3 * id + c];
  }
}

struct KeyHash {
  size_t operator()(const std::array<long long, 3>& k) const {
    uint64_t h = (uint64_t)k[0] * 0x9E3779B97F4A7C15ULL;
    h ^= (uint64_t)k[1] * 0xC2B2AE3D27D4EB4FULL + (h << 6) + (h >> 2);
    h ^= (uint64_t)k[2] * 0x165667B19E3779F9ULL + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};

// numbering: 0 = coordinate de-duplication (independent of the lattice),
//            1 = lattice key (fast path for large meshes; tests show 0 == 1).
// Both number unique points in order of first appearance while sweeping
// elements, then k, j, i (SPECFEM's get_global followed by its
// cache-friendly indirect renumbering).
void build_mesh(Mesh& m, const Gll& g, int nex, int ney, int nez, double lx, double ly,
                double lz, double perturb, int force_irregular, int numbering) {
  m.nex = nex; m.ney = ney; m.nez = nez;
  m.nspec = nex * ney * nez;
  m.lx = lx; m.ly = ly; m.lz = lz;
  double hx = lx / nex, hy = ly / ney, hz = lz / nez;
  m.anchor.resize((size_t)(nex + 1) * (ney + 1) * (nez + 1) * 3);
  for (int iz = 0; iz <= nez; ++iz)
    for (int iy = 0; iy <= ney; ++iy)
      for (int ix = 0; ix <= nex; ++ix) {
        size_t id = (size_t)ix + (size_t)(nex + 1) * (iy + (size_t)(ney + 1) * iz);
        double p[3] = {ix * hx, iy * hy, -lz + iz * hz};
        if (perturb != 0.0 && ix > 0 && ix < nex && iy > 0 && iy < ney && iz > 0 && iz < nez) {
          double d[3];
          anchor_offset(ix, iy, iz, hx, hy, hz, perturb, d);
          for (int c = 0; c < 3; ++c) p[c] += d[c];
        }
        for (int c = 0; c < 3; ++c) m.anchor[3 * id + c] = p[c];
      }
  const size_t nloc = (size_t)m.nspec * NG3;
  m.ibool.assign(nloc, -1);
  m.jinv.resize(nloc * 9);
  m.jac.resize(nloc);
  std::vector<double> lx_(nloc), ly_(nloc), lz_(nloc);
  for (int e = 0; e < m.nspec; ++e) {
    double X[24];
    element_anchors(m, e, X);
    for (int k = 0; k < NG; ++k)
      for (int j = 0; j < NG; ++j)
        for (int i = 0; i < NG; ++i) {
          size_t L = (size_t)e * NG3 + i + NG * j + NG2 * k;
          double x[3], J[3][3], Ji[3][3];
          trilinear(X, g.xi[i], g.xi[j], g.xi[k], x, J);
          double det = inv3(J, Ji);
          lx_[L] = x[0]; ly_[L] = x[1]; lz_[L] = x[2];
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) m.jinv[9 * L + 3 * r + c] = Ji[r][c];
          m.jac[L] = det;
        }
  }
  // global numbering
  int next = 0;
  if (numbering == 1) {
    const long long NX = 4LL * nex + 1, NY = 4LL * ney + 1, NZ = 4LL * nez + 1;
    std::vector<int> map((size_t)(NX * NY * NZ), -1);
    for (int e = 0; e < m.nspec; ++e) {
      int ex = e % nex, ey = (e / nex) % ney, ez = e / (nex * ney);
      for (int k = 0; k < NG; ++k)
        for (int j = 0; j < NG; ++j)
          for (int i = 0; i < NG; ++i) {
            long long key = (4LL * ex + i) + NX * ((4LL * ey + j) + NY * (4LL * ez + k));
            int& id = map[(size_t)key];
            if (id < 0) id = next++;
            m.ibool[(size_t)e * NG3 + i + NG * j + NG2 * k] = id;
          }
    }
  } else {
    double tol = 1e-6 * std::min(hx, std::min(hy, hz));
    std::unordered_map<std::array<long long, 3>, int, KeyHash> map;
    map.reserve(nloc / 2);
    for (size_t L = 0; L < nloc; ++L) {
      double f[3] = {lx_[L] / tol, ly_[L] / tol, lz_[L] / tol};
      std::array<long long, 3> q = {std::llround(f[0]), std::llround(f[1]), std::llround(f[2])};
      auto it = map.find(q);
      int found = -1;
      if (it != map.end()) {
        found = it->second;
      } else {
        bool near = false;
        for (int c = 0; c < 3; ++c)
          if (std::fabs(std::fabs(f[c] - (double)q[c]) - 0.5) < 1e-3) near = true;
        if (near) {
          for (int dx = -1; dx <= 1 && found < 0; ++dx)
            for (int dy = -1; dy <= 1 && found < 0; ++dy)
              for (int dz = -1; dz <= 1 && found < 0; ++dz) {
                auto it2 = map.find({q[0] + dx, q[1] + dy, q[2] + dz});
                if (it2 != map.end()) found = it2->second;
              }
        }
      }
      if (found < 0) {
        found = next++;
        map.emplace(q, found);
      }
      m.ibool[L] = found;
    }
  }
  m.nglob = next;
  m.gx.resize(next); m.gy.resize(next); m.gz.resize(next);
  for (size_t L = 0; L < nloc; ++L) {
    int gI = m.ibool[L];
    m.gx[gI] = lx_[L]; m.gy[gI] = ly_[L]; m.gz[gI] = lz_[L];
  }
  // regular / irregular classification: a regular element is an axis-aligned
  // cube whose edge equals the first such cube's edge; it stores no per-point
  // Jacobian (only the mesh-wide xix_regular / jacobian_regular scalars).
  m.irreg.assign(m.nspec, 0);
  m.nspec_irreg = 0;
  m.xix_reg = 0; m.jac_reg = 0;
  double href = -1.0;
  for (int e = 0; e < m.nspec; ++e) {
    bool cube = !force_irregular;
    double hloc = 0;
    if (cube) {
      double X[24];
      element_anchors(m, e, X);
      double ex_ = X[3 * 1 + 0] - X[0], ey_ = X[3 * 2 + 1] - X[1], ez_ = X[3 * 4 + 2] - X[2];
      hloc = ex_;
      double tolh = 1e-9 * ex_;
      for (int n = 0; n < 8 && cube; ++n) {
        double want[3] = {X[0] + ((n & 1) ? ex_ : 0), X[1] + ((n & 2) ? ey_ : 0),
                          X[2] + ((n & 4) ? ez_ : 0)};
        for (int c = 0; c < 3; ++c)
          if (std::fabs(X[3 * n + c] - want[c]) > tolh) cube = false;
      }
      if (std::fabs(ey_ - ex_) > tolh || std::fabs(ez_ - ex_) > tolh) cube = false;
      if (cube && href < 0) href = hloc;
      if (cube && std::fabs(hloc - href) > 1e-9 * href) cube = false;
    }
    if (!cube) m.irreg[e] = ++m.nspec_irreg;
  }
  if (href > 0) {
    m.xix_reg = 2.0 / href;
    m.jac_reg = href * href * href / 8.0;
  }
  // Stacey faces: xmin, xmax, ymin, ymax, bottom; elements ascending; face
  // points with the second tangential index outer.
  m.nfaces = 0;
  m.f_ispec.clear(); m.f_ijk.clear(); m.f_n.clear(); m.f_jw.clear();
  for (int side = 0; side < 5; ++side) {
    for (int e = 0; e < m.nspec; ++e) {
      int ix = e % nex, iy = (e / nex) % ney, iz = e / (nex * ney);
      bool on = (side == 0 && ix == 0) || (side == 1 && ix == nex - 1) ||
                (side == 2 && iy == 0) || (side == 3 && iy == ney - 1) ||
                (side == 4 && iz == 0);
      if (!on) continue;
      double X[24];
      element_anchors(m, e, X);
      m.f_ispec.push_back(e);
      for (int b = 0; b < NG; ++b)
        for (int a = 0; a < NG; ++a) {
          int i, j, k;
          if (side <= 1) { i = side == 0 ? 0 : 4; j = a; k = b; }
          else if (side <= 3) { i = a; j = side == 2 ? 0 : 4; k = b; }
          else { i = a; j = b; k = 0; }
          m.f_ijk.push_back(i); m.f_ijk.push_back(j); m.f_ijk.push_back(k);
          double x[3], J[3][3];
          trilinear(X, g.xi[i], g.xi[j], g.xi[k], x, J);
          int ta = side <= 1 ? 1 : 0, tb = side <= 3 ? 2 : 1;
          if (side >= 2 && side <= 3) { ta = 0; tb = 2; }
          double t1[3] = {J[0][ta], J[1][ta], J[2][ta]}, t2[3] = {J[0][tb], J[1][tb], J[2][tb]};
          double n[3] = {t1[1] * t2[2] - t1[2] * t2[1], t1[2] * t2[0] - t1[0] * t2[2],
                         t1[0] * t2[1] - t1[1] * t2[0]};
          double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
          double out[3] = {0, 0, 0};
          if (side == 0) out[0] = -1; else if (side == 1) out[0] = 1;
          else if (side == 2) out[1] = -1; else if (side == 3) out[1] = 1; else out[2] = -1;
          double sgn = (n[0] * out[0] + n[1] * out[1] + n[2] * out[2]) >= 0 ? 1.0 : -1.0;
          for (int c = 0; c < 3; ++c) m.f_n.push_back(sgn * n[c] / nn);
          double wa = g.w[side <= 1 ? j : i], wb = g.w[side <= 3 ? k : j];
          m.f_jw.push_back(nn * wa * wb);
        }
      ++m.nfaces;
    }
  }
  m.bbox.resize((size_t)m.nspec * 6);
  for (int e = 0; e < m.nspec; ++e) {
    double X[24];
    element_anchors(m, e, X);
    for (int c = 0; c < 3; ++c) { m.bbox[6 * e + c] = 1e300; m.bbox[6 * e + 3 + c] = -1e300; }
    for (int n = 0; n < 8; ++n)
      for (int c = 0; c < 3; ++c) {
        m.bbox[6 * e + c] = std::min(m.bbox[6 * e + c], X[3 * n + c]);
        m.bbox[6 * e + 3 + c] = std::max(m.bbox[6 * e + 3 + c], X[3 * n + c]);
      }
  }
  // greedy first-fit colouring in element order (used for race-free OpenMP
  // assembly here; the colour of element e is the lowest colour not used by an
  // earlier element sharing a point with it)
  m.color.assign(m.nspec, 0);
  {
    std::vector<uint32_t> used(m.nglob, 0u);
    int ncol = 0;
    for (int e = 0; e < m.nspec; ++e) {
      uint32_t mask = 0;
      for (int L = 0; L < NG3; ++L) mask |= used[m.ibool[(size_t)e * NG3 + L]];
      int c = 0;
      while (mask & (1u << c)) ++c;
      m.color[e] = c;
      ncol = std::max(ncol, c + 1);
      for (int L = 0; L < NG3; ++L) used[m.ibool[(size_t)e * NG3 + L]] |= (1u << c);
    }
    m.color_sets.assign(ncol, {});
    for (int e = 0; e < m.nspec; ++e) m.color_sets[m.color[e]].push_back(e);
  }
}

// Locate a point: the first element (ascending ispec) whose reference
// coordinates lie in [-1,1]^3 (Newton on the trilinear map).
bool locate_point(const Mesh& m, double x, double y, double z, int* ispec, double* ref) {
  const std::vector<double>& bbox = m.bbox;
  for (int e = 0; e < m.nspec; ++e) {
    double X[24];
    const double* lo = &bbox[6 * e];
    const double* hi = &bbox[6 * e + 3];
    double p[3] = {x, y, z};
    bool inbox = true;
    for (int c = 0; c < 3; ++c) {
      double t = 1e-9 * (hi[c] - lo[c]);
      if (p[c] < lo[c] - t || p[c] > hi[c] + t) inbox = false;
    }
    if (!inbox) continue;
    element_anchors(m, e, X);
    double r[3] = {0, 0, 0};
    for (int it = 0; it < 50; ++it) {
      double xx[3], J[3][3], Ji[3][3];
      trilinear(X, r[0], r[1], r[2], xx, J);
      inv3(J, Ji);
      double d[3] = {p[0] - xx[0], p[1] - xx[1], p[2] - xx[2]};
      double dr[3];
      for (int a = 0; a < 3; ++a) dr[a] = Ji[a][0] * d[0] + Ji[a][1] * d[1] + Ji[a][2] * d[2];
      for (int a = 0; a < 3; ++a) r[a] += dr[a];
      if (std::fabs(dr[0]) + std::fabs(dr[1]) + std::fabs(dr[2]) < 1e-15) break;
    }
    if (std::fabs(r[0]) <= 1 + 1e-9 && std::fabs(r[1]) <= 1 + 1e-9 && std::fabs(r[2]) <= 1 + 1e-9) {
      for (int a = 0; a < 3; ++a) ref[a] = std::max(-1.0, std::min(1.0, r[a]));
      *ispec = e;
      return true;
    }
  }
  return false;
}

double ricker(double t, double f0) {
  double a = M_PI * M_PI * f0 * f0;
  return (1.0 - 2.0 * a * t * t) * std::exp(-a * t * t);
}

}  // namespace

// ====================================================================== C API
extern "C" {

struct OrcParams {
  int nex, ney, nez;
  double lx, ly, lz;
  double perturb;
  int force_irregular;
  int numbering;        // 0 coordinate de-dup, 1 lattice key
  int aniso;            // 0: rho/vp/vs isotropic; 1: 21 c_ij per local point
  const float* rho;     // [nspec*125]
  const float* vp;
  const float* vs;
  const float* cij;     // [nspec*125][21] when aniso
  double dt;
  int nstep;
  int stacey;
  int nsrc;
  const double* src_xyz;    // [nsrc][3]
  const double* src_mt;     // [nsrc][6]  Mxx Myy Mzz Mxy Mxz Myz
  const double* src_f0;     // [nsrc]  Ricker dominant frequency
  const double* src_tshift; // [nsrc]
  int nrec;
  const double* rec_xyz;    // [nrec][3]
  int nthreads;
  // visiting order of the element loop that assembles the internal forces
  // (0-based, a permutation of the elements), or null: colour by colour
  const int* elem_order;
};

}  // extern "C"

namespace {
// Planted faults, set only by the tests that show a check fails on a
// plausible mistake: 1 = Stacey impedances exchanged (rho*vs on the normal
// term, rho*vp on the tangential term, in the traction and in the mass
// matrix), 2 = Mxz and Myz exchanged when a moment tensor is unpacked.
int g_plant = 0;


struct SimBase {
  virtual ~SimBase() {}
  Gll g;
  Mesh m;
  OrcParams p;
  int prec = 32;
  // source/receiver location
  std::vector<int> src_ispec, rec_ispec;
  std::vector<double> src_ref, rec_ref;
  std::vector<double> srcarr;  // [nsrc][125][3]
  std::vector<double> rech;    // [nrec][125]
  virtual void forward(float* seis, int save) = 0;
  virtual void adjoint(const float* adj, double* rho_kl, double* mu_kl, double* kappa_kl,
                       double* cijkl_kl, float* b_final) = 0;
  virtual void get_rmass(float* rx, float* ry, float* rz) = 0;
  virtual void get_final(float* u, float* v, float* a) = 0;
  virtual void set_final(const float* u, const float* v, const float* a) = 0;
  virtual void get_stacey(float* buf) = 0;
  virtual void energy_series(double* e) = 0;
  // FK injection of an incident field through the Stacey faces (forward runs):
  // velocity and traction [fk_nsamp][nfaces*25][3], sample k at step 1 + k*fk_resamp
  std::vector<float> fk_v, fk_t;
  int fk_nsamp = 0, fk_resamp = 1;
  // per-point assembly order of the element terms (contribution e*125 + L) when
  // the forces are assembled in an element visiting order (OrcParams::elem_order)
  std::vector<int> asm_off, asm_idx;
  virtual void apply_k(const double* u, double* f) = 0;
};

// Per-point lists of the element terms in the order a sequential loop over the
// elements in `visit` order adds them (each element's points in k, j, i order).
void assembly_order(const Mesh& m, const std::vector<int>& visit, std::vector<int>& off,
                    std::vector<int>& idx) {
  off.assign((size_t)m.nglob + 1, 0);
  for (size_t L = 0; L < m.ibool.size(); ++L) off[(size_t)m.ibool[L] + 1]++;
  for (int g = 0; g < m.nglob; ++g) off[(size_t)g + 1] += off[g];
  idx.assign(m.ibool.size(), -1);
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (int e : visit)
    for (int L = 0; L < NG3; ++L) {
      const size_t c = (size_t)e * NG3 + L;
      idx[fill[m.ibool[c]]++] = (int)c;
    }
}

template <class T>
struct Sim : SimBase {
  T hp[NG][NG], hpw[NG][NG], wxy[NG][NG];  // wxy[a][b] = w_a w_b
  std::vector<T> kappa, mu, rho, rvp, rvs, cij;      // per local point
  std::vector<T> jinv, jac;                           // per irregular local point
  T xr, jr;
  std::vector<T> rmx, rmy, rmz;                       // inverted mass (with Stacey term)
  std::vector<T> u, v, a, bu, bv, ba, au, av, aa;     // forward, backward, adjoint fields
  std::vector<T> store;                               // Stacey store [nstep][nfaces*25][3]
  std::vector<T> sarr;                                // [nsrc][125][3]
  std::vector<T> rh;                                  // [nrec][125]
  std::vector<double> energy;
  bool saved = false;
  std::vector<T> floc;                                // element terms [nspec*125][3] (ordered assembly)

  void setup() {
    size_t nloc = (size_t)m.nspec * NG3;
    for (int i = 0; i < NG; ++i)
      for (int j = 0; j < NG; ++j) {
        hp[i][j] = (T)g.hp[i][j];
        hpw[i][j] = (T)g.hpw[i][j];
        wxy[i][j] = (T)(g.w[i] * g.w[j]);
      }
    xr = (T)m.xix_reg;
    jr = (T)m.jac_reg;
    jinv.assign((size_t)m.nspec_irreg * NG3 * 9, 0);
    jac.assign((size_t)m.nspec_irreg * NG3, 0);
    for (int e = 0; e < m.nspec; ++e) {
      if (!m.irreg[e]) continue;
      size_t ie = (size_t)(m.irreg[e] - 1);
      for (int L = 0; L < NG3; ++L) {
        for (int c = 0; c < 9; ++c)
          jinv[(ie * NG3 + L) * 9 + c] = (T)m.jinv[((size_t)e * NG3 + L) * 9 + c];
        jac[ie * NG3 + L] = (T)m.jac[(size_t)e * NG3 + L];
      }
    }
    // materials: get_model_fwat.f90:305-317 (single precision arithmetic)
    kappa.resize(nloc); mu.resize(nloc); rho.resize(nloc); rvp.resize(nloc); rvs.resize(nloc);
    for (size_t L = 0; L < nloc; ++L) {
      T r = (T)p.rho[L], a_ = (T)p.vp[L], b_ = (T)p.vs[L];
      const T four_thirds = (T)(4.0 / 3.0);
      rho[L] = r;
      kappa[L] = r * (a_ * a_ - four_thirds * b_ * b_);
      mu[L] = r * b_ * b_;
      rvp[L] = r * a_;
      rvs[L] = r * b_;
      if (g_plant == 1) std::swap(rvp[L], rvs[L]);
    }
    if (p.aniso) {
      cij.resize(nloc * 21);
      for (size_t q = 0; q < nloc * 21; ++q) cij[q] = (T)p.cij[q];
    }
    // mass matrix
    std::vector<T> rmass(m.nglob, (T)0);
    for (int e = 0; e < m.nspec; ++e)
      for (int k = 0; k < NG; ++k)
        for (int j = 0; j < NG; ++j)
          for (int i = 0; i < NG; ++i) {
            size_t L = (size_t)e * NG3 + i + NG * j + NG2 * k;
            double weight = g.w[i] * g.w[j] * g.w[k];
            double jl = m.irreg[e] ? (double)jac[(size_t)(m.irreg[e] - 1) * NG3 + (L % NG3)]
                                   : (double)jr;
            rmass[m.ibool[L]] += (T)(jl * weight * (double)rho[L]);
          }
    rmx = rmass; rmy = rmass; rmz = rmass;
    if (p.stacey) {
      const T dto2 = (T)(0.5 * p.dt);
      for (int f = 0; f < m.nfaces; ++f) {
        int e = m.f_ispec[f];
        for (int q = 0; q < NG2; ++q) {
          int i = m.f_ijk[(f * NG2 + q) * 3], j = m.f_ijk[(f * NG2 + q) * 3 + 1],
              k = m.f_ijk[(f * NG2 + q) * 3 + 2];
          size_t L = (size_t)e * NG3 + i + NG * j + NG2 * k;
          int gI = m.ibool[L];
          T nx = (T)m.f_n[(f * NG2 + q) * 3], ny = (T)m.f_n[(f * NG2 + q) * 3 + 1],
            nz = (T)m.f_n[(f * NG2 + q) * 3 + 2];
          T jw = (T)m.f_jw[f * NG2 + q];
          T vn = dto2 * (nx + ny + nz);
          T tx = rvp[L] * vn * nx + rvs[L] * (dto2 - vn * nx);
          T ty = rvp[L] * vn * ny + rvs[L] * (dto2 - vn * ny);
          T tz = rvp[L] * vn * nz + rvs[L] * (dto2 - vn * nz);
          rmx[gI] += tx * jw;
          rmy[gI] += ty * jw;
          rmz[gI] += tz * jw;
        }
      }
    }
    for (int gI = 0; gI < m.nglob; ++gI) {
      rmx[gI] = (T)1 / rmx[gI];
      rmy[gI] = (T)1 / rmy[gI];
      rmz[gI] = (T)1 / rmz[gI];
    }
    if (!asm_off.empty()) floc.assign(nloc * 3, (T)0);
    sarr.resize(srcarr.size());
    for (size_t q = 0; q < srcarr.size(); ++q) sarr[q] = (T)srcarr[q];
    rh.resize(rech.size());
    for (size_t q = 0; q < rech.size(); ++q) rh[q] = (T)rech[q];
  }

  // Element internal forces: accel -= K_e u  (weak form); optionally the 6
  // strain components per local point (exx,eyy,ezz,exy,exz,eyz).
  // F: when given, the element's nodal force terms are stored at F[3 (125 e + L)]
  // (already negated) instead of being subtracted from A
  void elem(int e, const T* U, T* A, T* strain, T* F = nullptr) const {
    T ux[NG3], uy[NG3], uz[NG3];
    T t1x[NG3], t1y[NG3], t1z[NG3], t2x[NG3], t2y[NG3], t2z[NG3], t3x[NG3], t3y[NG3], t3z[NG3];
    const int* ib = &m.ibool[(size_t)e * NG3];
    for (int L = 0; L < NG3; ++L) {
      ux[L] = U[3 * (size_t)ib[L]];
      uy[L] = U[3 * (size_t)ib[L] + 1];
      uz[L] = U[3 * (size_t)ib[L] + 2];
    }
    const int ir = m.irreg[e];
    for (int k = 0; k < NG; ++k)
      for (int j = 0; j < NG; ++j)
        for (int i = 0; i < NG; ++i) {
          const int L = i + NG * j + NG2 * k;
          T ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0, cx = 0, cy = 0, cz = 0;
          for (int l = 0; l < NG; ++l) {
            const int La = l + NG * j + NG2 * k, Lb = i + NG * l + NG2 * k, Lc = i + NG * j + NG2 * l;
            ax += ux[La] * hp[i][l]; ay += uy[La] * hp[i][l]; az += uz[La] * hp[i][l];
            bx += ux[Lb] * hp[j][l]; by += uy[Lb] * hp[j][l]; bz += uz[Lb] * hp[j][l];
            cx += ux[Lc] * hp[k][l]; cy += uy[Lc] * hp[k][l]; cz += uz[Lc] * hp[k][l];
          }
          T xix, xiy, xiz, etx, ety, etz, gax, gay, gaz, jl;
          if (ir) {
            const T* J = &jinv[((size_t)(ir - 1) * NG3 + L) * 9];
            xix = J[0]; xiy = J[1]; xiz = J[2];
            etx = J[3]; ety = J[4]; etz = J[5];
            gax = J[6]; gay = J[7]; gaz = J[8];
            jl = jac[(size_t)(ir - 1) * NG3 + L];
          } else {
            xix = xr; ety = xr; gaz = xr;
            xiy = xiz = etx = etz = gax = gay = 0;
            jl = jr;
          }
          T duxdx = xix * ax + etx * bx + gax * cx;
          T duxdy = xiy * ax + ety * bx + gay * cx;
          T duxdz = xiz * ax + etz * bx + gaz * cx;
          T duydx = xix * ay + etx * by + gax * cy;
          T duydy = xiy * ay + ety * by + gay * cy;
          T duydz = xiz * ay + etz * by + gaz * cy;
          T duzdx = xix * az + etx * bz + gax * cz;
          T duzdy = xiy * az + ety * bz + gay * cz;
          T duzdz = xiz * az + etz * bz + gaz * cz;
          T exy2 = duxdy + duydx, exz2 = duzdx + duxdz, eyz2 = duzdy + duydz;
          const size_t Lg = (size_t)e * NG3 + L;
          T sxx, syy, szz, sxy, sxz, syz;
          if (!p.aniso) {
            T kl = kappa[Lg], ml = mu[Lg];
            T lp2m = kl + (T)(4.0 / 3.0) * ml;
            T lam = lp2m - (T)2 * ml;
            sxx = lp2m * duxdx + lam * (duydy + duzdz);
            syy = lp2m * duydy + lam * (duxdx + duzdz);
            szz = lp2m * duzdz + lam * (duxdx + duydy);
            sxy = ml * exy2;
            sxz = ml * exz2;
            syz = ml * eyz2;
          } else {
            const T* c = &cij[Lg * 21];
            // Voigt strain e = [exx, eyy, ezz, 2eyz, 2exz, 2exy]
            T ev[6] = {duxdx, duydy, duzdz, eyz2, exz2, exy2};
            T s[6];
            // c index of (I,J), I<=J, upper-triangle row-major
            static const int idx[6][6] = {{0, 1, 2, 3, 4, 5},       {1, 6, 7, 8, 9, 10},
                                          {2, 7, 11, 12, 13, 14},   {3, 8, 12, 15, 16, 17},
                                          {4, 9, 13, 16, 18, 19},   {5, 10, 14, 17, 19, 20}};
            for (int I = 0; I < 6; ++I) {
              T acc = 0;
              for (int J2 = 0; J2 < 6; ++J2) acc += c[idx[I][J2]] * ev[J2];
              s[I] = acc;
            }
            sxx = s[0]; syy = s[1]; szz = s[2]; syz = s[3]; sxz = s[4]; sxy = s[5];
          }
          if (strain) {
            T* st = &strain[(size_t)L * 6];
            st[0] = duxdx; st[1] = duydy; st[2] = duzdz;
            st[3] = (T)0.5 * exy2; st[4] = (T)0.5 * exz2; st[5] = (T)0.5 * eyz2;
          }
          t1x[L] = jl * (sxx * xix + sxy * xiy + sxz * xiz);
          t1y[L] = jl * (sxy * xix + syy * xiy + syz * xiz);
          t1z[L] = jl * (sxz * xix + syz * xiy + szz * xiz);
          t2x[L] = jl * (sxx * etx + sxy * ety + sxz * etz);
          t2y[L] = jl * (sxy * etx + syy * ety + syz * etz);
          t2z[L] = jl * (sxz * etx + syz * ety + szz * etz);
          t3x[L] = jl * (sxx * gax + sxy * gay + sxz * gaz);
          t3y[L] = jl * (sxy * gax + syy * gay + syz * gaz);
          t3z[L] = jl * (sxz * gax + syz * gay + szz * gaz);
        }
    for (int k = 0; k < NG; ++k)
      for (int j = 0; j < NG; ++j)
        for (int i = 0; i < NG; ++i) {
          const int L = i + NG * j + NG2 * k;
          T nx1 = 0, ny1 = 0, nz1 = 0, nx2 = 0, ny2 = 0, nz2 = 0, nx3 = 0, ny3 = 0, nz3 = 0;
          for (int l = 0; l < NG; ++l) {
            const int La = l + NG * j + NG2 * k, Lb = i + NG * l + NG2 * k, Lc = i + NG * j + NG2 * l;
            nx1 += t1x[La] * hpw[l][i]; ny1 += t1y[La] * hpw[l][i]; nz1 += t1z[La] * hpw[l][i];
            nx2 += t2x[Lb] * hpw[l][j]; ny2 += t2y[Lb] * hpw[l][j]; nz2 += t2z[Lb] * hpw[l][j];
            nx3 += t3x[Lc] * hpw[l][k]; ny3 += t3y[Lc] * hpw[l][k]; nz3 += t3z[Lc] * hpw[l][k];
          }
          T f1 = wxy[j][k], f2 = wxy[i][k], f3 = wxy[i][j];
          if (F) {
            T* o = &F[3 * ((size_t)e * NG3 + L)];
            o[0] = -(f1 * nx1 + f2 * nx2 + f3 * nx3);
            o[1] = -(f1 * ny1 + f2 * ny2 + f3 * ny3);
            o[2] = -(f1 * nz1 + f2 * nz2 + f3 * nz3);
            continue;
          }
          size_t gI = (size_t)ib[L];
          A[3 * gI] -= f1 * nx1 + f2 * nx2 + f3 * nx3;
          A[3 * gI + 1] -= f1 * ny1 + f2 * ny2 + f3 * ny3;
          A[3 * gI + 2] -= f1 * nz1 + f2 * nz2 + f3 * nz3;
        }
  }

  // Internal forces added into A.  Colour by colour (each colour race-free),
  // or, with an element visiting order, element terms first and then every
  // point's terms added in the order a sequential loop over that visiting
  // order would add them (the same per-point sums).
  void forces(const T* U, T* A) {
    if (asm_off.empty()) {
      for (auto& set : m.color_sets) {
#pragma omp parallel for schedule(static) num_threads(p.nthreads)
        for (size_t q = 0; q < set.size(); ++q) elem(set[q], U, A, nullptr);
      }
      return;
    }
#pragma omp parallel for schedule(static) num_threads(p.nthreads)
    for (int e = 0; e < m.nspec; ++e) elem(e, U, nullptr, nullptr, floc.data());
#pragma omp parallel for schedule(static) num_threads(p.nthreads)
    for (int gI = 0; gI < m.nglob; ++gI)
      for (int q = asm_off[gI]; q < asm_off[gI + 1]; ++q) {
        const T* f = &floc[3 * (size_t)asm_idx[q]];
        A[3 * (size_t)gI] += f[0];
        A[3 * (size_t)gI + 1] += f[1];
        A[3 * (size_t)gI + 2] += f[2];
      }
  }

  void predictor(std::vector<T>& U, std::vector<T>& V, std::vector<T>& A, T dt) {
    const T dt2 = (T)(0.5 * p.dt * p.dt), dto2 = (T)0.5 * dt;
    const size_t n = U.size();
#pragma omp parallel for schedule(static) num_threads(p.nthreads)
    for (size_t q = 0; q < n; ++q) {
      U[q] += dt * V[q] + dt2 * A[q];
      V[q] += dto2 * A[q];
      A[q] = 0;
    }
  }

  void corrector(std::vector<T>& V, std::vector<T>& A, T dt) {
    const T dto2 = (T)0.5 * dt;
#pragma omp parallel for schedule(static) num_threads(p.nthreads)
    for (int gI = 0; gI < m.nglob; ++gI) {
      A[3 * (size_t)gI] *= rmx[gI];
      A[3 * (size_t)gI + 1] *= rmy[gI];
      A[3 * (size_t)gI + 2] *= rmz[gI];
      V[3 * (size_t)gI] += dto2 * A[3 * (size_t)gI];
      V[3 * (size_t)gI + 1] += dto2 * A[3 * (size_t)gI + 1];
      V[3 * (size_t)gI + 2] += dto2 * A[3 * (size_t)gI + 2];
    }
  }

  // Stacey traction from the predicted velocity; mode 0: absorb, 1: absorb and
  // store, 2: replay the stored record (backward reconstruction).
  // it > 0: forward step it, whose incident field (linear between the stored
  // samples) is injected: the absorbing traction of v - v_inc minus t_inc
  // (SPECFEM's compute_stacey_viscoelastic under COUPLE_WITH_INJECTION_TECHNIQUE)
  void stacey(const std::vector<T>& V, std::vector<T>& A, int mode, T* rec, int it = 0) {
    if (!p.stacey) return;
    const bool inj = it > 0 && fk_nsamp > 0 && mode != 2;
    size_t k0 = 0, k1 = 0;
    T w1 = 0;
    if (inj) {
      k0 = (size_t)std::min((it - 1) / fk_resamp, fk_nsamp - 1);
      k1 = std::min(k0 + 1, (size_t)fk_nsamp - 1);
      w1 = k1 == k0 ? (T)0 : (T)((it - 1) - (int)k0 * fk_resamp) / (T)fk_resamp;
    }
    const T w0 = (T)1 - w1;
    const size_t nfp3 = (size_t)m.nfaces * NG2 * 3;
    for (int f = 0; f < m.nfaces; ++f) {
      int e = m.f_ispec[f];
      for (int q = 0; q < NG2; ++q) {
        size_t fq = (size_t)f * NG2 + q;
        int i = m.f_ijk[fq * 3], j = m.f_ijk[fq * 3 + 1], k = m.f_ijk[fq * 3 + 2];
        size_t L = (size_t)e * NG3 + i + NG * j + NG2 * k;
        size_t gI = (size_t)m.ibool[L];
        T tx, ty, tz;
        if (mode == 2) {
          tx = rec[fq * 3]; ty = rec[fq * 3 + 1]; tz = rec[fq * 3 + 2];
        } else {
          T nx = (T)m.f_n[fq * 3], ny = (T)m.f_n[fq * 3 + 1], nz = (T)m.f_n[fq * 3 + 2];
          T jw = (T)m.f_jw[fq];
          T vx = V[3 * gI], vy = V[3 * gI + 1], vz = V[3 * gI + 2];
          T ix = 0, iy = 0, iz = 0;
          if (inj) {
            const float *v0 = &fk_v[k0 * nfp3 + fq * 3], *v1 = &fk_v[k1 * nfp3 + fq * 3];
            const float *t0 = &fk_t[k0 * nfp3 + fq * 3], *t1 = &fk_t[k1 * nfp3 + fq * 3];
            vx -= w0 * (T)v0[0] + w1 * (T)v1[0];
            vy -= w0 * (T)v0[1] + w1 * (T)v1[1];
            vz -= w0 * (T)v0[2] + w1 * (T)v1[2];
            ix = w0 * (T)t0[0] + w1 * (T)t1[0];
            iy = w0 * (T)t0[1] + w1 * (T)t1[1];
            iz = w0 * (T)t0[2] + w1 * (T)t1[2];
          }
          T vn = vx * nx + vy * ny + vz * nz;
          tx = (rvp[L] * vn * nx + rvs[L] * (vx - vn * nx) - ix) * jw;
          ty = (rvp[L] * vn * ny + rvs[L] * (vy - vn * ny) - iy) * jw;
          tz = (rvp[L] * vn * nz + rvs[L] * (vz - vn * nz) - iz) * jw;
          if (mode == 1) { rec[fq * 3] = tx; rec[fq * 3 + 1] = ty; rec[fq * 3 + 2] = tz; }
        }
        A[3 * gI] -= tx;
        A[3 * gI + 1] -= ty;
        A[3 * gI + 2] -= tz;
      }
    }
  }

  void add_sources(std::vector<T>& A, int it /*1-based forward step*/) {
    for (int s = 0; s < p.nsrc; ++s) {
      double t = (it - 1) * p.dt - 1.2 / p.src_f0[s] - p.src_tshift[s];
      T stf = (T)ricker(t, p.src_f0[s]);
      for (int L = 0; L < NG3; ++L) {
        size_t gI = (size_t)m.ibool[(size_t)src_ispec[s] * NG3 + L];
        for (int c = 0; c < 3; ++c) A[3 * gI + c] += sarr[((size_t)s * NG3 + L) * 3 + c] * stf;
      }
    }
  }

  void record(const std::vector<T>& U, float* seis_row) {
    for (int r = 0; r < p.nrec; ++r) {
      T acc[3] = {0, 0, 0};
      for (int L = 0; L < NG3; ++L) {
        size_t gI = (size_t)m.ibool[(size_t)rec_ispec[r] * NG3 + L];
        T h = rh[(size_t)r * NG3 + L];
        for (int c = 0; c < 3; ++c) acc[c] += h * U[3 * gI + c];
      }
      for (int c = 0; c < 3; ++c) seis_row[r * 3 + c] = (float)acc[c];
    }
  }

  double energy_now(const std::vector<T>& U, const std::vector<T>& V) {
    // kinetic + strain energy: 0.5 v^T M v + 0.5 u^T K u (K u via the force routine)
    std::vector<T> f(U.size(), (T)0);
    forces(U.data(), f.data());
    double ek = 0, ep = 0;
    std::vector<double> rm(m.nglob, 0.0);
    for (int e = 0; e < m.nspec; ++e)
      for (int L = 0; L < NG3; ++L) {
        int i = L % NG, j = (L / NG) % NG, k = L / NG2;
        double jl = m.irreg[e] ? (double)jac[(size_t)(m.irreg[e] - 1) * NG3 + L] : (double)jr;
        rm[m.ibool[(size_t)e * NG3 + L]] += jl * g.w[i] * g.w[j] * g.w[k] * (double)rho[(size_t)e * NG3 + L];
      }
    for (int gI = 0; gI < m.nglob; ++gI)
      for (int c = 0; c < 3; ++c) {
        ek += 0.5 * rm[gI] * (double)V[3 * (size_t)gI + c] * (double)V[3 * (size_t)gI + c];
        ep += -0.5 * (double)U[3 * (size_t)gI + c] * (double)f[3 * (size_t)gI + c];
      }
    return ek + ep;
  }

  void forward(float* seis, int save) override {
    size_t n3 = (size_t)m.nglob * 3;
    u.assign(n3, 0); v.assign(n3, 0); a.assign(n3, 0);
    size_t rec_sz = (size_t)m.nfaces * NG2 * 3;
    if (save && p.stacey) store.assign(rec_sz * p.nstep, 0);
    energy.clear();
    const T dt = (T)p.dt;
    for (int it = 1; it <= p.nstep; ++it) {
      predictor(u, v, a, dt);
      forces(u.data(), a.data());
      stacey(v, a, save ? 1 : 0, save && p.stacey ? &store[rec_sz * (it - 1)] : nullptr, it);
      add_sources(a, it);
      corrector(v, a, dt);
      if (seis) record(u, &seis[(size_t)(it - 1) * p.nrec * 3]);
    }
    saved = save != 0;
  }

  void adjoint(const float* adj, double* rho_kl, double* mu_kl, double* kappa_kl,
               double* cijkl_kl, float* b_final) override {
    size_t n3 = (size_t)m.nglob * 3, nloc = (size_t)m.nspec * NG3;
    size_t rec_sz = (size_t)m.nfaces * NG2 * 3;
    // Reconstructed field: adjoint step it pairs with forward state
    // m = N + 1 - it (the discrete adjoint of the central-difference scheme
    // pairs lambda_n with u_n).  Step 1 uses the saved last state itself:
    // its half-step velocity v_N + dt/2 a_N is set here so the corrector of
    // step 1 returns v_N; steps it >= 2 take one step back from m + 1.
    bu = u; bv = v; ba = a;
    for (size_t q = 0; q < n3; ++q) bv[q] += (T)0.5 * (T)p.dt * ba[q];
    au.assign(n3, 0); av.assign(n3, 0); aa.assign(n3, 0);
    std::vector<double> krho(nloc, 0.0), kmu(nloc, 0.0), kkap(nloc, 0.0);
    std::vector<double> kc(p.aniso ? nloc * 21 : 0, 0.0);
    const T dt = (T)p.dt;
    const int N = p.nstep;
    std::vector<T> sa((size_t)NG3 * 6), sb((size_t)NG3 * 6);
    for (int it = 1; it <= N; ++it) {
      const int mstate = N + 1 - it;
      if (it >= 2) {
        predictor(bu, bv, ba, -dt);
      } else {
        std::fill(ba.begin(), ba.end(), (T)0);
      }
      forces(bu.data(), ba.data());
      if (p.stacey) stacey(bv, ba, 2, &store[rec_sz * (mstate - 1)]);
      add_sources(ba, mstate);
      corrector(bv, ba, -dt);
      // adjoint field, forward in reversed time
      predictor(au, av, aa, dt);
      forces(au.data(), aa.data());
      stacey(av, aa, 0, nullptr);
      const int isamp = N - it;  // 0-based trace sample
      for (int r = 0; r < p.nrec; ++r) {
        const float* f = &adj[((size_t)isamp * p.nrec + r) * 3];
        for (int L = 0; L < NG3; ++L) {
          size_t gI = (size_t)m.ibool[(size_t)rec_ispec[r] * NG3 + L];
          T h = rh[(size_t)r * NG3 + L];
          for (int c = 0; c < 3; ++c) aa[3 * gI + c] += h * (T)f[c];
        }
      }
      corrector(av, aa, dt);
      // kernels
#pragma omp parallel for schedule(static) num_threads(p.nthreads) firstprivate(sa, sb)
      for (int e = 0; e < m.nspec; ++e) {
        elem_strain_only(e, au.data(), sa.data());
        elem_strain_only(e, bu.data(), sb.data());
        for (int L = 0; L < NG3; ++L) {
          size_t Lg = (size_t)e * NG3 + L;
          size_t gI = (size_t)m.ibool[Lg];
          T dotp = aa[3 * gI] * bu[3 * gI] + aa[3 * gI + 1] * bu[3 * gI + 1] + aa[3 * gI + 2] * bu[3 * gI + 2];
          krho[Lg] += (double)p.dt * (double)dotp;
          const T* A_ = &sa[(size_t)L * 6];
          const T* B_ = &sb[(size_t)L * 6];
          if (!p.aniso) {
            T tra = A_[0] + A_[1] + A_[2], trb = B_[0] + B_[1] + B_[2];
            T t3a = tra / (T)3, t3b = trb / (T)3;
            T dxa = A_[0] - t3a, dya = A_[1] - t3a, dza = A_[2] - t3a;
            T dxb = B_[0] - t3b, dyb = B_[1] - t3b, dzb = B_[2] - t3b;
            T prod = dxa * dxb + dya * dyb + dza * dzb +
                     (T)2 * (A_[3] * B_[3] + A_[4] * B_[4] + A_[5] * B_[5]);
            kmu[Lg] += (double)p.dt * (double)prod;
            kkap[Lg] += (double)p.dt * (double)(tra * trb);
          } else {
            T ea[6] = {A_[0], A_[1], A_[2], (T)2 * A_[5], (T)2 * A_[4], (T)2 * A_[3]};
            T eb[6] = {B_[0], B_[1], B_[2], (T)2 * B_[5], (T)2 * B_[4], (T)2 * B_[3]};
            int q = 0;
            for (int I = 0; I < 6; ++I)
              for (int J = I; J < 6; ++J) {
                T pr = ea[I] * eb[J];
                if (J > I) pr += ea[J] * eb[I];
                kc[Lg * 21 + q] += (double)p.dt * (double)pr;
                ++q;
              }
          }
        }
      }
    }
    // save_adjoint_kernels scaling (Tromp, Tape & Liu 2005 relative kernels)
    for (size_t L = 0; L < nloc; ++L) {
      double r = (double)rho[L];
      if (rho_kl) rho_kl[L] = -r * krho[L];
      if (!p.aniso) {
        if (mu_kl) mu_kl[L] = -2.0 * (double)mu[L] * kmu[L];
        if (kappa_kl) kappa_kl[L] = -(double)kappa[L] * kkap[L];
      } else if (cijkl_kl) {
        for (int q = 0; q < 21; ++q) cijkl_kl[L * 21 + q] = -kc[L * 21 + q];
      }
    }
    if (b_final)
      for (size_t q = 0; q < n3; ++q) {
        b_final[q] = (float)bu[q];
        b_final[n3 + q] = (float)bv[q];
        b_final[2 * n3 + q] = (float)ba[q];
      }
  }

  void elem_strain_only(int e, const T* U, T* strain) const {
    T ux[NG3], uy[NG3], uz[NG3];
    const int* ib = &m.ibool[(size_t)e * NG3];
    for (int L = 0; L < NG3; ++L) {
      ux[L] = U[3 * (size_t)ib[L]];
      uy[L] = U[3 * (size_t)ib[L] + 1];
      uz[L] = U[3 * (size_t)ib[L] + 2];
    }
    const int ir = m.irreg[e];
    for (int k = 0; k < NG; ++k)
      for (int j = 0; j < NG; ++j)
        for (int i = 0; i < NG; ++i) {
          const int L = i + NG * j + NG2 * k;
          T ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0, cx = 0, cy = 0, cz = 0;
          for (int l = 0; l < NG; ++l) {
            const int La = l + NG * j + NG2 * k, Lb = i + NG * l + NG2 * k, Lc = i + NG * j + NG2 * l;
            ax += ux[La] * hp[i][l]; ay += uy[La] * hp[i][l]; az += uz[La] * hp[i][l];
            bx += ux[Lb] * hp[j][l]; by += uy[Lb] * hp[j][l]; bz += uz[Lb] * hp[j][l];
            cx += ux[Lc] * hp[k][l]; cy += uy[Lc] * hp[k][l]; cz += uz[Lc] * hp[k][l];
          }
          T xix, xiy, xiz, etx, ety, etz, gax, gay, gaz;
          if (ir) {
            const T* J = &jinv[((size_t)(ir - 1) * NG3 + L) * 9];
            xix = J[0]; xiy = J[1]; xiz = J[2];
            etx = J[3]; ety = J[4]; etz = J[5];
            gax = J[6]; gay = J[7]; gaz = J[8];
          } else {
            xix = xr; ety = xr; gaz = xr;
            xiy = xiz = etx = etz = gax = gay = 0;
          }
          T duxdx = xix * ax + etx * bx + gax * cx;
          T duxdy = xiy * ax + ety * bx + gay * cx;
          T duxdz = xiz * ax + etz * bx + gaz * cx;
          T duydx = xix * ay + etx * by + gax * cy;
          T duydy = xiy * ay + ety * by + gay * cy;
          T duydz = xiz * ay + etz * by + gaz * cy;
          T duzdx = xix * az + etx * bz + gax * cz;
          T duzdy = xiy * az + ety * bz + gay * cz;
          T duzdz = xiz * az + etz * bz + gaz * cz;
          T* st = &strain[(size_t)L * 6];
          st[0] = duxdx; st[1] = duydy; st[2] = duzdz;
          st[3] = (T)0.5 * (duxdy + duydx);
          st[4] = (T)0.5 * (duzdx + duxdz);
          st[5] = (T)0.5 * (duzdy + duydz);
        }
  }

  void get_rmass(float* rx, float* ry, float* rz) override {
    for (int gI = 0; gI < m.nglob; ++gI) {
      if (rx) rx[gI] = (float)rmx[gI];
      if (ry) ry[gI] = (float)rmy[gI];
      if (rz) rz[gI] = (float)rmz[gI];
    }
  }
  void get_final(float* U, float* V, float* A) override {
    for (size_t q = 0; q < u.size(); ++q) {
      if (U) U[q] = (float)u[q];
      if (V) V[q] = (float)v[q];
      if (A) A[q] = (float)a[q];
    }
  }
  void set_final(const float* U, const float* V, const float* A) override {
    size_t n3 = (size_t)m.nglob * 3;
    u.resize(n3); v.resize(n3); a.resize(n3);
    for (size_t q = 0; q < n3; ++q) { u[q] = (T)U[q]; v[q] = (T)V[q]; a[q] = (T)A[q]; }
  }
  void get_stacey(float* buf) override {
    for (size_t q = 0; q < store.size(); ++q) buf[q] = (float)store[q];
  }
  // closed-box energy history (no Stacey, no source): used for the O(dt^2)
  // energy-drift invariant of the Newmark + force pair
  void energy_series(double* e) override {
    for (int it = 1; it <= p.nstep; ++it) {
      const T dt = (T)p.dt;
      predictor(u, v, a, dt);
      forces(u.data(), a.data());
      corrector(v, a, dt);
      e[it - 1] = energy_now(u, v);
    }
  }
  void apply_k(const double* U, double* F) override {
    size_t n3 = (size_t)m.nglob * 3;
    std::vector<T> uu(n3), ff(n3, (T)0);
    for (size_t q = 0; q < n3; ++q) uu[q] = (T)U[q];
    forces(uu.data(), ff.data());
    for (size_t q = 0; q < n3; ++q) F[q] = (double)ff[q];
  }
};

void setup_sources(SimBase* s) {
  const OrcParams& p = s->p;
  s->src_ispec.assign(p.nsrc, -1);
  s->src_ref.assign(3 * p.nsrc, 0);
  s->srcarr.assign((size_t)p.nsrc * NG3 * 3, 0);
  for (int q = 0; q < p.nsrc; ++q) {
    int e;
    double r[3];
    if (!locate_point(s->m, p.src_xyz[3 * q], p.src_xyz[3 * q + 1], p.src_xyz[3 * q + 2], &e, r)) continue;
    s->src_ispec[q] = e;
    for (int c = 0; c < 3; ++c) s->src_ref[3 * q + c] = r[c];
    double X[24], x[3], J[3][3], Ji[3][3];
    element_anchors(s->m, e, X);
    trilinear(X, r[0], r[1], r[2], x, J);
    inv3(J, Ji);
    double hx[NG], hxd[NG], hy[NG], hyd[NG], hz[NG], hzd[NG];
    lagrange(s->g, r[0], hx, hxd);
    lagrange(s->g, r[1], hy, hyd);
    lagrange(s->g, r[2], hz, hzd);
    const double* M6 = &p.src_mt[6 * q];
    double M[3][3] = {{M6[0], M6[3], M6[4]}, {M6[3], M6[1], M6[5]}, {M6[4], M6[5], M6[2]}};
    if (g_plant == 2) {
      std::swap(M[0][2], M[1][2]);
      M[2][0] = M[0][2];
      M[2][1] = M[1][2];
    }
    // G[c][d] = sum_j M[c][j] dref_d/dx_j
    double G[3][3];
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < 3; ++d) G[c][d] = M[c][0] * Ji[d][0] + M[c][1] * Ji[d][1] + M[c][2] * Ji[d][2];
    for (int k = 0; k < NG; ++k)
      for (int j = 0; j < NG; ++j)
        for (int i = 0; i < NG; ++i) {
          int L = i + NG * j + NG2 * k;
          double d0 = hxd[i] * hy[j] * hz[k], d1 = hx[i] * hyd[j] * hz[k], d2 = hx[i] * hy[j] * hzd[k];
          for (int c = 0; c < 3; ++c)
            s->srcarr[((size_t)q * NG3 + L) * 3 + c] = G[c][0] * d0 + G[c][1] * d1 + G[c][2] * d2;
        }
  }
  s->rec_ispec.assign(p.nrec, -1);
  s->rec_ref.assign(3 * p.nrec, 0);
  s->rech.assign((size_t)p.nrec * NG3, 0);
  for (int q = 0; q < p.nrec; ++q) {
    int e;
    double r[3];
    if (!locate_point(s->m, p.rec_xyz[3 * q], p.rec_xyz[3 * q + 1], p.rec_xyz[3 * q + 2], &e, r)) continue;
    s->rec_ispec[q] = e;
    for (int c = 0; c < 3; ++c) s->rec_ref[3 * q + c] = r[c];
    double hx[NG], hxd[NG], hy[NG], hyd[NG], hz[NG], hzd[NG];
    lagrange(s->g, r[0], hx, hxd);
    lagrange(s->g, r[1], hy, hyd);
    lagrange(s->g, r[2], hz, hzd);
    for (int k = 0; k < NG; ++k)
      for (int j = 0; j < NG; ++j)
        for (int i = 0; i < NG; ++i) s->rech[(size_t)q * NG3 + i + NG * j + NG2 * k] = hx[i] * hy[j] * hz[k];
  }
}

}  // namespace

extern "C" {

void* orc_create(const OrcParams* p, int precision) {
  SimBase* s = precision == 64 ? (SimBase*)new Sim<double>() : (SimBase*)new Sim<float>();
  s->prec = precision;
  s->p = *p;
  if (s->p.nthreads <= 0) s->p.nthreads = omp_get_max_threads();
  s->g = make_gll();
  build_mesh(s->m, s->g, p->nex, p->ney, p->nez, p->lx, p->ly, p->lz, p->perturb,
             p->force_irregular, p->numbering);
  setup_sources(s);
  if (p->elem_order) {
    std::vector<int> visit(p->elem_order, p->elem_order + s->m.nspec), seen(s->m.nspec, 0);
    for (int e : visit) {
      if (e < 0 || e >= s->m.nspec || seen[e]++) { delete s; return nullptr; }
    }
    assembly_order(s->m, visit, s->asm_off, s->asm_idx);
  }
  if (precision == 64) static_cast<Sim<double>*>(s)->setup();
  else static_cast<Sim<float>*>(s)->setup();
  return s;
}

void orc_destroy(void* h) { delete static_cast<SimBase*>(h); }

// Per-point order in which the oracle adds the element terms of the internal
// forces: off[nglob+1], idx[nspec*125] (contribution e*125 + L).
void orc_assembly_order(void* h, int* off, int* idx) {
  SimBase* s = static_cast<SimBase*>(h);
  std::vector<int> o, i;
  if (s->asm_off.empty()) {
    std::vector<int> visit;
    for (auto& set : s->m.color_sets) visit.insert(visit.end(), set.begin(), set.end());
    assembly_order(s->m, visit, o, i);
  }
  const std::vector<int>& O = s->asm_off.empty() ? o : s->asm_off;
  const std::vector<int>& I = s->asm_off.empty() ? i : s->asm_idx;
  std::memcpy(off, O.data(), O.size() * sizeof(int));
  std::memcpy(idx, I.data(), I.size() * sizeof(int));
}

// Planted fault for later orc_create calls (0 = none); see g_plant.
void orc_set_plant(int code) { g_plant = code; }

int orc_nglob(void* h) { return static_cast<SimBase*>(h)->m.nglob; }
int orc_nfaces(void* h) { return static_cast<SimBase*>(h)->m.nfaces; }
int orc_nspec_irregular(void* h) { return static_cast<SimBase*>(h)->m.nspec_irreg; }
int orc_ncolors(void* h) { return (int)static_cast<SimBase*>(h)->m.color_sets.size(); }

void orc_get_mesh(void* h, int* ibool, double* xyz, int* irreg, double* regular /*xix,jac*/,
                  double* jinv, double* jac, int* color) {
  SimBase* s = static_cast<SimBase*>(h);
  const Mesh& m = s->m;
  size_t nloc = (size_t)m.nspec * NG3;
  if (ibool) for (size_t L = 0; L < nloc; ++L) ibool[L] = m.ibool[L];
  if (xyz)
    for (int gI = 0; gI < m.nglob; ++gI) {
      xyz[3 * (size_t)gI] = m.gx[gI]; xyz[3 * (size_t)gI + 1] = m.gy[gI]; xyz[3 * (size_t)gI + 2] = m.gz[gI];
    }
  if (irreg) for (int e = 0; e < m.nspec; ++e) irreg[e] = m.irreg[e];
  if (regular) { regular[0] = m.xix_reg; regular[1] = m.jac_reg; }
  if (jinv) std::memcpy(jinv, m.jinv.data(), nloc * 9 * sizeof(double));
  if (jac) std::memcpy(jac, m.jac.data(), nloc * sizeof(double));
  if (color) for (int e = 0; e < m.nspec; ++e) color[e] = m.color[e];
}

void orc_get_faces(void* h, int* ispec, int* ijk, double* normal, double* jw) {
  const Mesh& m = static_cast<SimBase*>(h)->m;
  for (int f = 0; f < m.nfaces; ++f) ispec[f] = m.f_ispec[f];
  std::memcpy(ijk, m.f_ijk.data(), m.f_ijk.size() * sizeof(int));
  std::memcpy(normal, m.f_n.data(), m.f_n.size() * sizeof(double));
  std::memcpy(jw, m.f_jw.data(), m.f_jw.size() * sizeof(double));
}

void orc_get_gll(double* xi, double* w, double* hp, double* hpw) {
  Gll g = make_gll();
  for (int i = 0; i < NG; ++i) {
    xi[i] = g.xi[i];
    w[i] = g.w[i];
    for (int j = 0; j < NG; ++j) { hp[i * NG + j] = g.hp[i][j]; hpw[i * NG + j] = g.hpw[i][j]; }
  }
}

void orc_lagrange(double x, double* h, double* hd) {
  Gll g = make_gll();
  lagrange(g, x, h, hd);
}

void orc_get_sources(void* h, int* src_ispec, double* src_ref, double* srcarr, int* rec_ispec,
                     double* rec_ref, double* rech) {
  SimBase* s = static_cast<SimBase*>(h);
  for (int q = 0; q < s->p.nsrc; ++q) {
    if (src_ispec) src_ispec[q] = s->src_ispec[q];
    for (int c = 0; c < 3; ++c) if (src_ref) src_ref[3 * q + c] = s->src_ref[3 * q + c];
  }
  if (srcarr) std::memcpy(srcarr, s->srcarr.data(), s->srcarr.size() * sizeof(double));
  for (int q = 0; q < s->p.nrec; ++q) {
    if (rec_ispec) rec_ispec[q] = s->rec_ispec[q];
    for (int c = 0; c < 3; ++c) if (rec_ref) rec_ref[3 * q + c] = s->rec_ref[3 * q + c];
  }
  if (rech) std::memcpy(rech, s->rech.data(), s->rech.size() * sizeof(double));
}

void orc_get_rmass(void* h, float* rx, float* ry, float* rz) { static_cast<SimBase*>(h)->get_rmass(rx, ry, rz); }

void orc_forward(void* h, float* seis, int save) { static_cast<SimBase*>(h)->forward(seis, save); }

void orc_get_final(void* h, float* u, float* v, float* a) { static_cast<SimBase*>(h)->get_final(u, v, a); }
void orc_set_final(void* h, const float* u, const float* v, const float* a) { static_cast<SimBase*>(h)->set_final(u, v, a); }
void orc_get_stacey(void* h, float* buf) { static_cast<SimBase*>(h)->get_stacey(buf); }
// incident field for FK injection in later forward runs (nsamp 0: off)
void orc_set_fk(void* h, int npt, int nsamp, int resamp, const float* v, const float* t) {
  SimBase* b = static_cast<SimBase*>(h);
  const size_t n = (size_t)nsamp * npt * 3;
  b->fk_v.assign(v, v + n);
  b->fk_t.assign(t, t + n);
  b->fk_nsamp = nsamp;
  b->fk_resamp = resamp;
}

void orc_adjoint(void* h, const float* adj, double* rho_kl, double* mu_kl, double* kappa_kl,
                 double* cijkl_kl, float* b_final) {
  static_cast<SimBase*>(h)->adjoint(adj, rho_kl, mu_kl, kappa_kl, cijkl_kl, b_final);
}

// initial displacement field then closed-box energy history
void orc_energy_series(void* h, const float* u0, double* e) {
  SimBase* s = static_cast<SimBase*>(h);
  size_t n3 = (size_t)s->m.nglob * 3;
  std::vector<float> z(n3, 0.0f);
  s->set_final(u0, z.data(), z.data());
  s->energy_series(e);
}

void orc_apply_k(void* h, const double* u, double* f) { static_cast<SimBase*>(h)->apply_k(u, f); }

double orc_ricker(double t, double f0) { return ricker(t, f0); }

// ------------------------------------------------ SpecFWAT in-tree leaves

// locate_bissection, src/shared/utils.f90:1490-1520 (1-based result, -1 outside)
int orc_locate_bissection(const double* valx, int n, double x) {
  if (x < valx[0] || x > valx[n - 1]) return -1;
  int jl = 0, ju = n + 1;
  while (ju - jl > 1) {
    int jm = (ju + jl) / 2;
    if (x >= valx[jm - 1]) jl = jm; else ju = jm;
  }
  if (x == valx[0]) return 1;
  if (x == valx[n - 1]) return n - 1;
  return jl;
}

float orc_interp3_pre(const double* dx, const float* vx, int nx, const double* dy, const float* vy, int ny,
                      const double* dz, const float* vz, int nz, const float* valin, float x, float y, float z);

// interp3_0_cr, src/shared/utils.f90:1162-1232; valin is Fortran (nx,ny,nz)
float orc_interp3(const float* vx, int nx, const float* vy, int ny, const float* vz, int nz,
                  const float* valin, float x, float y, float z) {
  std::vector<double> dx(vx, vx + nx), dy(vy, vy + ny), dz(vz, vz + nz);
  return orc_interp3_pre(dx.data(), vx, nx, dy.data(), vy, ny, dz.data(), vz, nz, valin, x, y, z);
}

// the same with the double copies of the axes made once by the caller
float orc_interp3_pre(const double* dx, const float* vx, int nx, const double* dy, const float* vy, int ny,
                      const double* dz, const float* vz, int nz, const float* valin, float x, float y, float z) {
  int indx = orc_locate_bissection(dx, nx, (double)x);
  int indy = orc_locate_bissection(dy, ny, (double)y);
  int indz = orc_locate_bissection(dz, nz, (double)z);
  const int m = 2;
  int kx = std::min(std::max(indx - (m - 1) / 2, 1), nx + 1 - m);
  int ky = std::min(std::max(indy - (m - 1) / 2, 1), ny + 1 - m);
  int kz = std::min(std::max(indz - (m - 1) / 2, 1), nz + 1 - m);
  auto V = [&](int a, int b, int c) { return valin[(size_t)(a - 1) + (size_t)nx * ((b - 1) + (size_t)ny * (c - 1))]; };
  float xix1 = x - vx[kx - 1], yiy1 = y - vy[ky - 1], ziz1 = z - vz[kz - 1];
  float xix2 = x - vx[kx], yiy2 = y - vy[ky], ziz2 = z - vz[kz];
  float x1x2 = 1.0f / (vx[kx - 1] - vx[kx]), y1y2 = 1.0f / (vy[ky - 1] - vy[ky]), z1z2 = 1.0f / (vz[kz - 1] - vz[kz]);
  float x2x1 = 1.0f / (vx[kx] - vx[kx - 1]), y2y1 = 1.0f / (vy[ky] - vy[ky - 1]), z2z1 = 1.0f / (vz[kz] - vz[kz - 1]);
  float f111 = xix2 * yiy2 * ziz2 * x1x2 * y1y2 * z1z2 * V(kx, ky, kz);
  float f112 = xix2 * yiy2 * ziz1 * x1x2 * y1y2 * z2z1 * V(kx, ky, kz + 1);
  float f121 = xix2 * yiy1 * ziz2 * x1x2 * y2y1 * z1z2 * V(kx, ky + 1, kz);
  float f122 = xix2 * yiy1 * ziz1 * x1x2 * y2y1 * z2z1 * V(kx, ky + 1, kz + 1);
  float f211 = xix1 * yiy2 * ziz2 * x2x1 * y1y2 * z1z2 * V(kx + 1, ky, kz);
  float f212 = xix1 * yiy2 * ziz1 * x2x1 * y1y2 * z2z1 * V(kx + 1, ky, kz + 1);
  float f221 = xix1 * yiy1 * ziz2 * x2x1 * y2y1 * z1z2 * V(kx + 1, ky + 1, kz);
  float f222 = xix1 * yiy1 * ziz1 * x2x1 * y2y1 * z2z1 * V(kx + 1, ky + 1, kz + 1);
  return f111 + f112 + f121 + f122 + f211 + f212 + f221 + f222;
}

// HTI -> 21 c_ij in the code frame, src/patch/aniso.f90:55-162 (single precision)
void orc_hti_cij(float vp, float vs, float rho, float gcp, float gsp, float* c21) {
  float A = rho * vp * vp, C = rho * vp * vp, AL = rho * vs * vs, AN = rho * vs * vs;
  float eta = 1.0f;
  float F = eta * (A - 2.0f * AL);
  float Gc = gcp * AL, Gs = gsp * AL;
  float Bc = 0, Bs = 0, Hc = 0, Hs = 0, Ec = 0, Es = 0, C1p = 0, C1sv = 0, C1sh = 0, S1p = 0,
        S1sv = 0, S1sh = 0, C3 = 0, S3 = 0;
  float d11 = A + Ec + Bc, d12 = A - 2.0f * AN - Ec, d13 = F + Hc;
  float d14 = S3 + 2.0f * S1sh + 2.0f * S1p, d15 = 2.0f * C1p + C3, d16 = -Bs / 2.0f - Es;
  float d22 = A + Ec - Bc, d23 = F - Hc, d24 = 2.0f * S1p - S3, d25 = 2.0f * C1p - 2.0f * C1sh - C3;
  float d26 = -Bs / 2.0f + Es, d33 = C, d34 = 2.0f * (S1p - S1sv), d35 = 2.0f * (C1p - C1sv);
  float d36 = -Hs, d44 = AL - Gc, d45 = -Gs, d46 = C1sh - C3, d55 = AL + Gc, d56 = S3 - S1sh;
  float d66 = AN - Ec;
  float out[21] = {d22, d12, d23, -d25, d24, -d26,  // c11..c16
                   d11, d13, -d15, d14, -d16,       // c22..c26
                   d33, -d35, d34, -d36,            // c33..c36
                   d55, -d45, d56,                  // c44..c46
                   d44, -d46,                       // c55, c56
                   d66};                            // c66
  for (int q = 0; q < 21; ++q) c21[q] = out[q];
}

// simpson, src/shared/utils.f90:1630-1663
double orc_simpson(const double* y, int n, double dx) {
  if (n < 2) return 0.0;
  double sum;
  if (n % 2 == 0) {
    sum = y[0] + y[n - 2];
    for (int i = 2; i <= n - 2; i += 2) sum += 4.0 * y[i - 1] + 2.0 * y[i];
    return (dx / 3.0) * sum + 0.5 * dx * (y[n - 2] + y[n - 1]);
  }
  sum = y[0] + y[n - 1];
  for (int i = 2; i <= n - 1; i += 2) sum += 4.0 * y[i - 1];
  for (int i = 3; i <= n - 2; i += 2) sum += 2.0 * y[i - 1];
  return (dx / 3.0) * sum;
}

// window_taper, src/shared/signal.f90:419-491
void orc_window_taper(double* sig, int npts, double pct, int itype) {
  int frac = (pct == 0.0 || pct == 1.0) ? (int)(npts * pct / 2.0) : (int)(npts * pct / 2.0 + 0.5);
  int idx1 = frac, idx2 = npts - frac;
  if (frac <= 0) return;
  std::vector<double> t(frac);
  auto fill = [&](int off) {
    for (int i = 1; i <= frac; ++i) {
      double r = (double)(off + i - 1);
      double v;
      switch (itype) {
        case 1: v = 0.5 - 0.5 * std::cos(2.0 * M_PI * r / (2 * frac - 1)); break;
        case 2: v = 0.54 - 0.46 * std::cos(2.0 * M_PI * r / (2 * frac - 1)); break;
        case 3: v = std::cos(M_PI * r / (2 * frac - 1) - M_PI / 2.0); break;
        default: v = 1.0 - std::pow(std::cos(M_PI * r / (2 * frac - 1)), 10.0); break;
      }
      t[i - 1] = v;
    }
  };
  fill(0);
  for (int i = 0; i < idx1; ++i) sig[i] *= t[i];
  fill(frac);
  for (int i = idx2; i < npts; ++i) sig[i] *= t[i - idx2];
}

// WaveformMisfit%calc_adjoint_source, waveform_misfit.f90:15-64.  windows
// [nwin][2] in seconds; returns the total misfit, fills adj[n] and misfits[nwin].
double orc_waveform_adjoint(const double* dat, const double* syn, int n, double dt,
                            const double* windows, int nwin, double taper_pct, int itaper,
                            double* adj, double* misfits, double* residuals) {
  double total = 0.0;
  for (int i = 0; i < n; ++i) adj[i] = 0.0;
  for (int w = 0; w < nwin; ++w) {
    int nb = (int)(windows[2 * w] / dt) + 1, ne = (int)(windows[2 * w + 1] / dt);
    int nl = ne - nb + 1;
    if (nl <= 0) return NAN;
    std::vector<double> s(syn + nb - 1, syn + ne), d(dat + nb - 1, dat + ne), tw(nl), sq(nl);
    orc_window_taper(s.data(), nl, taper_pct, itaper);
    orc_window_taper(d.data(), nl, taper_pct, itaper);
    for (int i = 0; i < nl; ++i) tw[i] = s[i] - d[i];
    orc_window_taper(tw.data(), nl, taper_pct, itaper);
    double rs = 0.0;
    for (int i = 0; i < nl; ++i) { sq[i] = (s[i] - d[i]) * (s[i] - d[i]); rs += s[i] - d[i]; }
    double mf = 0.5 * orc_simpson(sq.data(), nl, dt);
    if (misfits) misfits[w] = mf;
    if (residuals) residuals[w] = rs / nl;
    total += mf;
    for (int i = 0; i < nl; ++i) adj[nb - 1 + i] = tw[i];
  }
  return total;
}

// ------------------------------------------------ smooth_sem_pde CPU loop
// src/postproc/smooth.f90:7-437 with get_gradient_element(_regular) :603-689.
// faithful=1 keeps the reference's regular-element arithmetic (xix_regular in
// all nine inverse-Jacobian slots, and the gradient of an irregular element
// reading xixstore(:,:,:,ispec)); faithful=0 uses the diagonal map.
// ibool is 1-based (Fortran), arrays are Fortran-ordered per element.
// Returns ntstep.
int orc_smooth_pde(int nspec, int nglob, const int* ibool, const float* xstore,
                   const float* ystore, const float* zstore, int nspec_irreg,
                   const float* xix, const float* xiy, const float* xiz, const float* etax,
                   const float* etay, const float* etaz, const float* gamx, const float* gamy,
                   const float* gamz, const float* jacobian, const int* irreg,
                   float xix_regular, float jacobian_regular, int nspec_outer, int nspec_inner,
                   int num_phase_ispec, const int* phase_ispec /*(num_phase,2) 1-based*/,
                   const int* is_cpml, const float* dat_in, float sigma_h, float sigma_v,
                   float distance_min_glob, int is_sph, int faithful, int max_steps,
                   float* dat_out, float* dat_glob_out) {
  (void)nspec_irreg;
  Gll g = make_gll();
  float hpT[NG][NG], hpw[NG][NG], wyz[NG][NG];
  for (int i = 0; i < NG; ++i)
    for (int j = 0; j < NG; ++j) {
      hpT[j][i] = (float)g.hp[i][j];  // hprime_xxT(j,i) = hprime_xx(i,j)
      hpw[i][j] = (float)g.hpw[i][j];  // hprimewgll_xx(i,j)
      wyz[i][j] = (float)(g.w[i] * g.w[j]);
    }
  const size_t nloc = (size_t)nspec * NG3;
  std::vector<float> rot(is_sph ? nloc * 3 : 0);
  if (is_sph)
    for (size_t L = 0; L < nloc; ++L) {
      int gI = ibool[L] - 1;
      float xl = xstore[gI], yl = ystore[gI], zl = zstore[gI];
      float rl = std::sqrt(xl * xl + yl * yl + zl * zl);
      rot[3 * L] = xl / rl; rot[3 * L + 1] = yl / rl; rot[3 * L + 2] = zl / rl;
    }
  float cmax = distance_min_glob * distance_min_glob / 9.0f;
  float cv, ch;
  if (sigma_v >= sigma_h) { cv = cmax; ch = cv * (sigma_h * sigma_h) / (sigma_v * sigma_v); }
  else { ch = cmax; cv = ch * (sigma_v * sigma_v) / (sigma_h * sigma_h); }
  float mx = std::max(sigma_h, sigma_v);
  int ntstep = (int)std::ceil((mx * mx) / (2.0f * cmax));
  std::vector<float> dat(nloc), dat_bak(nloc, 0.0f), rvol_local(nloc), rvol(nglob, 0.0f);
  std::vector<float> dat_glob(nglob), ddat(nglob);
  auto jac_of = [&](int e, int L) -> float {
    int ir = irreg[e];
    return ir ? jacobian[(size_t)(ir - 1) * NG3 + L] : jacobian_regular;
  };
  for (int e = 0; e < nspec; ++e)
    for (int k = 0; k < NG; ++k)
      for (int j = 0; j < NG; ++j)
        for (int i = 0; i < NG; ++i) {
          int L = i + NG * j + NG2 * k;
          double weight = g.w[i] * g.w[j] * g.w[k];
          float jl = jac_of(e, L);
          rvol_local[(size_t)e * NG3 + L] = (float)((double)jl * weight);
          rvol[ibool[(size_t)e * NG3 + L] - 1] += rvol_local[(size_t)e * NG3 + L];
        }
  for (int gI = 0; gI < nglob; ++gI) rvol[gI] = 1.0f / rvol[gI];
  for (size_t L = 0; L < nloc; ++L) dat[L] = dat_in[L];
  std::fill(dat_glob.begin(), dat_glob.end(), 0.0f);
  for (size_t L = 0; L < nloc; ++L) dat_glob[ibool[L] - 1] += dat[L] * rvol_local[L];
  for (int gI = 0; gI < nglob; ++gI) dat_glob[gI] *= rvol[gI];
  for (size_t L = 0; L < nloc; ++L) dat[L] = dat_glob[ibool[L] - 1];
  int nsteps_run = (max_steps > 0 && max_steps < ntstep) ? max_steps : ntstep;
  float s1[NG3], s2[NG3], s3[NG3], dx[NG3], dy[NG3], dz[NG3];
  for (int istep = 0; istep < nsteps_run; ++istep) {
    std::fill(ddat.begin(), ddat.end(), 0.0f);
    for (int iphase = 1; iphase <= 2; ++iphase) {
      int nel = iphase == 1 ? nspec_outer : nspec_inner;
      for (int ip = 0; ip < nel; ++ip) {
        int e = phase_ispec[ip + (size_t)(iphase - 1) * num_phase_ispec] - 1;
        int ir = irreg[e];
        const float* s = &dat[(size_t)e * NG3];
        // gradient (get_gradient_element / _regular)
        for (int k = 0; k < NG; ++k)
          for (int j = 0; j < NG; ++j)
            for (int i = 0; i < NG; ++i) {
              int L = i + NG * j + NG2 * k;
              float t1 = 0, t2 = 0, t3 = 0;
              for (int l = 0; l < NG; ++l) t1 = t1 + s[l + NG * j + NG2 * k] * hpT[l][i];
              for (int l = 0; l < NG; ++l) t2 = t2 + s[i + NG * l + NG2 * k] * hpT[l][j];
              for (int l = 0; l < NG; ++l) t3 = t3 + s[i + NG * j + NG2 * l] * hpT[l][k];
              if (ir) {
                // faithful: the reference passes xixstore(:,:,:,ispec)
                size_t Lj = (size_t)((faithful ? e + 1 : ir) - 1) * NG3 + L;
                dx[L] = t1 * xix[Lj] + t2 * etax[Lj] + t3 * gamx[Lj];
                dy[L] = t1 * xiy[Lj] + t2 * etay[Lj] + t3 * gamy[Lj];
                dz[L] = t1 * xiz[Lj] + t2 * etaz[Lj] + t3 * gamz[Lj];
              } else if (faithful) {
                dx[L] = t1 * xix_regular + t2 * xix_regular + t3 * xix_regular;
                dy[L] = dx[L];
                dz[L] = dx[L];
              } else {
                dx[L] = t1 * xix_regular;
                dy[L] = t2 * xix_regular;
                dz[L] = t3 * xix_regular;
              }
            }
        for (int L = 0; L < NG3; ++L) {
          float xixl, xiyl, xizl, etaxl, etayl, etazl, gxl, gyl, gzl, jl;
          if (ir) {
            size_t Lj = (size_t)(ir - 1) * NG3 + L;
            xixl = xix[Lj]; xiyl = xiy[Lj]; xizl = xiz[Lj];
            etaxl = etax[Lj]; etayl = etay[Lj]; etazl = etaz[Lj];
            gxl = gamx[Lj]; gyl = gamy[Lj]; gzl = gamz[Lj];
            jl = jacobian[Lj];
          } else if (faithful) {
            xixl = xiyl = xizl = etaxl = etayl = etazl = gxl = gyl = gzl = xix_regular;
            jl = jacobian_regular;
          } else {
            xixl = etayl = gzl = xix_regular;
            xiyl = xizl = etaxl = etazl = gxl = gyl = 0.0f;
            jl = jacobian_regular;
          }
          if (is_sph) {
            const float* r = &rot[((size_t)e * NG3 + L) * 3];
            float rd = r[0] * dx[L] + r[1] * dy[L] + r[2] * dz[L];
            s1[L] = ((cv - ch) * (r[0] * xixl + r[1] * xiyl + r[2] * xizl) * rd +
                     ch * (xixl * dx[L] + xiyl * dy[L] + xizl * dz[L])) * jl;
            s2[L] = ((cv - ch) * (r[0] * etaxl + r[1] * etayl + r[2] * etazl) * rd +
                     ch * (etaxl * dx[L] + etayl * dy[L] + etazl * dz[L])) * jl;
            s3[L] = ((cv - ch) * (r[0] * gxl + r[1] * gyl + r[2] * gzl) * rd +
                     ch * (gxl * dx[L] + gyl * dy[L] + gzl * dz[L])) * jl;
          } else {
            s1[L] = ((cv - ch) * xizl * dz[L] + ch * (xixl * dx[L] + xiyl * dy[L] + xizl * dz[L])) * jl;
            s2[L] = ((cv - ch) * etazl * dz[L] + ch * (etaxl * dx[L] + etayl * dy[L] + etazl * dz[L])) * jl;
            s3[L] = ((cv - ch) * gzl * dz[L] + ch * (gxl * dx[L] + gyl * dy[L] + gzl * dz[L])) * jl;
          }
        }
        for (int k = 0; k < NG; ++k)
          for (int j = 0; j < NG; ++j)
            for (int i = 0; i < NG; ++i) {
              int L = i + NG * j + NG2 * k;
              float n1 = 0, n2 = 0, n3 = 0;
              for (int l = 0; l < NG; ++l) {
                n1 = n1 + s1[l + NG * j + NG2 * k] * hpw[l][i];
                n2 = n2 + s2[i + NG * l + NG2 * k] * hpw[l][j];
                n3 = n3 + s3[i + NG * j + NG2 * l] * hpw[l][k];
              }
              float f1 = wyz[j][k], f2 = wyz[i][k], f3 = wyz[i][j];
              int gI = ibool[(size_t)e * NG3 + L] - 1;
              ddat[gI] = ddat[gI] - (f1 * n1 + f2 * n2 + f3 * n3);
            }
      }
    }
    for (int gI = 0; gI < nglob; ++gI) {
      ddat[gI] = ddat[gI] * rvol[gI];
      dat_glob[gI] = dat_glob[gI] + ddat[gI];
    }
    for (int e = 0; e < nspec; ++e)
      if (is_cpml && is_cpml[e])
        for (int L = 0; L < NG3; ++L) dat_glob[ibool[(size_t)e * NG3 + L] - 1] = dat_bak[(size_t)e * NG3 + L];
    for (size_t L = 0; L < nloc; ++L) dat[L] = dat_glob[ibool[L] - 1];
  }
  if (dat_out) for (size_t L = 0; L < nloc; ++L) dat_out[L] = dat[L];
  if (dat_glob_out) for (int gI = 0; gI < nglob; ++gI) dat_glob_out[gI] = dat_glob[gI];
  return ntstep;
}

}  // extern "C"
