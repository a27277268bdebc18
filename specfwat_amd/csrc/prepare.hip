// Fortran-callable device preparation, transfer and clean-up entry points.
//
// Names follow SPECFEM3D's GPU C-ABI that SpecFWAT's prepare_GPU /
// FinalizeSpecfem path reaches (prepare_timerun_fwat.f90:120,
// specfem_api.f90:303).  Calling convention as in the reference's own GPU
// code (smooth_sem_pde.cu:298-316): lowercase name + '_', every argument by
// reference, Fortran column-major arrays, 1-based element/point indices,
// logical flags as int (any non-zero is true).  The argument lists below are
// this library's own (SPECFEM's are not in the SpecFWAT tree); they carry
// the same arrays under the same names.
#include "sem_device.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

uint16_t f2bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf/nan
  uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(r >> 16);
}

uint16_t f2h16(float f) {
  _Float16 h = (_Float16)f;
  uint16_t x;
  std::memcpy(&x, &h, 2);
  return x;
}

uint16_t f2lp(float f, int sp) { return sp == SEM_STATIC_BF16 ? f2bf16(f) : f2h16(f); }

// Element visiting order of SPECFEM's CPU element loop: phase_ispec_inner_elastic
// (num_phase_ispec, 2), 1-based, outer elements (iphase 1) then inner ones
// (iphase 2) -- the loop shape of smooth.f90:243-247.  False (visit empty)
// unless the two lists together hold every element exactly once.
bool phase_visit_order(int nspec, int num_phase_ispec, const int* phase_ispec_inner, int nspec_outer,
                       int nspec_inner, std::vector<int>& visit) {
  visit.clear();
  if (nspec_outer < 0 || nspec_inner < 0 || nspec_outer + nspec_inner != nspec ||
      std::max(nspec_outer, nspec_inner) > num_phase_ispec)
    return false;
  std::vector<int> seen(nspec, 0);
  visit.reserve(nspec);
  for (int iphase = 0; iphase < 2; ++iphase) {
    const int n = iphase == 0 ? nspec_outer : nspec_inner;
    for (int q = 0; q < n; ++q) {
      const int e = phase_ispec_inner[(size_t)iphase * num_phase_ispec + q] - 1;
      if (e < 0 || e >= nspec || seen[e]++) {
        visit.clear();
        return false;
      }
      visit.push_back(e);
    }
  }
  return true;
}

// Host copy of the contribution owners of a prepared mesh (see build_assembly_csr).
std::vector<int> contribution_owners(SemMesh* m) {
  const size_t nloc = (size_t)m->nspec * 125, nfp = (size_t)m->nfaces * 25;
  const size_t nsp = (size_t)m->nsources * 125, nrp = (size_t)m->nrec * 125;
  std::vector<int> owner(nloc + nfp + nsp + nrp);
  hipStream_t s = m->stream;
  HIP_CHECK(hipMemcpyAsync(owner.data(), m->d_ibool, nloc * sizeof(int), hipMemcpyDeviceToHost, s));
  if (nfp) HIP_CHECK(hipMemcpyAsync(owner.data() + nloc, m->d_face_iglob, nfp * sizeof(int), hipMemcpyDeviceToHost, s));
  if (nsp) HIP_CHECK(hipMemcpyAsync(owner.data() + nloc + nfp, m->d_src_iglob, nsp * sizeof(int), hipMemcpyDeviceToHost, s));
  if (nrp)
    HIP_CHECK(hipMemcpyAsync(owner.data() + nloc + nfp + nsp, m->d_rec_iglob, nrp * sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return owner;
}

// The device CSR addresses element-local terms by their place in the element rows of
// floc: with grouped rows q = ispec * 125 + L becomes ispec * 125 + sem_floc_slot(L).
void upload_assembly_csr(SemMesh* m, const std::vector<int>& off, std::vector<int> idx) {
  const int nloc = m->nspec * 125;
  if (m->floc_grouped)
    for (int& q : idx)
      if (q < nloc) q = q - q % 125 + sem_floc_slot(q % 125);
  dev_free(m->d_asm_off);
  dev_free(m->d_asm_idx);
  m->n_contrib = idx.size();
  m->d_asm_off = dev_upload(off.data(), off.size(), m->stream);
  m->d_asm_idx = dev_upload(idx.data(), idx.size(), m->stream);
  HIP_CHECK(hipStreamSynchronize(m->stream));
}

}  // namespace

// Gather-assembly CSR.  owner[q] is the global point of contribution q: the
// element-local points (125 per element, element-major) first, then the Stacey
// face points, the source points and the receiver points.  Every point's row
// lists its element terms in the order a sequential loop over the elements in
// `visit` order adds them, then its face, source and receiver terms.  Summing
// a row in this order makes the assembled force independent of scheduling.
void sem_build_assembly_csr(int nglob, int nspec, const std::vector<int>& owner, const std::vector<int>& visit,
                        std::vector<int>& off, std::vector<int>& idx) {
  const size_t nloc = (size_t)nspec * 125;
  if (owner.size() >= (size_t)INT32_MAX) SEM_FAIL("too many contributions for 32-bit CSR indices");
  off.assign((size_t)nglob + 1, 0);
  for (size_t q = 0; q < owner.size(); ++q) off[(size_t)owner[q] + 1]++;
  for (int g = 0; g < nglob; ++g) off[(size_t)g + 1] += off[g];
  idx.assign(owner.size(), 0);
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (int e : visit)
    for (int L = 0; L < 125; ++L) {
      const size_t q = (size_t)e * 125 + L;
      idx[fill[owner[q]]++] = (int)q;
    }
  for (size_t q = nloc; q < owner.size(); ++q) idx[fill[owner[q]]++] = (int)q;
}

extern "C" {

// initialize_GPU (initialize_simulation.f90:300): pick the device of this rank.
void initialize_gpu_device_(int* runtime_type, int* myrank, int* ncuda_devices) {
  (void)runtime_type;
  int n = 0;
  HIP_CHECK(hipGetDeviceCount(&n));
  if (n <= 0) SEM_FAIL("no HIP device visible");
  HIP_CHECK(hipSetDevice(*myrank % n));
  *ncuda_devices = n;
}

// Mesh constants, geometry, GLL tables, Stacey faces, sources and receivers.
//   h_xix..h_gammaz, h_jacobian : (NGLLX,NGLLY,NGLLZ,NSPEC_IRREGULAR)
//   h_ibool                     : (NGLLX,NGLLY,NGLLZ,NSPEC_AB), 1-based
//   h_irregular_element_number  : (NSPEC_AB), 0 = regular
//   h_hprime_xx, h_hprimewgll_xx, h_wgllwgll_xy/xz/yz : (NGLLX,NGLLX)
//   h_abs_boundary_ispec (nfaces), h_abs_boundary_ijk (3,NGLLSQUARE,nfaces),
//   h_abs_boundary_normal (3,NGLLSQUARE,nfaces), h_abs_boundary_jacobian2Dw (NGLLSQUARE,nfaces)
//   h_sourcearrays (NDIM,NGLLX,NGLLY,NGLLZ,NSOURCES), h_ispec_selected_source (NSOURCES)
//   h_ispec_selected_rec (nrec), h_hlagrange_rec (NGLLX*NGLLY*NGLLZ, nrec)
void prepare_constants_device_(long* Mesh_pointer, int* h_NGLLX, int* NSPEC_AB, int* NGLOB_AB,
                               int* NSPEC_IRREGULAR, realw* h_xix, realw* h_xiy, realw* h_xiz,
                               realw* h_etax, realw* h_etay, realw* h_etaz, realw* h_gammax,
                               realw* h_gammay, realw* h_gammaz, realw* h_jacobian,
                               realw* xix_regular, realw* jacobian_regular, int* h_ibool,
                               int* h_irregular_element_number, realw* h_hprime_xx,
                               realw* h_hprimewgll_xx, realw* h_wgllwgll_xy, realw* h_wgllwgll_xz,
                               realw* h_wgllwgll_yz, int* ABSORBING_CONDITIONS,
                               int* h_num_abs_boundary_faces, int* h_abs_boundary_ispec,
                               int* h_abs_boundary_ijk, realw* h_abs_boundary_normal,
                               realw* h_abs_boundary_jacobian2Dw, int* NSOURCES,
                               realw* h_sourcearrays, int* h_ispec_selected_source, int* nrec,
                               int* h_ispec_selected_rec, realw* h_hlagrange_rec, int* NSTEP,
                               int* SIMULATION_TYPE, int* SAVE_FORWARD) {
  if (*h_NGLLX != SEM_NGLL) SEM_FAIL("NGLLX = %d not supported (built for 5)", *h_NGLLX);
  SemMesh* m = new SemMesh();
  *Mesh_pointer = reinterpret_cast<long>(m);
  HIP_CHECK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  HIP_CHECK(hipEventCreate(&m->ev_a));
  HIP_CHECK(hipEventCreate(&m->ev_b));
  hipStream_t s = m->stream;
  m->nspec = *NSPEC_AB;
  m->nglob = *NGLOB_AB;
  m->nspec_irregular = *NSPEC_IRREGULAR;
  m->xix_regular = *xix_regular;
  m->jacobian_regular = *jacobian_regular;
  m->simulation_type = *SIMULATION_TYPE;
  m->save_forward = *SAVE_FORWARD != 0;
  m->nstep = *NSTEP;
  const size_t nloc = (size_t)m->nspec * SEM_NGLL3;

  std::vector<int> ib(nloc);
  for (size_t L = 0; L < nloc; ++L) {
    int g = h_ibool[L] - 1;
    if (g < 0 || g >= m->nglob) SEM_FAIL("ibool(%zu) = %d outside 1..NGLOB_AB=%d", L, g + 1, m->nglob);
    ib[L] = g;
  }
  {
    // library numbering: a point is "interior" when exactly one element-local slot
    // refers to it and no Stacey face, source or receiver term does
    std::vector<uint8_t> touch(m->nglob, 0);  // element-local slots, saturating at 2
    for (size_t L = 0; L < nloc; ++L) touch[ib[L]] = (uint8_t)std::min(2, touch[ib[L]] + 1);
    const int nf = *ABSORBING_CONDITIONS ? *h_num_abs_boundary_faces : 0;
    for (int f = 0; f < nf; ++f) {
      const int e = h_abs_boundary_ispec[f] - 1;
      if (e < 0 || e >= m->nspec) SEM_FAIL("abs_boundary_ispec(%d) = %d out of range", f + 1, e + 1);
      for (int q = 0; q < 25; ++q) {
        const size_t fq = (size_t)f * 25 + q;
        const int i = h_abs_boundary_ijk[fq * 3] - 1, j = h_abs_boundary_ijk[fq * 3 + 1] - 1,
                  k = h_abs_boundary_ijk[fq * 3 + 2] - 1;
        if (i < 0 || i > 4 || j < 0 || j > 4 || k < 0 || k > 4) SEM_FAIL("abs_boundary_ijk out of range");
        touch[ib[(size_t)e * 125 + i + 5 * j + 25 * k]] = 2;
      }
    }
    const auto whole_element = [&](int e) {
      if (e < 0 || e >= m->nspec) return;  // checked with a message below
      for (int L = 0; L < 125; ++L) touch[ib[(size_t)e * 125 + L]] = 2;
    };
    for (int q = 0; q < *NSOURCES; ++q) whole_element(h_ispec_selected_source[q] - 1);
    for (int q = 0; q < *nrec; ++q) whole_element(h_ispec_selected_rec[q] - 1);
    int n_sh = 0;
    std::vector<int> perm(m->nglob, -1);
    for (int g = 0; g < m->nglob; ++g)
      if (touch[g] != 1) perm[g] = n_sh++;
    int k = n_sh;
    for (size_t L = 0; L < nloc; ++L)
      if (touch[ib[L]] == 1) perm[ib[L]] = k++;
    if (k != m->nglob) SEM_FAIL("point numbering: %d of %d points placed", k, m->nglob);
    m->n_shared = n_sh;
    if (n_sh < m->nglob) {
      uint64_t h = 1469598103934665603ull;  // FNV-1a over the numbering
      for (int g = 0; g < m->nglob; ++g) { h ^= (uint32_t)perm[g]; h *= 1099511628211ull; }
      m->perm_hash = h;
      for (size_t L = 0; L < nloc; ++L) ib[L] = perm[ib[L]];
      m->h_perm.swap(perm);
    }
  }
  m->d_ibool = dev_upload(ib.data(), nloc, s);
  std::vector<int> irr(h_irregular_element_number, h_irregular_element_number + m->nspec);
  for (int e = 0; e < m->nspec; ++e)
    if (irr[e] < 0 || irr[e] > m->nspec_irregular)
      SEM_FAIL("irregular_element_number(%d) = %d outside 0..%d", e + 1, irr[e], m->nspec_irregular);
  m->d_irregular = dev_upload(irr.data(), (size_t)m->nspec, s);
  if (m->nspec_irregular > 0) {
    const size_t ni = (size_t)m->nspec_irregular;
    std::vector<realw> j10(ni * 1250);
    const realw* src[10] = {h_xix, h_xiy, h_xiz, h_etax, h_etay, h_etaz, h_gammax, h_gammay, h_gammaz, h_jacobian};
    for (size_t e = 0; e < ni; ++e)
      for (int c = 0; c < 10; ++c)
        std::memcpy(&j10[(e * 10 + c) * 125], &src[c][e * 125], 125 * sizeof(realw));
    m->d_jinv10 = dev_upload(j10.data(), j10.size(), s);
  }
  for (int q = 0; q < 25; ++q) {
    m->h_tables[q] = h_hprime_xx[q];
    m->h_tables[25 + q] = h_hprimewgll_xx[q];
    m->h_tables[50 + q] = h_wgllwgll_yz[q];
    m->h_tables[75 + q] = h_wgllwgll_xz[q];
    m->h_tables[100 + q] = h_wgllwgll_xy[q];
  }
  m->d_tables = dev_upload(m->h_tables, 125, s);

  // Stacey faces: the per-point geometry; material coefficients are filled by
  // prepare_fields_elastic_device_ (rho_vp, rho_vs live there).
  m->absorbing = *ABSORBING_CONDITIONS != 0;
  m->nfaces = m->absorbing ? *h_num_abs_boundary_faces : 0;
  if (m->nfaces > 0) {
    const size_t nfp = (size_t)m->nfaces * 25;
    std::vector<int> fi(nfp);
    std::vector<realw> fc(nfp * 8, 0.0f);
    for (int f = 0; f < m->nfaces; ++f) {
      int e = h_abs_boundary_ispec[f] - 1;
      if (e < 0 || e >= m->nspec) SEM_FAIL("abs_boundary_ispec(%d) = %d out of range", f + 1, e + 1);
      for (int q = 0; q < 25; ++q) {
        size_t fq = (size_t)f * 25 + q;
        int i = h_abs_boundary_ijk[fq * 3] - 1, j = h_abs_boundary_ijk[fq * 3 + 1] - 1,
            k = h_abs_boundary_ijk[fq * 3 + 2] - 1;
        if (i < 0 || i > 4 || j < 0 || j > 4 || k < 0 || k > 4) SEM_FAIL("abs_boundary_ijk out of range");
        fi[fq] = ib[(size_t)e * 125 + i + 5 * j + 25 * k];
        fc[fq * 8 + 0] = h_abs_boundary_normal[fq * 3];
        fc[fq * 8 + 1] = h_abs_boundary_normal[fq * 3 + 1];
        fc[fq * 8 + 2] = h_abs_boundary_normal[fq * 3 + 2];
        fc[fq * 8 + 3] = h_abs_boundary_jacobian2Dw[fq];
        // 4, 5: local index, replaced by rho_vp / rho_vs in prepare_fields_elastic_device_
        int L = e * 125 + i + 5 * j + 25 * k;
        std::memcpy(&fc[fq * 8 + 6], &L, sizeof(int));
      }
    }
    m->d_face_iglob = dev_upload(fi.data(), nfp, s);
    m->d_face_coef = dev_upload(fc.data(), fc.size(), s);
  }

  // sources
  m->nsources = *NSOURCES;
  if (m->nsources > 0) {
    std::vector<int> si((size_t)m->nsources * 125);
    for (int q = 0; q < m->nsources; ++q) {
      int e = h_ispec_selected_source[q] - 1;
      if (e < 0 || e >= m->nspec) SEM_FAIL("ispec_selected_source(%d) = %d out of range", q + 1, e + 1);
      for (int L = 0; L < 125; ++L) si[(size_t)q * 125 + L] = ib[(size_t)e * 125 + L];
    }
    m->d_src_iglob = dev_upload(si.data(), si.size(), s);
    m->d_sourcearrays = dev_upload(h_sourcearrays, (size_t)m->nsources * 375, s);
    m->d_stf_step = dev_alloc<double>((size_t)m->nsources);
  }
  // receivers
  m->nrec = *nrec;
  if (m->nrec > 0) {
    std::vector<int> ri((size_t)m->nrec * 125);
    for (int q = 0; q < m->nrec; ++q) {
      int e = h_ispec_selected_rec[q] - 1;
      if (e < 0 || e >= m->nspec) SEM_FAIL("ispec_selected_rec(%d) = %d out of range", q + 1, e + 1);
      for (int L = 0; L < 125; ++L) ri[(size_t)q * 125 + L] = ib[(size_t)e * 125 + L];
    }
    m->d_rec_iglob = dev_upload(ri.data(), ri.size(), s);
    m->d_rec_h = dev_upload(h_hlagrange_rec, (size_t)m->nrec * 125, s);
    m->d_seismograms = dev_alloc<realw>((size_t)std::max(m->nstep, 1) * m->nrec * 3);
    HIP_CHECK(hipMemsetAsync(m->d_seismograms, 0, sizeof(realw) * (size_t)std::max(m->nstep, 1) * m->nrec * 3, s));
    m->d_adj_step = dev_alloc<realw>((size_t)m->nrec * 3);
  }

  // Gather-assembly CSR, element terms in element order until the phase lists
  // (prepare_fields_elastic_device_) give the element loop's visiting order.
  {
    const size_t nfp = (size_t)m->nfaces * 25;
    std::vector<int> owner = contribution_owners(m);
    std::vector<int> visit(m->nspec), off, idx;
    for (int e = 0; e < m->nspec; ++e) visit[e] = e;
    sem_build_assembly_csr(m->nglob, m->nspec, owner, visit, off, idx);
    upload_assembly_csr(m, off, idx);
    m->d_floc = dev_alloc<realw>(nloc * 3);
    if (nfp) m->d_face_scratch = dev_alloc<realw>(nfp * 3);
    HIP_CHECK(hipStreamSynchronize(s));
  }
  HIP_CHECK(hipStreamSynchronize(s));
}

// Elastic fields and materials.
//   rmassx/y/z : (NGLOB_AB), already inverted (1/(M + dt/2 C_stacey))
//   rho_vp, rho_vs, kappav, muv, c11store..c66store : (NGLLX,NGLLY,NGLLZ,NSPEC_AB)
//   phase_ispec_inner_elastic : (num_phase_ispec_elastic, 2), 1-based
void prepare_fields_elastic_device_(long* Mesh_pointer, realw* rmassx, realw* rmassy,
                                    realw* rmassz, realw* rho_vp, realw* rho_vs, realw* kappav,
                                    realw* muv, int* ANISOTROPY, realw* c11store, realw* c12store,
                                    realw* c13store, realw* c14store, realw* c15store,
                                    realw* c16store, realw* c22store, realw* c23store,
                                    realw* c24store, realw* c25store, realw* c26store,
                                    realw* c33store, realw* c34store, realw* c35store,
                                    realw* c36store, realw* c44store, realw* c45store,
                                    realw* c46store, realw* c55store, realw* c56store,
                                    realw* c66store, int* num_phase_ispec_elastic,
                                    int* phase_ispec_inner_elastic, int* nspec_outer_elastic,
                                    int* nspec_inner_elastic, int* COMPUTE_AND_STORE_STRAIN) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  hipStream_t s = m->stream;
  const size_t nloc = (size_t)m->nspec * 125, n3 = (size_t)m->nglob * 3;
  std::vector<realw> rm(n3);
  for (int g = 0; g < m->nglob; ++g) {
    const size_t p = 3 * (size_t)sem_point(m, g);
    rm[p] = rmassx[g];
    rm[p + 1] = rmassy[g];
    rm[p + 2] = rmassz[g];
  }
  m->d_rmass_inv = dev_upload(rm.data(), n3, s);
  m->anisotropy = *ANISOTROPY != 0;
  m->d_kappa = dev_upload(kappav, nloc, s);
  m->d_mu = dev_upload(muv, nloc, s);
  if (m->anisotropy) {
    const realw* c[21] = {c11store, c12store, c13store, c14store, c15store, c16store, c22store,
                          c23store, c24store, c25store, c26store, c33store, c34store, c35store,
                          c36store, c44store, c45store, c46store, c55store, c56store, c66store};
    std::vector<realw> cij(nloc * 21);
    for (size_t e = 0; e < (size_t)m->nspec; ++e)
      for (int q = 0; q < 21; ++q)
        std::memcpy(&cij[(e * 21 + q) * 125], &c[q][e * 125], 125 * sizeof(realw));
    m->d_cij = dev_upload(cij.data(), cij.size(), s);
  }
  if (m->nfaces > 0) {
    const size_t nfp = (size_t)m->nfaces * 25;
    std::vector<realw> fc(nfp * 8);
    HIP_CHECK(hipMemcpyAsync(fc.data(), m->d_face_coef, fc.size() * sizeof(realw), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    for (size_t fq = 0; fq < nfp; ++fq) {
      int L;
      std::memcpy(&L, &fc[fq * 8 + 6], sizeof(int));
      fc[fq * 8 + 4] = rho_vp[L];
      fc[fq * 8 + 5] = rho_vs[L];
      fc[fq * 8 + 6] = 0.0f;
    }
    HIP_CHECK(hipMemcpyAsync(m->d_face_coef, fc.data(), fc.size() * sizeof(realw), hipMemcpyHostToDevice, s));
  }
  m->num_phase_ispec = *num_phase_ispec_elastic;
  m->nspec_outer = *nspec_outer_elastic;
  m->nspec_inner = *nspec_inner_elastic;
  if (m->num_phase_ispec > 0) {
    std::vector<int> ph((size_t)m->num_phase_ispec * 2);
    for (size_t q = 0; q < ph.size(); ++q) ph[q] = phase_ispec_inner_elastic[q] - 1;
    m->d_phase_ispec = dev_upload(ph.data(), ph.size(), s);
    // element terms of every point in the order of the phase loop (outer, then
    // inner elements); lists that do not cover every element once keep element order
    std::vector<int> visit;
    if (phase_visit_order(m->nspec, m->num_phase_ispec, phase_ispec_inner_elastic, m->nspec_outer, m->nspec_inner,
                          visit)) {
      bool identity = true;
      for (int e = 0; e < m->nspec && identity; ++e) identity = visit[e] == e;
      if (!identity) m->h_visit = visit;
      if (!identity) {
        std::vector<int> off, idx;
        sem_build_assembly_csr(m->nglob, m->nspec, contribution_owners(m), visit, off, idx);
        upload_assembly_csr(m, off, idx);
      }
    }
  }
  m->compute_and_store_strain = *COMPUTE_AND_STORE_STRAIN != 0;
  m->d_displ = dev_alloc<realw>(n3);
  m->d_veloc = dev_alloc<realw>(n3);
  m->d_accel = dev_alloc<realw>(n3);
  HIP_CHECK(hipMemsetAsync(m->d_displ, 0, n3 * sizeof(realw), s));
  HIP_CHECK(hipMemsetAsync(m->d_veloc, 0, n3 * sizeof(realw), s));
  HIP_CHECK(hipMemsetAsync(m->d_accel, 0, n3 * sizeof(realw), s));
  if (m->compute_and_store_strain) {
    m->d_epsilon = dev_alloc<realw>(nloc * 6);
    HIP_CHECK(hipMemsetAsync(m->d_epsilon, 0, nloc * 6 * sizeof(realw), s));
  }
  // Stacey store for the forward run that feeds a later reconstruction
  if (m->absorbing && m->nfaces > 0 && (m->save_forward || m->simulation_type == 3)) {
    m->stacey_store_steps = (size_t)m->nstep;
    size_t n = m->stacey_store_steps * m->nfaces * 25 * 3;
    m->d_stacey_store = dev_alloc<realw>(n);
  }
  HIP_CHECK(hipStreamSynchronize(s));
}

// Adjoint-run fields: reconstructed b_ wavefield, strain stores and kernels.
//   rhostore : (NGLLX,NGLLY,NGLLZ,NSPEC_AB)
void prepare_fields_elastic_adj_dev_(long* Mesh_pointer, int* size, int* COMPUTE_AND_STORE_STRAIN,
                                     realw* rhostore, int* ANISOTROPIC_KL,
                                     int* APPROXIMATE_HESS_KL) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  hipStream_t s = m->stream;
  const size_t n3 = (size_t)*size, nloc = (size_t)m->nspec * 125;
  if (n3 != (size_t)m->nglob * 3) SEM_FAIL("prepare_fields_elastic_adj_dev_: size %zu != 3*NGLOB_AB", n3);
  m->d_b_displ = dev_alloc<realw>(n3);
  m->d_b_veloc = dev_alloc<realw>(n3);
  m->d_b_accel = dev_alloc<realw>(n3);
  HIP_CHECK(hipMemsetAsync(m->d_b_displ, 0, n3 * sizeof(realw), s));
  HIP_CHECK(hipMemsetAsync(m->d_b_veloc, 0, n3 * sizeof(realw), s));
  HIP_CHECK(hipMemsetAsync(m->d_b_accel, 0, n3 * sizeof(realw), s));
  m->d_rho = dev_upload(rhostore, nloc, s);
  m->anisotropic_kl = *ANISOTROPIC_KL != 0;
  m->approximate_hess_kl = *APPROXIMATE_HESS_KL != 0;
  if (*COMPUTE_AND_STORE_STRAIN) {
    m->compute_and_store_strain = 1;
    if (!m->d_epsilon) m->d_epsilon = dev_alloc<realw>(nloc * 6);
    m->d_b_epsilon = dev_alloc<realw>(nloc * 6);
  }
  m->d_rho_kl_glob = dev_alloc<double>((size_t)m->nglob);
  m->d_floc_b = dev_alloc<realw>(nloc * 3);
  if (m->anisotropic_kl) {
    if (!m->anisotropy) SEM_FAIL("ANISOTROPIC_KL requires ANISOTROPY");
    m->d_cijkl_kl = dev_alloc<double>(nloc * 21);
  } else {
    m->d_mu_kl = dev_alloc<double>(nloc);
    m->d_kappa_kl = dev_alloc<double>(nloc);
  }
  if (m->approximate_hess_kl) m->d_hess_kl_glob = dev_alloc<double>((size_t)m->nglob);
  sem_zero_kernels(m);
  HIP_CHECK(hipStreamSynchronize(s));
}

// Adjoint sources: nadj_rec_local receivers share the receiver interpolants
// given to prepare_constants_device_.
void prepare_sim2_or_3_const_device_(long* Mesh_pointer, int* nadj_rec_local, int* NTSTEP_BETWEEN_READ_ADJSRC) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (*nadj_rec_local != m->nrec) SEM_FAIL("nadj_rec_local (%d) must equal nrec (%d)", *nadj_rec_local, m->nrec);
  (void)NTSTEP_BETWEEN_READ_ADJSRC;
}

// Place of each local point L = i + 5j + 25k in the grouped element row
// (sem_floc_slot): slot[L] for L = 0..124.  Host only; no counterpart in the reference.
void fwat_floc_row_layout_(int* slot) {
  for (int L = 0; L < 125; ++L) slot[L] = sem_floc_slot(L);
}

// Layout of the element-local force rows of the gather path: 1 = grouped by reader
// (sem_floc_slot, the default), 0 = local-point order.  Rebuilds the device CSR; results
// are the same bit for bit.  No counterpart in the reference.
void fwat_set_floc_layout_(long* Mesh_pointer, int* grouped) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  const int g = *grouped != 0;
  if (g == m->floc_grouped) return;
  m->floc_grouped = g;
  std::vector<int> visit = m->h_visit, off, idx;
  if (visit.empty()) {
    visit.resize(m->nspec);
    for (int e = 0; e < m->nspec; ++e) visit[e] = e;
  }
  HIP_CHECK(hipStreamSynchronize(m->stream));
  sem_build_assembly_csr(m->nglob, m->nspec, contribution_owners(m), visit, off, idx);
  upload_assembly_csr(m, off, idx);
}

// Store the static material/geometry arrays in BF16 (precision=1) or FP16
// with per-array scaling (precision=2); 0 restores FP32.  No counterpart in
// the reference: it implements the mixed-precision variant.
void fwat_set_static_precision_(long* Mesh_pointer, int* precision) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  hipStream_t s = m->stream;
  const int sp = *precision;
  if (sp != SEM_STATIC_FP32 && sp != SEM_STATIC_BF16 && sp != SEM_STATIC_FP16)
    SEM_FAIL("unknown static precision %d", sp);
  if (sp != SEM_STATIC_FP32 && m->anisotropy)
    SEM_FAIL("reduced-precision statics are implemented for the isotropic path only");
  m->static_precision = sp;
  dev_free(m->d_kmu_lp);
  dev_free(m->d_kmu_base);
  dev_free(m->d_jinv10_lp);
  dev_free(m->d_jinv10_base);
  if (sp == SEM_STATIC_FP32) return;
  // Encoding: value = fp32 mean over the element's 125 points + deviation
  // stored in bf16 (or fp16 scaled by a per-array factor that maps the
  // largest deviation to 1024).  A homogeneous element is exact; otherwise
  // the error is the deviation's rounding, not the value's.
  const auto encode = [sp](const float* v, size_t nelem, int ncomp, std::vector<float>& base,
                           std::vector<uint16_t>& dev, float& scale) {
    // v: [nelem][ncomp][125]
    base.assign(nelem * ncomp, 0.0f);
    std::vector<float> d(nelem * ncomp * 125);
    float dmax = 0.0f;
    for (size_t e = 0; e < nelem; ++e)
      for (int c = 0; c < ncomp; ++c) {
        const float* x = v + (e * ncomp + c) * 125;
        double sum = 0.0;
        for (int L = 0; L < 125; ++L) sum += x[L];
        const float b = (float)(sum / 125.0);
        base[e * ncomp + c] = b;
        for (int L = 0; L < 125; ++L) {
          const float r = x[L] - b;
          d[(e * ncomp + c) * 125 + L] = r;
          dmax = std::max(dmax, std::fabs(r));
        }
      }
    scale = (sp == SEM_STATIC_FP16 && dmax > 0.0f) ? dmax / 1024.0f : 1.0f;
    dev.resize(d.size());
    for (size_t q = 0; q < d.size(); ++q) dev[q] = f2lp(d[q] / scale, sp);
  };
  const size_t nloc = (size_t)m->nspec * 125;
  {
    std::vector<float> k(nloc), mu(nloc);
    HIP_CHECK(hipMemcpyAsync(k.data(), m->d_kappa, nloc * 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(mu.data(), m->d_mu, nloc * 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::vector<float> kb, mb;
    std::vector<uint16_t> kd, md;
    encode(k.data(), m->nspec, 1, kb, kd, m->static_scale_kappa);
    encode(mu.data(), m->nspec, 1, mb, md, m->static_scale_mu);
    std::vector<uint32_t> pk(nloc);
    for (size_t L = 0; L < nloc; ++L) pk[L] = (uint32_t)kd[L] | ((uint32_t)md[L] << 16);
    std::vector<float> base(2 * (size_t)m->nspec);
    for (int e = 0; e < m->nspec; ++e) { base[2 * e] = kb[e]; base[2 * e + 1] = mb[e]; }
    m->d_kmu_lp = dev_upload(pk.data(), nloc, s);
    m->d_kmu_base = dev_upload(base.data(), base.size(), s);
  }
  if (m->nspec_irregular > 0) {
    // jinv10 [nirr][10][125]: the nine inverse-Jacobian terms share one fp16
    // scale, the jacobian its own
    const size_t n = (size_t)m->nspec_irregular * 1250;
    std::vector<float> j(n);
    HIP_CHECK(hipMemcpyAsync(j.data(), m->d_jinv10, n * 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::vector<float> g9((size_t)m->nspec_irregular * 1125), jac((size_t)m->nspec_irregular * 125);
    for (size_t e = 0; e < (size_t)m->nspec_irregular; ++e) {
      std::memcpy(&g9[e * 1125], &j[e * 1250], 1125 * 4);
      std::memcpy(&jac[e * 125], &j[e * 1250 + 1125], 125 * 4);
    }
    std::vector<float> gb, jb;
    std::vector<uint16_t> gd, jd;
    encode(g9.data(), m->nspec_irregular, 9, gb, gd, m->static_scale_jinv);
    encode(jac.data(), m->nspec_irregular, 1, jb, jd, m->static_scale_jac);
    std::vector<uint16_t> lp(n);
    std::vector<float> base((size_t)m->nspec_irregular * 10);
    for (size_t e = 0; e < (size_t)m->nspec_irregular; ++e) {
      std::memcpy(&lp[e * 1250], &gd[e * 1125], 1125 * 2);
      std::memcpy(&lp[e * 1250 + 1125], &jd[e * 125], 125 * 2);
      for (int c = 0; c < 9; ++c) base[e * 10 + c] = gb[e * 9 + c];
      base[e * 10 + 9] = jb[e];
    }
    m->d_jinv10_lp = dev_upload(lp.data(), n, s);
    m->d_jinv10_base = dev_upload(base.data(), base.size(), s);
  }
  HIP_CHECK(hipStreamSynchronize(s));
}

// Host-only: the gather-assembly CSR that prepare_constants_device_ and
// prepare_fields_elastic_device_ build for these arrays (no device work), so
// its per-point order can be checked without a GPU.  Arguments as there;
// outputs off (NGLOB_AB+1) and idx (contributions: 125 NSPEC_AB element-local
// points, then 25 per Stacey face, 125 per source, 125 per receiver; 0-based).
// Phase lists that do not cover every element once give element order.
void fwat_assembly_csr_(int* NSPEC_AB, int* NGLOB_AB, int* h_ibool, int* num_abs_boundary_faces,
                        int* h_abs_boundary_ispec, int* h_abs_boundary_ijk, int* NSOURCES, int* h_ispec_selected_source,
                        int* nrec, int* h_ispec_selected_rec, int* num_phase_ispec_elastic,
                        int* phase_ispec_inner_elastic, int* nspec_outer_elastic, int* nspec_inner_elastic, int* off,
                        int* idx) {
  const int nspec = *NSPEC_AB, nglob = *NGLOB_AB;
  const size_t nloc = (size_t)nspec * 125;
  std::vector<int> owner;
  owner.reserve(nloc + (size_t)*num_abs_boundary_faces * 25 + (size_t)(*NSOURCES + *nrec) * 125);
  for (size_t L = 0; L < nloc; ++L) {
    const int g = h_ibool[L] - 1;
    if (g < 0 || g >= nglob) SEM_FAIL("ibool(%zu) = %d outside 1..NGLOB_AB=%d", L, g + 1, nglob);
    owner.push_back(g);
  }
  const auto elem = [nspec](int e1, const char* what, int q) {
    if (e1 < 1 || e1 > nspec) SEM_FAIL("%s(%d) = %d out of range", what, q + 1, e1);
    return e1 - 1;
  };
  for (int f = 0; f < *num_abs_boundary_faces; ++f) {
    const int e = elem(h_abs_boundary_ispec[f], "abs_boundary_ispec", f);
    for (int q = 0; q < 25; ++q) {
      const int* ijk = &h_abs_boundary_ijk[((size_t)f * 25 + q) * 3];
      if (ijk[0] < 1 || ijk[0] > 5 || ijk[1] < 1 || ijk[1] > 5 || ijk[2] < 1 || ijk[2] > 5)
        SEM_FAIL("abs_boundary_ijk out of range");
      owner.push_back(owner[(size_t)e * 125 + (ijk[0] - 1) + 5 * (ijk[1] - 1) + 25 * (ijk[2] - 1)]);
    }
  }
  for (int q = 0; q < *NSOURCES; ++q) {
    const int e = elem(h_ispec_selected_source[q], "ispec_selected_source", q);
    for (int L = 0; L < 125; ++L) owner.push_back(owner[(size_t)e * 125 + L]);
  }
  for (int q = 0; q < *nrec; ++q) {
    const int e = elem(h_ispec_selected_rec[q], "ispec_selected_rec", q);
    for (int L = 0; L < 125; ++L) owner.push_back(owner[(size_t)e * 125 + L]);
  }
  std::vector<int> visit;
  if (!phase_visit_order(nspec, *num_phase_ispec_elastic, phase_ispec_inner_elastic, *nspec_outer_elastic,
                         *nspec_inner_elastic, visit)) {
    visit.resize(nspec);
    for (int e = 0; e < nspec; ++e) visit[e] = e;
  }
  std::vector<int> o, i;
  sem_build_assembly_csr(nglob, nspec, owner, visit, o, i);
  std::memcpy(off, o.data(), o.size() * sizeof(int));
  std::memcpy(idx, i.data(), i.size() * sizeof(int));
}

}  // extern "C"

// transfers of the forward / adjoint field (Fortran displ(3,NGLOB_AB) layout, the
// caller's point numbering; the device holds the library numbering, sem_point)
void sem_fields_to_device(SemMesh* m, int size, realw* const host[3], realw* const dev[3]) {
  if (size != 3 * m->nglob) SEM_FAIL("field transfer: size %d != 3*NGLOB_AB", size);
  const size_t n = (size_t)size;
  std::vector<realw> buf;
  for (int q = 0; q < 3; ++q) {
    const realw* src = host[q];
    if (!m->h_perm.empty()) {
      buf.resize(n);
      for (int g = 0; g < m->nglob; ++g) {
        const size_t p = 3 * (size_t)m->h_perm[g];
        buf[p] = src[3 * (size_t)g]; buf[p + 1] = src[3 * (size_t)g + 1]; buf[p + 2] = src[3 * (size_t)g + 2];
      }
      src = buf.data();
    }
    HIP_CHECK(hipMemcpyAsync(dev[q], src, n * sizeof(realw), hipMemcpyHostToDevice, m->stream));
    HIP_CHECK(hipStreamSynchronize(m->stream));
  }
}

void sem_fields_from_device(SemMesh* m, int size, realw* const host[3], realw* const dev[3]) {
  if (size != 3 * m->nglob) SEM_FAIL("field transfer: size %d != 3*NGLOB_AB", size);
  const size_t n = (size_t)size;
  std::vector<realw> buf(m->h_perm.empty() ? 0 : n);
  for (int q = 0; q < 3; ++q) {
    realw* dst = m->h_perm.empty() ? host[q] : buf.data();
    HIP_CHECK(hipMemcpyAsync(dst, dev[q], n * sizeof(realw), hipMemcpyDeviceToHost, m->stream));
    HIP_CHECK(hipStreamSynchronize(m->stream));
    if (!m->h_perm.empty())
      for (int g = 0; g < m->nglob; ++g) {
        const size_t p = 3 * (size_t)m->h_perm[g];
        host[q][3 * (size_t)g] = buf[p]; host[q][3 * (size_t)g + 1] = buf[p + 1]; host[q][3 * (size_t)g + 2] = buf[p + 2];
      }
  }
}

extern "C" {

void transfer_fields_el_to_device_(int* size, realw* displ, realw* veloc, realw* accel, long* Mesh_pointer) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  realw* const h[3] = {displ, veloc, accel};
  realw* const d[3] = {m->d_displ, m->d_veloc, m->d_accel};
  sem_fields_to_device(m, *size, h, d);
}

void transfer_fields_el_from_device_(int* size, realw* displ, realw* veloc, realw* accel, long* Mesh_pointer) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  realw* const h[3] = {displ, veloc, accel};
  realw* const d[3] = {m->d_displ, m->d_veloc, m->d_accel};
  sem_fields_from_device(m, *size, h, d);
}

void transfer_b_fields_to_device_(int* size, realw* b_displ, realw* b_veloc, realw* b_accel, long* Mesh_pointer) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (!m->d_b_displ) SEM_FAIL("transfer_b_fields_to_device_: adjoint fields not prepared");
  realw* const h[3] = {b_displ, b_veloc, b_accel};
  realw* const d[3] = {m->d_b_displ, m->d_b_veloc, m->d_b_accel};
  sem_fields_to_device(m, *size, h, d);
}

void transfer_b_fields_from_device_(int* size, realw* b_displ, realw* b_veloc, realw* b_accel, long* Mesh_pointer) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (!m->d_b_displ) SEM_FAIL("transfer_b_fields_from_device_: adjoint fields not prepared");
  realw* const h[3] = {b_displ, b_veloc, b_accel};
  realw* const d[3] = {m->d_b_displ, m->d_b_veloc, m->d_b_accel};
  sem_fields_from_device(m, *size, h, d);
}

void transfer_stacey_store_from_device_(long* Mesh_pointer, realw* h_store) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (!m->d_stacey_store) SEM_FAIL("no Stacey store allocated");
  size_t n = m->stacey_store_steps * m->nfaces * 75 * sizeof(realw);
  HIP_CHECK(hipMemcpyAsync(h_store, m->d_stacey_store, n, hipMemcpyDeviceToHost, m->stream));
  HIP_CHECK(hipStreamSynchronize(m->stream));
}

void transfer_stacey_store_to_device_(long* Mesh_pointer, realw* h_store) {
  SemMesh* m = sem_mesh(Mesh_pointer);
  if (!m->d_stacey_store) SEM_FAIL("no Stacey store allocated");
  size_t n = m->stacey_store_steps * m->nfaces * 75 * sizeof(realw);
  HIP_CHECK(hipMemcpyAsync(m->d_stacey_store, h_store, n, hipMemcpyHostToDevice, m->stream));
  HIP_CHECK(hipStreamSynchronize(m->stream));
}

// Hand the forward run's device-resident Stacey store to a later adjoint run
// on the same device (the in-HBM replacement of absorb_field.bin).  The
// source handle gives up ownership.
void fwat_move_stacey_store_(long* Mesh_pointer_from, long* Mesh_pointer_to) {
  SemMesh* a = sem_mesh(Mesh_pointer_from);
  SemMesh* b = sem_mesh(Mesh_pointer_to);
  if (!a->d_stacey_store) SEM_FAIL("fwat_move_stacey_store_: source has no store");
  if (a->nfaces != b->nfaces) SEM_FAIL("fwat_move_stacey_store_: face counts differ");
  if (a->stacey_store_steps < (size_t)b->nstep) SEM_FAIL("fwat_move_stacey_store_: store shorter than NSTEP");
  dev_free(b->d_stacey_store);
  b->d_stacey_store = a->d_stacey_store;
  b->stacey_store_steps = a->stacey_store_steps;
  a->d_stacey_store = nullptr;
  a->stacey_store_steps = 0;
}

// prepare_cleanup_device (called from FinalizeSpecfem, specfem_api.f90:303)
void prepare_cleanup_device_(long* Mesh_pointer, int* ACOUSTIC_SIMULATION, int* ELASTIC_SIMULATION,
                             int* NOISE_TOMOGRAPHY) {
  (void)ACOUSTIC_SIMULATION; (void)ELASTIC_SIMULATION; (void)NOISE_TOMOGRAPHY;
  SemMesh* m = reinterpret_cast<SemMesh*>(*Mesh_pointer);
  if (!m) return;
  HIP_CHECK(hipStreamSynchronize(m->stream));
  dev_free(m->d_ibool); dev_free(m->d_irregular); dev_free(m->d_jinv10); dev_free(m->d_tables);
  dev_free(m->d_jinv10_lp);
  dev_free(m->d_jinv10_base);
  dev_free(m->d_kmu_lp);
  dev_free(m->d_kmu_base);
  dev_free(m->d_brick_special); dev_free(m->d_brick_part0); dev_free(m->d_brick_part1);
  dev_free(m->d_halo_g); dev_free(m->d_halo_off); dev_free(m->d_halo_idx);
  dev_free(m->d_mu_part); dev_free(m->d_kappa_part); dev_free(m->d_cij_part); dev_free(m->d_rho_part);
  dev_free(m->d_slab_g);
  dev_free(m->d_kappa); dev_free(m->d_mu); dev_free(m->d_rho); dev_free(m->d_cij); dev_free(m->d_rmass_inv);
  dev_free(m->d_phase_ispec);
  dev_free(m->d_displ); dev_free(m->d_veloc); dev_free(m->d_accel);
  dev_free(m->d_b_displ); dev_free(m->d_b_veloc); dev_free(m->d_b_accel);
  dev_free(m->d_face_iglob); dev_free(m->d_face_coef); dev_free(m->d_stacey_store);
  dev_free(m->d_fk_veloc); dev_free(m->d_fk_tract);
  dev_free(m->d_src_iglob); dev_free(m->d_sourcearrays); dev_free(m->d_stf); dev_free(m->d_stf_step);
  dev_free(m->d_rec_iglob); dev_free(m->d_rec_h); dev_free(m->d_seismograms); dev_free(m->d_adj);
  dev_free(m->d_adj_step);
  dev_free(m->d_epsilon); dev_free(m->d_b_epsilon);
  dev_free(m->d_rho_kl_glob); dev_free(m->d_mu_kl); dev_free(m->d_kappa_kl); dev_free(m->d_cijkl_kl);
  dev_free(m->d_hess_kl_glob);
  dev_free(m->d_floc); dev_free(m->d_floc_b); dev_free(m->d_face_scratch);
  dev_free(m->d_asm_off); dev_free(m->d_asm_idx);
  HIP_CHECK(hipEventDestroy(m->ev_a));
  HIP_CHECK(hipEventDestroy(m->ev_b));
  HIP_CHECK(hipStreamDestroy(m->stream));
  delete m;
  *Mesh_pointer = 0;
}

}  // extern "C"
