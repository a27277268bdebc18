// Device-side state of one SEM simulation on one MI355X (gfx950).
//
// The Fortran side keeps an opaque `long Mesh_pointer` (SPECFEM's convention,
// kept by SpecFWAT in specfem_api.f90:303); it holds a `SemMesh*`.
//
// HBM layout (all arrays 0-based):
//   ibool            int32 [nspec][125]            local point L = i + 5j + 25k
//   kappa, mu        fp32 (or bf16/fp16) [nspec][125]
//   rho              fp32 [nspec][125]               (adjoint runs: rho kernel scale)
//   jinv10           fp32 [nspec_irregular][10][125] 9 inverse-Jacobian terms + jacobian, SoA
//                                                    per element so a 125-lane load is contiguous
//   cij              fp32 [nspec][21][125]           anisotropic runs
//   displ/veloc/accel  fp32 [nglob][3] (AoS, as Fortran displ(3,nglob))
//   rmass_inv        fp32 [nglob][3]                 1/(M + dt/2 C_stacey), per component
//   face_iglob       int32 [nfaces*25]
//   face_coef        fp32 [nfaces*25][8]             nx ny nz jw rho_vp rho_vs pad pad
//   stacey_store     fp32 [nstep][nfaces*25][3]      forward absorbed traction, replayed backwards
//   mu_kl, kappa_kl  fp64 [nspec][125]; rho_kl_glob fp64 [nglob]; cijkl_kl fp64 [nspec][21][125]
//   floc             fp32 [nspec][125][3]            element-local forces (gather assembly)
//   asm_off/asm_idx  int32 CSR: global point -> its local/face/source/receiver contributions
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

typedef float realw;

#define SEM_NGLL 5
#define SEM_NGLL2 25
#define SEM_NGLL3 125

#define HIP_CHECK(call)                                                                  \
  do {                                                                                   \
    hipError_t err_ = (call);                                                            \
    if (err_ != hipSuccess) {                                                            \
      std::fprintf(stderr, "specfwat_hip: HIP error '%s' at %s:%d in %s\n",              \
                   hipGetErrorString(err_), __FILE__, __LINE__, #call);                  \
      std::fflush(stderr);                                                               \
      std::exit(1);                                                                      \
    }                                                                                    \
  } while (0)

#define SEM_FAIL(...)                                                                    \
  do {                                                                                   \
    std::fprintf(stderr, "specfwat_hip: ");                                              \
    std::fprintf(stderr, __VA_ARGS__);                                                   \
    std::fprintf(stderr, "\n");                                                          \
    std::fflush(stderr);                                                                 \
    std::exit(1);                                                                        \
  } while (0)

// storage precision of the static material / geometry arrays
enum SemStaticPrecision { SEM_STATIC_FP32 = 0, SEM_STATIC_BF16 = 1, SEM_STATIC_FP16 = 2 };

struct SemMesh {
  int nspec = 0, nglob = 0, nspec_irregular = 0;
  realw xix_regular = 0, jacobian_regular = 0;
  int simulation_type = 1;
  int save_forward = 0;
  int anisotropy = 0;
  int anisotropic_kl = 0;
  int approximate_hess_kl = 0;
  int compute_and_store_strain = 0;
  int static_precision = SEM_STATIC_FP32;
  int regular_fast_path = 1;  // folded-table element pass on all-regular meshes
  int blocks_per_cu_fwd = 0;  // element-pass workgroups per CU (0: one per element)
  int blocks_per_cu_adj = 8;
  // reduced-precision statics hold value/scale (FP16 needs it for range)
  float static_scale_kappa = 1.0f, static_scale_mu = 1.0f;
  float static_scale_jinv = 1.0f, static_scale_jac = 1.0f;
  int nstep = 0;

  hipStream_t stream = nullptr;

  // geometry / indexing
  int* d_ibool = nullptr;
  int* d_irregular = nullptr;  // [nspec], 0 regular else 1-based
  realw* d_jinv10 = nullptr;
  void* d_jinv10_lp = nullptr;  // bf16/fp16 deviations from the per-element means [nspec_irr][10][125]
  float* d_jinv10_base = nullptr;  // fp32 per-element means [nspec_irr][10]
  realw* d_tables = nullptr;    // hprime[25], hprimewgll[25], wgllwgll_yz[25], _xz[25], _xy[25]
  realw h_tables[125];

  // materials
  realw* d_kappa = nullptr;
  realw* d_mu = nullptr;
  void* d_kmu_lp = nullptr;     // packed (kappa,mu) deviations in bf16/fp16 [nspec][125][2]
  float* d_kmu_base = nullptr;  // fp32 per-element means of kappa, mu [nspec][2]
  realw* d_rho = nullptr;
  realw* d_cij = nullptr;
  realw* d_rmass_inv = nullptr;  // [nglob][3]

  // inner/outer element phases (1-based element ids), as SPECFEM's phase_ispec_inner_elastic
  int num_phase_ispec = 0, nspec_outer = 0, nspec_inner = 0;
  int* d_phase_ispec = nullptr;  // 0-based [2][num_phase_ispec]

  // wavefields
  realw *d_displ = nullptr, *d_veloc = nullptr, *d_accel = nullptr;
  realw *d_b_displ = nullptr, *d_b_veloc = nullptr, *d_b_accel = nullptr;

  // Stacey
  int absorbing = 0;
  int nfaces = 0;
  int* d_face_iglob = nullptr;
  realw* d_face_coef = nullptr;
  realw* d_stacey_store = nullptr;
  size_t stacey_store_steps = 0;

  // point sources
  int nsources = 0;
  int* d_src_iglob = nullptr;      // [nsrc][125]
  realw* d_sourcearrays = nullptr;  // [nsrc][125][3]
  double* d_stf = nullptr;          // [nstep][nsrc]  (whole-loop path)
  double* d_stf_step = nullptr;     // [nsrc]         (per-step API)

  // receivers
  int nrec = 0;
  int* d_rec_iglob = nullptr;   // [nrec][125]
  realw* d_rec_h = nullptr;     // [nrec][125]
  realw* d_seismograms = nullptr;  // [nstep][nrec][3]
  realw* d_adj = nullptr;          // [nstep][nrec][3] adjoint source traces, forward time order
  realw* d_adj_step = nullptr;     // [nrec][3]

  // strain stores (SPECFEM-style kernel path)
  realw* d_epsilon = nullptr;    // [nspec][6][125] exx eyy ezz exy exz eyz
  realw* d_b_epsilon = nullptr;

  // kernels (FP64 accumulators)
  double* d_rho_kl_glob = nullptr;
  double* d_mu_kl = nullptr;
  double* d_kappa_kl = nullptr;
  double* d_cijkl_kl = nullptr;  // [nspec][21][125] internally
  double* d_hess_kl_glob = nullptr;

  // gather assembly: element-local forces and the static CSR point -> contributions.
  // Contribution index q: [0,nloc) element-local point, [nloc, nloc+nfp) Stacey
  // face point, then nsources*125 source points, then nrec*125 receiver points.
  realw* d_floc = nullptr;      // [nspec][125][3]
  realw* d_floc_b = nullptr;    // second field of the adjoint pass
  realw* d_face_scratch = nullptr;  // [nfaces*25][3] absorbing term of a non-stored field
  int* d_asm_off = nullptr;     // [nglob+1]
  int* d_asm_idx = nullptr;     // [ncontrib]
  size_t n_contrib = 0;

  // timing (filled by the whole-loop driver)
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  double last_force_ms = 0.0;
  int last_force_launches = 0;
  double last_asm_ms = 0.0;
  int last_asm_launches = 0;
};

inline SemMesh* sem_mesh(long* Mesh_pointer) {
  SemMesh* m = reinterpret_cast<SemMesh*>(*Mesh_pointer);
  if (!m) SEM_FAIL("Mesh_pointer is null (prepare_constants_device_ not called?)");
  return m;
}

template <class T>
inline T* dev_alloc(size_t n) {
  T* p = nullptr;
  if (n == 0) return nullptr;
  HIP_CHECK(hipMalloc((void**)&p, n * sizeof(T)));
  return p;
}

template <class T>
inline T* dev_upload(const T* h, size_t n, hipStream_t s) {
  T* p = dev_alloc<T>(n);
  if (n) HIP_CHECK(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  return p;
}

template <class T>
inline void dev_free(T*& p) {
  if (p) HIP_CHECK(hipFree((void*)p));
  p = nullptr;
}

// XCD-aware block remap: workgroups are dispatched round-robin over the 8
// XCDs (blockIdx % 8).  Give each XCD a contiguous chunk of elements so that
// neighbouring elements, which share GLL points, meet in the same L2.
__device__ __forceinline__ int xcd_remap(int b, int nblocks_padded) {
  const int per = nblocks_padded >> 3;
  return (b & 7) * per + (b >> 3);
}

inline int xcd_pad(int n) { return (n + 7) & ~7; }

// internal launchers shared between the C-ABI files
void sem_launch_forces(SemMesh* m, int nfields, int kernel_mode, int elem_begin, int elem_count,
                       realw deltat, const int* elem_list, bool gather);
void sem_zero_kernels(SemMesh* m);
