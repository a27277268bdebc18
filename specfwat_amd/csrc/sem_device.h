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
#include <hip/hip_ext.h>
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

// Timing of one kernel launch by the dispatch's own start and end timestamps
// (hipExtLaunchKernelGGL), the interval a profiler reports: a timer arms an
// event pair before a sampled launch, and the next SEM_LAUNCH on this thread
// consumes it.  Event markers recorded around a launch instead add the marker
// packets' latency to the interval.
struct SemLaunchEvents {
  hipEvent_t start = nullptr, stop = nullptr;
};
inline SemLaunchEvents& sem_launch_events() {
  static thread_local SemLaunchEvents e;
  return e;
}
inline void sem_arm_launch_events(hipEvent_t start, hipEvent_t stop) {
  SemLaunchEvents& e = sem_launch_events();
  if (e.start) SEM_FAIL("timed launch armed twice");
  e.start = start;
  e.stop = stop;
}
// true when the armed pair was taken by a launch (and clears a stale pair)
inline bool sem_launch_events_consumed() {
  SemLaunchEvents& e = sem_launch_events();
  const bool ok = e.start == nullptr;
  e.start = e.stop = nullptr;
  return ok;
}
#define SEM_LAUNCH(kernel, grid, block, shmem, stream, ...)                                          \
  do {                                                                                               \
    SemLaunchEvents& sle_ = sem_launch_events();                                                     \
    if (sle_.start) {                                                                                \
      hipEvent_t s0_ = sle_.start, s1_ = sle_.stop;                                                  \
      sle_.start = sle_.stop = nullptr;                                                              \
      hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, s0_, s1_, 0u, __VA_ARGS__);          \
    } else {                                                                                         \
      hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);                           \
    }                                                                                                \
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
  int graph = 0;              // whole-run loops captured into a HIP graph (fwat_set_graph_)
  // reduced-precision statics hold value/scale (FP16 needs it for range)
  float static_scale_kappa = 1.0f, static_scale_mu = 1.0f;
  float static_scale_jinv = 1.0f, static_scale_jac = 1.0f;
  int nstep = 0;

  hipStream_t stream = nullptr;

  // geometry / indexing
  // Library point numbering (prepare_constants_device_): points touched by more than one
  // element-local slot or by a Stacey face, source or receiver term come first, in the
  // caller's order, [0, n_shared); points touched by one element-local slot only follow,
  // element by element.  h_perm[caller point] = library point (empty: the caller's
  // numbering, every point shared); perm_hash identifies the numbering.
  int n_shared = 0;
  std::vector<int> h_perm;
  uint64_t perm_hash = 0;
  int interior_newmark = 1;  // fused loops: the element pass finishes points [n_shared, nglob)
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
  // FK injection of an incident wavefield (forward runs): velocity and
  // traction per face point, [fk_nsamp][nfaces*25][3] each; sample k belongs
  // to step 1 + k * fk_resamp, steps between samples interpolate linearly
  realw* d_fk_veloc = nullptr;
  realw* d_fk_tract = nullptr;
  int fk_nsamp = 0, fk_resamp = 1;

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

  // kernels (FP64 accumulators).  The element-pass kernels (mu/kappa or the 21
  // c_ij) reach them through fp32 partial sums over blocks of kernel_block steps
  // (kernel_acc): part += x each step, K += dt * part at the block's last step.
  int kernel_block = 16;  // 1: every step straight into fp64
  int kernel_flush = 1;   // set per step by the time loop: this step closes a block
  float* d_mu_part = nullptr;
  float* d_kappa_part = nullptr;
  float* d_cij_part = nullptr;  // [nspec][21][125]
  float* d_rho_part = nullptr;  // [nglob] partial sums of the density kernel
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
  int asm_points_per_lane = 0;  // gather assembly: 1 or 2 points per lane, 0 = by pass (see stage_assemble)
  int asm_block = 256;          // gather assembly: lanes per workgroup (64..1024)
  int floc_grouped = 1;         // floc rows grouped by reader (sem_floc_slot), 0: local-point order
  int* d_asm_off = nullptr;     // [nglob+1]
  int* d_asm_idx = nullptr;     // [ncontrib]
  size_t n_contrib = 0;
  // element visiting order of the phase lists when it is not element order
  // (empty: element order); every point's element terms are summed in it
  std::vector<int> h_visit;

  // element slabs: the fused loops run element pass + assembly slab after slab
  // (contiguous element ranges of slab_elems elements); a point is assembled
  // right after the slab holding its last contributing element, while that
  // slab's element-local forces are still in the memory-side cache
  int slab_elems = 0;                // 0: one pass over the whole mesh
  int nslabs = 0;
  std::vector<int> h_slab_goff;      // [nslabs+1] offsets into d_slab_g
  int* d_slab_g = nullptr;           // points of each slab, ascending

  // brick-fused pass on regular lattice meshes (element + assembly in one
  // kernel for brick-interior points; see brick.hip)
  int lat_nex = 0, lat_ney = 0, lat_nez = 0;  // validated element lattice, 0 = none
  int brick_x = 0, brick_y = 0, brick_z = 0;  // 0 = brick pass off
  int brick_bpc = 0;                          // persistent workgroups per CU (0: one per brick)
  int nbricks = 0;
  uint8_t* d_brick_special = nullptr;  // [nbricks] 1: every point goes through the halo pass
  realw* d_brick_part0 = nullptr;      // [nbricks][npts][3] partial sums of brick-boundary points
  realw* d_brick_part1 = nullptr;      // second field of the adjoint pass
  int nhalo = 0;
  int* d_halo_g = nullptr;    // [nhalo] points assembled by the halo pass
  int* d_halo_off = nullptr;  // [nhalo+1]
  int* d_halo_idx = nullptr;  // slots b*npts+bl, then nslot + (face, source, receiver contributions)

  // timing (filled by the whole-loop driver)
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  double last_force_ms = 0.0;
  int last_force_launches = 0;
  double last_asm_ms = 0.0;
  int last_asm_launches = 0;
};

// Position of local point L = i + 5j + 25k in an element's row of element-local forces
// (floc[ispec][slot][3]).  A gather reads a row in pieces, each at the time the point
// numbering reaches the points that piece belongs to, so the row is grouped by piece: the
// k = 0 face (25 slots, shared with the element below), the j = 0 face above it (20, with
// the element in front), the i = 0 face of the rest (16, with the element to the left),
// then the element's other points (i, j, k >= 1): the 37 on its k, j or i = 4 faces, and
// last the 27 inner ones (i, j, k in 1..3), which the gather never reads when the element
// pass finishes them.  A bijection of 0..124.
__host__ __device__ inline int sem_floc_slot(int L) {
  const int i = L % 5, j = (L / 5) % 5, k = L / 25;
  if (k == 0) return i + 5 * j;
  if (j == 0) return 25 + i + 5 * (k - 1);
  if (i == 0) return 45 + (j - 1) + 4 * (k - 1);
  if (i <= 3 && j <= 3 && k <= 3) return 98 + (i - 1) + 3 * (j - 1) + 9 * (k - 1);
  // rank among the (i, j, k >= 1) points not inner, in L order: own-index minus the inner
  // points before it
  const int o = (i - 1) + 4 * (j - 1) + 16 * (k - 1);
  const int inner_before = 9 * (k - 1 < 3 ? k - 1 : 3) + (k <= 3 ? 3 * (j - 1 < 3 ? j - 1 : 3) + (j <= 3 ? (i - 1 < 3 ? i - 1 : 3) : 0) : 0);
  return 61 + o - inner_before;
}

// library number of the caller's point g (0-based)
inline int sem_point(const SemMesh* m, int g) { return m->h_perm.empty() ? g : m->h_perm[g]; }

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

// Time integration of one kernel density x (this step's integrand, fp32):
// with partial sums, part += x and, at the block's last step,
// K += dt * (part + x), part = 0; without (part == null), K += dt * x.
__device__ __forceinline__ void kernel_acc(double* __restrict__ K, float* __restrict__ part, size_t i, float x,
                                           double dt, bool flush) {
  if (!part) {
    K[i] = fma(dt, (double)x, K[i]);
  } else if (flush) {
    K[i] = fma(dt, (double)(part[i] + x), K[i]);
    part[i] = 0.0f;
  } else {
    part[i] += x;
  }
}

// One field's contributions to the Newmark update of a point (assembly passes).
struct AsmField {
  const realw* floc;
  const realw* face;   // [nfp][3] traction*jw to subtract, or null
  const double* stf;   // [nsources] this step, or null
  const realw* adj;    // [nrec][3] this step, or null
  realw* u;
  realw* v;
  realw* a;
  realw dt;            // signed time step
};

// Newmark finish of point g once its force sums are complete: a = f * rmass_inv,
// corrector v += dt/2 a of this step and, unless last, the predictor
// u += dt v + dt^2/2 a, v += dt/2 a of the next one (last: a and v stored).
// TWO: the second field too, and the density kernel a0 . u1 (fp64, through
// fp32 block partial sums when rho_part is given) and the approximate Hessian
// a0 . a1 (fp64) accumulated.  Every assembly pass calls this, so they share
// one expression order.
// The point's state (inverse mass, u and v of each field) as loaded before its
// force sums, so that these loads overlap the gather's dependent chain.
struct PointState {
  realw r[3], u0[3], v0[3], u1[3], v1[3];
  float part;  // the density kernel's fp32 partial sum (two fields, partial sums on)
};

template <bool TWO>
__device__ __forceinline__ PointState load_point(size_t g, const realw* __restrict__ rm, const AsmField& F0,
                                                 const AsmField& F1, const float* __restrict__ rho_part = nullptr) {
  PointState P;
  const size_t b = 3 * g;
  P.part = (TWO && rho_part) ? rho_part[g] : 0.0f;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    P.r[c] = rm[b + c];
    P.u0[c] = F0.u[b + c];
    P.v0[c] = F0.v[b + c];
    if (TWO) {
      P.u1[c] = F1.u[b + c];
      P.v1[c] = F1.v[b + c];
    }
  }
  return P;
}

__device__ __forceinline__ void store3(realw* p, realw x, realw y, realw z) {
  p[0] = x;
  p[1] = y;
  p[2] = z;
}

template <bool TWO>
__device__ __forceinline__ void newmark_point(size_t g, const PointState& P, realw f0x, realw f0y, realw f0z,
                                              realw f1x, realw f1y, realw f1z, const AsmField& F0,
                                              const AsmField& F1, bool last, double* __restrict__ rho_kl,
                                              float* __restrict__ rho_part, bool flush,
                                              double* __restrict__ hess_kl, double dtk) {
  // every multiply-add is an explicit fma, so the result does not depend on how
  // the compiler contracts the surrounding code (the assembly variants agree bitwise)
  const size_t b = 3 * g;
  const realw r0 = P.r[0], r1 = P.r[1], r2 = P.r[2];
  const realw A[3] = {f0x * r0, f0y * r1, f0z * r2};
  const realw d0 = F0.dt, h0 = 0.5f * F0.dt, s0 = 0.5f * F0.dt * F0.dt;
  realw V[3], U[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) { V[c] = fmaf(h0, A[c], P.v0[c]); U[c] = P.u0[c]; }
  if (TWO) {
    const realw BA[3] = {f1x * r0, f1y * r1, f1z * r2};
    const realw d1 = F1.dt, h1 = 0.5f * F1.dt, s1 = 0.5f * F1.dt * F1.dt;
    realw BV[3], BU[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) { BV[c] = fmaf(h1, BA[c], P.v1[c]); BU[c] = P.u1[c]; }
    const float x = fmaf(A[2], BU[2], fmaf(A[1], BU[1], A[0] * BU[0]));
    if (!rho_part) {
      kernel_acc(rho_kl, rho_part, g, x, dtk, flush);
    } else if (flush) {  // kernel_acc with the partial sum read before the gather
      rho_kl[g] = fma(dtk, (double)(P.part + x), rho_kl[g]);
      rho_part[g] = 0.0f;
    } else {
      rho_part[g] = P.part + x;
    }
    if (hess_kl) hess_kl[g] = fma(dtk, (double)fmaf(A[2], BA[2], fmaf(A[1], BA[1], A[0] * BA[0])), hess_kl[g]);
    // each point's three components are stored together (one 12-byte store per array)
    if (last) {
      store3(F1.a + b, BA[0], BA[1], BA[2]);
      store3(F1.v + b, BV[0], BV[1], BV[2]);
    } else {
      store3(F1.u + b, fmaf(s1, BA[0], fmaf(d1, BV[0], BU[0])), fmaf(s1, BA[1], fmaf(d1, BV[1], BU[1])),
             fmaf(s1, BA[2], fmaf(d1, BV[2], BU[2])));
      store3(F1.v + b, fmaf(h1, BA[0], BV[0]), fmaf(h1, BA[1], BV[1]), fmaf(h1, BA[2], BV[2]));
    }
  }
  if (last) {
    store3(F0.a + b, A[0], A[1], A[2]);
    store3(F0.v + b, V[0], V[1], V[2]);
  } else {
    store3(F0.u + b, fmaf(s0, A[0], fmaf(d0, V[0], U[0])), fmaf(s0, A[1], fmaf(d0, V[1], U[1])),
           fmaf(s0, A[2], fmaf(d0, V[2], U[2])));
    store3(F0.v + b, fmaf(h0, A[0], V[0]), fmaf(h0, A[1], V[1]), fmaf(h0, A[2], V[2]));
  }
}

template <bool TWO>
__device__ __forceinline__ void newmark_point(size_t g, const realw* __restrict__ rm, realw f0x, realw f0y,
                                              realw f0z, realw f1x, realw f1y, realw f1z, const AsmField& F0,
                                              const AsmField& F1, bool last, double* __restrict__ rho_kl,
                                              float* __restrict__ rho_part, bool flush,
                                              double* __restrict__ hess_kl, double dtk) {
  newmark_point<TWO>(g, load_point<TWO>(g, rm, F0, F1, rho_part), f0x, f0y, f0z, f1x, f1y, f1z, F0, F1, last,
                     rho_kl, rho_part, flush, hess_kl, dtk);
}

// Gather-assembly CSR (prepare.hip): owner[q] = global point of contribution q
// (element-local points first, then face, source, receiver points); every
// row lists its element terms in the `visit` order of the elements, then the
// rest in contribution order.
void sem_build_assembly_csr(int nglob, int nspec, const std::vector<int>& owner, const std::vector<int>& visit,
                            std::vector<int>& off, std::vector<int>& idx);

// Newmark finish of the points [n_shared, nglob) inside the element pass of a fused
// loop (each such point has one element-local term, so its force is complete there):
// the fields and time steps of the pass, this step's last-step flag and the kernel step.
struct SemInterior {
  AsmField F0, F1;
  bool last;
  double dtk;
};

// internal launchers shared between the C-ABI files
// (in: the element pass also finishes the element-interior points when
// sem_interior_active(m, nfields); the assembly then covers [0, n_shared) only)
void sem_launch_forces(SemMesh* m, int nfields, int kernel_mode, int elem_begin, int elem_count,
                       realw deltat, const int* elem_list, bool gather, const SemInterior* in = nullptr);
// true when a whole-mesh fused element pass of nfields fields finishes the element-interior points
bool sem_interior_active(const SemMesh* m, int nfields);
// host <-> device copies of three point fields (u, v, a) of size 3 * nglob in the
// caller's numbering (the device arrays hold the library numbering)
void sem_fields_to_device(SemMesh* m, int size, realw* const host[3], realw* const dev[3]);
void sem_fields_from_device(SemMesh* m, int size, realw* const host[3], realw* const dev[3]);

void sem_zero_kernels(SemMesh* m);
// brick-fused pass (brick.hip): true when the mesh handle runs it
bool sem_brick_active(const SemMesh* m, int nfields);
// elements + Newmark of brick-interior points (field 0, and field 1 when nfields = 2)
void sem_launch_brick(SemMesh* m, int nfields, const AsmField& F0, const AsmField& F1, bool last, double dtk);
// Newmark of the halo points: brick partial sums + Stacey, source and receiver terms
void sem_launch_halo(SemMesh* m, int nfields, const AsmField& F0, const AsmField& F1, bool last, double dtk);
