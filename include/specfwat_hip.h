/*
 * specfwat_hip.h -- C ABI of libspecfwat_hip.so, the MI355X (gfx950) engine for
 * SpecFWAT's spectral-element forward/adjoint/kernel path and its PDE kernel
 * smoother.
 *
 * Calling convention (the reference's own, smooth_sem_pde.cu:298-316 and
 * include/config.h:13-16): Fortran-callable names = lowercase + '_', every
 * argument passed by reference, arrays in Fortran column-major order,
 * element / point indices 1-based, logical flags as int (non-zero = true).
 * realw is float (CUSTOM_REAL single precision).  All functions return void
 * and abort the process with a message on error, as the reference's GPU code
 * and exit_MPI do.
 *
 * Device state lives behind `long *Mesh_pointer`, SPECFEM's opaque handle
 * that SpecFWAT stores and hands to prepare_cleanup_device
 * (src/specfem_interface/specfem_api.f90:303).
 *
 * Two groups of entry points:
 *  (1) names of the SPECFEM3D GPU C-ABI that SpecFWAT's solver path reaches
 *      (prepare_GPU at src/specfem_interface/prepare_timerun_fwat.f90:120,
 *      iterate_time at src/preproc/prepare_fwd.f90:364, save_adjoint_kernels
 *      at src/preproc/prepare_fwd.f90:381).  The SPECFEM3D sources are an
 *      un-vendored submodule (.gitmodules:1-4), so their exact argument lists
 *      are not in the reference tree; the lists below are this library's and
 *      carry SPECFEM's arrays under SPECFEM's names.
 *  (2) fwat_* additions with no counterpart in the reference: the whole-run
 *      device loop, in-HBM hand-off between the forward and adjoint runs,
 *      mixed-precision statics and batched smoothing.
 * smooth_pde_cuda_ keeps the reference's exact 42-argument signature.
 */
#ifndef SPECFWAT_HIP_H
#define SPECFWAT_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef float realw;

/* ---- device selection: initialize_GPU, src/specfem_interface/initialize_simulation.f90:300 */
void initialize_gpu_device_(int *runtime_type, int *myrank, int *ncuda_devices);

/* ---- preparation: prepare_GPU, prepare_timerun_fwat.f90:120 */
/* Mesh constants, GLL tables, Stacey faces (read_mesh_databases.f90:766-792),
 * sources (setup_sources_receivers_fwat.f90:41-56) and receivers.
 * h_sourcearrays (NDIM,NGLLX,NGLLY,NGLLZ,NSOURCES); h_hlagrange_rec (125,nrec). */
void prepare_constants_device_(long *Mesh_pointer, int *h_NGLLX, int *NSPEC_AB, int *NGLOB_AB,
                               int *NSPEC_IRREGULAR, realw *h_xix, realw *h_xiy, realw *h_xiz,
                               realw *h_etax, realw *h_etay, realw *h_etaz, realw *h_gammax,
                               realw *h_gammay, realw *h_gammaz, realw *h_jacobian,
                               realw *xix_regular, realw *jacobian_regular, int *h_ibool,
                               int *h_irregular_element_number, realw *h_hprime_xx,
                               realw *h_hprimewgll_xx, realw *h_wgllwgll_xy, realw *h_wgllwgll_xz,
                               realw *h_wgllwgll_yz, int *ABSORBING_CONDITIONS,
                               int *h_num_abs_boundary_faces, int *h_abs_boundary_ispec,
                               int *h_abs_boundary_ijk, realw *h_abs_boundary_normal,
                               realw *h_abs_boundary_jacobian2Dw, int *NSOURCES,
                               realw *h_sourcearrays, int *h_ispec_selected_source, int *nrec,
                               int *h_ispec_selected_rec, realw *h_hlagrange_rec, int *NSTEP,
                               int *SIMULATION_TYPE, int *SAVE_FORWARD);

/* Mass matrices (inverted, read_mesh_databases.f90:323-333,802-804), Stacey
 * impedances and moduli (get_model_fwat.f90:305-317), 21 c_ij when
 * ANISOTROPY, inner/outer phases (read_mesh_databases.f90:1129-1143). */
void prepare_fields_elastic_device_(long *Mesh_pointer, realw *rmassx, realw *rmassy, realw *rmassz,
                                    realw *rho_vp, realw *rho_vs, realw *kappav, realw *muv,
                                    int *ANISOTROPY, realw *c11store, realw *c12store,
                                    realw *c13store, realw *c14store, realw *c15store,
                                    realw *c16store, realw *c22store, realw *c23store,
                                    realw *c24store, realw *c25store, realw *c26store,
                                    realw *c33store, realw *c34store, realw *c35store,
                                    realw *c36store, realw *c44store, realw *c45store,
                                    realw *c46store, realw *c55store, realw *c56store,
                                    realw *c66store, int *num_phase_ispec_elastic,
                                    int *phase_ispec_inner_elastic, int *nspec_outer_elastic,
                                    int *nspec_inner_elastic, int *COMPUTE_AND_STORE_STRAIN);

/* Adjoint-run fields: b_ wavefield (read_mesh_databases.f90:1616-1622) and
 * kernels rho/mu/kappa or cijkl (read_mesh_databases.f90:1631-1677). */
void prepare_fields_elastic_adj_dev_(long *Mesh_pointer, int *size, int *COMPUTE_AND_STORE_STRAIN,
                                     realw *rhostore, int *ANISOTROPIC_KL, int *APPROXIMATE_HESS_KL);
void prepare_sim2_or_3_const_device_(long *Mesh_pointer, int *nadj_rec_local,
                                     int *NTSTEP_BETWEEN_READ_ADJSRC);

/* ---- per-step solver stages (iterate_time, prepare_fwd.f90:364) */
/* FORWARD_OR_ADJOINT: 1 forward/adjoint field, 3 reconstructed b_ field */
void update_displacement_cuda_(long *Mesh_pointer, realw *deltat, realw *deltatsqover2,
                               realw *deltatover2, int *FORWARD_OR_ADJOINT);
void compute_forces_viscoelastic_cuda_(long *Mesh_pointer, int *iphase, realw *deltat,
                                       int *nspec_outer_elastic, int *nspec_inner_elastic,
                                       int *COMPUTE_AND_STORE_STRAIN, int *ATTENUATION,
                                       int *FORWARD_OR_ADJOINT);
/* forward SAVE_FORWARD run stores record it_record; the b_ field replays it
 * (absorb_field.bin, specfem_api.f90:359-361) */
void compute_stacey_viscoelastic_cuda_(long *Mesh_pointer, int *iphase, int *it_record,
                                       int *FORWARD_OR_ADJOINT);
void compute_add_sources_viscoelastic_cuda_(long *Mesh_pointer, int *NSOURCES,
                                            double *h_stf_pre_compute, int *FORWARD_OR_ADJOINT);
/* h_source_adjoint (NDIM,nrec): adjoint sources of SEM/NET.STA.BX?.adj (syn_data.f90:232-251) */
void add_sources_el_sim_type_2_or_3_(long *Mesh_pointer, int *nrec, realw *h_source_adjoint, int *it);
void kernel_3_a_cuda_(long *Mesh_pointer, realw *deltatover2, int *FORWARD_OR_ADJOINT);
void compute_seismograms_cuda_(long *Mesh_pointer, int *it);
/* seismograms (NDIM,nrec,NSTEP), consumed as seismograms_d (syn_data.f90:64-66) */
void transfer_seismograms_from_device_(long *Mesh_pointer, realw *h_seismograms);

/* ---- kernels (compute_kernels / save_adjoint_kernels, prepare_fwd.f90:381) */
void compute_kernels_elastic_cuda_(long *Mesh_pointer, realw *deltat);
void transfer_kernels_el_to_host_(long *Mesh_pointer, realw *h_rho_kl, realw *h_mu_kl,
                                  realw *h_kappa_kl, int *NSPEC_AB);
void transfer_kernels_el_to_host_dp_(long *Mesh_pointer, double *h_rho_kl, double *h_mu_kl,
                                     double *h_kappa_kl, int *NSPEC_AB);
/* cijkl_kl (21,NGLLX,NGLLY,NGLLZ,NSPEC_AB), read as c11..c66 kernels by kernel_io.f90:100-156 */
void transfer_kernels_ani_to_host_dp_(long *Mesh_pointer, double *h_cijkl_kl, int *NSPEC_AB);
void transfer_kernels_hess_el_tohost_dp_(long *Mesh_pointer, double *h_hess_kl, int *NSPEC_AB);

/* ---- field transfers (save_forward_arrays.bin, specfem_api.f90:312-358) */
void transfer_fields_el_to_device_(int *size, realw *displ, realw *veloc, realw *accel, long *Mesh_pointer);
void transfer_fields_el_from_device_(int *size, realw *displ, realw *veloc, realw *accel, long *Mesh_pointer);
void transfer_b_fields_to_device_(int *size, realw *b_displ, realw *b_veloc, realw *b_accel, long *Mesh_pointer);
void transfer_b_fields_from_device_(int *size, realw *b_displ, realw *b_veloc, realw *b_accel, long *Mesh_pointer);
/* Stacey store (NDIM,NGLLSQUARE,num_abs_boundary_faces,NSTEP) <-> host */
void transfer_stacey_store_from_device_(long *Mesh_pointer, realw *h_store);
void transfer_stacey_store_to_device_(long *Mesh_pointer, realw *h_store);

/* ---- clean-up: FinalizeSpecfem, specfem_api.f90:303 */
void prepare_cleanup_device_(long *Mesh_pointer, int *ACOUSTIC_SIMULATION, int *ELASTIC_SIMULATION,
                             int *NOISE_TOMOGRAPHY);

/* ---- additions (no counterpart in the reference) */
/* whole-run source time functions (NSOURCES,NSTEP) and adjoint traces (NDIM,nrec,NSTEP) */
void fwat_set_stf_(long *Mesh_pointer, double *h_stf, int *NSTEP);
void fwat_set_adjoint_sources_(long *Mesh_pointer, realw *h_adj, int *NSTEP);
/* FK injection of a teleseismic incident field into the forward run's Stacey term; replaces the
 * Veloc_FK / Tract_FK arrays that couple_with_injection_prepare_boundary_fwat builds
 * (src/shared/fk_coupling.f90:71-297) and SPECFEM's Stacey routine injects: veloc_fk, tract_fk
 * [nsamp][npt][NDIM], npt = num_abs_boundary_faces*NGLLSQUARE in face order, sample k at step
 * 1 + k*resamp, linear in between; nsamp = 0 switches it off */
void fwat_set_fk_injection_(long *Mesh_pointer, int *npt, int *nsamp, int *resamp, const realw *veloc_fk,
                            const realw *tract_fk);
/* a whole SIMULATION_TYPE 1 or 3 run on the device; fused 1 = fused Newmark and
 * two-field force pass with in-pass fp64 kernel accumulation, 0 = SPECFEM stage order */
void fwat_iterate_time_(long *Mesh_pointer, int *SIMULATION_TYPE, realw *DT, int *fused, int *timed);
void fwat_last_force_time_(long *Mesh_pointer, double *ms, int *launches);
void fwat_last_assembly_time_(long *Mesh_pointer, double *ms, int *launches);
void fwat_synchronize_(long *Mesh_pointer);
/* plain device copy of nbytes repeated reps times (counter calibration probe) */
void fwat_copy_probe_(long *nbytes, int *reps, double *avg_ms);
/* the same copy with unroll (1, 2, 4, 8) float4 loads in flight per lane on ncu * blocks_per_cu
 * workgroups (0: one float4 per lane over the array): what a streaming kernel reaches here */
void fwat_copy_probe_unrolled_(long *nbytes, int *reps, int *unroll, int *blocks_per_cu, double *avg_ms);
/* in-HBM replacements of save_forward_arrays.bin and absorb_field.bin */
void fwat_copy_forward_to_backward_(long *Mesh_pointer_from, long *Mesh_pointer_to);
void fwat_move_stacey_store_(long *Mesh_pointer_from, long *Mesh_pointer_to);
/* element-pass launch shape (autotuner knobs): regular_path 1 = folded-table kernel on
 * all-regular meshes, 2 = its software-pipelined form (next element's gathers in flight during
 * the current one's contractions; isotropic), 3 = the pipelined form with two barriers per
 * element (both fields through ping-pong LDS buffers), 4 = form 3 with non-temporal hints on the
 * streamed indices, moduli and kernel partial sums, 5 = form 3 visiting the elements of a lattice
 * mesh (fwat_set_lattice_) in column blocks, layer by layer (element order elsewhere);
 * blocks_per_cu_* = persistent workgroups per CU, 0 = one per element */
void fwat_set_tuning_(long *Mesh_pointer, int *regular_path, int *blocks_per_cu_fwd, int *blocks_per_cu_adj);
/* The column-block element order regular_path 5 uses for a nex x ney x nez lattice and per_xcd
 * workgroups per XCD: order[nex*ney*nez], 0-based elements in traversal order.  Host only; no
 * counterpart in the reference (the reference's element loop visits elements in order). */
void fwat_column_block_order_(int *nex, int *ney, int *nez, int *per_xcd, int *order);
/* Element-interior points finished by the element pass of the fused loops (on = 1, the default;
 * the pipelined passes 3 and 5 on isotropic all-regular meshes without slabs or bricks): points
 * that one element-local slot alone refers to (no Stacey face, source or receiver term) are
 * numbered after all others inside the library, the element pass applies their Newmark update
 * and density-kernel term, and the gather assembly covers the others only.  Results are those
 * of the gather path bit for bit.  No counterpart in the reference (its assembly is the
 * element loop's scatter-add, prepare_fwd.f90:364 -> iterate_time). */
void fwat_set_interior_(long *Mesh_pointer, int *on);
/* Library numbering of a prepared mesh: points assembled by the gather (n_shared) and
 * element-interior points (n_interior); whether a forward / two-field fused pass with the
 * current launch shape finishes the interior points in the element pass. */
void fwat_interior_info_(long *Mesh_pointer, int *n_shared, int *n_interior, int *active_fwd, int *active_adj);
/* Element lattice of the mesh: element (ix,iy,iz) = ix + nex*(iy + ney*iz).  Checked against
 * ibool (shared faces, point multiplicities); enables fwat_set_brick_.  No counterpart in the
 * reference (SPECFEM meshes carry no lattice; this box mesher's do). */
void fwat_set_lattice_(long *Mesh_pointer, int *nex, int *ney, int *nez);
/* Brick-fused element + assembly pass of the whole-run loop on all-regular isotropic lattice
 * meshes: bricks of bx x by x bz elements (2x2x2, 4x2x2, 4x4x2; bx = 0: off), blocks_per_cu
 * persistent workgroups per CU (0: one per brick).  Replaces compute_forces + assembly of
 * iterate_time (prepare_fwd.f90:364) for that mesh class; results within rounding of the
 * gather path. */
void fwat_set_brick_(long *Mesh_pointer, int *bx, int *by, int *bz, int *blocks_per_cu);
/* brick decomposition sizes: halo points, partial-sum slots per field and step, halo CSR length */
void fwat_brick_info_(long *Mesh_pointer, int *nhalo, long *nslots, long *ncontrib);
/* Element slabs of the whole-run loop: element pass + assembly run slab after slab over
 * contiguous ranges of elems_per_slab elements (0: one pass over the mesh); each point is
 * assembled right after the slab holding its last contributing element, so the element-local
 * forces are read back while in the memory-side cache.  Results are bitwise those of the
 * single pass (same per-point summation order).  No counterpart in the reference. */
void fwat_set_slabs_(long *Mesh_pointer, int *elems_per_slab);
/* Steps per fp32 partial-sum block of the element-pass kernels (mu/kappa, c_ij) in the fused
 * adjoint loop before they are added into the fp64 accumulators (default 16; 1 = every step in
 * fp64).  The rho kernel (summed in the assembly pass) uses the same blocks. */
void fwat_set_kernel_block_(long *Mesh_pointer, int *nsteps);
/* points per lane of the gather assembly: 1, 2 (twice the loads in flight per lane; the same
 * per-point sums) or 0 (default: 2 for the forward one-field pass, 1 for the adjoint two-field
 * pass); launch-shape knob, no counterpart in the reference */
void fwat_set_assembly_lanes_(long *Mesh_pointer, int *points_per_lane);
/* lanes per workgroup of the gather assembly (64..1024, a multiple of 64; default 256): larger
 * workgroups keep longer runs of consecutive points on one XCD; launch-shape knob, no
 * counterpart in the reference */
void fwat_set_assembly_block_(long *Mesh_pointer, int *lanes);
/* layout of the element-local force rows of the gather path: 1 = grouped by the neighbour that
 * reads each piece (k = 0, j = 0, i = 0 faces, the element's other points, its inner points last;
 * the default), 0 = local-point order; the same results bit for bit; launch-shape knob, no
 * counterpart in the reference */
void fwat_set_floc_layout_(long *Mesh_pointer, int *grouped);
/* the grouped row's slot of each local point L = i + 5j + 25k (slot[125]); host only, no
 * counterpart in the reference */
void fwat_floc_row_layout_(int *slot);
/* HIP graph capture of the whole-run loop (fused gather path): on = 1 makes every
 * fwat_iterate_time_ call capture its step loop from the simulation stream into one graph and
 * launch it as a unit (the GPU then runs the steps' kernels back to back, independent of the
 * host's launch rate); captured runs are not timed.  Replaces the per-step launches of
 * iterate_time (prepare_fwd.f90:364); no counterpart in the reference. */
void fwat_set_graph_(long *Mesh_pointer, int *on);
/* Host-only (no device work): the gather-assembly CSR the preparation entry points build for
 * these arrays -- for every point its element terms in the order of SPECFEM's CPU element loop
 * over phase_ispec_inner_elastic (outer elements, then inner; the loop shape of
 * smooth.f90:243-247; phase arrays as read at read_mesh_databases.f90:1129-1143), then its
 * Stacey face, source and receiver terms.  off (NGLOB_AB+1), idx (0-based contributions:
 * 125 per element, 25 per face, 125 per source, 125 per receiver).  No counterpart in the
 * reference: it exposes the summation order so it can be checked without a GPU. */
void fwat_assembly_csr_(int *NSPEC_AB, int *NGLOB_AB, int *h_ibool, int *num_abs_boundary_faces,
                        int *h_abs_boundary_ispec, int *h_abs_boundary_ijk, int *NSOURCES,
                        int *h_ispec_selected_source, int *nrec, int *h_ispec_selected_rec,
                        int *num_phase_ispec_elastic, int *phase_ispec_inner_elastic,
                        int *nspec_outer_elastic, int *nspec_inner_elastic, int *off, int *idx);
/* 0 fp32, 1 bf16, 2 fp16 with per-array scaling, for kappa/mu and the Jacobian */
void fwat_set_static_precision_(long *Mesh_pointer, int *precision);

/* ---- multi-event batches: E = 1, 2, 4 or 8 events advanced together on one regular
 *      (all-cube) mesh, isotropic or with 21 c_ij.  Replaces SpecFWAT's serial event loop
 *      (src/preproc/fwat_fwd_measure_adj.f90:43-51, one PrepareFWD%simulation per event,
 *      src/preproc/prepare_fwd.f90:297-326); per-event arithmetic is fwat_iterate_time_'s.
 *      The mesh handle holds the model, receivers and Stacey faces (its own sources are
 *      not used).  Events may have different NSTEP: forward runs stop at NSTEP_e, adjoint
 *      runs are aligned at the end. */
/* nstep(nevents); h_sourcearrays(NDIM,NGLLX,NGLLY,NGLLZ,nevents), one source per event;
 * h_ispec_selected_source(nevents); h_stf(NSTEP_max,nevents) */
void fwat_batch_create_(long *Mesh_pointer, long *Batch_pointer, int *nevents, int *nstep, realw *DT,
                        realw *h_sourcearrays, int *h_ispec_selected_source, double *h_stf);
void fwat_batch_set_tuning_(long *Batch_pointer, int *blocks_per_cu);
/* 1: an isotropic batch's element pass finishes the element-interior points (one element term,
 * no source of the batch in their element: Newmark of every active event and the density kernel)
 * and the assembly covers the other points; 0 (default): every point in the assembly.  Bitwise
 * the same results.  No reference counterpart (the reference runs one event at a time,
 * fwat_fwd_measure_adj.f90:43-51). */
void fwat_batch_set_interior_(long *Batch_pointer, int *on);
/* Stacey store of the batch's next forward run: 0 fp32 (12 bytes per face point, event and step),
 * 1 fp16 with one scale per (step, event) record (6 bytes): the record's largest value maps to
 * 16384, so each stored value keeps 11 significant bits; halves the reconstruction store that
 * replaces absorb_field.bin (specfem_api.f90:359-361). */
void fwat_batch_set_store_precision_(long *Batch_pointer, int *precision);
/* forward runs with SAVE_FORWARD (per-event Stacey stores kept in HBM) */
void fwat_batch_forward_(long *Batch_pointer, int *timed);
/* h_adj(NDIM,nrec,NSTEP_max,nevents), forward time order; event e uses its first NSTEP_e rows */
void fwat_batch_set_adjoint_sources_(long *Batch_pointer, realw *h_adj);
/* adjoint + reconstruction + rho/mu/kappa kernels of every event */
void fwat_batch_adjoint_(long *Batch_pointer, int *timed);
/* h_seis(NDIM,nrec,NSTEP_max,nevents) */
void fwat_batch_seismograms_(long *Batch_pointer, realw *h_seis);
/* save_adjoint_kernels-scaled kernels, (NGLLX,NGLLY,NGLLZ,NSPEC_AB,nevents) each, fp64;
 * h_rhostore(NGLLX,NGLLY,NGLLZ,NSPEC_AB) scales rho_kl */
void fwat_batch_kernels_(long *Batch_pointer, realw *h_rhostore, double *h_rho_kl, double *h_mu_kl,
                         double *h_kappa_kl);
/* kernels of an anisotropic batch (a mesh prepared with ANISOTROPY, 21 c_ij): rho_kl
 * (NGLLX,NGLLY,NGLLZ,NSPEC_AB,nevents) scaled by rhostore and cijkl_kl
 * (21,NGLLX,NGLLY,NGLLZ,NSPEC_AB,nevents) in c11..c66 order (kernel_io.f90:100-156), as
 * transfer_kernels_ani_to_host_dp_ returns them for one event */
void fwat_batch_kernels_ani_(long *Batch_pointer, realw *h_rhostore, double *h_rho_kl, double *h_cijkl_kl);
void fwat_batch_last_time_(long *Batch_pointer, double *elem_ms, int *elem_n, double *asm_ms, int *asm_n);
void fwat_batch_destroy_(long *Batch_pointer);

/* ---- PDE smoothing: smooth_pde_cuda_, src/postproc/smooth_sem_pde.cu:298-316,
 *      called from src/postproc/smooth.f90:392-402 (exact reference signature) */
void smooth_pde_cuda_(int *h_is_sph, int *h_nspec, int *h_nglob, int *nprocs, int *h_nstep,
                      int *nspec_max, int *phase_ispec_inner_elastic, int *h_nspec_outer,
                      int *h_nspec_inner, const realw *xix, const realw *xiy, const realw *xiz,
                      const realw *etax, const realw *etay, const realw *etaz, const realw *gamx,
                      const realw *gamy, const realw *gamz, const realw *jaco, const realw *rotate0,
                      const realw *wgllwgll_xy, const realw *wgllwgll_xz, const realw *wgllwgll_yz,
                      const realw *hprimeT, const realw *hprime_wgll, const int *ibool,
                      const int *is_CPML, const realw *h_cv, const realw *h_ch, int *h_num_intfs,
                      int *h_max_nibool, const int *my_neighbors, const int *nibool_intf,
                      const int *ibool_intf, const realw *dat_bak, realw *dat, const realw *rvol,
                      realw *dat_glob, realw *ddat_glob, const int *irregular_element_number,
                      const realw *h_xix_regular, const int *h_nspec_irregular);
/* 1 (default) = reference arithmetic on regular elements, 0 = diagonal inverse Jacobian */
void fwat_smooth_set_faithful_(int *faithful);
/* nfield kernels smoothed together; dat_glob (NGLOB_AB,nfield) */
void fwat_smooth_pde_batch_(int *h_is_sph, int *h_nspec, int *h_nglob, int *h_nstep, int *h_nfield,
                            const realw *xix, const realw *xiy, const realw *xiz, const realw *etax,
                            const realw *etay, const realw *etaz, const realw *gamx, const realw *gamy,
                            const realw *gamz, const realw *jaco, const realw *rotate0,
                            const realw *wgllwgll_xy, const realw *wgllwgll_xz, const realw *wgllwgll_yz,
                            const realw *hprimeT, const realw *hprime_wgll, const int *ibool,
                            const realw *h_cv, const realw *h_ch, const realw *rvol, realw *dat_glob,
                            const int *irregular_element_number, const realw *h_xix_regular,
                            const int *h_nspec_irregular, int *timed);
/* element-pass and gather/update time of the last timed fwat_smooth_pde_batch_ (ms) */
void fwat_smooth_last_time_(double *elem_ms, double *update_ms);

/* ---- kernel post-processing (xfwat_post_proc, src/postproc/fwat_post_proc.f90:22-47).
 *      A post handle holds one data type's fp64 kernel sums and Hessian sum in HBM
 *      (PostFlow, src/postproc/post_processing.f90:20-32).  model_type 1: alpha, beta, rhop
 *      (KERNEL_ISO, src/shared/config.f90:35); 2: alpha, beta, rhop, gcp, gsp
 *      (KERNEL_AZI_ANI, config.f90:36).  use_rho_scaling leaves rhop out of the sums
 *      (post_processing.f90:131). */
void fwat_post_create_(long *Post_pointer, int *nspec, int *model_type, int *use_rho_scaling);
/* model of kernel_cijkl2hti (src/postproc/kernel_io.f90:107-118), each (NGLLX,NGLLY,NGLLZ,NSPEC) */
void fwat_post_set_model_(long *Post_pointer, realw *c44store, realw *c45store, realw *c55store,
                          realw *rho_vp, realw *rho_vs, realw *rhostore);
/* sum_kernel's per-event step (post_processing.f90:108-149) fed straight from a finished
 * adjoint run's device kernels: save_adjoint_kernels scaling + alpha/beta/rhop, or
 * kernel_cijkl2hti (kernel_io.f90:100-156); with precond_type 0/1 also the Hessian sum of
 * apply_precond (post_processing.f90:171-183; 1 = absolute values) */
void fwat_post_add_event_(long *Post_pointer, long *Mesh_pointer, int *precond_type);
/* the same from one event's EKERNEL files as read by read_event_kernel (kernel_io.f90:76-98):
 * h_ker (NGLLX,NGLLY,NGLLZ,NSPEC,3) alpha, beta, rhop or (...,11) rho, c11, c12, c13, c22, c23,
 * c33, c44, c45, c55, c66; h_hess (NGLLX,NGLLY,NGLLZ,NSPEC) or NULL */
void fwat_post_add_event_files_(long *Post_pointer, realw *h_ker, realw *h_hess, int *precond_type);
/* h_ker (NGLLX,NGLLY,NGLLZ,NSPEC,nker) fp64 */
void fwat_post_get_kernels_(long *Post_pointer, double *h_ker);
void fwat_post_get_hess_(long *Post_pointer, double *h_hess);
/* replaces the Hessian sum (run_mode 2 reads a stored one, post_processing.f90:184-186) */
void fwat_post_set_hess_(long *Post_pointer, double *h_hess);
/* invert_hess (post_processing.f90:344-377) on the Hessian sum, in place; precond_type 0
 * takes |sum| first (:181) */
void fwat_post_invert_hess_(long *Post_pointer, int *precond_type);
/* kernels *= inverted Hessian (apply_precond, post_processing.f90:189-191) */
void fwat_post_apply_precond_(long *Post_pointer);
void fwat_post_info_(long *Post_pointer, int *nker, int *nevents, int *nhess);
/* device time (ms) of the last add / invert / apply call */
void fwat_post_last_time_(long *Post_pointer, double *ms);
void fwat_post_destroy_(long *Post_pointer);

/* ---- gll2grid (src/grid/model_grid_data.f90:192-211): point location and interpolation
 *      coefficients of compute_interpolation_coeff_FD_SEM
 *      (src/shared/projection_on_FD_grid_mod_fwat.f90:186-506, locate_point_in_element
 *      :935-1125) for the grid ox + (i-1) hx (create_grid, model_grid_data.f90:18-53);
 *      ibool 1-based (NGLLX,NGLLY,NGLLZ,NSPEC); returns nb_fd_point */
void fwat_gll2grid_prepare_(long *Grid_pointer, int *nspec, int *nglob, int *ibool, realw *xstore,
                            realw *ystore, realw *zstore, realw *ox, realw *oy, realw *oz, realw *hx,
                            realw *hy, realw *hz, int *nx, int *ny, int *nz, int *num_iter,
                            long *nb_fd_point);
/* Project_model_SEM2FD_grid (projection_on_FD_grid_mod_fwat.f90:512-596) of nfield fields:
 * gll_data (NGLLX,NGLLY,NGLLZ,NSPEC,nfield) -> grid_data (nx,ny,nz,nfield) */
void fwat_gll2grid_(long *Grid_pointer, int *nfield, realw *gll_data, realw *grid_data);
/* owner(nx,ny,nz): 1-based element of each grid point (0 = none); ref(3,nx,ny,nz) xi,eta,gamma */
void fwat_gll2grid_owner_(long *Grid_pointer, int *owner, double *ref);
void fwat_gll2grid_last_time_(long *Grid_pointer, double *ms);
void fwat_gll2grid_destroy_(long *Grid_pointer);
/* taper_kernel_grid (post_processing.f90:296-333): taper_cls over the mesh box with spacing
 * (distance_max_glob + distance_min_glob)/2 (src/postproc/taper_3D_mod.f90:16-141), sampled at
 * every grid point; data (nx,ny,nz,nfield) is multiplied in place; mask (nx,ny,nz) or NULL */
void fwat_taper_grid_(realw *x_min_glob, realw *x_max_glob, realw *y_min_glob, realw *y_max_glob,
                      realw *z_min_glob, realw *z_max_glob, realw *distance_min_glob,
                      realw *distance_max_glob, realw *h_suppress, realw *h_buffer, realw *v_suppress,
                      realw *v_buffer, realw *ox, realw *oy, realw *oz, realw *hx, realw *hy, realw *hz,
                      int *nx, int *ny, int *nz, int *nfield, realw *data, realw *mask);

/* ---- grid optimiser (xfwat_optimize, src/optimization/optimize_grid.f90).  Grids are
 *      (nx,ny,nz,nkernel) single precision, n = nx*ny*nz*nkernel, npg = nx*ny*nz.  Every inner
 *      product is an fp64 sum in a fixed order (the reference's inner_product,
 *      src/shared/utils.f90:1527-1534, is a single-precision sum). */
/* get_SD_direction (optimize_grid.f90:229-241): direction = -gradient*hess / max|.| */
void fwat_opt_sd_direction_(long *n, realw *gradient, realw *hess, realw *direction);
/* get_lbfgs_direction (optimize_grid.f90:243-339): grads and models (nx,ny,nz,nkernel,nstore+1)
 * hold iterations iter_store .. iter_current as read_gradient_grid / read_model_grid return them
 * (models linear; the first three parameters enter as logs, opt_io.f90:79-81); hess the inverse
 * Hessian grid; returns the normalised direction and its angle to -gradient in degrees */
void fwat_opt_lbfgs_direction_(long *n, long *npg, int *nstore, realw *grads, realw *models, realw *hess,
                               realw *dx, realw *dy, realw *dz, realw *direction, double *angle);
/* get_CG_direction (optimize_grid.f90:341-389): g1 current, g0 previous gradient; alpha returned */
void fwat_opt_cg_direction_(long *n, long *npg, realw *g1, realw *g0, realw *hess, realw *direction,
                            double *alpha);
/* model_update + alpha_scaling (optimize_grid.f90:176-203, 391-404): model_type 1 or 2;
 * update_iso / update_ani select the alternating-inversion halves of model type 2;
 * vpvs_ratio_range(2) = VPVS_RATIO_RANGE; model (nx,ny,nz,nkernel) updated in place */
void fwat_opt_model_update_(long *npg, int *nkernel, int *model_type, int *update_iso, int *update_ani,
                            realw *step_len, realw *vpvs_ratio_range, realw *model, realw *direction);
/* device time (ms) of the last fwat_opt_* call, uploads excluded */
void fwat_opt_last_time_(double *ms);

/* ---- waveform misfit and adjoint sources of every trace of an event in one device pass:
 *      the per-receiver loop of leq_data.f90:148 over WaveformMisfit%calc_adjoint_source
 *      (src/preproc/adjoint_sources/waveform_misfit.f90:15-64; simpson, utils.f90:1630-1663;
 *      window_taper, signal.f90:419-491) and filter_adj (adjoint_source.f90:92-114).
 *      obs, syn, adj: (NDIM,nrec,NSTEP) = ntrace traces sample-major; windows(2,nwin) in seconds;
 *      sos(6,nsect) = b0 b1 b2 1 a1 a2 of the bandpass sections (bandpass_dp's Butterworth,
 *      signal.f90:7-36), nsect = 0 for no filter; misfits(ntrace) = per-trace misfit summed over
 *      windows; residuals(nwin,ntrace) or NULL */
void fwat_waveform_adjoint_batch_(int *nstep, int *ntrace, double *dt, realw *obs, realw *syn, int *nwin,
                                  double *windows, double *taper_percentage, int *itaper_type, int *nsect,
                                  realw *sos, double *misfits, double *residuals, realw *adj);
/* device time (ms) of the last fwat_waveform_adjoint_batch_, fwat_cc_adjoint_batch_ or
 * fwat_mt_adjoint_batch_, copies excluded */
void fwat_waveform_adjoint_last_time_(double *ms);
/* No counterpart in the reference: fwat_waveform_adjoint_batch_ for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  obs, syn (in) and adj, misfits, residuals (out; residuals may be NULL) are the
 * caller's device buffers, windows and sos host arrays as above, *stream a hipStream_t (0: null
 * stream).  Intermediates live in a static per-GPU scratch of 3.1 MB: ntrace <= 16384,
 * nwin <= 8, nsect <= 8, window lengths summed <= 65536 samples, filter window <= 65536 samples;
 * larger inputs are refused with a message.  Returns when the outputs are written; calls that
 * overlap on one GPU share the scratch and race. */
void fwat_waveform_adjoint_batch_noalloc_(int *nstep, int *ntrace, double *dt, realw *d_obs, realw *d_syn,
                                          int *nwin, double *windows, double *taper_percentage,
                                          int *itaper_type, int *nsect, realw *sos, double *d_misfits,
                                          double *d_residuals, realw *d_adj, long *stream);
/* cross-correlation traveltime (imeas 11) / amplitude (12) measurement of every trace: CCTTMisfit%
 * calc_adjoint_source (src/preproc/adjoint_sources/cc_tt_misfit.f90:17-146 with utils/cc.f90:10-160),
 * or the zero-lag correlation coefficient (imeas 5, CCCMisfit, ccc_misfit.f90:15-71), and filter_adj; cfg(7) = dt_sigma_min, dlna_sigma_min, TSHIFT_MIN, TSHIFT_MAX, DLNA_MIN, DLNA_MAX,
 * CC_MIN (adj_config.f90:18-23); meas(8,nwin,ntrace) = misfit, residual, error, cc_max, tshift, dlna,
 * selected, ishift (or NULL) */
void fwat_cc_adjoint_batch_(int *nstep, int *ntrace, double *dt, realw *obs, realw *syn, int *nwin,
                            double *windows, double *taper_percentage, int *itaper_type, int *imeas,
                            double *cfg, int *nsect, realw *sos, double *misfits, double *meas, realw *adj);
/* multitaper traveltime (imeas 13) / amplitude (14) measurement of every trace with the reference's
 * fall back to the cross-correlation measurement: MTTTMisfit%calc_adjoint_source
 * (src/preproc/adjoint_sources/mt_tt_misfit.f90:29-791 with utils/dpss.f90:137-262 and
 * process_cycle_skipping, signal.f90:510-598) and filter_adj; cfg(19) = the 7 values of
 * fwat_cc_adjoint_batch_, transfunc_waterlevel, water_threshold, mt_nw, num_taper, dt_fac, err_fac,
 * dt_max_scale, min_cycle_in_window, use_cc_error (0/1), min_period, max_period, use_mt_error (0/1:
 * the jackknife errors of calculate_mt_error, mt_tt_misfit.f90:302-409; adj_config.f90:18-26 defaults
 * to .false.); meas(10,nwin,ntrace) = misfit, residual, error, cc_max, tshift, dlna,
 * selected, ishift, measurement used (13/14, 11/12 after the fall back, 0 unselected), the other misfit
 * (or NULL).  32 <= nstep <= 16383 */
/* exponentiated phase measurement of every trace (imeas 4): ExponentiatedPhaseMisfit%calc_adjoint_source
 * (src/preproc/adjoint_sources/exponentiated_phase_misfit.f90:21-104 with hilbert, fftpack.f90:135-201)
 * and filter_adj; wtr_env = the envelope water level of adj_config (adj_config.f90:20);
 * meas(8,nwin,ntrace) = misfit, residual, misfit_real, misfit_imag, residual_real, residual_imag,
 * selected (1), 0 (or NULL).  Windows of 32 to 16383 samples */
void fwat_ep_adjoint_batch_(int *nstep, int *ntrace, double *dt, realw *obs, realw *syn, int *nwin,
                            double *windows, double *taper_percentage, int *itaper_type, double *wtr_env,
                            int *nsect, realw *sos, double *misfits, double *meas, realw *adj);
/* window selection of every trace (WIN_SELECTOR_TYPE): win_sel_type%select_windows
 * (src/preproc/window_selector/win_sel.f90:35-469 with sta_lta.f90:6-104, called at
 * src/preproc/leq_data.f90:108-114); obs, syn (NDIM,nrec,NSTEP) filtered as the caller does; tp(ntrace), dist(ntrace) the
 * first-arrival time and distance of each trace's receiver; cfg(10) = sliding_win_len_fac, min_velocity,
 * threshold_corr, threshold_shift_fac, jump_fac, min_win_len_fac, min_snr_window, min_energy_ratio,
 * min_peaks_troughs, is_split_phases (0/1; splitting takes nstep <= 16383); nwin(ntrace) out;
 * windows(2,maxwin,ntrace) out (s) */
void fwat_win_sel_batch_(int *nstep, int *ntrace, double *dt, realw *obs, realw *syn, double *t0, double *tp,
                         double *dist, double *min_period, double *cfg, int *maxwin, int *nwin, double *windows);
/* FLEXWIN window selection (WIN_TYPE 2, flexwin_type%select_windows, src/preproc/window_selector/
 * flexwin.f90:75-734 with schedule_weighted_intervals, interval_scheduling.f90:13-165, as
 * leq_data.f90:116-121 calls it) of every trace: obs, syn (NDIM, nrec, NSTEP) single precision
 * (filtered), tp(ntrace), dist(ntrace) first-arrival time and distance, t0, min/max period of the
 * band; cfg(16) = min_velocity, min_win_len_fac, threshold_corr, threshold_shift_fac,
 * min_peaks_troughs, threshold_dlna, min_snr_window, is_split_phases (0/1),
 * max_time_before_first_arrival, stalta_water_level, c_0, c_2, c_3a, c_3b, c_4a, c_4b
 * (input_params.f90:400-418).  nwin(ntrace); windows(2, maxwin, ntrace) in s relative to t0 (twin);
 * info(4, maxwin, ntrace) = weight, cc, time shift, dlnA.  3 <= nstep <= 16383; at most 16384
 * candidates per trace survive the per-candidate passes (refused beyond) */
void fwat_flexwin_batch_(int *nstep, int *ntrace, double *dt, realw *obs, realw *syn, double *t0, double *tp,
                         double *dist, double *min_period, double *max_period, double *cfg, int *maxwin, int *nwin,
                         double *windows, double *info);
/* cross-convolution misfit of every receiver (imeas 2): WaveformConvMisfit%calc_adjoint_source
 * (src/preproc/adjoint_sources/waveform_conv_misfit.f90:17-64 with myconvolution_dp part 0,
 * src/shared/signal.f90:146-189) as src/preproc/telecc_data.f90:175-178 calls it; dat_r, dat_z, syn_r,
 * syn_z (nrec,NSTEP) = [NSTEP][nrec]; windows(2,nrec) in seconds; misfits(nrec), residuals(nrec) (or
 * NULL) out; adj_r, adj_z out (zero outside each window) */
void fwat_conv_adjoint_batch_(int *nstep, int *nrec, double *dt, realw *dat_r, realw *dat_z, realw *syn_r,
                              realw *syn_z, double *windows, double *taper_percentage, int *itaper_type,
                              double *misfits, double *residuals, realw *adj_r, realw *adj_z);
void fwat_mt_adjoint_batch_(int *nstep, int *ntrace, double *dt, realw *obs, realw *syn, int *nwin,
                            double *windows, double *taper_percentage, int *itaper_type, int *imeas,
                            double *cfg, int *nsect, realw *sos, double *misfits, double *meas, realw *adj);

#ifdef __cplusplus
}
#endif
#endif /* SPECFWAT_HIP_H */
