! specfwat_hip.f90 -- iso_c_binding interfaces of libspecfwat_hip.so for the
! Fortran side of SpecFWAT (the C declarations are in specfwat_hip.h).  Every
! argument is passed by reference, as SpecFWAT's own GPU calls pass them
! (smooth.f90:392-402 <-> smooth_sem_pde.cu:298-316).  A maintainer adds this
! module and calls it from prepare_GPU / iterate_time / save_adjoint_kernels,
! or directly from PrepareFWD%run_simulation (src/preproc/prepare_fwd.f90:328-368).
! smooth_pde_cuda keeps its implicit interface in smooth.f90 and needs no entry here.
module specfwat_hip
  use iso_c_binding
  implicit none
  interface
    subroutine prepare_constants_device(Mesh_pointer, NGLLX, NSPEC_AB, NGLOB_AB, NSPEC_IRREGULAR, &
        xix, xiy, xiz, etax, etay, etaz, gammax, gammay, gammaz, jacobian, xix_regular, &
        jacobian_regular, ibool, irregular_element_number, hprime_xx, hprimewgll_xx, &
        wgllwgll_xy, wgllwgll_xz, wgllwgll_yz, ABSORBING_CONDITIONS, num_abs_boundary_faces, &
        abs_boundary_ispec, abs_boundary_ijk, abs_boundary_normal, abs_boundary_jacobian2Dw, &
        NSOURCES, sourcearrays, ispec_selected_source, nrec, ispec_selected_rec, hlagrange_rec, &
        NSTEP, SIMULATION_TYPE, SAVE_FORWARD) bind(C, name="prepare_constants_device_")
      import :: c_long, c_int, c_float
      integer(c_long) :: Mesh_pointer
      integer(c_int) :: NGLLX, NSPEC_AB, NGLOB_AB, NSPEC_IRREGULAR, ABSORBING_CONDITIONS
      integer(c_int) :: num_abs_boundary_faces, NSOURCES, nrec, NSTEP, SIMULATION_TYPE, SAVE_FORWARD
      real(c_float) :: xix(*), xiy(*), xiz(*), etax(*), etay(*), etaz(*), gammax(*), gammay(*), gammaz(*)
      real(c_float) :: jacobian(*), xix_regular, jacobian_regular
      integer(c_int) :: ibool(*), irregular_element_number(*), abs_boundary_ispec(*), abs_boundary_ijk(*)
      real(c_float) :: hprime_xx(*), hprimewgll_xx(*), wgllwgll_xy(*), wgllwgll_xz(*), wgllwgll_yz(*)
      real(c_float) :: abs_boundary_normal(*), abs_boundary_jacobian2Dw(*), sourcearrays(*), hlagrange_rec(*)
      integer(c_int) :: ispec_selected_source(*), ispec_selected_rec(*)
    end subroutine

    subroutine prepare_fields_elastic_device(Mesh_pointer, rmassx, rmassy, rmassz, rho_vp, rho_vs, &
        kappav, muv, ANISOTROPY, c11store, c12store, c13store, c14store, c15store, c16store, &
        c22store, c23store, c24store, c25store, c26store, c33store, c34store, c35store, c36store, &
        c44store, c45store, c46store, c55store, c56store, c66store, num_phase_ispec_elastic, &
        phase_ispec_inner_elastic, nspec_outer_elastic, nspec_inner_elastic, COMPUTE_AND_STORE_STRAIN) &
        bind(C, name="prepare_fields_elastic_device_")
      import :: c_long, c_int, c_float
      integer(c_long) :: Mesh_pointer
      real(c_float) :: rmassx(*), rmassy(*), rmassz(*), rho_vp(*), rho_vs(*), kappav(*), muv(*)
      integer(c_int) :: ANISOTROPY, num_phase_ispec_elastic, nspec_outer_elastic, nspec_inner_elastic
      integer(c_int) :: COMPUTE_AND_STORE_STRAIN, phase_ispec_inner_elastic(*)
      real(c_float) :: c11store(*), c12store(*), c13store(*), c14store(*), c15store(*), c16store(*)
      real(c_float) :: c22store(*), c23store(*), c24store(*), c25store(*), c26store(*)
      real(c_float) :: c33store(*), c34store(*), c35store(*), c36store(*)
      real(c_float) :: c44store(*), c45store(*), c46store(*), c55store(*), c56store(*), c66store(*)
    end subroutine

    subroutine prepare_fields_elastic_adj_dev(Mesh_pointer, size, COMPUTE_AND_STORE_STRAIN, rhostore, &
        ANISOTROPIC_KL, APPROXIMATE_HESS_KL) bind(C, name="prepare_fields_elastic_adj_dev_")
      import :: c_long, c_int, c_float
      integer(c_long) :: Mesh_pointer
      integer(c_int) :: size, COMPUTE_AND_STORE_STRAIN, ANISOTROPIC_KL, APPROXIMATE_HESS_KL
      real(c_float) :: rhostore(*)
    end subroutine

    subroutine prepare_sim2_or_3_const_device(Mesh_pointer, nadj_rec_local, NTSTEP_BETWEEN_READ_ADJSRC) &
        bind(C, name="prepare_sim2_or_3_const_device_")
      import :: c_long, c_int
      integer(c_long) :: Mesh_pointer
      integer(c_int) :: nadj_rec_local, NTSTEP_BETWEEN_READ_ADJSRC
    end subroutine

    subroutine fwat_set_stf(Mesh_pointer, stf, NSTEP) bind(C, name="fwat_set_stf_")
      import :: c_long, c_int, c_double
      integer(c_long) :: Mesh_pointer
      real(c_double) :: stf(*)          ! (NSOURCES, NSTEP)
      integer(c_int) :: NSTEP
    end subroutine

    subroutine fwat_set_adjoint_sources(Mesh_pointer, adj, NSTEP) bind(C, name="fwat_set_adjoint_sources_")
      import :: c_long, c_int, c_float
      integer(c_long) :: Mesh_pointer
      real(c_float) :: adj(*)           ! (NDIM, nrec, NSTEP), SEM/*.adj in forward time order
      integer(c_int) :: NSTEP
    end subroutine

    subroutine fwat_set_fk_injection(Mesh_pointer, npt, nsamp, resamp, veloc_fk, tract_fk) &
        bind(C, name="fwat_set_fk_injection_")
      import :: c_long, c_int, c_float
      integer(c_long) :: Mesh_pointer
      integer(c_int) :: npt, nsamp, resamp
      real(c_float) :: veloc_fk(*), tract_fk(*)   ! (NDIM, npt, nsamp) each
    end subroutine

    subroutine fwat_iterate_time(Mesh_pointer, SIMULATION_TYPE, DT, fused, timed) bind(C, name="fwat_iterate_time_")
      import :: c_long, c_int, c_float
      integer(c_long) :: Mesh_pointer
      integer(c_int) :: SIMULATION_TYPE, fused, timed
      real(c_float) :: DT
    end subroutine

    subroutine fwat_set_tuning(Mesh_pointer, regular_path, blocks_per_cu_fwd, blocks_per_cu_adj) &
        bind(C, name="fwat_set_tuning_")
      import :: c_long, c_int
      integer(c_long) :: Mesh_pointer
      integer(c_int) :: regular_path, blocks_per_cu_fwd, blocks_per_cu_adj
    end subroutine

    subroutine fwat_set_interior(Mesh_pointer, on) bind(C, name="fwat_set_interior_")
      import :: c_long, c_int
      integer(c_long) :: Mesh_pointer
      integer(c_int) :: on
    end subroutine

    subroutine fwat_copy_forward_to_backward(Mesh_from, Mesh_to) bind(C, name="fwat_copy_forward_to_backward_")
      import :: c_long
      integer(c_long) :: Mesh_from, Mesh_to
    end subroutine

    subroutine fwat_move_stacey_store(Mesh_from, Mesh_to) bind(C, name="fwat_move_stacey_store_")
      import :: c_long
      integer(c_long) :: Mesh_from, Mesh_to
    end subroutine

    subroutine transfer_seismograms_from_device(Mesh_pointer, seismograms) &
        bind(C, name="transfer_seismograms_from_device_")
      import :: c_long, c_float
      integer(c_long) :: Mesh_pointer
      real(c_float) :: seismograms(*)   ! (NDIM, nrec, NSTEP)
    end subroutine

    subroutine transfer_kernels_el_to_host_dp(Mesh_pointer, rho_kl, mu_kl, kappa_kl, NSPEC_AB) &
        bind(C, name="transfer_kernels_el_to_host_dp_")
      import :: c_long, c_int, c_double
      integer(c_long) :: Mesh_pointer
      real(c_double) :: rho_kl(*), mu_kl(*), kappa_kl(*)
      integer(c_int) :: NSPEC_AB
    end subroutine

    subroutine transfer_kernels_ani_to_host_dp(Mesh_pointer, cijkl_kl, NSPEC_AB) &
        bind(C, name="transfer_kernels_ani_to_host_dp_")
      import :: c_long, c_int, c_double
      integer(c_long) :: Mesh_pointer
      real(c_double) :: cijkl_kl(*)     ! (21, NGLLX, NGLLY, NGLLZ, NSPEC_AB)
      integer(c_int) :: NSPEC_AB
    end subroutine

    subroutine prepare_cleanup_device(Mesh_pointer, ACOUSTIC_SIMULATION, ELASTIC_SIMULATION, NOISE_TOMOGRAPHY) &
        bind(C, name="prepare_cleanup_device_")
      import :: c_long, c_int
      integer(c_long) :: Mesh_pointer
      integer(c_int) :: ACOUSTIC_SIMULATION, ELASTIC_SIMULATION, NOISE_TOMOGRAPHY
    end subroutine

    subroutine fwat_assembly_csr(NSPEC_AB, NGLOB_AB, ibool, num_abs_boundary_faces, abs_boundary_ispec, &
        abs_boundary_ijk, NSOURCES, ispec_selected_source, nrec, ispec_selected_rec, num_phase_ispec_elastic, &
        phase_ispec_inner_elastic, nspec_outer_elastic, nspec_inner_elastic, off, idx) &
        bind(C, name="fwat_assembly_csr_")
      import :: c_int
      integer(c_int) :: NSPEC_AB, NGLOB_AB, num_abs_boundary_faces, NSOURCES, nrec
      integer(c_int) :: num_phase_ispec_elastic, nspec_outer_elastic, nspec_inner_elastic
      integer(c_int) :: ibool(*), abs_boundary_ispec(*), abs_boundary_ijk(*), ispec_selected_source(*)
      integer(c_int) :: ispec_selected_rec(*), phase_ispec_inner_elastic(*), off(*), idx(*)
    end subroutine

    subroutine fwat_smooth_set_faithful(faithful) bind(C, name="fwat_smooth_set_faithful_")
      import :: c_int
      integer(c_int) :: faithful
    end subroutine
  end interface
end module specfwat_hip
