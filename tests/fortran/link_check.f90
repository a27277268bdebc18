! Links against libspecfwat_hip.so through the specfwat_hip module and calls
! its host-only entry point fwat_assembly_csr on two elements stacked in z
! (no GPU needed): the 25 points of the shared face get two element terms, in
! the order of the phase lists (element 2 outer, element 1 inner: element 2's
! term first).  The addresses of the solver-path entry points are taken so
! that the link resolves every one of them.  Prints OK and exits 0 on success.
program link_check
  use iso_c_binding
  use specfwat_hip
  implicit none
  integer(c_int), parameter :: nspec = 2, nglob = 225
  integer(c_int) :: ibool(5, 5, 5, nspec), phase(nspec, 2), off(nglob + 1), idx(nspec * 125)
  integer(c_int) :: dummy(1), i, j, k, e, g, nshared, zero, one, ntwo
  type(c_funptr) :: fp(16)
  ! point numbering: element 1 owns layers k = 1..5 of its 25 columns (1..125),
  ! element 2's bottom layer is element 1's top layer
  do k = 1, 5
    do j = 1, 5
      do i = 1, 5
        ibool(i, j, k, 1) = i + 5 * (j - 1) + 25 * (k - 1)
        ibool(i, j, k, 2) = i + 5 * (j - 1) + 25 * (k + 3)
      end do
    end do
  end do
  phase = 0
  phase(1, 1) = 2     ! outer: element 2
  phase(1, 2) = 1     ! inner: element 1
  zero = 0
  one = 1
  dummy = 1
  call fwat_assembly_csr(nspec, nglob, ibool, zero, dummy, dummy, zero, dummy, zero, dummy, nspec, phase, &
                         one, one, off, idx)
  nshared = 0
  ntwo = 0
  do g = 1, nglob
    if (off(g + 1) - off(g) == 2) then
      ntwo = ntwo + 1
      ! the two terms are 0-based contributions e*125 + L: element 2's (e = 1) first
      e = idx(off(g) + 1) / 125
      if (e /= 1 .or. idx(off(g) + 2) / 125 /= 0) stop 2
      nshared = nshared + 1
    else if (off(g + 1) - off(g) /= 1) then
      stop 3
    end if
  end do
  if (ntwo /= 25 .or. nshared /= 25 .or. off(nglob + 1) /= nspec * 125) stop 4
  fp(1) = c_funloc(prepare_constants_device)
  fp(2) = c_funloc(prepare_fields_elastic_device)
  fp(3) = c_funloc(prepare_fields_elastic_adj_dev)
  fp(4) = c_funloc(prepare_sim2_or_3_const_device)
  fp(5) = c_funloc(fwat_set_stf)
  fp(6) = c_funloc(fwat_set_adjoint_sources)
  fp(7) = c_funloc(fwat_set_fk_injection)
  fp(8) = c_funloc(fwat_iterate_time)
  fp(9) = c_funloc(fwat_copy_forward_to_backward)
  fp(10) = c_funloc(fwat_move_stacey_store)
  fp(11) = c_funloc(transfer_seismograms_from_device)
  fp(12) = c_funloc(transfer_kernels_el_to_host_dp)
  fp(13) = c_funloc(transfer_kernels_ani_to_host_dp)
  fp(14) = c_funloc(prepare_cleanup_device)
  fp(15) = c_funloc(fwat_smooth_set_faithful)
  fp(16) = c_funloc(fwat_assembly_csr)
  do i = 1, 16
    if (.not. c_associated(fp(i))) stop 5
  end do
  print '(a)', 'OK'
end program link_check
