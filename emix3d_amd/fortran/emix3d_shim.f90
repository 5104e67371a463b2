! Fortran drop-in for the hot path of eMix3d's RUNGEK sub-step.
!
! The reference's own GPU wrappers need no change: solver/deriv_vel_cuda.f90
! (DERIV_VEL -> cudaDeriv), solver/divergencia_cuda.f90 (divergencia ->
! cuDivergence) and solver/advectLS.f90 (advectLevelSet -> advect_CUDA)
! bind C symbols that libemix3d_hip.so exports with the same signatures.
! This file replaces the two hot-path routines the reference runs in
! Fortran on the CPU with the same names, arguments and module reads and
! writes, computed by the library's HIP kernels:
!   flujos(neq)     solver/flujos_compact.f90:1-220
!   viscosidad_t()  solver/sgdm_incomp.f90:117-162
!   viscosidad()    solver/viscosidad_incomp.f90:1-23
! Link: replace flujos_compact.o, viscosidad_incomp.o and the
! viscosidad_t of sgdm_incomp.o (keep its INISGDM) with this object, and the
! CUDA objects with -lemix3d_hip.

module emx_capi
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function emx_host_flujos(neq, um, u, pres, vis, jbn, dvel, &
        e, f, g, ev, fv, gv, nx, ny, nz) bind(C, name='emx_host_flujos')
      import :: c_int, c_double
      integer(c_int), value :: neq, nx, ny, nz
      real(c_double) :: um(*), u(*), pres(*), vis(*), jbn(*), dvel(*)
      real(c_double) :: e(*), f(*), g(*), ev(*), fv(*), gv(*)
    end function emx_host_flujos
    integer(c_int) function emx_host_sgdm(u, dvel, jbn, sgd, c4, amut, vis, &
        vis5, nx, ny, nz) bind(C, name='emx_host_sgdm')
      import :: c_int, c_double
      real(c_double) :: u(*), dvel(*), jbn(*), sgd(*)
      real(c_double), value :: c4
      real(c_double) :: amut(*), vis(*), vis5(*)
      integer(c_int), value :: nx, ny, nz
    end function emx_host_sgdm
    function emx_last_error() bind(C, name='emx_last_error')
      import :: c_ptr
      type(c_ptr) :: emx_last_error
    end function emx_last_error
  end interface
  ! the seven SGS geometry fields of INISGDM, gathered once per run
  real(c_double), allocatable, save :: emx_sgd(:,:,:,:)
contains
  subroutine emx_fail(who, rc)
    character(*), intent(in) :: who
    integer(c_int), intent(in) :: rc
    print *, who, ' failed on the GPU, status ', rc
    stop
  end subroutine emx_fail
end module emx_capi

SUBROUTINE flujos(neq)
  use mflujos
  use viscosidades
  use dimensiones
  use variables
  use velocidades
  use jacobtools
  use derivvel
  use emx_capi
  IMPLICIT NONE
  integer :: neq
  integer(c_int) :: rc
  rc = emx_host_flujos(neq, um, u, pres, vis, jbn, dvel, e, f, g, ev, fv, gv, &
                       nx, ny, nz)
  if (rc /= 0) call emx_fail('flujos', rc)
END SUBROUTINE flujos

! Structure-function eddy viscosity with the selective switch.  The library
! forms amut and also vis = c4 + amut and vis5 = amut/0.7 in the same pass.
SUBROUTINE viscosidad_t()
  use dimensiones
  use sgdmodel
  use velocidades
  use derivvel
  use jacobtools
  use viscosidades
  use consadim
  use emx_capi
  IMPLICIT NONE
  integer(c_int) :: rc
  ! packed on every call: INISGDM may have run again (new grid or sizes)
  ! since the last one, and the copy is small next to the transfer
  if (allocated(emx_sgd)) then
    if (size(emx_sgd, 1) /= nx .or. size(emx_sgd, 2) /= ny .or. &
        size(emx_sgd, 3) /= nz) deallocate(emx_sgd)
  end if
  if (.not. allocated(emx_sgd)) allocate(emx_sgd(nx, ny, nz, 7))
  emx_sgd(:,:,:,1) = dxyzsgd
  emx_sgd(:,:,:,2) = dxpsgd
  emx_sgd(:,:,:,3) = dxmsgd
  emx_sgd(:,:,:,4) = dypsgd
  emx_sgd(:,:,:,5) = dymsgd
  emx_sgd(:,:,:,6) = dzpsgd
  emx_sgd(:,:,:,7) = dzmsgd
  rc = emx_host_sgdm(u, dvel, jbn, emx_sgd, c4, amut, vis, vis5, nx, ny, nz)
  if (rc /= 0) call emx_fail('viscosidad_t', rc)
END SUBROUTINE viscosidad_t

! vis and vis5 from amut, pointwise (the same arithmetic as the reference).
SUBROUTINE viscosidad()
  use dimensiones
  use viscosidades
  use sgdmodel
  use consadim
  IMPLICIT NONE
  vis = c4 * 1.0 + amut
  vis5 = amut / 0.7d0
END SUBROUTINE viscosidad
