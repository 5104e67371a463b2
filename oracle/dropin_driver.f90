! Drop-in check program (test infrastructure, built into oracle/_ref/).
!
! Linked from the reference's own GPU-facing Fortran -- solver/
! deriv_vel_cuda.f90, solver/divergencia_cuda.f90, solver/advectLS.f90 and
! the module files -- plus emix3d_amd/fortran/emix3d_shim.f90, against
! libemix3d_hip.so in place of the reference's CUDA objects.  It runs the
! hot path of one RUNGEK sub-step in the reference's order
! (solver/rungek_ico.f90:79-95 and :125) through those unmodified call
! sites and writes the results for the test to compare with the oracle.
!
! Usage: emx_dropin <input file> <output file> [heviside]
! Input (stream, native endianness): nx, ny, nz (int32); deltamin, c4
! (float64); then u(3N), temp(N), um(5N), pres(N), jbn(11N), sgd(7N),
! sobject(N).  Output: dvel(9N), dtemp(3N), amut(N), vis(N), vis5(N),
! rs(5N), rsv(5N), pres(N) after advectLevelSet.
! With a third argument "heviside" it runs the reference's heviside
! (solver/advectLS.f90:29-59, CPU Fortran, no device use) instead: input
! nx, ny, nz (int32), deltamax (float64), phi(N), hevi(N) before the call;
! output hevi(N).
program emx_dropin
  use iso_c_binding
  use dimensiones
  use deltas
  use velocidades
  use derivvel
  use jacobtools
  use sgdmodel
  use viscosidades
  use variables
  use mflujos
  use right
  use multiphase
  use consadim
  implicit none
  character(len=512) :: fin, fout, mode
  integer(c_int) :: n1, n2, n3
  real(c_double) :: dmin_in, c4_in
  real(8), allocatable :: sgd(:,:,:,:), rs_all(:,:,:,:), rsv_all(:,:,:,:)
  integer :: m

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, mode)
  if (trim(mode) == 'heviside') then
    open(10, file=trim(fin), access='stream', form='unformatted', status='old')
    read(10) n1, n2, n3
    nx = n1; ny = n2; nz = n3
    allocate(phi(nx,ny,nz), hevi(nx,ny,nz))
    read(10) deltamax, phi, hevi
    close(10)
    call heviside()
    open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(11) hevi
    close(11)
    stop
  end if
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) n1, n2, n3
  read(10) dmin_in, c4_in
  nx = n1; ny = n2; nz = n3; nd = 5; ne = 5
  deltax = float(nx-1)
  deltay = float(ny-1)
  deltaz = float(nz-1)
  allocate(u(nx,ny,nz,3), temp(nx,ny,nz), pres(nx,ny,nz))
  allocate(dvel(nx,ny,nz,9), dtemp(nx,ny,nz,3), jbn(nx,ny,nz,11))
  allocate(um(nx,ny,nz,5), amut(nx,ny,nz), dxyzsgd(nx,ny,nz))
  allocate(dxpsgd(nx,ny,nz), dxmsgd(nx,ny,nz), dypsgd(nx,ny,nz))
  allocate(dymsgd(nx,ny,nz), dzpsgd(nx,ny,nz), dzmsgd(nx,ny,nz))
  allocate(vis(nx,ny,nz), vis5(nx,ny,nz))
  allocate(e(nx,ny,nz), f(nx,ny,nz), g(nx,ny,nz))
  allocate(ev(nx,ny,nz), fv(nx,ny,nz), gv(nx,ny,nz))
  allocate(rs(nx,ny,nz), rsv(nx,ny,nz), sobject(nx,ny,nz))
  allocate(sgd(nx,ny,nz,7), rs_all(nx,ny,nz,5), rsv_all(nx,ny,nz,5))
  read(10) u, temp, um, pres, jbn, sgd, sobject
  close(10)
  dxyzsgd = sgd(:,:,:,1)
  dxpsgd = sgd(:,:,:,2)
  dxmsgd = sgd(:,:,:,3)
  dypsgd = sgd(:,:,:,4)
  dymsgd = sgd(:,:,:,5)
  dzpsgd = sgd(:,:,:,6)
  dzmsgd = sgd(:,:,:,7)
  deltamin = dmin_in
  c4 = c4_in
  amut = 0.0

  call deriv_vel()
  call viscosidad_t()
  call viscosidad()
  do m = 1, 5
    call flujos(m)
    call divergencia(m)
    rs_all(:,:,:,m) = rs
    rsv_all(:,:,:,m) = rsv
  end do
  call advectLevelSet(0)

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) dvel, dtemp, amut, vis, vis5, rs_all, rsv_all, pres
  close(11)
end program emx_dropin
