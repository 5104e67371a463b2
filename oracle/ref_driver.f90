! Test-infrastructure driver for the reference Fortran hot path.
!
! This file is part of the oracle: it is compiled together with the
! reference's own Fortran sources (listed in oracle/build_ref.sh) into
! oracle/_ref/libemx_ref_<scheme>.so and is only ever loaded by tests,
! the golden-vector generator and the CPU-baseline leg of bench.py.
! It never ships in the product path.
!
! Every routine copies its C-side inputs into the reference's module
! globals, calls the unmodified reference subroutine, and copies the
! module outputs back.  Arrays are Fortran column-major (nx,ny,nz[,c]),
! FP64 (the reference is built with -fdefault-real-8, the flang spelling
! of ifort -r8 used by solver/mix3d.mk:3).

module emx_ref_state
  implicit none
  logical :: ready = .false.
end module emx_ref_state

! Allocate the module arrays the hot path touches and initialise the
! compact-scheme coefficients, mirroring the program set-up order of
! src/MSVSingle/atl_incomp.f90:29-33 (deltax=float(nx-1), then INIDERIV).
subroutine ref_init(n1, n2, n3) bind(C, name='ref_init')
  use iso_c_binding
  use dimensiones
  use deltas
  use consderper
  use consdernper
  use derivtools
  use velocidades
  use derivvel
  use mflujos
  use dmflujos
  use jacobtools
  use right
  use variables
  use viscosidades
  use sgdmodel
  use vorticidad
  use mallagrid
  use multiphase
  use consadim
  use emx_ref_state
  implicit none
  integer(c_int), value :: n1, n2, n3
  interface
    subroutine ref_free() bind(C, name='ref_free')
    end subroutine ref_free
  end interface

  if (ready) call ref_free()
  nx = n1; ny = n2; nz = n3; nd = 5; ne = 5
  allocate(axnp(nx), bxnp(nx), cxnp(nx), aynp(ny), bynp(ny), cynp(ny))
  allocate(aznp(nz), bznp(nz), cznp(nz))
  allocate(axp(nx), bxp(nx), cxp(nx), ayp(ny), byp(ny), cyp(ny))
  allocate(azp(nz), bzp(nz), czp(nz))
  allocate(du(nx), dv(ny), dw(nz))
  allocate(u(nx,ny,nz,3), temp(nx,ny,nz), pres(nx,ny,nz))
  allocate(dvel(nx,ny,nz,9), dtemp(nx,ny,nz,3))
  allocate(e(nx,ny,nz), f(nx,ny,nz), g(nx,ny,nz))
  allocate(ev(nx,ny,nz), fv(nx,ny,nz), gv(nx,ny,nz))
  allocate(dere(nx,ny,nz), derf(nx,ny,nz), derg(nx,ny,nz))
  allocate(derev(nx,ny,nz), derfv(nx,ny,nz), dergv(nx,ny,nz))
  allocate(jbn(nx,ny,nz,11))
  allocate(rs(nx,ny,nz), rsv(nx,ny,nz))
  allocate(um(nx,ny,nz,5))
  allocate(vis(nx,ny,nz), vis5(nx,ny,nz))
  allocate(amut(nx,ny,nz), dxyzsgd(nx,ny,nz))
  allocate(dxmsgd(nx,ny,nz), dxpsgd(nx,ny,nz), dymsgd(nx,ny,nz))
  allocate(dypsgd(nx,ny,nz), dzmsgd(nx,ny,nz), dzpsgd(nx,ny,nz))
  allocate(wx(nx,ny,nz), wy(nx,ny,nz), wz(nx,ny,nz))
  allocate(x(nx,ny,nz,3), dmins(nx,ny,nz))
  deltax = float(nx-1)
  deltay = float(ny-1)
  deltaz = float(nz-1)
  call inideriv()
  ready = .true.
end subroutine ref_init

subroutine ref_free() bind(C, name='ref_free')
  use dimensiones
  use consderper
  use consdernper
  use derivtools
  use velocidades
  use derivvel
  use mflujos
  use dmflujos
  use jacobtools
  use right
  use variables
  use viscosidades
  use sgdmodel
  use vorticidad
  use mallagrid
  use multiphase
  use emx_ref_state
  implicit none
  if (.not. ready) return
  deallocate(axnp, bxnp, cxnp, aynp, bynp, cynp, aznp, bznp, cznp)
  deallocate(axp, bxp, cxp, ayp, byp, cyp, azp, bzp, czp)
  deallocate(du, dv, dw, u, temp, pres, dvel, dtemp)
  deallocate(e, f, g, ev, fv, gv, dere, derf, derg, derev, derfv, dergv)
  deallocate(jbn, rs, rsv, um, vis, vis5, amut, dxyzsgd)
  deallocate(dxmsgd, dxpsgd, dymsgd, dypsgd, dzmsgd, dzpsgd)
  deallocate(wx, wy, wz, x, dmins)
  ready = .false.
end subroutine ref_free

! One line through the reference derivnper with the coefficient vectors
! of direction idir (1=xi, 2=eta, 3=zeta) that INIDERIV set up.
subroutine ref_derivnper(n, line, idir) bind(C, name='ref_derivnper')
  use iso_c_binding
  use deltas
  use consdernper
  implicit none
  integer(c_int), value :: n, idir
  real(c_double) :: line(n)
  if (idir == 1) then
    call derivnper(n, line, axnp, bxnp, cxnp, deltax)
  else if (idir == 2) then
    call derivnper(n, line, aynp, bynp, cynp, deltay)
  else
    call derivnper(n, line, aznp, bznp, cznp, deltaz)
  end if
end subroutine ref_derivnper

! Grid metrics: solver/jacob_malla.f90 JACOBEANO.
subroutine ref_jacobeano(xin, jout) bind(C, name='ref_jacobeano')
  use iso_c_binding
  use dimensiones
  use mallagrid
  use jacobtools
  implicit none
  real(c_double), intent(in) :: xin(nx,ny,nz,3)
  real(c_double), intent(out) :: jout(nx,ny,nz,11)
  x = xin
  call jacobeano()
  jout = jbn
end subroutine ref_jacobeano

! Velocity and temperature gradients: solver/deriv_vel.f90 DERIV_VEL.
subroutine ref_deriv_vel(uin, tin, dvout, dtout) bind(C, name='ref_deriv_vel')
  use iso_c_binding
  use dimensiones
  use velocidades
  use derivvel
  implicit none
  real(c_double), intent(in) :: uin(nx,ny,nz,3), tin(nx,ny,nz)
  real(c_double), intent(out) :: dvout(nx,ny,nz,9), dtout(nx,ny,nz,3)
  u = uin
  temp = tin
  call deriv_vel()
  dvout = dvel
  dtout = dtemp
end subroutine ref_deriv_vel

! Pointwise fluxes: solver/flujos_compact.f90 flujos(neq).
! fout(:,:,:,1..6) = e, f, g, ev, fv, gv.
subroutine ref_flujos(neq, umin, uin, pin, visin, jin, dvin, fout) &
    bind(C, name='ref_flujos')
  use iso_c_binding
  use dimensiones
  use velocidades
  use variables
  use viscosidades
  use jacobtools
  use derivvel
  use mflujos
  implicit none
  integer(c_int), value :: neq
  real(c_double), intent(in) :: umin(nx,ny,nz,5), uin(nx,ny,nz,3)
  real(c_double), intent(in) :: pin(nx,ny,nz), visin(nx,ny,nz)
  real(c_double), intent(in) :: jin(nx,ny,nz,11), dvin(nx,ny,nz,9)
  real(c_double), intent(out) :: fout(nx,ny,nz,6)
  um = umin
  u = uin
  pres = pin
  vis = visin
  jbn = jin
  dvel = dvin
  call flujos(neq)
  fout(:,:,:,1) = e
  fout(:,:,:,2) = f
  fout(:,:,:,3) = g
  fout(:,:,:,4) = ev
  fout(:,:,:,5) = fv
  fout(:,:,:,6) = gv
end subroutine ref_flujos

! Flux divergence: solver/divergencia.f90 divergencia(neq).
! fin(:,:,:,1..6) = e, f, g, ev, fv, gv.
subroutine ref_divergencia(neq, fin, jin, rsout, rsvout) &
    bind(C, name='ref_divergencia')
  use iso_c_binding
  use dimensiones
  use mflujos
  use jacobtools
  use right
  implicit none
  integer(c_int), value :: neq
  real(c_double), intent(in) :: fin(nx,ny,nz,6), jin(nx,ny,nz,11)
  real(c_double), intent(out) :: rsout(nx,ny,nz), rsvout(nx,ny,nz)
  e = fin(:,:,:,1)
  f = fin(:,:,:,2)
  g = fin(:,:,:,3)
  ev = fin(:,:,:,4)
  fv = fin(:,:,:,5)
  gv = fin(:,:,:,6)
  jbn = jin
  call divergencia(neq)
  rsout = rs
  rsvout = rsv
end subroutine ref_divergencia

! SGS geometry: solver/sgdm_incomp.f90 INISGDM.
! sgd(:,:,:,1..7) = dxyzsgd, dxpsgd, dxmsgd, dypsgd, dymsgd, dzpsgd, dzmsgd.
subroutine ref_inisgdm(xin, sgd, dminsout, dmaxout) bind(C, name='ref_inisgdm')
  use iso_c_binding
  use dimensiones
  use mallagrid
  use sgdmodel
  use multiphase
  implicit none
  real(c_double), intent(in) :: xin(nx,ny,nz,3)
  real(c_double), intent(out) :: sgd(nx,ny,nz,7), dminsout(nx,ny,nz)
  real(c_double), intent(out) :: dmaxout
  x = xin
  call inisgdm()
  sgd(:,:,:,1) = dxyzsgd
  sgd(:,:,:,2) = dxpsgd
  sgd(:,:,:,3) = dxmsgd
  sgd(:,:,:,4) = dypsgd
  sgd(:,:,:,5) = dymsgd
  sgd(:,:,:,6) = dzpsgd
  sgd(:,:,:,7) = dzmsgd
  dminsout = dmins
  dmaxout = deltamax
end subroutine ref_inisgdm

! Eddy viscosity: solver/sgdm_incomp.f90 viscosidad_t (structure function,
! VORT, SELECTIVO) followed by solver/viscosidad_incomp.f90 viscosidad.
! The SGS geometry sgd (layout of ref_inisgdm) and the model constants
! come from INISGDM, which the caller runs first through ref_inisgdm.
subroutine ref_sgdm(uin, dvin, jin, sgd, c4in, amutout, visout, vis5out) &
    bind(C, name='ref_sgdm')
  use iso_c_binding
  use dimensiones
  use velocidades
  use derivvel
  use jacobtools
  use sgdmodel
  use viscosidades
  use consadim
  implicit none
  real(c_double), intent(in) :: uin(nx,ny,nz,3), dvin(nx,ny,nz,9)
  real(c_double), intent(in) :: jin(nx,ny,nz,11), sgd(nx,ny,nz,7)
  real(c_double), value :: c4in
  real(c_double), intent(out) :: amutout(nx,ny,nz), visout(nx,ny,nz)
  real(c_double), intent(out) :: vis5out(nx,ny,nz)
  u = uin
  dvel = dvin
  jbn = jin
  dxyzsgd = sgd(:,:,:,1)
  dxpsgd = sgd(:,:,:,2)
  dxmsgd = sgd(:,:,:,3)
  dypsgd = sgd(:,:,:,4)
  dymsgd = sgd(:,:,:,5)
  dzpsgd = sgd(:,:,:,6)
  dzmsgd = sgd(:,:,:,7)
  c4 = c4in
  call viscosidad_t()
  call viscosidad()
  amutout = amut
  visout = vis
  vis5out = vis5
end subroutine ref_sgdm

! One whole RK-stage spatial operator as RUNGEK sub-step 1 runs it
! (solver/rungek_ico.f90:79-95): DERIV_VEL, viscosidad_t, viscosidad,
! then flujos(m)+divergencia(m) for m=1..5.  Inputs are loaded once by
! ref_stage_load; ref_stage_run is the timed part; rs/rsv of every
! equation are returned by ref_stage_fetch.
module emx_ref_stage
  implicit none
  real(8), allocatable :: rs_all(:,:,:,:), rsv_all(:,:,:,:)
end module emx_ref_stage

subroutine ref_stage_load(umin, uin, tin, pin, jin, sgd, c4in) &
    bind(C, name='ref_stage_load')
  use iso_c_binding
  use dimensiones
  use velocidades
  use variables
  use jacobtools
  use sgdmodel
  use consadim
  use emx_ref_stage
  implicit none
  real(c_double), intent(in) :: umin(nx,ny,nz,5), uin(nx,ny,nz,3)
  real(c_double), intent(in) :: tin(nx,ny,nz), pin(nx,ny,nz)
  real(c_double), intent(in) :: jin(nx,ny,nz,11), sgd(nx,ny,nz,7)
  real(c_double), value :: c4in
  um = umin
  u = uin
  temp = tin
  pres = pin
  jbn = jin
  dxyzsgd = sgd(:,:,:,1)
  dxpsgd = sgd(:,:,:,2)
  dxmsgd = sgd(:,:,:,3)
  dypsgd = sgd(:,:,:,4)
  dymsgd = sgd(:,:,:,5)
  dzpsgd = sgd(:,:,:,6)
  dzmsgd = sgd(:,:,:,7)
  c4 = c4in
  ! isgm, amu0 and cutoff were set by INISGDM (ref_inisgdm runs first).
  if (allocated(rs_all)) deallocate(rs_all, rsv_all)
  allocate(rs_all(nx,ny,nz,5), rsv_all(nx,ny,nz,5))
end subroutine ref_stage_load

subroutine ref_stage_run() bind(C, name='ref_stage_run')
  use dimensiones
  use right
  use emx_ref_stage
  implicit none
  integer :: m
  call deriv_vel()
  call viscosidad_t()
  call viscosidad()
  do m = 1, 5
    call flujos(m)
    call divergencia(m)
    rs_all(:,:,:,m) = rs
    rsv_all(:,:,:,m) = rsv
  end do
end subroutine ref_stage_run

subroutine ref_stage_fetch(rsout, rsvout, amutout, visout, dvout) &
    bind(C, name='ref_stage_fetch')
  use iso_c_binding
  use dimensiones
  use sgdmodel
  use viscosidades
  use derivvel
  use emx_ref_stage
  implicit none
  real(c_double), intent(out) :: rsout(nx,ny,nz,5), rsvout(nx,ny,nz,5)
  real(c_double), intent(out) :: amutout(nx,ny,nz), visout(nx,ny,nz)
  real(c_double), intent(out) :: dvout(nx,ny,nz,9)
  rsout = rs_all
  rsvout = rsv_all
  amutout = amut
  visout = vis
  dvout = dvel
end subroutine ref_stage_fetch

! Flush the Fortran runtime's buffered standard output (the reference prints
! from divergencia and viscosidad_t), so callers can redirect it.
subroutine ref_flush() bind(C, name='ref_flush')
  flush(6)
end subroutine ref_flush

! ---- RUNGEK glue: boundary values and variable conversions -----------------

! frontera_z_val then frontera_y_val on um1, RUNGEK's first two boundary
! calls (solver/rungek_ico.f90:96-97).  frontera_x_val is not built: its
! source file also holds a routine that needs Intel's ifport module.
subroutine ref_frontera_zy(umio) bind(C, name='ref_frontera_zy')
  use iso_c_binding
  use dimensiones
  use variables
  implicit none
  real(c_double), intent(inout) :: umio(nx,ny,nz,5)
  if (allocated(um1)) deallocate(um1)
  allocate(um1(nx,ny,nz,5))
  um1 = umio
  call frontera_z_val()
  call frontera_y_val()
  umio = um1
  deallocate(um1)
end subroutine ref_frontera_zy

subroutine ref_varc_to_varnc(umin, machin, uout, pout, tout, cout) &
    bind(C, name='ref_varc_to_varnc')
  use iso_c_binding
  use dimensiones
  use variables
  use velocidades
  use flow
  implicit none
  real(c_double), intent(in) :: umin(nx,ny,nz,5)
  real(c_double), value :: machin
  real(c_double), intent(out) :: uout(nx,ny,nz,3), pout(nx,ny,nz), tout(nx,ny,nz)
  real(c_double), intent(out) :: cout(nx,ny,nz)
  if (allocated(um1)) deallocate(um1)
  allocate(um1(nx,ny,nz,5))
  if (.not. allocated(conc)) allocate(conc(nx,ny,nz))
  um1 = umin
  mach = machin
  call varc_to_varnc()
  uout = u
  pout = pres
  tout = temp
  cout = conc
  deallocate(um1)
end subroutine ref_varc_to_varnc

subroutine ref_varnc_to_varc(uin, pin, tin, machin, umout, p0out) &
    bind(C, name='ref_varnc_to_varc')
  use iso_c_binding
  use dimensiones
  use variables
  use velocidades
  use flow
  implicit none
  real(c_double), intent(in) :: uin(nx,ny,nz,3), pin(nx,ny,nz), tin(nx,ny,nz)
  real(c_double), value :: machin
  real(c_double), intent(out) :: umout(nx,ny,nz,5), p0out(nx,ny,nz)
  if (allocated(um0)) deallocate(um0)
  allocate(um0(nx,ny,nz,5))
  if (.not. allocated(pres0)) allocate(pres0(nx,ny,nz))
  u = uin
  pres = pin
  temp = tin
  mach = machin
  call varnc_to_varc()
  umout = um0
  p0out = pres0
  deallocate(um0)
end subroutine ref_varnc_to_varc

! ---- compact filter ---------------------------------------------------------

! INIFILTRO with the given cfilt (src/MSVSingle/filtro_incomp.f90:2-166),
! then FILTRADO(itf) on um1 (herramientas_incomp2.f90:227-410); returns the
! filtered um1 and the u, temp, pres it recomputes.
subroutine ref_filtrado(itf, umio, jin, cfiltin, machin, uout, tout, pout) &
    bind(C, name='ref_filtrado')
  use iso_c_binding
  use dimensiones
  use variables
  use velocidades
  use jacobtools
  use consfiltro
  use consderper_f
  use consdernper_f
  use flow
  implicit none
  integer(c_int), value :: itf
  real(c_double), intent(inout) :: umio(nx,ny,nz,5)
  real(c_double), intent(in) :: jin(nx,ny,nz,11)
  real(c_double), value :: cfiltin, machin
  real(c_double), intent(out) :: uout(nx,ny,nz,3), tout(nx,ny,nz), pout(nx,ny,nz)
  integer :: nmcit
  if (allocated(um1)) deallocate(um1)
  allocate(um1(nx,ny,nz,5))
  if (allocated(axnpf)) deallocate(axnpf, aynpf, aznpf, bxnpf, bynpf, bznpf, &
                                   cxnpf, cynpf, cznpf)
  if (allocated(axpf)) deallocate(axpf, aypf, azpf, bxpf, bypf, bzpf, cxpf, cypf, czpf)
  allocate(axnpf(nx), bxnpf(nx), cxnpf(nx), aynpf(ny), bynpf(ny), cynpf(ny))
  allocate(aznpf(nz), bznpf(nz), cznpf(nz))
  allocate(axpf(nx), bxpf(nx), cxpf(nx), aypf(ny), bypf(ny), cypf(ny))
  allocate(azpf(nz), bzpf(nz), czpf(nz))
  cfilt = cfiltin
  mach = machin
  call inifiltro()
  um1 = umio
  jbn = jin
  nmcit = 1
  call filtrado(itf, nmcit)
  umio = um1
  uout = u
  tout = temp
  pout = pres
  deallocate(um1)
end subroutine ref_filtrado

! ---- on-disk formats (run in the caller's working directory) ---------------

! GRAVA (solver/gravar_incomp.f90:2-53): writes field_SS.NNN, time_SS.NNN
! and impellermask_SS.NNN for serie SS = n_serie, sample NNN = i_sample.
subroutine ref_grava(uin, tin, pin, maskin, dtoin, itin, serie_in, sample_in) &
    bind(C, name='ref_grava')
  use iso_c_binding
  use dimensiones
  use velocidades
  use fieldrec
  use tiempo
  use maskvar
  implicit none
  real(c_double), intent(in) :: uin(nx,ny,nz,3), tin(nx,ny,nz), pin(nx,ny,nz)
  real(c_double), intent(in) :: maskin(nx,ny,nz)
  real(c_double), value :: dtoin
  integer(c_int), value :: itin, serie_in, sample_in
  if (allocated(impellermask)) deallocate(impellermask)
  allocate(impellermask(nx,ny,nz))
  u = uin
  temp = tin
  pres = pin
  impellermask = maskin
  dto = dtoin
  it = itin
  n_serie = serie_in
  i_sample = sample_in
  dgrava1 = 1.0
  call grava()
  deallocate(impellermask)
end subroutine ref_grava

! LEER_CAMPOS (solver/leer_incomp.f90:45-116): reads the field file named
! fname and sobject.field, then VARNC_TO_VARC.
subroutine ref_leer_campos(fname, n, machin, uout, tout, pout, umout, sout) &
    bind(C, name='ref_leer_campos')
  use iso_c_binding
  use dimensiones
  use velocidades
  use variables
  use inputname
  use multiphase
  use flow
  implicit none
  integer(c_int), value :: n
  character(kind=c_char), intent(in) :: fname(n)
  real(c_double), value :: machin
  real(c_double), intent(out) :: uout(nx,ny,nz,3), tout(nx,ny,nz), pout(nx,ny,nz)
  real(c_double), intent(out) :: umout(nx,ny,nz,5), sout(nx,ny,nz)
  integer :: i
  inputfield = ' '
  do i = 1, n
    inputfield(i:i) = fname(i)
  end do
  mach = machin
  if (allocated(phi)) deallocate(phi)
  if (allocated(dmins)) deallocate(dmins)
  if (allocated(hevi)) deallocate(hevi)
  if (allocated(rhoTwoPhase)) deallocate(rhoTwoPhase)
  if (allocated(muTwoPhase)) deallocate(muTwoPhase)
  if (allocated(rhoint)) deallocate(rhoint)
  if (allocated(muint)) deallocate(muint)
  if (allocated(sobject)) deallocate(sobject)
  if (allocated(um0)) deallocate(um0)
  if (.not. allocated(pres0)) allocate(pres0(nx,ny,nz))
  allocate(um0(nx,ny,nz,5))
  call leer_campos()
  uout = u
  tout = temp
  pout = pres
  umout = um0
  sout = sobject
  deallocate(phi, hevi, rhoTwoPhase, muTwoPhase, rhoint, muint, sobject, um0)
end subroutine ref_leer_campos

! LEER_MALLA (solver/leer_incomp.f90:1-41): reads the grid file fname.
subroutine ref_leer_malla(fname, n, xout) bind(C, name='ref_leer_malla')
  use iso_c_binding
  use dimensiones
  use mallagrid
  use inputname
  implicit none
  integer(c_int), value :: n
  character(kind=c_char), intent(in) :: fname(n)
  real(c_double), intent(out) :: xout(nx,ny,nz,3)
  integer :: i
  inputgrid = ' '
  do i = 1, n
    inputgrid(i:i) = fname(i)
  end do
  call leer_malla()
  xout = x
end subroutine ref_leer_malla

! LEER_TIME (solver/leer_incomp.f90:245-262).
subroutine ref_leer_time(fname, n, dtoout, itout) bind(C, name='ref_leer_time')
  use iso_c_binding
  use inputname
  use tiempo
  implicit none
  integer(c_int), value :: n
  character(kind=c_char), intent(in) :: fname(n)
  real(c_double), intent(out) :: dtoout
  integer(c_int), intent(out) :: itout
  integer :: i
  inputtime = ' '
  do i = 1, n
    inputtime(i:i) = fname(i)
  end do
  call leer_time()
  dtoout = dto
  itout = it
end subroutine ref_leer_time
