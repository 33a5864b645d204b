! Test program of the drop-in shim: a host that fills the reference's
! globals (here read from a file the test writes: the set-up after
! INITIALISE_RHS, the CPML faces, the initial state) and runs the time loop
! of driver.f90:123-202 with the shim's STEP and FILTER_VARS, then writes a.
!
!   drive_quiet <input file> <output file>
!
! Input (stream, native doubles and 32-bit ints): int header(24), real
! header(32), then the arrays named below in Fortran order.
program drive_quiet
  use iso_c_binding
  use all_modules
  use sparc_gpu
  implicit none
  integer :: hi(24), nv, npsi, nmax, nn, f, u, w
  real(dp) :: hr(32)
  character(len=512) :: fin, fout
  logical :: fail
  integer, parameter :: SPARC_ZBOT = 4, SPARC_ZTOP = 5

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hi, hr
  nx = hi(1); ny = hi(2); nz = hi(3); option = hi(4)
  PERIODIC = hi(5) /= 0; USE_PML = hi(6) /= 0; USE_HPML = .false.
  nxpmlbot = 0; nxpmltop = 0; nypmlbot = 0; nypmltop = 0
  nzpmlbot = hi(7); nzpmltop = hi(8)
  g_xy_filt = hi(9); g_z_filt = hi(10)
  xyfreq_filtering = hi(11); zfreq_filtering = hi(12)
  PULSE = hi(13) /= 0; pulse_dir = hi(14)
  nmax = hi(15)
  nv = 5; npsi = 2
  s_xy = hr(1); s_z = hr(2); deltat = hr(3); betas = hr(4:8)
  stretchx = hr(9); stretchy = hr(10); timestep = hr(11)
  pulse_amp = hr(12); pulse_t0 = hr(13); pulse_t1 = hr(14); pulse_st = hr(15)
  pulse_px = hr(16); pulse_py = hr(17); pulse_pz = hr(18); pulse_sr = hr(19); pulse_sz = hr(20)
  dimc = hr(21); diml = hr(22); dimrho = hr(23); xlength = hr(24); ylength = hr(25)

  allocate(x(nx), y(ny), z(nz), stretch(nz), g(nz))
  allocate(rho0(nx,ny,nz), rhoinv(nx,ny,nz), c2rho0(nx,ny,nz), gradrho0_z(nx,ny,nz), &
       gradp0_z(nx,ny,nz), spongexyz(nx,ny,nz), p0(nx,ny,nz), a(nx,ny,nz,nv))
  read(u) x, y, z, stretch, g
  read(u) rho0, rhoinv, c2rho0, gradrho0_z, gradp0_z, spongexyz, p0
  read(u) a
  ! the two z faces, one after the other through the per-rank globals, as
  ! INITIALISE_RHS would leave them on the bottom and the top rank
  do f = SPARC_ZBOT, SPARC_ZTOP
    w = merge(nzpmlbot, nzpmltop, f == SPARC_ZBOT)
    if (.not. USE_PML .or. w == 0) cycle
    if (allocated(az)) deallocate(az, bzpml, kappaz, zpmlvars)
    allocate(az(nx,ny,w), bzpml(nx,ny,w), kappaz(nx,ny,w), zpmlvars(nx,ny,w,npsi))
    read(u) az, bzpml, kappaz, zpmlvars
    call sparc_gpu_set_face(f)
  end do
  close(u)

  ! driver.f90:123-202 with the device-side failure checks
  do nn = 0, nmax
    time = nn
    call step()
    fail = .false.
    call sparc_gpu_guard(fail)
    if (fail) then
      print *, 'A crash has occured at iteration ', time
      stop 2
    end if
    if (time /= 0) call filter_vars()
  end do

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) a
  do f = SPARC_ZBOT, SPARC_ZTOP
    w = merge(nzpmlbot, nzpmltop, f == SPARC_ZBOT)
    if (.not. USE_PML .or. w == 0) cycle
    deallocate(zpmlvars)
    allocate(zpmlvars(nx,ny,w,npsi))
    call sparc_gpu_get_face(f)
    write(u) zpmlvars
  end do
  close(u)
  call sparc_gpu_finalize()
  print *, 'done ', nmax + 1, ' steps'
end program drive_quiet
