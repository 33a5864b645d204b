! Stand-in for the reference's module all_modules (all_modules.f90:1-173) for
! the test program of the drop-in shim: the global names, kinds and shapes
! STEP and FILTER_VARS read and write, without MPI, HDF5 or FFTW.  A host
! built from the reference uses its own all_modules instead.
module all_modules
  implicit none
  integer, parameter :: dp = kind(0.d0)
  real(dp), parameter :: rsun = 69598946770.0_dp, pi = 3.14159265358979_dp

  integer nx, ny, nz, nxpmlbot, nxpmltop, nypmlbot, nypmltop, nzpmlbot, nzpmltop, pulse_dir
  integer time, option, xyfreq_filtering, zfreq_filtering, g_xy_filt, g_z_filt
  logical PULSE, USE_PML, USE_HPML, PERIODIC
  real(dp) xlength, ylength, timestep, pulse_t0, pulse_t1, pulse_st, pulse_px, pulse_py, &
       pulse_pz, pulse_sr, pulse_sz, pulse_amp, s_z, s_xy, deltat, dimc, diml, dimrho, &
       stretchx, stretchy
  real(dp), dimension(5) :: betas
  real(dp), dimension(:), allocatable :: x, y, z, stretch, g
  real(dp), allocatable, dimension(:,:,:) :: p0, gradp0_z, rho0, gradrho0_z, rhoinv, c2rho0, &
       spongexyz, reduction
  real(dp), allocatable, dimension(:,:,:) :: box, boy, boz, gradp0_x, gradp0_y, curlbox, curlboy, &
       curlboz, gradrho0_x, gradrho0_y
  real(dp), allocatable, target, dimension(:,:,:,:) :: a
  real(dp), allocatable, dimension(:,:,:) :: az, bzpml, ax, bxpml, ay, bypml, kappax, kappay, kappaz
  real(dp), target, allocatable, dimension(:,:,:,:) :: zpmlvars, xpmlvars, ypmlvars
end module all_modules
