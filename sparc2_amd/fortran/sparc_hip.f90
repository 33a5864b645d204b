! sparc_hip.f90 -- ISO_C_BINDING interface to libsparc_hip.so for the SPARC2
! Fortran host.  Types mirror the C structs of include/sparc_hip.h member by
! member (bind(C) gives them the C layout); every function returns an
! integer(c_int) status, 0 on success.
!
! The host keeps its own READ_SPARC_PARAMETERS, READ_IN_HDF5 and
! INITIALISE_RHS (initialize.f90, all_modules.f90) and hands their results to
! the context once; STEP (step.f90:20-63) and FILTER_VARS
! (derivatives.f90:659-703) then become sparc_hip_step / sparc_hip_filter.
module sparc_hip
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: SPARC_OK = 0, SPARC_E_INVALID = -1, &
       SPARC_E_HIP = -2, SPARC_E_NOMEM = -3, SPARC_E_STATE = -4
  integer(c_int), parameter, public :: SPARC_PREC_FP64 = 0, SPARC_PREC_FP32 = 1
  integer(c_int), parameter, public :: SPARC_XBOT = 0, SPARC_XTOP = 1, SPARC_YBOT = 2, &
       SPARC_YTOP = 3, SPARC_ZBOT = 4, SPARC_ZTOP = 5

  type, bind(C), public :: sparc_cfg
    integer(c_int) :: nx, ny, nz
    integer(c_int) :: option          ! 1 MP_QUIET_3D, 3 MP_MHD_3D
    integer(c_int) :: precision
    integer(c_int) :: periodic
    integer(c_int) :: use_pml, use_hpml
    integer(c_int) :: npml(6)         ! nxpmlbot, nxpmltop, nypmlbot, nypmltop, nzpmlbot, nzpmltop
    integer(c_int) :: g_xy_filt, g_z_filt
    real(c_double) :: s_xy, s_z
    integer(c_int) :: xyfreq, zfreq
    integer(c_int) :: pulse_dir
    integer(c_int) :: bg3d
    real(c_double) :: deltat
    real(c_double) :: betas(5)
    real(c_double) :: stretchx, stretchy
    type(c_ptr) :: stretch            ! c_loc(stretch(1)), nz values
    integer(c_int) :: device
    integer(c_int) :: tile(3)
  end type sparc_cfg

  type, bind(C), public :: sparc_bg
    type(c_ptr) :: rho0, rhoinv, c2rho0, gradrho0_z, gradp0_z, spongexyz, p0
    type(c_ptr) :: g
    type(c_ptr) :: box, boy, boz, curlbox, curlboy, curlboz
    type(c_ptr) :: gradrho0_x, gradrho0_y, gradp0_x, gradp0_y, reduction
    type(c_ptr) :: pml_a(6), pml_b(6), pml_kappa(6)
    type(c_ptr) :: src_x, src_y, src_z
    real(c_double) :: dimc, dimrho
  end type sparc_bg

  public :: sparc_hip_create, sparc_hip_destroy, sparc_hip_set_background, &
       sparc_hip_put_state, sparc_hip_get_state, sparc_hip_step, sparc_hip_filter, &
       sparc_hip_advance, sparc_hip_diag, sparc_hip_ddx, sparc_hip_ddy, sparc_hip_ddz, &
       sparc_hip_filtx, sparc_hip_filty, sparc_hip_filtz, sparc_hip_sync, &
       sparc_hip_timing, sparc_hip_kernel_time, sparc_hip_tile, sparc_hip_version, &
       sparc_hip_alfven_search, sparc_hip_get_slab, sparc_hip_batch_create, &
       sparc_hip_batch_advance, sparc_hip_batch_info, sparc_hip_batch_destroy, &
       sparc_hip_autotune, sparc_hip_set_tile, sparc_hip_batch_time_stage, sparc_hip_set_damping, &
       sparc_hip_set_forcing, sparc_hip_set_guard, sparc_hip_guard, sparc_hip_batch_filter, &
       sparc_hip_stream_bench, sparc_hip_deriv_noalloc, sparc_hip_device_memory, &
       sparc_error_message

  interface
    integer(c_int) function sparc_hip_create(cfg, ctx) bind(C, name="sparc_hip_create")
      import :: c_int, c_ptr, sparc_cfg
      type(sparc_cfg), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function sparc_hip_destroy(ctx) bind(C, name="sparc_hip_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function sparc_hip_set_background(ctx, bg) bind(C, name="sparc_hip_set_background")
      import :: c_int, c_ptr, sparc_bg
      type(c_ptr), value :: ctx
      type(sparc_bg), intent(in) :: bg
    end function

    ! rates: c_loc of damping_rates(nx/2+1, ny) (INIT_DAMPING), or c_null_ptr
    ! to switch the FFT damping off
    integer(c_int) function sparc_hip_set_damping(ctx, rates) bind(C, name="sparc_hip_set_damping")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: rates
    end function

    ! PULSE = 0: map = c_loc(forcing(nx,ny)) after LAGRANGE_INTERP and
    ! source_damp, sdep = c_loc(source_dep(nz)) or c_null_ptr to keep it;
    ! map = c_null_ptr returns to the pulse
    integer(c_int) function sparc_hip_set_forcing(ctx, map, sdep) bind(C, name="sparc_hip_set_forcing")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: map
      type(c_ptr), value :: sdep
    end function

    ! a: the reference's a(dim1,dim2,dim3,dimfive); psi(f): c_loc of the PML
    ! bases (zpmlvars & co.) or c_null_ptr
    integer(c_int) function sparc_hip_put_state(ctx, a, psi) bind(C, name="sparc_hip_put_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: a(*)
      type(c_ptr), intent(in) :: psi(6)
    end function

    integer(c_int) function sparc_hip_get_state(ctx, a, psi) bind(C, name="sparc_hip_get_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: a(*)
      type(c_ptr), intent(in) :: psi(6)
    end function

    integer(c_int) function sparc_hip_step(ctx, src_amp) bind(C, name="sparc_hip_step")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: src_amp
    end function

    integer(c_int) function sparc_hip_filter(ctx, time) bind(C, name="sparc_hip_filter")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: time
    end function

    integer(c_int) function sparc_hip_advance(ctx, n0, n1, src_amp) bind(C, name="sparc_hip_advance")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n0, n1
      real(c_double), intent(in) :: src_amp(*)
    end function

    integer(c_int) function sparc_hip_diag(ctx, out) bind(C, name="sparc_hip_diag")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(3)
    end function

    integer(c_int) function sparc_hip_ddx(ctx, input, output, bc) bind(C, name="sparc_hip_ddx")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: input(*)
      real(c_double), intent(out) :: output(*)
      integer(c_int), value :: bc
    end function

    integer(c_int) function sparc_hip_ddy(ctx, input, output, bc) bind(C, name="sparc_hip_ddy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: input(*)
      real(c_double), intent(out) :: output(*)
      integer(c_int), value :: bc
    end function

    integer(c_int) function sparc_hip_ddz(ctx, input, output, bc) bind(C, name="sparc_hip_ddz")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: input(*)
      real(c_double), intent(out) :: output(*)
      integer(c_int), value :: bc
    end function

    integer(c_int) function sparc_hip_filtx(ctx, input, output) bind(C, name="sparc_hip_filtx")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: input(*)
      real(c_double), intent(out) :: output(*)
    end function

    integer(c_int) function sparc_hip_filty(ctx, input, output) bind(C, name="sparc_hip_filty")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: input(*)
      real(c_double), intent(out) :: output(*)
    end function

    integer(c_int) function sparc_hip_filtz(ctx, input, output) bind(C, name="sparc_hip_filtz")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: input(*)
      real(c_double), intent(out) :: output(*)
    end function

    integer(c_int) function sparc_hip_sync(ctx) bind(C, name="sparc_hip_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function sparc_hip_timing(ctx, enable) bind(C, name="sparc_hip_timing")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
    end function

    integer(c_int) function sparc_hip_kernel_time(ctx, kind, count, total_ms) &
         bind(C, name="sparc_hip_kernel_time")
      import :: c_int, c_ptr, c_long, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: kind
      integer(c_long), intent(out) :: count
      real(c_double), intent(out) :: total_ms
    end function

    integer(c_int) function sparc_hip_tile(ctx, tile, nblocks) bind(C, name="sparc_hip_tile")
      import :: c_int, c_ptr, c_long
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: tile(3)
      integer(c_long), intent(out) :: nblocks
    end function

    ! Alfven-limiter search temperature (initialize.f90:962-986)
    integer(c_int) function sparc_hip_alfven_search(ctx, red, b2, rho0, dimb, dimrho, maxva, temp) &
         bind(C, name="sparc_hip_alfven_search")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: red(*), b2(*), rho0(*)
      real(c_double), value :: dimb, dimrho, maxva
      real(c_double), intent(out) :: temp
    end function

    ! planes k0..k1 (0-based) of one state field: the WRITE_OUT_CHUNK read-out
    integer(c_int) function sparc_hip_get_slab(ctx, field, k0, k1, out) bind(C, name="sparc_hip_get_slab")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: field, k0, k1
      real(c_double), intent(out) :: out(*)
    end function

    ! batched form: runs(nrun) are contexts created on one device
    integer(c_int) function sparc_hip_batch_create(runs, nrun, use_graph, batch) &
         bind(C, name="sparc_hip_batch_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: runs(*)
      integer(c_int), value :: nrun, use_graph
      type(c_ptr), intent(out) :: batch
    end function

    ! amps(nsteps, nrun): amps(n, r) is run r's source amplitude at step n;
    ! flags = 1 (SPARC_DEFER_FILTER) leaves the last step's filters to
    ! sparc_hip_batch_filter, so outputs read the pre-filter state
    integer(c_int) function sparc_hip_batch_advance(batch, nsteps, time0, amps, flags) &
         bind(C, name="sparc_hip_batch_advance")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch
      integer(c_int), value :: nsteps
      integer(c_int), intent(in) :: time0(*)
      real(c_double), intent(in) :: amps(*)
      integer(c_int), value :: flags
    end function

    integer(c_int) function sparc_hip_batch_filter(batch, time) bind(C, name="sparc_hip_batch_filter")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), intent(in) :: time(*)
    end function

    ! ddx/ddy/ddz (axis 0/1/2) with no device allocation (static staging,
    ! grids of at most 4194304 points; overlapping calls on a device race)
    integer(c_int) function sparc_hip_deriv_noalloc(ctx, axis, input, output, bc) &
         bind(C, name="sparc_hip_deriv_noalloc")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: axis
      real(c_double), intent(in) :: input(*)
      real(c_double), intent(out) :: output(*)
      integer(c_int), value :: bc
    end function

    integer(c_int) function sparc_hip_device_memory(device, free_bytes, total_bytes) &
         bind(C, name="sparc_hip_device_memory")
      import :: c_int, c_long
      integer(c_int), value :: device
      integer(c_long), intent(out) :: free_bytes, total_bytes
    end function

    ! bandwidth probe: nread arrays of n doubles summed into one more
    integer(c_int) function sparc_hip_stream_bench(device, n, nread, reps, gbps) &
         bind(C, name="sparc_hip_stream_bench")
      import :: c_int, c_long, c_double
      integer(c_int), value :: device
      integer(c_long), value :: n
      integer(c_int), value :: nread, reps
      real(c_double), intent(out) :: gbps
    end function

    ! failure checks of driver.f90:175-196 on the device: mode 0 off,
    ! 1 record, 2 halt at the failing step (default)
    integer(c_int) function sparc_hip_set_guard(ctx, mode) bind(C, name="sparc_hip_set_guard")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
    end function

    integer(c_int) function sparc_hip_guard(ctx, tripped, at_step, steps, diag) &
         bind(C, name="sparc_hip_guard")
      import :: c_int, c_ptr, c_long, c_double
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: tripped
      integer(c_long), intent(out) :: at_step, steps
      real(c_double), intent(out) :: diag(3)
    end function

    integer(c_int) function sparc_hip_batch_info(batch, ngroups, nlaunch_blocks) &
         bind(C, name="sparc_hip_batch_info")
      import :: c_int, c_ptr, c_long
      type(c_ptr), value :: batch
      integer(c_int), intent(out) :: ngroups
      integer(c_long), intent(out) :: nlaunch_blocks
    end function

    integer(c_int) function sparc_hip_batch_time_stage(batch, reps, ms_out) &
         bind(C, name="sparc_hip_batch_time_stage")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch
      integer(c_int), value :: reps
      real(c_double), intent(out) :: ms_out
    end function

    integer(c_int) function sparc_hip_batch_destroy(batch) bind(C, name="sparc_hip_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function

    ! tile autotuner: tile_out = (tx, ty, kc), ms_out = time per middle stage
    integer(c_int) function sparc_hip_autotune(ctx, reps, tile_out, ms_out) &
         bind(C, name="sparc_hip_autotune")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: reps
      integer(c_int), intent(out) :: tile_out(3)
      real(c_double), intent(out) :: ms_out
    end function

    integer(c_int) function sparc_hip_set_tile(ctx, tile) bind(C, name="sparc_hip_set_tile")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: tile(3)
    end function

    integer(c_int) function sparc_hip_version() bind(C, name="sparc_hip_version")
      import :: c_int
    end function

    type(c_ptr) function sparc_hip_last_error_c() bind(C, name="sparc_hip_last_error")
      import :: c_ptr
    end function

    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! Message of the most recent failure, as a Fortran string.
  function sparc_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = sparc_hip_last_error_c()
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function sparc_error_message

end module sparc_hip
