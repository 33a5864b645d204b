! sparc_gpu.f90 -- drop-in replacements of the reference's time-step routines
! with the reference's names and argument-less signatures:
!
!   STEP          step.f90:20-63        (called at driver.f90:157)
!   FILTER_VARS   derivatives.f90:659-703 (called at driver.f90:200)
!
! They work on the globals of module all_modules, as the routines they replace
! do, and advance the run on the GPU through the C ABI (module sparc_hip,
! include/sparc_hip.h).  A host built from the reference uses this module in
! place of Time_Step's STEP and DERIVATIVES' FILTER_VARS; its initialisation
! (READ_SPARC_PARAMETERS, INITANDALLOCATE, INITIALISE_RHS) is unchanged.
!
! What the host adds:
!  * the reference keeps the CPML arrays of one face per MPI rank (az, bzpml,
!    kappaz, zpmlvars for its z face; ax/bxpml/kappax/xpmlvars, ay/...).  One
!    GPU holds the whole domain, so after INITIALISE_RHS has filled a face's
!    arrays the host calls sparc_gpu_set_face(face) (SPARC_ZBOT, SPARC_ZTOP,
!    ...) to hand them over, for every active face;
!  * the failure checks of driver.f90:175-183 run on the device: the host
!    calls sparc_gpu_guard(fail) after STEP (fail as driver.f90:186 uses it);
!  * sparc_gpu_finalize() releases the device context.
!
! The first STEP creates the device context from the globals (grid, switches,
! deltat, betas, the metric, the backgrounds after INITIALISE_RHS, the pulse
! parameters) and uploads a and the CPML bases -- after a restart that is the
! state READ_IN_FULL_STATE loaded.  Every STEP and FILTER_VARS then leaves a
! (and the CPML bases) updated on the host, so norm(), the outputs and the
! restart files of the host read current values.  Hosts that read them only
! at output steps call sparc_gpu_pull_every_step(.false.) and sparc_gpu_pull()
! before reading.  Errors from the library stop the program with its message,
! as the reference stops on its fatal checks.
module sparc_gpu
  use iso_c_binding
  use all_modules
  use sparc_hip
  implicit none
  private

  public :: step, filter_vars, sparc_gpu_set_face, sparc_gpu_get_face, sparc_gpu_guard, &
       sparc_gpu_pull, sparc_gpu_push, sparc_gpu_pull_every_step, sparc_gpu_finalize

  type face_store
    logical :: set = .false.
    real(c_double), allocatable :: a(:,:,:), b(:,:,:), kappa(:,:,:), psi(:,:,:,:)
  end type face_store

  type(c_ptr), save :: ctx = c_null_ptr
  logical, save :: pull_each = .true.
  type(face_store), target, save :: faces(0:5)
  real(c_double), allocatable, target, save :: srcx(:), srcy(:), srcz(:), stretch_c(:)

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    if (rc /= SPARC_OK) then
      print *, 'sparc_gpu: ', what, ' failed (', rc, '): ', sparc_error_message()
      stop 1
    end if
  end subroutine check

  ! Hand over the CPML arrays of one face as INITIALISE_RHS left them in the
  ! reference's per-rank globals (initialize.f90:1036-1199; bases
  ! zpmlvars & co., initialize.f90:541-585).
  subroutine sparc_gpu_set_face(face)
    integer, intent(in) :: face
    select case (face)
    case (SPARC_ZBOT, SPARC_ZTOP)
      call keep(faces(face), az, bzpml, kappaz, zpmlvars)
    case (SPARC_XBOT, SPARC_XTOP)
      call keep(faces(face), ax, bxpml, kappax, xpmlvars)
    case (SPARC_YBOT, SPARC_YTOP)
      call keep(faces(face), ay, bypml, kappay, ypmlvars)
    case default
      print *, 'sparc_gpu_set_face: face must be 0..5'
      stop 1
    end select
  end subroutine sparc_gpu_set_face

  subroutine keep(f, a_, b_, k_, psi_)
    type(face_store), intent(inout) :: f
    real(c_double), intent(in) :: a_(:,:,:), b_(:,:,:), k_(:,:,:), psi_(:,:,:,:)
    f%a = a_
    f%b = b_
    f%kappa = k_
    f%psi = psi_
    f%set = .true.
  end subroutine keep

  ! Copy face `face`'s current CPML bases into the per-rank global the
  ! reference writes to its full snapshot (zpmlvars, xpmlvars or ypmlvars).
  subroutine sparc_gpu_get_face(face)
    integer, intent(in) :: face
    if (.not. faces(face)%set) return
    select case (face)
    case (SPARC_ZBOT, SPARC_ZTOP)
      zpmlvars = faces(face)%psi
    case (SPARC_XBOT, SPARC_XTOP)
      xpmlvars = faces(face)%psi
    case (SPARC_YBOT, SPARC_YTOP)
      ypmlvars = faces(face)%psi
    end select
  end subroutine sparc_gpu_get_face

  logical function face_active(face)
    integer, intent(in) :: face
    integer :: w(0:5)
    w = (/ nxpmlbot, nxpmltop, nypmlbot, nypmltop, nzpmlbot, nzpmltop /)
    if (face < 4) then
      face_active = USE_HPML .and. w(face) > 0
    else
      face_active = USE_PML .and. w(face) > 0
    end if
  end function face_active

  subroutine init()
    type(sparc_cfg), target :: cfg
    type(sparc_bg), target :: bg
    real(c_double), allocatable, target :: bgc(:,:,:,:), gc(:)
    integer :: f, i, nbg

    cfg%nx = nx
    cfg%ny = ny
    cfg%nz = nz
    cfg%option = option
    cfg%precision = SPARC_PREC_FP64
    cfg%periodic = merge(1, 0, PERIODIC)
    cfg%use_pml = merge(1, 0, USE_PML)
    cfg%use_hpml = merge(1, 0, USE_HPML)
    cfg%npml = (/ nxpmlbot, nxpmltop, nypmlbot, nypmltop, nzpmlbot, nzpmltop /)
    cfg%g_xy_filt = g_xy_filt
    cfg%g_z_filt = g_z_filt
    cfg%s_xy = s_xy
    cfg%s_z = s_z
    cfg%xyfreq = xyfreq_filtering
    cfg%zfreq = zfreq_filtering
    cfg%pulse_dir = merge(pulse_dir, 0, PULSE)
    cfg%bg3d = 1
    cfg%deltat = deltat
    cfg%betas = betas
    cfg%stretchx = stretchx
    cfg%stretchy = stretchy
    stretch_c = stretch
    cfg%stretch = c_loc(stretch_c(1))
    cfg%device = 0
    cfg%tile = 0
    call check(sparc_hip_create(cfg, ctx), 'sparc_hip_create')

    ! the background arrays go through target copies (c_loc needs TARGET;
    ! the reference's module arrays do not carry it), released after upload
    nbg = merge(18, 7, option == 3)
    allocate(bgc(nx, ny, nz, nbg), gc(nz))
    bgc(:,:,:,1) = rho0
    bgc(:,:,:,2) = rhoinv
    bgc(:,:,:,3) = c2rho0
    bgc(:,:,:,4) = gradrho0_z
    bgc(:,:,:,5) = gradp0_z
    bgc(:,:,:,6) = spongexyz
    bgc(:,:,:,7) = p0
    gc = g
    bg%rho0 = c_loc(bgc(1,1,1,1))
    bg%rhoinv = c_loc(bgc(1,1,1,2))
    bg%c2rho0 = c_loc(bgc(1,1,1,3))
    bg%gradrho0_z = c_loc(bgc(1,1,1,4))
    bg%gradp0_z = c_loc(bgc(1,1,1,5))
    bg%spongexyz = c_loc(bgc(1,1,1,6))
    bg%p0 = c_loc(bgc(1,1,1,7))
    bg%g = c_loc(gc(1))
    bg%box = c_null_ptr
    bg%boy = c_null_ptr
    bg%boz = c_null_ptr
    bg%curlbox = c_null_ptr
    bg%curlboy = c_null_ptr
    bg%curlboz = c_null_ptr
    bg%gradrho0_x = c_null_ptr
    bg%gradrho0_y = c_null_ptr
    bg%gradp0_x = c_null_ptr
    bg%gradp0_y = c_null_ptr
    bg%reduction = c_null_ptr
    if (option == 3) then
      bgc(:,:,:,8) = box
      bgc(:,:,:,9) = boy
      bgc(:,:,:,10) = boz
      bgc(:,:,:,11) = curlbox
      bgc(:,:,:,12) = curlboy
      bgc(:,:,:,13) = curlboz
      bgc(:,:,:,14) = gradrho0_x
      bgc(:,:,:,15) = gradrho0_y
      bgc(:,:,:,16) = gradp0_x
      bgc(:,:,:,17) = gradp0_y
      bgc(:,:,:,18) = reduction
      bg%box = c_loc(bgc(1,1,1,8))
      bg%boy = c_loc(bgc(1,1,1,9))
      bg%boz = c_loc(bgc(1,1,1,10))
      bg%curlbox = c_loc(bgc(1,1,1,11))
      bg%curlboy = c_loc(bgc(1,1,1,12))
      bg%curlboz = c_loc(bgc(1,1,1,13))
      bg%gradrho0_x = c_loc(bgc(1,1,1,14))
      bg%gradrho0_y = c_loc(bgc(1,1,1,15))
      bg%gradp0_x = c_loc(bgc(1,1,1,16))
      bg%gradp0_y = c_loc(bgc(1,1,1,17))
      bg%reduction = c_loc(bgc(1,1,1,18))
    end if
    do f = 0, 5
      bg%pml_a(f + 1) = c_null_ptr
      bg%pml_b(f + 1) = c_null_ptr
      bg%pml_kappa(f + 1) = c_null_ptr
      if (face_active(f)) then
        if (.not. faces(f)%set) then
          print *, 'sparc_gpu: CPML face ', f, ' is active but sparc_gpu_set_face was not called'
          stop 1
        end if
        bg%pml_a(f + 1) = c_loc(faces(f)%a(1,1,1))
        bg%pml_b(f + 1) = c_loc(faces(f)%b(1,1,1))
        bg%pml_kappa(f + 1) = c_loc(faces(f)%kappa(1,1,1))
      end if
    end do
    ! the pulse of driver.f90:140-145 in separable form: the space factor
    ! exp(-(X^2+Y^2)/2s^2) exp(-Z^2/2s_z^2) = srcx(i) srcy(j) srcz(k), the
    ! time factor per step in STEP
    bg%src_x = c_null_ptr
    bg%src_y = c_null_ptr
    bg%src_z = c_null_ptr
    if (PULSE) then
      allocate(srcx(nx), srcy(ny), srcz(nz))
      do i = 1, nx
        srcx(i) = exp(-1.0_dp*(x(i)*xlength/rsun - pulse_px/diml)**2/(2.0_dp*(pulse_sr/diml)**2))
      end do
      do i = 1, ny
        srcy(i) = exp(-1.0_dp*(y(i)*ylength/rsun - pulse_py/diml)**2/(2.0_dp*(pulse_sr/diml)**2))
      end do
      do i = 1, nz
        srcz(i) = exp(-1.0_dp*((z(i)-(pulse_pz/diml+1.0_dp))**2)/(2.0_dp*(pulse_sz/diml)**2))
      end do
      bg%src_x = c_loc(srcx(1))
      bg%src_y = c_loc(srcy(1))
      bg%src_z = c_loc(srcz(1))
    end if
    bg%dimc = dimc
    bg%dimrho = dimrho
    call check(sparc_hip_set_background(ctx, bg), 'sparc_hip_set_background')
    deallocate(bgc, gc)
    call sparc_gpu_push()
  end subroutine init

  ! psi(f): c_loc of face f's bases (the reference's zpmlvars & co.)
  subroutine psi_ptrs(p)
    type(c_ptr), intent(out) :: p(6)
    integer :: f
    do f = 0, 5
      p(f + 1) = c_null_ptr
      if (face_active(f)) p(f + 1) = c_loc(faces(f)%psi(1,1,1,1))
    end do
  end subroutine psi_ptrs

  ! Upload the host's a and CPML bases (after the host changed them, e.g. a
  ! restart after the first STEP).
  subroutine sparc_gpu_push()
    type(c_ptr) :: p(6)
    if (.not. c_associated(ctx)) then
      call init()
      return
    end if
    call psi_ptrs(p)
    call check(sparc_hip_put_state(ctx, a, p), 'sparc_hip_put_state')
  end subroutine sparc_gpu_push

  ! Download a and the CPML bases (the shim's copies; sparc_gpu_get_face puts
  ! one face's bases back into the reference's per-rank globals).
  subroutine sparc_gpu_pull()
    type(c_ptr) :: p(6)
    if (.not. c_associated(ctx)) return
    call psi_ptrs(p)
    call check(sparc_hip_get_state(ctx, a, p), 'sparc_hip_get_state')
  end subroutine sparc_gpu_pull

  subroutine sparc_gpu_pull_every_step(on)
    logical, intent(in) :: on
    pull_each = on
  end subroutine sparc_gpu_pull_every_step

  ! STEP (step.f90:20-63): one LDDRK-5 step of a and the CPML bases with the
  ! source of time index `time` (driver.f90:137-145).
  subroutine step()
    real(c_double) :: amp
    if (.not. c_associated(ctx)) call init()
    amp = 0.0_dp
    if (PULSE) amp = pulse_amp/0.56458272_dp/dimc* &
         sin(2.0_dp*pi*(time*timestep-pulse_t1+2.0_dp*pulse_t0)/pulse_t0)* &
         exp(-1.0_dp*(time*timestep-pulse_t1)**2/(2.0_dp*pulse_st**2))
    call check(sparc_hip_step(ctx, amp), 'sparc_hip_step')
    if (pull_each) call sparc_gpu_pull()
  end subroutine step

  ! FILTER_VARS (derivatives.f90:659-703) at time index `time`.
  subroutine filter_vars()
    if (.not. c_associated(ctx)) call init()
    call check(sparc_hip_filter(ctx, int(time, c_int)), 'sparc_hip_filter')
    if (pull_each) call sparc_gpu_pull()
  end subroutine filter_vars

  ! The failure checks of the last STEP (driver.f90:175-183), evaluated on
  ! the device: fail becomes .true. if norm_KE > 1e3 or a density or
  ! pressure went non-positive (the device then holds the state of that
  ! step for the dump of driver.f90:189-196).
  subroutine sparc_gpu_guard(fail)
    logical, intent(inout) :: fail
    integer(c_int) :: tripped
    integer(c_long) :: at, steps
    real(c_double) :: d(3)
    if (.not. c_associated(ctx)) return
    call check(sparc_hip_guard(ctx, tripped, at, steps, d), 'sparc_hip_guard')
    if (tripped /= 0) then
      print *, 'norm_KE or minimum out of range: ', d
      fail = .true.
    end if
  end subroutine sparc_gpu_guard

  subroutine sparc_gpu_finalize()
    integer(c_int) :: rc
    integer :: f
    if (c_associated(ctx)) rc = sparc_hip_destroy(ctx)
    ctx = c_null_ptr
    do f = 0, 5
      faces(f)%set = .false.
    end do
  end subroutine sparc_gpu_finalize

end module sparc_gpu
