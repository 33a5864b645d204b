! Fills the bind(C) types of the sparc_hip module with distinct values so the
! C side (check_layout.c, compiled against include/sparc_hip.h) can verify the
! Fortran layout member by member.
subroutine fill_structs(cfg, bg, stretch) bind(C, name="fill_structs")
  use iso_c_binding
  use sparc_hip
  implicit none
  type(sparc_cfg), intent(inout) :: cfg
  type(sparc_bg), intent(inout) :: bg
  real(c_double), target, intent(in) :: stretch(*)
  integer :: f
  cfg%nx = 101; cfg%ny = 102; cfg%nz = 103
  cfg%option = 3; cfg%precision = SPARC_PREC_FP32; cfg%periodic = 1
  cfg%use_pml = 1; cfg%use_hpml = 0
  cfg%npml = [11, 12, 13, 14, 15, 16]
  cfg%g_xy_filt = 5; cfg%g_z_filt = 3
  cfg%s_xy = 0.25d0; cfg%s_z = 0.75d0
  cfg%xyfreq = 30; cfg%zfreq = 10; cfg%pulse_dir = 2; cfg%bg3d = 1
  cfg%deltat = 1.5d-4
  cfg%betas = [0d0, 1d0, 2d0, 3d0, 4d0]
  cfg%stretchx = 7.25d0; cfg%stretchy = 8.5d0
  cfg%stretch = c_loc(stretch(1))
  cfg%device = 2
  cfg%tile = [64, 8, 16]
  bg%dimc = 4.5d6; bg%dimrho = 2.5d-4
  do f = 1, 6
    bg%pml_a(f) = c_loc(stretch(f))
  end do
  bg%src_z = c_loc(stretch(7))
end subroutine fill_structs
