! shark_gpu_mod.f90 — ISO_C_BINDING interface to libshark_gpu.so (include/shark_gpu.h).
!
! The Fortran side of the drop-in boundary: a host code built like the reference (module state,
! parameterless solve, src/solver/solve.f90:10-181) keeps its grid and setup and replaces the
! body of solve by shark_gpu_step on a context created once after setup.  Every call returns 0
! or a positive code; shark_check prints shark_gpu_last_error() and stops, the reference's error
! convention (print + stop, src/commons/read_params.f90:18-33).
module shark_gpu_mod
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: SHARK_CFG_VERSION = 2
  integer, parameter :: SHARK_MAX_DUST = 64
  integer(c_int32_t), parameter :: SHARK_GEOM_CARTESIAN = 0, SHARK_GEOM_POLAR = 2
  integer(c_int32_t), parameter :: SHARK_SOLVER_LLF = 0, SHARK_SOLVER_HLL = 1, SHARK_SOLVER_HLLC = 2
  integer(c_int32_t), parameter :: SHARK_DUST_HUANG_BAI = 0
  integer(c_int32_t), parameter :: SHARK_BC_PERIODIC = 0, SHARK_BC_SHEAR_SI = 1, SHARK_BC_DISK = 2
  integer(c_int32_t), parameter :: SHARK_BC_SOD = 3, SHARK_BC_ZERO_GRADIENT = 4
  integer(c_int32_t), parameter :: SHARK_TSTOP_DUSTYWAVE = 0, SHARK_TSTOP_SI = 1, SHARK_TSTOP_EPSTEIN_DISK = 2, &
       SHARK_TSTOP_EPSTEIN = 3
  integer(c_int32_t), parameter :: SHARK_FORCE_NONE = 0, SHARK_INLOOP_NONE = 0
  integer(c_int32_t), parameter :: SHARK_TSTOP_EPSTEIN_CDISK = 4, SHARK_FORCE_CDISK = 3, SHARK_INLOOP_CDISK = 3

  ! same layout as struct shark_cfg
  type, bind(C) :: shark_cfg
     integer(c_int32_t) :: version, nx, ny, ndust, nghost, geom, solver, solverdust, slope_type
     integer(c_int32_t) :: iso_cs, force_kick, fargo, drag, dust_back_reaction, growth, viscosity
     integer(c_int32_t) :: bc_kind, tstop_kind, force_kind, inloop_kind, disk_bc_mode, shock_dir
     real(c_double) :: gamma, cfl, smallr, smallc, smallp, dust_ratio_min
     real(c_double) :: K(SHARK_MAX_DUST)
     real(c_double) :: q_shear, omega_shear, box_l, rad0, eta_stream, rho_init, hoverr
     real(c_double) :: rhograin
     real(c_double) :: mstar, r_relax, n_rel, smooth_r
     real(c_double) :: rho_l, rho_r, p_l, p_r
     real(c_double) :: box_l_y, mplanet, rplanet, sigma_r0, r0_disk, frac_r_disk, buffer_in, buffer_out
     real(c_double) :: alpha_visc, dust2gas
     real(c_double) :: reserved(2)
     integer(c_int32_t) :: accretion, gravity
     real(c_double) :: m_acc, phi_mom, l_out
     integer(c_int32_t) :: eos_baro, self_gravity
     real(c_double) :: l_soft, m_cent, t0_cloud, mu_gas, unit_v, unit_nh
     real(c_double) :: a_smooth_pl
  end type shark_cfg

  ! same layout as struct shark_tune (launch shapes of the step kernels)
  type, bind(C) :: shark_tune
     integer(c_int32_t) :: strip_h, post_tw, xcd_order, strip_gas
     integer(c_int32_t) :: reserved_i(4)
     real(c_double) :: hydro_ms, post_ms
     real(c_double) :: reserved(6)
  end type shark_tune

  ! same layout as struct shark_output_info (include/shark_output.h)
  type, bind(C) :: shark_output_info
     integer(c_int32_t) :: version, gridspace, mhd, turb, ndustpscal, charging
     real(c_double) :: time
     real(c_double) :: unit_t, unit_d, unit_l, unit_v, unit_p, unit_dcol
     type(c_ptr) :: base_dir
  end type shark_output_info

  interface
     integer(c_int) function shark_gpu_create(cfg, ctx) bind(C, name="shark_gpu_create")
       import :: c_int, c_ptr, shark_cfg
       type(shark_cfg), intent(in) :: cfg
       type(c_ptr), intent(out) :: ctx
     end function shark_gpu_create
     integer(c_int) function shark_gpu_set_geometry(ctx, dx, surf, vol, radii_c, position) &
          bind(C, name="shark_gpu_set_geometry")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: dx(*), surf(*), vol(*), radii_c(*), position(*)
     end function shark_gpu_set_geometry
     integer(c_int) function shark_gpu_set_dvol(ctx, dvol) bind(C, name="shark_gpu_set_dvol")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: dvol(*)
     end function shark_gpu_set_dvol
     integer(c_int) function shark_gpu_upload(ctx, u_prim, cs, sdust, condinit, fargo_velocity) &
          bind(C, name="shark_gpu_upload")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, u_prim, cs, sdust, condinit, fargo_velocity
     end function shark_gpu_upload
     integer(c_int) function shark_gpu_step(ctx, dt_out) bind(C, name="shark_gpu_step")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: dt_out
     end function shark_gpu_step
     integer(c_int) function shark_gpu_run(ctx, nsteps, tend, steps_done, time_out) &
          bind(C, name="shark_gpu_run")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: nsteps
       real(c_double), value :: tend
       integer(c_int64_t), intent(out) :: steps_done
       real(c_double), intent(out) :: time_out
     end function shark_gpu_run
     integer(c_int) function shark_gpu_download(ctx, u_prim, q) bind(C, name="shark_gpu_download")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, u_prim, q
     end function shark_gpu_download
     integer(c_int) function shark_gpu_get_time(ctx, time, dt_last) bind(C, name="shark_gpu_get_time")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: time, dt_last
     end function shark_gpu_get_time
     integer(c_int) function shark_gpu_autotune(ctx, reps, chosen) bind(C, name="shark_gpu_autotune")
       import :: c_int, c_ptr, shark_tune
       type(c_ptr), value :: ctx
       integer(c_int), value :: reps
       type(shark_tune), intent(out) :: chosen
     end function shark_gpu_autotune
     ! output(iout) of the reference (write_output.f90:2), from the device state
     integer(c_int) function shark_gpu_write_output(ctx, info, iout) bind(C, name="shark_gpu_write_output")
       import :: c_int, c_ptr, shark_output_info
       type(c_ptr), value :: ctx
       type(shark_output_info), intent(in) :: info
       integer(c_int), value :: iout
     end function shark_gpu_write_output
     ! restart dumps (path: NUL-terminated)
     integer(c_int) function shark_gpu_save_state(ctx, path) bind(C, name="shark_gpu_save_state")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
     end function shark_gpu_save_state
     integer(c_int) function shark_gpu_load_state(ctx, path) bind(C, name="shark_gpu_load_state")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
     end function shark_gpu_load_state
     integer(c_int) function shark_gpu_destroy(ctx) bind(C, name="shark_gpu_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function shark_gpu_destroy
     type(c_ptr) function shark_gpu_last_error() bind(C, name="shark_gpu_last_error")
       import :: c_ptr
     end function shark_gpu_last_error
  end interface

contains

  subroutine shark_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == 0) return
    call c_f_pointer(shark_gpu_last_error(), msg, [512])
    text = ' '
    do i = 1, 512
       if (msg(i) == c_null_char) exit
       text(i:i) = msg(i)
    end do
    print *, 'shark_gpu: ', what, ' failed (', rc, '): ', trim(text)
    stop 1
  end subroutine shark_check

  ! reference defaults (src/solver/hydro_commons.f90, src/dust/dust_commons.f90) for the step
  subroutine shark_cfg_defaults(c)
    type(shark_cfg), intent(out) :: c
    c%version = SHARK_CFG_VERSION
    c%nx = 1; c%ny = 1; c%ndust = 0; c%nghost = 2
    c%geom = SHARK_GEOM_CARTESIAN; c%solver = SHARK_SOLVER_HLL; c%solverdust = SHARK_DUST_HUANG_BAI
    c%slope_type = 1; c%iso_cs = 1; c%force_kick = 0; c%fargo = 0; c%drag = 1
    c%dust_back_reaction = 1; c%growth = 0; c%viscosity = 0
    c%bc_kind = SHARK_BC_PERIODIC; c%tstop_kind = SHARK_TSTOP_DUSTYWAVE
    c%force_kind = SHARK_FORCE_NONE; c%inloop_kind = SHARK_INLOOP_NONE; c%disk_bc_mode = 0; c%shock_dir = 0
    c%gamma = 1.66667d0; c%cfl = 0.5d0; c%smallr = 1d-27; c%smallc = 1d-27
    c%smallp = 1d-27 * 1d-27**2; c%dust_ratio_min = 1d-11
    c%K = 0.0d0
    c%q_shear = 0.0d0; c%omega_shear = 0.0d0; c%box_l = 0.0d0; c%rad0 = 0.0d0
    c%eta_stream = 0.0d0; c%rho_init = 0.0d0; c%hoverr = 0.0d0; c%rhograin = 0.0d0
    c%mstar = 0.0d0; c%r_relax = 0.0d0; c%n_rel = 0.0d0; c%smooth_r = 0.0d0
    c%rho_l = 0.0d0; c%rho_r = 0.0d0; c%p_l = 0.0d0; c%p_r = 0.0d0
    c%box_l_y = 0.0d0; c%mplanet = 0.0d0; c%rplanet = 0.0d0; c%sigma_r0 = 0.0d0; c%r0_disk = 0.0d0
    c%frac_r_disk = 0.0d0; c%buffer_in = 0.0d0; c%buffer_out = 0.0d0; c%alpha_visc = 0.0d0
    c%dust2gas = 0.0d0
    c%reserved = 0.0d0
    c%accretion = 0; c%gravity = 0; c%m_acc = 0.0d0; c%phi_mom = 2.0d0; c%l_out = 1.0d0
    c%eos_baro = 0; c%self_gravity = 0; c%l_soft = 0.0d0; c%m_cent = 0.0d0; c%t0_cloud = 0.0d0
    c%mu_gas = 0.0d0; c%unit_v = 0.0d0; c%unit_nh = 0.0d0; c%a_smooth_pl = 0.0d0
  end subroutine shark_cfg_defaults
end module shark_gpu_mod
