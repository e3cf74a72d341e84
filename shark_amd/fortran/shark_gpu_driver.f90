! shark_gpu_driver.f90 — Fortran host driving the GPU step through shark_gpu_mod.
!
! Mirrors the reference program flow (src/shark.f90:3-119: parameters, grid, setup, time loop
! calling solve until tend, output) with solve's body on the GPU.
!
!   shark_gpu_driver dustywave NSTEPS OUT [NX]   1D dusty wave built here in Fortran: periodic
!        box L=1 (gridinit, src/grid/grid.f90), isothermal gas cs=1, rho=1, one dust fluid with
!        dust-to-gas 1 and K=10 (t_s = rho_d/K), sine velocity 1e-4 at k=2 pi, MUSCL + HLL,
!        CFL 0.5 (the DUSTYWAVE setup, src/setups/DUSTYWAVE/setup.f90)
!   shark_gpu_driver state IN OUT                 configuration, grid and initial state read from
!        a state file (shark_amd/statefile.py), nsteps and tend from the file
! OUT receives: the initial u_prim, the final u_prim (ncells x nvar, column-major), the step
! count and the final time.
program shark_gpu_driver
  use iso_c_binding
  use shark_gpu_mod
  implicit none
  type(shark_cfg), target :: cfg
  type(c_ptr) :: ctx
  type(shark_tune) :: tune
  type(shark_output_info) :: oinfo
  character(len=512) :: odir
  character(kind=c_char, len=513), target :: odir_c
  integer :: ostat
  character(len=512) :: mode, arg, outfile
  integer :: nx, ncells, ndim, nvar, i, u_out
  integer(c_int64_t) :: nsteps, steps_done
  real(c_double) :: tend, time_out
  real(c_double), allocatable, target :: dx(:,:), surf(:,:), vol(:), radii_c(:), position(:,:)
  real(c_double), allocatable, target :: u(:,:), u0(:,:), cs(:), sdust(:,:), cond(:,:), fargo_v(:)
  logical :: has_sdust, has_cond, has_fargo

  call get_command_argument(1, mode)
  select case (trim(mode))
  case ('dustywave')
     call get_command_argument(2, arg)
     read(arg, *) nsteps
     call get_command_argument(3, outfile)
     nx = 512
     if (command_argument_count() >= 4) then
        call get_command_argument(4, arg)
        read(arg, *) nx
     end if
     tend = huge(1.0d0)
     call setup_dustywave(nx)
  case ('state')
     call get_command_argument(2, arg)
     call get_command_argument(3, outfile)
     call read_state(trim(arg))
  case default
     print *, 'usage: shark_gpu_driver dustywave NSTEPS OUT [NX] | state IN OUT'
     stop 2
  end select

  u0 = u
  call shark_check(shark_gpu_create(cfg, ctx), 'shark_gpu_create')
  call shark_check(shark_gpu_set_geometry(ctx, dx, surf, vol, radii_c, position), 'shark_gpu_set_geometry')
  call shark_check(shark_gpu_upload(ctx, c_loc(u), c_loc(cs), opt(has_sdust, sdust), &
       opt(has_cond, cond), opt1(has_fargo, fargo_v)), 'shark_gpu_upload')
  ! launch shapes picked on this state (the state is restored, so the run is unchanged)
  call shark_check(shark_gpu_autotune(ctx, 3_c_int, tune), 'shark_gpu_autotune')
  print '(A,I0,A,I0,A,I0,A,I0)', 'shark_gpu_driver: tuned strip_h=', tune%strip_h, ' strip_gas=', &
       tune%strip_gas, ' post_tw=', tune%post_tw, ' xcd_order=', tune%xcd_order
  call shark_check(shark_gpu_run(ctx, nsteps, tend, steps_done, time_out), 'shark_gpu_run')
  call shark_check(shark_gpu_download(ctx, c_loc(u), c_null_ptr), 'shark_gpu_download')
  ! the reference's output(1) of the final state (code units: the dusty wave's units are all 1)
  call get_environment_variable('SHARK_OUTPUT_DIR', odir, status=ostat)
  if (ostat == 0 .and. len_trim(odir) > 0) then
     odir_c = trim(odir)//c_null_char
     oinfo%version = 1; oinfo%gridspace = 0; oinfo%mhd = 0; oinfo%turb = 0
     oinfo%ndustpscal = 0; oinfo%charging = 0; oinfo%time = time_out
     oinfo%unit_t = 1d0; oinfo%unit_d = 1d0; oinfo%unit_l = 1d0; oinfo%unit_v = 1d0
     oinfo%unit_p = 1d0; oinfo%unit_dcol = 1d0
     oinfo%base_dir = c_loc(odir_c)
     call shark_check(shark_gpu_write_output(ctx, oinfo, 1_c_int), 'shark_gpu_write_output')
  end if
  call shark_check(shark_gpu_destroy(ctx), 'shark_gpu_destroy')

  open(newunit=u_out, file=trim(outfile), form='unformatted', access='stream', status='replace')
  write(u_out) int(ncells, c_int64_t), int(nvar, c_int64_t)
  write(u_out) u0
  write(u_out) u
  write(u_out) steps_done, time_out
  close(u_out)
  print '(A,I0,A,ES24.16)', 'shark_gpu_driver: ', steps_done, ' steps, time ', time_out

contains

  type(c_ptr) function opt(present_, a)
    logical, intent(in) :: present_
    real(c_double), intent(in), target :: a(:,:)
    if (present_) then
       opt = c_loc(a)
    else
       opt = c_null_ptr
    end if
  end function opt

  type(c_ptr) function opt1(present_, a)
    logical, intent(in) :: present_
    real(c_double), intent(in), target :: a(:)
    if (present_) then
       opt1 = c_loc(a)
    else
       opt1 = c_null_ptr
    end if
  end function opt1

  subroutine setup_dustywave(n)
    integer, intent(in) :: n
    real(c_double), parameter :: pi = 3.141592653589793238462643383279d0
    real(c_double) :: box_l, delta, pos, s, cs0, rho0, vx0, d2g
    integer :: ix
    call shark_cfg_defaults(cfg)
    cfg%nx = n; cfg%ny = 1; cfg%ndust = 1
    cfg%gamma = 1.6666667d0
    cfg%K(1) = 10.0d0
    box_l = 1.0d0; cs0 = 1.0d0; rho0 = 1.0d0; vx0 = 1.0d-4; d2g = 1.0d0
    ndim = 1
    ncells = n + 4
    nvar = 5 + 4 * cfg%ndust
    allocate(dx(ncells, 1), surf(ncells, 1), vol(ncells), radii_c(ncells), position(ncells, 1))
    allocate(u(ncells, nvar), cs(ncells))
    delta = box_l / float(n)
    dx = delta; surf = 1.0d0; vol = delta; radii_c = 1.0d0; position = 0.0d0
    pos = 0.0d0
    do ix = 1, ncells
       if (ix >= 3 .and. ix <= n + 2) then
          pos = pos + delta
          position(ix, 1) = pos
       end if
    end do
    cs = cs0
    u = 0.0d0
    do ix = 1, ncells
       s = sin(((2.0d0 * pi) * position(ix, 1)) / box_l * 1.0d0)
       ! primitives rho, vx, P = cs^2 rho; dust rho_d = d2g rho, vdx = vx; then primtoc
       u(ix, 1) = rho0 * (1.0d0 + 0.0d0 * s)
       u(ix, 2) = u(ix, 1) * (vx0 * s)
       u(ix, 5) = (cs(ix)**2 * u(ix, 1)) / (cfg%gamma - 1.0d0) + 0.5d0 * u(ix, 1) * (vx0 * s)**2
       u(ix, 6) = d2g * u(ix, 1)
       u(ix, 7) = u(ix, 6) * (vx0 * s)
    end do
    has_sdust = .false.; has_cond = .false.; has_fargo = .false.
  end subroutine setup_dustywave

  subroutine read_state(path)
    character(len=*), intent(in) :: path
    integer :: un
    integer(c_int32_t) :: magic, fver, nd32, nv32, hs, hc, hf, pad
    integer(c_int64_t) :: nc64
    open(newunit=un, file=path, form='unformatted', access='stream', status='old')
    read(un) magic, fver
    if (magic /= int(z'53484B31', c_int32_t) .or. fver /= 1) then
       print *, 'shark_gpu_driver: ', path, ' is not a version-1 state file'
       stop 2
    end if
    read(un) cfg
    read(un) nc64, nd32, nv32
    read(un) hs, hc, hf, pad
    ncells = int(nc64); ndim = nd32; nvar = nv32
    has_sdust = hs /= 0; has_cond = hc /= 0; has_fargo = hf /= 0
    allocate(dx(ncells, ndim), surf(ncells, ndim), vol(ncells), radii_c(ncells), position(ncells, ndim))
    allocate(u(ncells, nvar), cs(ncells))
    read(un) dx, surf, vol, radii_c, position, u, cs
    allocate(sdust(ncells, max(cfg%ndust, 1)), cond(ncells, nvar), fargo_v(ncells))
    if (has_sdust) read(un) sdust(:, 1:cfg%ndust)
    if (has_cond) read(un) cond
    if (has_fargo) read(un) fargo_v
    read(un) nsteps, tend
    close(un)
  end subroutine read_state
end program shark_gpu_driver
