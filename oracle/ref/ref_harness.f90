! Dump harness for the reference build (test infrastructure, never shipped).
! Calls the reference's own read_params (src/commons/read_params.f90:3), setup
! (src/setups/<SETUP>/setup.f90:1) and solve (src/solver/solve.f90:10) exactly as
! time_loop does (src/shark.f90:71-108), for a fixed number of steps, and writes:
!   <out>/ic.bin     : u_prim after setup (ncells*nvar, column-major FP64)
!   <out>/aux.bin    : cs, dx(:,1..ndim), surf(:,1..ndim), vol, radii_c, position(:,1..ndim),
!                      active_cell, active_cell_predictor, [sdust], [fargo_velocity], [uprim_condinit]
!   <out>/final.bin  : u_prim after nsteps
!   <out>/dt.bin     : dt of every step
!   <out>/meta.txt   : sizes and scalars the restatement needs
!   ./output_NNNNN/  : with a fourth argument iout, the reference's own output(iout)
!                      (src/commons/write_output.f90:2) of the final state, after ctoprim
!   <out>/fargo_v.bin: with REF_FARGO_SHIFTS=<file> (one signed integer per column, nx_max of
!                      them), the fargo_velocity of every step: before each step the harness
!                      takes the step's dt as solve does (apply_boundaries, ctoprim,
!                      update_force_setup, courant: solve.f90:20-75) and gives column ix the
!                      velocity sign(s)*v with |v dt| = |s| dy exactly (dy = dx(:,2)*radii_c),
!                      so that the remap shifts that ring by exactly |s| cells (eps = 0)
! usage: shark_ref <namelist> <nsteps> <outdir> [iout]
program ref_harness
  use parameters
  use commons
  use units
#ifdef REF_DISK
  use setup_parameters, only: uprim_condinit, mstar, r_relax, n_rel, smooth_r, hoverr, alpha_visc, &
       box_l, accretion, M_acc, phi_mom, L_out, Mplanet, Rplanet, a_smooth_pl
  use gravity_commons, only: self_gravity
#endif
#ifdef REF_SI
  use setup_parameters, only: box_l, rho_init, omega_shear, q_shear, hoverr, eta_stream, rad0
#endif
#ifdef REF_SOD
  use setup_parameters, only: rho_l, rho_r, P_l, P_r, direction_shock
#endif
#ifdef REF_KH
  use setup_parameters, only: box_l, box_l_y, rho_dense, rho_diffuse, T_dense, Mach, vy0, kx
#endif
#ifdef REF_CDISK
  use setup_parameters, only: box_l, box_l_y, sigma_R0, R0_disk, Mstar, Mplanet, Rplanet, &
       frac_R_disk, buffer_in, buffer_out, HoverR, smooth_r, alpha_visc
#endif
#ifdef REF_COLLAPSE
  use setup_parameters, only: M_cloud, alpha_cloud, T0_cloud, mu_gas
#endif
  implicit none
  integer :: nsteps, istep, ilun, iout, fst
  integer(8) :: clk0, clk1, clk_rate
  character(len=256) :: arg, outdir, fpath
  real(dp), allocatable :: dts(:)
  integer, allocatable :: shifts(:)
  logical :: fargo_shifts

  call read_params
  call setup
  call getarg(2, arg)
  read(arg, *) nsteps
  call getarg(3, outdir)
  ! nsteps < 0: run until time >= tend, as flag_continue does (e.g. setups/DUSTYWAVE/setup.f90:136)
  allocate(dts(max(abs(nsteps),1)*1 + 200000))

  ilun = 31
  open(ilun, file=trim(outdir)//'/ic.bin', form='unformatted', access='stream', status='replace')
  write(ilun) u_prim
  close(ilun)

  open(ilun, file=trim(outdir)//'/aux.bin', form='unformatted', access='stream', status='replace')
  write(ilun) cs
  write(ilun) dx
  write(ilun) surf
  write(ilun) vol
  write(ilun) radii_c
  write(ilun) position
  write(ilun) active_cell
  write(ilun) active_cell_predictor
#if NDUST>0
  write(ilun) sdust
#endif
  if (fargo) write(ilun) fargo_velocity
#ifdef REF_DISK
  write(ilun) uprim_condinit
#endif
  close(ilun)

  call get_environment_variable('REF_FARGO_SHIFTS', fpath, status=fst)
  fargo_shifts = fst == 0 .and. len_trim(fpath) > 0
  if (fargo_shifts) then
     allocate(shifts(nx_max))
     open(33, file=trim(fpath), status='old')
     read(33, *) shifts
     close(33)
     open(34, file=trim(outdir)//'/fargo_v.bin', form='unformatted', access='stream', status='replace')
  end if

  time = 0.0d0
  call system_clock(clk0, clk_rate)
  if (nsteps >= 0) then
     do istep = 1, nsteps
        if (fargo_shifts) call integer_fargo_velocities
        call solve(.false., .false.)
        time = time + dt
        dts(istep) = dt
     end do
  else
     istep = 0
     do while (time < tend)
        istep = istep + 1
        call solve(.false., .false.)
        time = time + dt
        dts(istep) = dt
     end do
     nsteps = istep
  end if

  call system_clock(clk1)
  if (fargo_shifts) close(34)
  open(ilun, file=trim(outdir)//'/final.bin', form='unformatted', access='stream', status='replace')
  write(ilun) u_prim
  close(ilun)
  open(ilun, file=trim(outdir)//'/dt.bin', form='unformatted', access='stream', status='replace')
  write(ilun) dts(1:nsteps)
  close(ilun)

  if (command_argument_count() >= 4) then
     call getarg(4, arg)
     read(arg, *) iout
     call ctoprim
     call output(iout)
  end if

  open(ilun, file=trim(outdir)//'/meta.txt', status='replace')
  write(ilun, '(A,I12)') 'ncells ', ncells
  write(ilun, '(A,I12)') 'nvar ', nvar
  write(ilun, '(A,I12)') 'nx ', nx
  write(ilun, '(A,I12)') 'ny ', ny
  write(ilun, '(A,I12)') 'ndim ', ndim
  write(ilun, '(A,I12)') 'ndust ', ndust
  write(ilun, '(A,I12)') 'nsteps ', nsteps
  write(ilun, '(A,ES25.17)') 'loop_seconds ', real(clk1 - clk0, dp) / real(clk_rate, dp)
  write(ilun, '(A,I12)') 'slope_type ', slope_type
  write(ilun, '(A,I12)') 'iso_cs ', iso_cs
  write(ilun, '(A,L2)') 'force_kick ', force_kick
  write(ilun, '(A,L2)') 'fargo ', fargo
#if NDUST>0
  write(ilun, '(A,L2)') 'drag ', drag
  write(ilun, '(A,L2)') 'dust_back_reaction ', dust_back_reaction
  write(ilun, '(A,ES25.17)') 'dust_ratio_min ', dust_ratio_min
  write(ilun, '(A,ES25.17)') 'rhograin ', rhograin
  write(ilun, '(A,ES25.17)') 'dust2gas ', dust2gas
#endif
  write(ilun, '(A,ES25.17)') 'gamma ', gamma
  write(ilun, '(A,ES25.17)') 'CFL ', CFL
  write(ilun, '(A,ES25.17)') 'smallr ', smallr
  write(ilun, '(A,ES25.17)') 'smallc ', smallc
  write(ilun, '(A,ES25.17)') 'smallp ', smallp
  write(ilun, '(A,ES25.17)') 'time ', time
  write(ilun, '(A,ES25.17)') 'unit_d ', unit_d
  write(ilun, '(A,ES25.17)') 'unit_l ', unit_l
  write(ilun, '(A,ES25.17)') 'unit_t ', unit_t
  write(ilun, '(A,ES25.17)') 'unit_v ', unit_v
  write(ilun, '(A,ES25.17)') 'unit_p ', unit_P
  write(ilun, '(A,ES25.17)') 'unit_dcol ', unit_dcol
#ifdef REF_DISK
  write(ilun, '(A,ES25.17)') 'mstar ', mstar
  write(ilun, '(A,ES25.17)') 'r_relax ', r_relax
  write(ilun, '(A,ES25.17)') 'n_rel ', n_rel
  write(ilun, '(A,ES25.17)') 'smooth_r ', smooth_r
  write(ilun, '(A,ES25.17)') 'hoverr ', hoverr
  write(ilun, '(A,ES25.17)') 'alpha_visc ', alpha_visc
  write(ilun, '(A,L2)') 'viscosity ', viscosity
  write(ilun, '(A,ES25.17)') 'box_l ', box_l
  write(ilun, '(A,L2)') 'accretion ', accretion
  write(ilun, '(A,ES25.17)') 'M_acc ', M_acc
  write(ilun, '(A,ES25.17)') 'phi_mom ', phi_mom
  write(ilun, '(A,ES25.17)') 'L_out ', L_out
  write(ilun, '(A,ES25.17)') 'mplanet ', Mplanet
  write(ilun, '(A,ES25.17)') 'rplanet ', Rplanet
  write(ilun, '(A,ES25.17)') 'a_smooth_pl ', a_smooth_pl
  write(ilun, '(A,L2)') 'self_gravity ', self_gravity
  open(32, file=trim(outdir)//'/eta_visc.bin', form='unformatted', access='stream', status='replace')
  write(32) eta_visc
  close(32)
#endif
#ifdef REF_SI
  write(ilun, '(A,ES25.17)') 'box_l ', box_l
  write(ilun, '(A,ES25.17)') 'rho_init ', rho_init
  write(ilun, '(A,ES25.17)') 'omega_shear ', omega_shear
  write(ilun, '(A,ES25.17)') 'q_shear ', q_shear
  write(ilun, '(A,ES25.17)') 'hoverr ', hoverr
  write(ilun, '(A,ES25.17)') 'eta_stream ', eta_stream
  write(ilun, '(A,ES25.17)') 'rad0 ', rad0
#endif
#ifdef REF_SOD
  write(ilun, '(A,ES25.17)') 'rho_l ', rho_l
  write(ilun, '(A,ES25.17)') 'rho_r ', rho_r
  write(ilun, '(A,ES25.17)') 'p_l ', P_l
  write(ilun, '(A,ES25.17)') 'p_r ', P_r
  write(ilun, '(A,I12)') 'shock_dir ', direction_shock
#endif
#ifdef REF_CDISK
  write(ilun, '(A,ES25.17)') 'box_l ', box_l
  write(ilun, '(A,ES25.17)') 'box_l_y ', box_l_y
  write(ilun, '(A,ES25.17)') 'sigma_R0 ', sigma_R0
  write(ilun, '(A,ES25.17)') 'R0_disk ', R0_disk
  write(ilun, '(A,ES25.17)') 'mstar ', Mstar
  write(ilun, '(A,ES25.17)') 'mplanet ', Mplanet
  write(ilun, '(A,ES25.17)') 'rplanet ', Rplanet
  write(ilun, '(A,ES25.17)') 'frac_R_disk ', frac_R_disk
  write(ilun, '(A,ES25.17)') 'buffer_in ', buffer_in
  write(ilun, '(A,ES25.17)') 'buffer_out ', buffer_out
  write(ilun, '(A,ES25.17)') 'hoverr ', HoverR
  write(ilun, '(A,ES25.17)') 'smooth_r ', smooth_r
  write(ilun, '(A,ES25.17)') 'alpha_visc ', alpha_visc
  write(ilun, '(A,L2)') 'viscosity ', viscosity
  open(32, file=trim(outdir)//'/eta_visc.bin', form='unformatted', access='stream', status='replace')
  write(32) eta_visc
  close(32)
#endif
#ifdef REF_COLLAPSE
  write(ilun, '(A,ES25.17)') 'M_cloud ', M_cloud
  write(ilun, '(A,ES25.17)') 'alpha_cloud ', alpha_cloud
  write(ilun, '(A,ES25.17)') 't0_cloud ', T0_cloud
  write(ilun, '(A,ES25.17)') 'mu_gas ', mu_gas
  write(ilun, '(A,ES25.17)') 'unit_nh ', unit_nH
  write(ilun, '(A,ES25.17)') 'unit_m ', unit_m
  write(ilun, '(A,ES25.17)') 'l_soft ', l_soft
  write(ilun, '(A,ES25.17)') 'm_cent ', M_cent
  write(ilun, '(A,ES25.17)') 'rin ', rin
  write(ilun, '(A,I12)') 'non_standard_eos ', non_standard_eos
#if NDUST>0
  write(ilun, '(A,ES25.17)') 'rhograin_c ', rhograin
#endif
  open(32, file=trim(outdir)//'/dvol.bin', form='unformatted', access='stream', status='replace')
  write(32) dvol
  close(32)
#endif
#ifdef REF_KH
  write(ilun, '(A,ES25.17)') 'box_l ', box_l
  write(ilun, '(A,ES25.17)') 'box_l_y ', box_l_y
  write(ilun, '(A,ES25.17)') 'rho_dense ', rho_dense
  write(ilun, '(A,ES25.17)') 'rho_diffuse ', rho_diffuse
  write(ilun, '(A,ES25.17)') 'T_dense ', T_dense
  write(ilun, '(A,ES25.17)') 'Mach ', Mach
  write(ilun, '(A,ES25.17)') 'vy0 ', vy0
  write(ilun, '(A,ES25.17)') 'kx ', kx
  write(ilun, '(A,ES25.17)') 'tend ', tend
#if NDUST>0
  write(ilun, '(A,ES25.17)') 'scut ', scut
#endif
#endif
  close(ilun)

contains

  subroutine integer_fargo_velocities
    integer, external :: icell
    integer :: ix, iy, k, m, i
    real(dp) :: dyl, v, sh, target, sgn
    call apply_boundaries
    call ctoprim
    if (force_kick) call update_force_setup
    call courant
    do ix = 1, nx_max
       m = abs(shifts(ix))
       sgn = 1.0d0
       if (shifts(ix) < 0) sgn = -1.0d0
       i = icell(ix, nghost + 1)
       dyl = dx(i, 2) * radii_c(i)
       v = 0.0d0
       if (m > 0) then
          target = real(m, dp) * dyl
          v = target / dt
          do k = 1, 64
             sh = abs(v * dt)
             if (sh == target .and. int(sh / dyl) == m .and. mod(sh, dyl) == 0.0d0) exit
             if (sh < target) then
                v = nearest(v, 1.0d0)
             else
                v = nearest(v, -1.0d0)
             end if
          end do
          if (k > 64) v = 0.0d0
       end if
       do iy = 1, ny_max
          fargo_velocity(icell(ix, iy)) = sgn * v
       end do
    end do
    write(34) fargo_velocity
  end subroutine integer_fargo_velocities

end program ref_harness
