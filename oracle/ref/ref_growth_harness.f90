! Growth dump harness for the reference build (test infrastructure, never shipped).
! After the reference's read_params and setup (Collapse_1D), it rescales the density of cell i
! by 10**(3*(i-first_active)) (so cells span a range of temperatures and Stokes numbers),
! recomputes primitives and stopping times with the reference's ctoprim and compute_tstop,
! and calls the reference's dust_growth (src/dust/dust_growth.f90:3) once with dt = argv(2).
! eta_a and gamma_d, read by dust_growth.f90:75,106 but allocated only when charging is on,
! are allocated here with zeros (their value then does not enter any active velocity term).
! Writes ic.bin / final.bin (u_prim), aux.bin (q(:,irho), tstop, sdust, mdust, cs,
! active_cell, aplus, aminus, mminus) and meta.txt.
program ref_growth_harness
  use parameters
  use commons
  use units
  implicit none
  integer :: ilun, i, ix, idust
  real(dp) :: dtg, fac
  character(len=256) :: arg, outdir

  call read_params
  call setup
  call getarg(2, arg)
  read(arg, *) dtg
  call getarg(3, outdir)
  if (.not. allocated(eta_a)) then
     allocate(eta_a(ncells))
     eta_a = 0.0d0
  end if
  if (.not. allocated(gamma_d)) then
     allocate(gamma_d(ncells, ndust))
     gamma_d = 0.0d0
  end if
  do i = 1, ncells
     ix = i - first_active
     if (ix < 0) ix = 0
     fac = 10.0d0**(3*ix)
     u_prim(i, irho) = u_prim(i, irho) * fac
     u_prim(i, ivx) = u_prim(i, ivx) * fac
     u_prim(i, iP) = u_prim(i, iP) * fac
     do idust = 1, ndust
        u_prim(i, irhod(idust)) = u_prim(i, irhod(idust)) * fac
        u_prim(i, ivdx(idust)) = u_prim(i, irhod(idust)) * 1.0d-3 * real(idust, dp) / unit_v
     end do
  end do
  call ctoprim
  call compute_tstop

  ilun = 31
  open(ilun, file=trim(outdir)//'/ic.bin', form='unformatted', access='stream', status='replace')
  write(ilun) u_prim
  close(ilun)
  open(ilun, file=trim(outdir)//'/aux.bin', form='unformatted', access='stream', status='replace')
  write(ilun) q(:, irho)
  write(ilun) tstop
  write(ilun) sdust
  write(ilun) mdust
  write(ilun) cs
  write(ilun) active_cell
  write(ilun) aplus
  write(ilun) aminus
  write(ilun) mminus
  write(ilun) q
  close(ilun)

  dt = dtg
  call dust_growth(.false.)

  open(ilun, file=trim(outdir)//'/final.bin', form='unformatted', access='stream', status='replace')
  write(ilun) u_prim
  close(ilun)
  open(ilun, file=trim(outdir)//'/meta.txt', status='replace')
  write(ilun, '(A,I12)') 'ncells ', ncells
  write(ilun, '(A,I12)') 'nvar ', nvar
  write(ilun, '(A,I12)') 'ndust ', ndust
  write(ilun, '(A,ES25.17)') 'dt ', dtg
  write(ilun, '(A,ES25.17)') 'unit_l ', unit_l
  write(ilun, '(A,ES25.17)') 'unit_d ', unit_d
  write(ilun, '(A,ES25.17)') 'unit_t ', unit_t
  write(ilun, '(A,ES25.17)') 'unit_m ', unit_m
  write(ilun, '(A,ES25.17)') 'unit_v ', unit_v
  write(ilun, '(A,ES25.17)') 'unit_nH ', unit_nH
  write(ilun, '(A,ES25.17)') 'gamma ', gamma
  write(ilun, '(A,ES25.17)') 'smin ', smin
  write(ilun, '(A,ES25.17)') 'smax ', smax
  write(ilun, '(A,ES25.17)') 'rhograin ', rhograin
  write(ilun, '(A,ES25.17)') 'ice_mantle ', ice_mantle
  write(ilun, '(A,ES25.17)') 'slope_mono ', slope_mono
  write(ilun, '(A,ES25.17)') 'size_mono ', size_mono
  write(ilun, '(A,ES25.17)') 'abr ', abr
  write(ilun, '(A,ES25.17)') 'gamma_grains ', gamma_grains
  write(ilun, '(A,ES25.17)') 'estar_grains ', estar_grains
  write(ilun, '(A,ES25.17)') 'alpha_turb ', alpha_turb
  write(ilun, '(A,ES25.17)') 'vfrag ', vfrag
  write(ilun, '(A,ES25.17)') 'cfl_growth ', CFL_growth
  write(ilun, '(A,ES25.17)') 'eps_threshold ', eps_threshold
  write(ilun, '(A,ES25.17)') 'eps_threshold_frag ', eps_threshold_frag
  write(ilun, '(A,ES25.17)') 'dust_ratio_min ', dust_ratio_min
  write(ilun, '(A,ES25.17)') 'sticking ', sticking_efficiency
  write(ilun, '(A,ES25.17)') 'clustered ', clustered_fraction
  write(ilun, '(A,ES25.17)') 'delta_vambi ', delta_vambi
  write(ilun, '(A,L2)') 'fragmentation ', fragmentation
  write(ilun, '(A,L2)') 'turb ', turb_in_growth
  write(ilun, '(A,L2)') 'brownian ', brownian_in_growth
  write(ilun, '(A,L2)') 'drift ', drift_in_growth
  write(ilun, '(A,L2)') 'ambipolar ', ambipolar_in_growth
  write(ilun, '(A,I12)') 'kernel_type ', kernel_type
  close(ilun)
end program ref_growth_harness
