! Marchand charge dump harness for the reference build (test infrastructure, never shipped).
! After the reference's read_params and setup (Collapse_1D, charging off in the namelist so the
! setup calls no charge routine), it rescales the density of cell i by 10**(2*(i-first_active))
! (dust follows, each species by 1 + 0.1 k), recomputes primitives with ctoprim, allocates the
! charging arrays the namelist left unallocated, and calls the reference's `charge` (Marchand et
! al. 2021, src/chemistry/ionisation.f90:4-214, with its resistivities) twice: first from
! psi_old = 0 (the global initial guess), then again from the psi_old the first call left.
! Writes aux.bin (inputs: u_prim(:,irho), u_prim(:,irhod(:)), T = barotrop, cs_eos(T), sdust,
! mdust; then per call: zd, ni, ne, psi_old, eta_o, eta_h, eta_a, sigma_o, sigma_p, sigma_h,
! gamma_d) and meta.txt.
program ref_marchand_harness
  use parameters
  use commons
  use units
  implicit none
  integer :: ilun, i, ix, idust, icall
  real(dp) :: fac, barotrop, cs_eos
  real(dp), allocatable :: tt(:), cc(:)
  character(len=256) :: outdir

  call read_params
  call setup
  call getarg(2, outdir)
  do i = 1, ncells
     ix = i - first_active
     if (ix < 0) ix = 0
     fac = 10.0d0**(2*ix)
     u_prim(i, irho) = u_prim(i, irho) * fac
     u_prim(i, ivx) = u_prim(i, ivx) * fac
     u_prim(i, iP) = u_prim(i, iP) * fac
     do idust = 1, ndust
        u_prim(i, irhod(idust)) = u_prim(i, irhod(idust)) * fac * (1.0d0 + 0.1d0 * idust)
     end do
  end do
  call ctoprim
  if (.not. allocated(zd)) allocate(zd(ncells, ndust))
  if (.not. allocated(gamma_d)) allocate(gamma_d(ncells, ndust))
  if (.not. allocated(ni)) allocate(ni(ncells))
  if (.not. allocated(ne)) allocate(ne(ncells))
  if (.not. allocated(psi_old)) allocate(psi_old(ncells))
  if (.not. allocated(eta_o)) allocate(eta_o(ncells), eta_h(ncells), eta_a(ncells))
  if (.not. allocated(sigma_o)) allocate(sigma_o(ncells), sigma_p(ncells), sigma_h(ncells))
  zd = 0.0d0; gamma_d = 0.0d0; ni = 0.0d0; ne = 0.0d0; psi_old = 0.0d0
  eta_o = 0.0d0; eta_h = 0.0d0; eta_a = 0.0d0; sigma_o = 0.0d0; sigma_p = 0.0d0; sigma_h = 0.0d0
  allocate(tt(ncells), cc(ncells))
  do i = 1, ncells
     tt(i) = barotrop(u_prim(i, irho))
     cc(i) = cs_eos(tt(i))
  end do

  ilun = 31
  open(ilun, file=trim(outdir)//'/aux.bin', form='unformatted', access='stream', status='replace')
  write(ilun) u_prim(:, irho)
  do idust = 1, ndust
     write(ilun) u_prim(:, irhod(idust))
  end do
  write(ilun) tt
  write(ilun) cc
  write(ilun) sdust
  write(ilun) mdust
  do icall = 1, 2
     call charge
     write(ilun) zd
     write(ilun) ni
     write(ilun) ne
     write(ilun) psi_old
     write(ilun) eta_o
     write(ilun) eta_h
     write(ilun) eta_a
     write(ilun) sigma_o
     write(ilun) sigma_p
     write(ilun) sigma_h
     write(ilun) gamma_d
  end do
  close(ilun)
  open(ilun, file=trim(outdir)//'/meta.txt', status='replace')
  write(ilun, '(A,I12)') 'ncells ', ncells
  write(ilun, '(A,I12)') 'ndust ', ndust
  write(ilun, '(A,ES25.17)') 'zeta ', x
  write(ilun, '(A,ES25.17)') 'mu_ions ', mu_ions
  write(ilun, '(A,ES25.17)') 'mu_gas ', mu_gas
  write(ilun, '(A,ES25.17)') 'stickeff_el ', stickeff_el
  write(ilun, '(A,ES25.17)') 'epsilon_ionis ', epsilon_ionis
  write(ilun, '(A,I12)') 'nitermax_ionis ', nitermax_ionis
  write(ilun, '(A,ES25.17)') 'B_0_lee ', B_0_lee
  write(ilun, '(A,ES25.17)') 'rhograin ', rhograin
  write(ilun, '(A,ES25.17)') 'gamma ', gamma
  write(ilun, '(A,L2)') 'res_Marchand ', res_Marchand
  write(ilun, '(A,ES25.17)') 'unit_l ', unit_l
  write(ilun, '(A,ES25.17)') 'unit_d ', unit_d
  write(ilun, '(A,ES25.17)') 'unit_m ', unit_m
  write(ilun, '(A,ES25.17)') 'unit_v ', unit_v
  write(ilun, '(A,ES25.17)') 'unit_nh ', unit_nH
  close(ilun)
end program ref_marchand_harness
