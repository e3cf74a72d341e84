! Charge dump harness for the reference build (test infrastructure, never shipped).
! After the reference's read_params and setup (Collapse_1D, charging off in the namelist so the
! setup calls no charge routine), it rescales the density of cell i by 10**(2*(i-first_active))
! (dust follows), recomputes primitives with ctoprim, allocates zd/ni/ne when the namelist left
! them unallocated, and calls the reference's analytical_charge
! (src/chemistry/ionisation.f90:809-922).  Writes aux.bin (q(:,irho), q(:,irhod(:)), mdust,
! sdust, zd, ni, ne) and meta.txt.
program ref_charge_harness
  use parameters
  use commons
  use units
  implicit none
  integer :: ilun, i, ix, idust
  real(dp) :: fac
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
  if (.not. allocated(ni)) allocate(ni(ncells))
  if (.not. allocated(ne)) allocate(ne(ncells))
  zd = 0.0d0
  call analytical_charge

  ilun = 31
  open(ilun, file=trim(outdir)//'/aux.bin', form='unformatted', access='stream', status='replace')
  write(ilun) q(:, irho)
  do idust = 1, ndust
     write(ilun) q(:, irhod(idust))
  end do
  write(ilun) mdust
  write(ilun) sdust
  write(ilun) zd
  write(ilun) ni
  write(ilun) ne
  close(ilun)
  open(ilun, file=trim(outdir)//'/meta.txt', status='replace')
  write(ilun, '(A,I12)') 'ncells ', ncells
  write(ilun, '(A,I12)') 'ndust ', ndust
  write(ilun, '(A,ES25.17)') 'zeta ', x
  write(ilun, '(A,ES25.17)') 'mu_ions ', mu_ions
  write(ilun, '(A,ES25.17)') 'mu_gas ', mu_gas
  write(ilun, '(A,ES25.17)') 'T0_cloud ', T0_cloud
  write(ilun, '(A,L2)') 'electrons ', electrons
  close(ilun)
end program ref_charge_harness
