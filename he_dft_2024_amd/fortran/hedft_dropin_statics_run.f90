!> A driver-shaped program over the He_statics drop-in: the reference's own
!> module state (He_statics/modules.f90), then the iteration loop of
!> BCN4HeDFT.f90:1086-1160 with Lrkpc: STEPRKR for the first four iterations,
!> STEPPCR after, each followed by the driver's refresh and poten.
!>
!> usage: hedft_dropin_statics_run <input.bin> <output.bin>
!> input  (stream): n, niter, n4 (int32); xmax, dt (real64); den(n,n,n)
!> output (stream): mu4 after every iteration (niter), ErrPC, etot, ekin4
!>        (real64); psi(n,n,n)
program hedft_dropin_statics_run
  use grid, only: nx, ny, nz, xmax, ymax, zmax
  use rho, only: den, psi
  use field, only: pot4
  use he4, only: lsolid
  use lenard4, only: core4, h4, h4ot
  use rkpc, only: ErrPC
  use energies, only: etot, ekin4
  implicit none
  character(len=512) :: fin, fout
  integer(4) :: n, niter, n4, it
  real(8) :: dt, mu4, mu4err
  real(8), allocatable :: mus(:)
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n, niter, n4
  read(u) xmax, dt
  nx = n; ny = n; nz = n; ymax = xmax; zmax = xmax
  allocate(den(nx, ny, nz), psi(nx, ny, nz), pot4(nx, ny, nz), mus(niter))
  read(u) den
  close(u)
  core4 = 'OTC'
  lsolid = .false.
  h4 = h4ot
  call hedft_dropin_statics_start(n4, mu4)
  ErrPC = 0.d0
  do it = 1, niter
    if (it <= 4) then
      call steprkr(dt, mu4, mu4err, n4)
    else
      call steppcr(dt, mu4, mu4err, n4)
    end if
    call hedft_dropin_statics_refresh()
    mus(it) = mu4
  end do
  call hedft_dropin_statics_pull()
  call energy()
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) mus, ErrPC, etot, ekin4, psi
  close(u)
  write(6, '(A,ES24.16)') 'hedft_dropin_statics_run etot ', etot
end program hedft_dropin_statics_run
