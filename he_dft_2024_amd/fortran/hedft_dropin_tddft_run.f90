!> A driver-shaped program over the TDDFT drop-in: it sets the reference's own
!> module state (Dynamics_TDDFT/modules.f90) as dimen.f90 and the input stage
!> leave it, then runs the time loop of DFT4He3d.f90:671-696 calling STEPRK /
!> STEPPC / potenimp / poten / forceimp by their reference names, with the
!> driver's own aimp = F/m on the host.
!>
!> usage: hedft_dropin_tddft_run <input.bin> <output.bin>
!> input  (stream): n, niter (int32); xmax, dt, delinter, rmaxinterpol, mass
!>        (real64); npot (int32); potion(npot); rimp(3), vimp(3); den(n,n,n)
!> output (stream): errHe of the last iteration, etot, ekin4, eHeX (real64);
!>        rimp(3), vimp(3), F(3); psi(n,n,n) complex(8)
program hedft_dropin_tddft_run
  use grid, only: nx, ny, nz, xmax, ymax, zmax
  use rho, only: den, psi
  use field, only: pot4
  use he4, only: lsolid
  use lenard4, only: core4, h4, h4ot
  use classicimp, only: N_imp, rimp, vimp, aimp, F, m_imp
  use interpol, only: DelInter, potion, rmaxinterpol, npot
  use energies, only: etot, ekin4, eHeX
  implicit none
  character(len=512) :: fin, fout
  integer(4) :: n, niter, it, np
  real(8) :: dt, mass, errHe, errimp, errvimp, r0(3), v0(3)
  real(8), allocatable :: tab(:)
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n, niter
  read(u) xmax, dt, DelInter, rmaxinterpol, mass
  read(u) np
  allocate(tab(np))
  read(u) tab, r0, v0
  nx = n; ny = n; nz = n; ymax = xmax; zmax = xmax
  allocate(den(nx, ny, nz), psi(nx, ny, nz), pot4(nx, ny, nz))
  read(u) den
  close(u)
  core4 = 'OTC'
  lsolid = .false.
  h4 = h4ot   ! the core radius the driver picks for 'OT'/'OTC' (DFT4He3d.f90:464-470)
  N_imp = 1
  npot = np
  allocate(potion(N_imp, npot), rimp(N_imp, 3), vimp(N_imp, 3), aimp(N_imp, 3), F(N_imp, 3), m_imp(N_imp))
  potion(1, :) = tab
  rimp(1, :) = r0
  vimp(1, :) = v0
  aimp = 0.d0
  m_imp(1) = mass
  ! mode 0 start: psi = sqrt(den), den = |psi|^2 (DFT4He3d.f90:541-556)
  psi = sqrt(den)
  den = abs(psi)**2
  call hedft_dropin_start(2.d0, 1000.d0, 1000.d0, 1000.d0, 2.d0, 2.d0, 2.d0)
  call potenimp()
  call poten()
  call forceimp()
  aimp(:, 1) = F(:, 1) / m_imp(:)
  aimp(:, 2) = F(:, 2) / m_imp(:)
  aimp(:, 3) = F(:, 3) / m_imp(:)
  errHe = 0.d0
  do it = 1, niter
    if (it <= 3) then
      call steprk(dt)
    else
      call steppc(dt, errHe, errimp, errvimp)
    end if
    call potenimp()
    call poten()
    call forceimp()
    aimp(:, 1) = F(:, 1) / m_imp(:)
    aimp(:, 2) = F(:, 2) / m_imp(:)
    aimp(:, 3) = F(:, 3) / m_imp(:)
  end do
  call hedft_dropin_pull()
  call energy()
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) errHe, etot, ekin4, eHeX, rimp(1, :), vimp(1, :), F(1, :), psi
  close(u)
  write(6, '(A,ES24.16)') 'hedft_dropin_tddft_run etot ', etot
end program hedft_dropin_tddft_run
