!> A driver-shaped program over the quantum-impurity drop-in: it sets the
!> reference's own module state (Dynamics_quantum_impurities/modules.f90) for
!> N_imp impurities as dimen.f90 and potenimpini.f90 leave it (uimp_k, uimp_k_k,
!> Selec_gs_k_k, m_imp_u, psix), then runs the time loop of DFT4He3d.f90:667-677
!> calling STEPRK / STEPPC / potenimp / poten by their reference names, and
!> energy at the end.
!>
!> usage: hedft_dropin_quantum_run <input.bin> <output.bin>
!> input  (stream): n, niter, nq (int32); xmax, dt, mu4 (real64); m_imp_u(nq)
!>        (real64); pair(nq,nq) (int32: 1 = a uimp_k_k kernel, 0 = 'null');
!>        uimp_k(n/2+1,n,n,nq), uimp_k_k(n/2+1,n,n,nq,nq), den(n,n,n) (real64);
!>        psix(n,n,n,nq) (complex(8))
!> output (stream): errHe, errpsix of the last iteration, etot, ekinx, eHeX,
!>        eimpu_impu (real64); psi(n,n,n), psix(n,n,n,nq) (complex(8));
!>        upotx(n,n,n,nq), potx4(n,n,n) (real64)
program hedft_dropin_quantum_run
  use grid, only: nx, ny, nz, xmax, ymax, zmax
  use rho, only: den, psi
  use field, only: pot4
  use he4, only: lsolid
  use lenard4, only: core4, h4, h4ot
  use impur, only: psix, potx4, upotx, mu4
  use quantal_imp, only: N_imp, uimp_k, uimp_k_k, m_imp_u, selec_gs_k_k
  use energies, only: etot, ekinx, eHeX, eimpu_impu
  implicit none
  character(len=512) :: fin, fout
  integer(4) :: n, niter, nq, it, i, j
  integer(4), allocatable :: pair(:, :)
  real(8) :: dt, errHe, errpsix
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n, niter, nq
  read(u) xmax, dt, mu4
  nx = n; ny = n; nz = n; ymax = xmax; zmax = xmax
  N_imp = nq
  allocate(m_imp_u(N_imp), pair(N_imp, N_imp), selec_gs_k_k(N_imp, N_imp))
  allocate(uimp_k(nx / 2 + 1, ny, nz, N_imp), uimp_k_k(nx / 2 + 1, ny, nz, N_imp, N_imp))
  allocate(den(nx, ny, nz), psi(nx, ny, nz), pot4(nx, ny, nz), potx4(nx, ny, nz))
  allocate(psix(nx, ny, nz, N_imp), upotx(nx, ny, nz, N_imp))
  read(u) m_imp_u, pair, uimp_k, uimp_k_k, den, psix
  close(u)
  do j = 1, N_imp
    do i = 1, N_imp
      selec_gs_k_k(i, j) = 'null'
      if (pair(i, j) /= 0) selec_gs_k_k(i, j) = 'LJ'
    end do
  end do
  core4 = 'OTC'
  lsolid = .false.
  h4 = h4ot   ! the core radius the driver picks for 'OT'/'OTC'
  ! mode 0 start at rest: psi = sqrt(den), den = |psi|^2 (DFT4He3d.f90:523-541)
  psi = sqrt(den)
  den = abs(psi)**2
  call hedft_dropin_start(0.d0, 1000.d0, 1000.d0, 1000.d0, 2.d0, 2.d0, 2.d0)
  call potenimp()
  call poten()
  errHe = 0.d0
  errpsix = 0.d0
  do it = 1, niter
    if (it <= 3) then
      call steprk(dt)
    else
      call steppc(dt, errHe, errpsix)
    end if
    call potenimp()
    call poten()
  end do
  call hedft_dropin_pull()
  call energy()
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) errHe, errpsix, etot, ekinx, eHeX, eimpu_impu, psi, psix, upotx, potx4
  close(u)
  write(6, '(A,ES24.16)') 'hedft_dropin_quantum_run etot ', etot
end program hedft_dropin_quantum_run
