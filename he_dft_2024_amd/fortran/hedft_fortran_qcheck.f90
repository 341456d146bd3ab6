!> Drives the quantum-impurity path of libhedft from Fortran through module
!> hedft_gpu the way the quantum driver loop does (Dynamics_quantum_impurities/
!> DFT4He3d.f90:661-700): create the context, set uimp_k and the impurity mass,
!> load den and psix, start (tddft_init: psi = sqrt(den), qx = 0, potenimp,
!> poten), nrk RK-Gill iterations, then PC iterations, each followed by
!> potenimp/poten, and the impurity density's norm and centre.
!>
!> usage: hedft_fortran_qcheck <input.bin> <output.bin>
!> input  (stream, native): n, niter, nrk (int32); xmax, dt, m_imp_u, mu4 (real64);
!>        uimp_k(n/2+1,n,n), den(n,n,n) (real64); psix(n,n,n) (complex(8))
!> output (stream): errHe, errpsix of the last iteration, etot, ekinx, eHeX, norm
!>        of psix, rcm(3) (real64); psix(n,n,n) complex(8)
program hedft_fortran_qcheck
  use iso_c_binding
  use hedft_gpu
  implicit none
  character(len=512) :: fin, fout
  integer(c_int32_t) :: n, niter, nrk, it
  real(c_double) :: xmax, dt, m_imp_u, mu4, errHe, errpsix, errv, xnorm, rcm(3)
  real(c_double), allocatable :: uk(:,:,:), den(:,:,:)
  complex(c_double_complex), allocatable :: psix(:,:,:)
  type(hedft_grid_t) :: g
  type(hedft_params_t) :: p
  type(hedft_energies_t) :: e
  type(c_ptr) :: ctx
  integer(c_size_t) :: nn
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n, niter, nrk
  read(u) xmax, dt, m_imp_u, mu4
  allocate(uk(n / 2 + 1, n, n), den(n, n, n), psix(n, n, n))
  read(u) uk, den, psix
  close(u)

  g%nx = n; g%ny = n; g%nz = n; g%pad_ = 0
  g%xmax = xmax; g%ymax = xmax; g%zmax = xmax
  call hedft_default_params(p)
  call hedft_check(hedft_create(g, p, HEDFT_MODE_TDDFT, 0_c_int32_t, ctx), 'create')
  nn = int(n, c_size_t) * n * n
  call hedft_check(hedft_set_quantum_impurity(ctx, uk, 8_c_size_t * size(uk, kind=c_size_t), m_imp_u, mu4), &
                   'set_quantum_impurity')
  call hedft_check(hedft_upload(ctx, HEDFT_DEN, den, 8_c_size_t * nn), 'upload den')
  call hedft_check(hedft_upload(ctx, HEDFT_PSIX, psix, 16_c_size_t * nn), 'upload psix')
  call hedft_check(hedft_tddft_init(ctx), 'tddft_init')
  do it = 1, niter
    call hedft_check(hedft_tddft_iteration(ctx, dt, it, nrk, errHe, errpsix, errv), 'tddft_iteration')
  end do
  call hedft_check(hedft_energy(ctx, e), 'energy')
  call hedft_check(hedft_quantum_centroid(ctx, xnorm, rcm), 'centroid')
  call hedft_check(hedft_download(ctx, HEDFT_PSIX, psix, 16_c_size_t * nn), 'download psix')
  call hedft_check(hedft_destroy(ctx), 'destroy')

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) errHe, errpsix, e%etot, e%ekinx, e%eHeX, xnorm, rcm, psix
  close(u)
  write(6, '(A,ES24.16)') 'hedft_fortran_qcheck etot ', e%etot
end program hedft_fortran_qcheck
