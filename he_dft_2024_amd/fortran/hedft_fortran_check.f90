!> Drives libhedft from Fortran through module hedft_gpu the way the TDDFT
!> driver loop does (DFT4He3d.f90:671-696): create the context, load den and
!> the impurity table, start (tddft_init), nrk RK-Gill iterations, then PC
!> iterations, each followed by potenimp/poten/forceimp/aimp.
!>
!> usage: hedft_fortran_check <input.bin> <output.bin>
!> input  (stream, native doubles/int32): n, niter, nrk (int32); xmax, n4, dt,
!>        delinter, rmaxinterpol, mass (real64); npot (int32); potion(npot);
!>        rimp(3), vimp(3); den(n,n,n)
!> output (stream): errHe of the last iteration, etot, ekin4, norm (real64);
!>        psi(n,n,n) complex(8)
program hedft_fortran_check
  use iso_c_binding
  use hedft_gpu
  implicit none
  character(len=512) :: fin, fout
  integer(c_int32_t) :: n, niter, nrk, npot, it
  real(c_double) :: xmax, n4, dt, delinter, rmaxinterpol, mass(1), errHe, errimp, errvimp
  real(c_double) :: rimp(3), vimp(3), aimp(3)
  real(c_double), allocatable :: potion(:), den(:,:,:)
  complex(c_double_complex), allocatable :: psi(:,:,:)
  type(hedft_grid_t) :: g
  type(hedft_params_t) :: p
  type(hedft_energies_t) :: e
  type(c_ptr) :: ctx
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n, niter, nrk
  read(u) xmax, n4, dt, delinter, rmaxinterpol, mass(1)
  read(u) npot
  allocate(potion(npot), den(n, n, n), psi(n, n, n))
  read(u) potion, rimp, vimp, den
  close(u)

  g%nx = n; g%ny = n; g%nz = n; g%pad_ = 0
  g%xmax = xmax; g%ymax = xmax; g%zmax = xmax
  call hedft_default_params(p)
  p%core_otc = 1
  call hedft_check(hedft_create(g, p, HEDFT_MODE_TDDFT, 0_c_int32_t, ctx), 'create')
  call hedft_check(hedft_upload(ctx, HEDFT_DEN, den, int(8, c_size_t) * n * n * n), 'upload den')
  call hedft_check(hedft_set_impurities(ctx, 1_c_int32_t, npot, delinter, rmaxinterpol, potion, mass), &
                   'set_impurities')
  aimp = 0.d0
  call hedft_check(hedft_set_impurity_state(ctx, rimp, vimp, aimp), 'set_impurity_state')
  call hedft_check(hedft_tddft_init(ctx), 'tddft_init')
  do it = 1, niter
    call hedft_check(hedft_tddft_iteration(ctx, dt, it, nrk, errHe, errimp, errvimp), 'tddft_iteration')
  end do
  call hedft_check(hedft_energy(ctx, e), 'energy')
  call hedft_check(hedft_download(ctx, HEDFT_PSI, psi, int(16, c_size_t) * n * n * n), 'download psi')
  call hedft_check(hedft_destroy(ctx), 'destroy')

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) errHe, e%etot, e%ekin4, e%norm, psi
  close(u)
  write(6, '(A,ES24.16)') 'hedft_fortran_check etot ', e%etot
end program hedft_fortran_check
