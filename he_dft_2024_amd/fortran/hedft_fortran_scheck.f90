!> Drives the statics path of libhedft from Fortran through module hedft_gpu
!> with the entry points a driver that tests convergence every k iterations
!> uses: hedft_statics_init, then hedft_statics_run over the iterations of
!> BCN4HeDFT.f90:1113-1160 in batches of k (STEPRKR for iterrk <= 4, STEPPCR
!> after, each followed by the driver's refresh and poten; PC iterations from
!> a graph), and hedft_get_tuning / hedft_last_elapsed.
!>
!> usage: hedft_fortran_scheck <input.bin> <output.bin>
!> input  (stream, native): n, nbatch, k (int32); xmax, dt, n4 (real64); den(n,n,n) (real64)
!> output (stream): mu4, mu4err, errpc after every batch (real64, 3 nbatch);
!>        elapsed ms of the last batch > 0 (1 or 0, real64); psi(n,n,n) (real64)
program hedft_fortran_scheck
  use iso_c_binding
  use hedft_gpu
  implicit none
  character(len=512) :: fin, fout
  integer(c_int32_t) :: n, nbatch, k, b, it
  real(c_double) :: xmax, dt, n4, mu4, mu4err, errpc, ms, ms2, ok
  real(c_double), allocatable :: den(:,:,:), psi(:,:,:), hist(:)
  type(hedft_grid_t) :: g
  type(hedft_params_t) :: p
  type(hedft_tuning_t) :: t
  type(c_ptr) :: ctx
  integer(c_size_t) :: nn
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n, nbatch, k
  read(u) xmax, dt, n4
  allocate(den(n, n, n), psi(n, n, n), hist(3 * nbatch))
  read(u) den
  close(u)

  g%nx = n; g%ny = n; g%nz = n; g%pad_ = 0
  g%xmax = xmax; g%ymax = xmax; g%zmax = xmax
  call hedft_default_params(p)
  p%core_otc = 1
  call hedft_check(hedft_create(g, p, HEDFT_MODE_STATICS, 0_c_int32_t, ctx), 'create')
  nn = int(n, c_size_t) * n * n
  call hedft_check(hedft_upload(ctx, HEDFT_DEN, den, 8_c_size_t * nn), 'upload den')
  call hedft_check(hedft_statics_init(ctx, n4, mu4), 'statics_init')
  call hedft_check(hedft_get_tuning(ctx, t), 'get_tuning')
  if (t%lap25 /= 0) stop 'statics contexts take no 2.5-D stepper'
  it = 1
  do b = 1, nbatch
    call hedft_check(hedft_statics_run(ctx, dt, it, k, n4, 1_c_int32_t, mu4, mu4err, errpc, ms), 'statics_run')
    hist(3 * b - 2) = mu4
    hist(3 * b - 1) = mu4err
    hist(3 * b) = errpc
    it = it + k
  end do
  call hedft_check(hedft_last_elapsed(ctx, ms2), 'last_elapsed')
  ok = merge(1.d0, 0.d0, ms > 0.d0 .and. ms2 == ms)
  call hedft_check(hedft_download(ctx, HEDFT_PSI, psi, 8_c_size_t * nn), 'download psi')
  call hedft_check(hedft_destroy(ctx), 'destroy')
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) hist, ok, psi
  close(u)
  write(6, '(A,ES24.16)') 'hedft_fortran_scheck mu4 ', mu4
end program hedft_fortran_scheck
