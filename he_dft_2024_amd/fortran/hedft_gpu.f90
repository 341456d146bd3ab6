!> hedft_gpu — ISO_C_BINDING interface of libhedft (include/hedft.h) for the
!> Fortran drivers of BCN-4He-DFT.  Every interface binds one C entry point;
!> the derived types mirror hedft_grid, hedft_params and hedft_energies
!> field by field.  hedft_check() stops with the library's message, the
!> reference's own error convention (He_statics/BCN4HeDFT.f90:233-235).
module hedft_gpu
  use iso_c_binding
  implicit none
  private :: c_strlen

  ! status codes and modes
  integer(c_int), parameter :: HEDFT_OK = 0, HEDFT_EINVAL = -1, HEDFT_EGRID = -2, &
       HEDFT_EHIP = -3, HEDFT_ENOMEM = -4, HEDFT_ESTATE = -5, HEDFT_EBOUNDS = -6
  integer(c_int32_t), parameter :: HEDFT_MODE_STATICS = 0, HEDFT_MODE_TDDFT = 1
  integer(c_int32_t), parameter :: HEDFT_PREC_FP64 = 0, HEDFT_PREC_MIXED = 1
  integer(c_int32_t), parameter :: HEDFT_MAX_IMP = 4, HEDFT_MAX_QIMP = 4

  ! field identifiers (enum hedft_field)
  integer(c_int32_t), parameter :: HEDFT_DEN = 0, HEDFT_PSI = 1, HEDFT_POT4 = 2, &
       HEDFT_DENCG = 3, HEDFT_DELJ4 = 4, HEDFT_DENALF = 5, HEDFT_FALFS = 6, &
       HEDFT_INTXALF = 7, HEDFT_INTYALF = 8, HEDFT_INTZALF = 9, HEDFT_UALPHAS = 10, &
       HEDFT_DXDEN = 11, HEDFT_DYDEN = 12, HEDFT_DZDEN = 13, HEDFT_UIMP = 14, &
       HEDFT_UEXT = 15, HEDFT_TIMEC = 16, HEDFT_PSIOLD1 = 17, HEDFT_PSIOLD2 = 18, &
       HEDFT_PSIOLD3 = 19, HEDFT_HPSIOLD1 = 20, HEDFT_HPSIOLD2 = 21, HEDFT_PC = 22, &
       HEDFT_Q = 23, HEDFT_FDEN = 24, HEDFT_FVLJ4 = 25, HEDFT_WCGK = 26, &
       HEDFT_KALFS = 27, HEDFT_WORK1 = 28, HEDFT_WORK2 = 29, HEDFT_WORKC1 = 30, &
       HEDFT_WORKK1 = 31
  ! quantum impurity (module quantal_imp of Dynamics_quantum_impurities)
  integer(c_int32_t), parameter :: HEDFT_PSIX = 32, HEDFT_POTX4 = 33, HEDFT_UPOTX = 34, &
       HEDFT_DENX = 35, HEDFT_UIMPK = 36, HEDFT_PSIXOLD1 = 37, HEDFT_PSIXOLD2 = 38, &
       HEDFT_PSIXOLD3 = 39, HEDFT_HPSIXOLD1 = 40, HEDFT_HPSIXOLD2 = 41, HEDFT_PCX = 42, &
       HEDFT_QX = 43
  ! statics log-psi steppers: psil = log|psi| (rkpc: Psil)
  integer(c_int32_t), parameter :: HEDFT_PSIL = 44

  type, bind(C) :: hedft_grid_t
    integer(c_int32_t) :: nx, ny, nz, pad_
    real(c_double) :: xmax, ymax, zmax
  end type hedft_grid_t

  type, bind(C) :: hedft_params_t
    real(c_double) :: h2o2m4, cp4, cpp4, alphas, l, den0s
    real(c_double) :: csolid, beta_solid, den_m
    real(c_double) :: h4, eps4, sigma4
    integer(c_int32_t) :: core_otc, lsolid, npd, icon
    real(c_double) :: dtmin, sdmu, dmumax, psimin
    real(c_double) :: lambdah, tmean(3), tsurf(3)
    integer(c_int32_t) :: precision, lexternal
  end type hedft_params_t

  type, bind(C) :: hedft_energies_t
    real(c_double) :: ekin4, elj4, ecor4, ealphas, esolid, eext, etot4
    real(c_double) :: ekinx, eHeX, eimpu_impu, eimpu, etot, norm
  end type hedft_energies_t

  ! observables block of the TDDFT driver (Dynamics_TDDFT/DFT4He3d.f90:709-836)
  type, bind(C) :: hedft_observables_t
    real(c_double) :: norm, rcm(3), l(3), vcom(3), r2(3)
  end type hedft_observables_t

  type, bind(C) :: hedft_tuning_t
    integer(c_int32_t) :: xy_split_fwd, xy_split_inv, no_side, prefer_eager
    integer(c_int32_t) :: lap_late, lap25, fused
    real(c_double) :: ms_per_step
  end type hedft_tuning_t

  type, bind(C) :: hedft_denfile_t
    real(c_double) :: xmax, ymax, zmax, hx, hy, hz
    integer(c_int32_t) :: nx, ny, nz, limp
    real(c_double) :: ximp, yimp, zimp
  end type hedft_denfile_t

  interface
    function hedft_strerror_c(rc) bind(C, name="hedft_strerror")
      import :: c_int, c_ptr
      integer(c_int), value :: rc
      type(c_ptr) :: hedft_strerror_c
    end function
    subroutine hedft_default_params(p) bind(C, name="hedft_default_params")
      import :: hedft_params_t
      type(hedft_params_t), intent(out) :: p
    end subroutine
    integer(c_int) function hedft_grid_supported(n) bind(C, name="hedft_grid_supported")
      import :: c_int, c_int32_t
      integer(c_int32_t), value :: n
    end function
    integer(c_int) function hedft_create(g, p, mode, device, ctx) bind(C, name="hedft_create")
      import :: c_int, c_int32_t, c_ptr, hedft_grid_t, hedft_params_t
      type(hedft_grid_t), intent(in) :: g
      type(hedft_params_t), intent(in) :: p
      integer(c_int32_t), value :: mode, device
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function hedft_destroy(ctx) bind(C, name="hedft_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    ! host: any contiguous array in the module layout (assumed type -> void*)
    integer(c_int) function hedft_upload(ctx, field, host, bytes) bind(C, name="hedft_upload")
      import :: c_int, c_int32_t, c_ptr, c_size_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: field
      type(*), intent(in) :: host(*)
      integer(c_size_t), value :: bytes
    end function
    integer(c_int) function hedft_download(ctx, field, host, bytes) bind(C, name="hedft_download")
      import :: c_int, c_int32_t, c_ptr, c_size_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: field
      type(*), intent(inout) :: host(*)
      integer(c_size_t), value :: bytes
    end function
    integer(c_size_t) function hedft_field_bytes(ctx, field) bind(C, name="hedft_field_bytes")
      import :: c_int32_t, c_ptr, c_size_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: field
    end function
    integer(c_int) function hedft_synchronize(ctx) bind(C, name="hedft_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hedft_derivnD(ctx, id, iv, fin, fout) bind(C, name="hedft_derivnD")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: id, iv, fin, fout
    end function
    integer(c_int) function hedft_fftfw(ctx, fin, fout) bind(C, name="hedft_fftfw")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: fin, fout
    end function
    integer(c_int) function hedft_fftbk(ctx, fin, fout) bind(C, name="hedft_fftbk")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: fin, fout
    end function
    integer(c_int) function hedft_derden(ctx) bind(C, name="hedft_derden")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hedft_poten(ctx) bind(C, name="hedft_poten")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hedft_term_alfa(ctx) bind(C, name="hedft_term_alfa")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hedft_energy(ctx, e) bind(C, name="hedft_energy")
      import :: c_int, c_ptr, hedft_energies_t
      type(c_ptr), value :: ctx
      type(hedft_energies_t), intent(out) :: e
    end function
    ! n4: the nominal atom number r_cm divides by; mhe in K^-1 A^-2
    integer(c_int) function hedft_compute_observables(ctx, n4, mhe, o) bind(C, name="hedft_compute_observables")
      import :: c_int, c_ptr, c_int32_t, c_double, hedft_observables_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n4
      real(c_double), value :: mhe
      type(hedft_observables_t), intent(out) :: o
    end function
    integer(c_int) function hedft_steprkr(ctx, deltat, mu4, mu4err, n4) bind(C, name="hedft_steprkr")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat, n4
      real(c_double), intent(inout) :: mu4
      real(c_double), intent(out) :: mu4err
    end function
    integer(c_int) function hedft_steppcr(ctx, deltat, mu4, mu4err, n4, errpc) bind(C, name="hedft_steppcr")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat, n4
      real(c_double), intent(inout) :: mu4
      real(c_double), intent(out) :: mu4err, errpc
    end function
    integer(c_int) function hedft_steprk(ctx, deltat) bind(C, name="hedft_steprk")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat
    end function
    integer(c_int) function hedft_steppc(ctx, deltat, errHe, errimp, errvimp) bind(C, name="hedft_steppc")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat
      real(c_double), intent(out) :: errHe, errimp, errvimp
    end function
    integer(c_int) function hedft_set_impurities(ctx, nimp, npot, delinter, rmaxinterpol, potion, mass) &
         bind(C, name="hedft_set_impurities")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nimp, npot
      real(c_double), value :: delinter, rmaxinterpol
      real(c_double), intent(in) :: potion(*), mass(*)
    end function
    ! (N_imp,3) arrays; an absent argument is passed as NULL (left unchanged)
    integer(c_int) function hedft_set_impurity_state(ctx, rimp, vimp, aimp) bind(C, name="hedft_set_impurity_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      ! every argument is required: an absent optional argument of a bind(C)
      ! interface is not passed as a null pointer by every Fortran compiler
      real(c_double), intent(in) :: rimp(*), vimp(*), aimp(*)
    end function
    integer(c_int) function hedft_get_impurity_state(ctx, rimp, vimp, aimp, force, lstopimp) &
         bind(C, name="hedft_get_impurity_state")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: rimp(*), vimp(*), aimp(*), force(*)
      integer(c_int32_t), intent(out) :: lstopimp
    end function
    integer(c_int) function hedft_potenimp(ctx) bind(C, name="hedft_potenimp")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hedft_forceimp(ctx) bind(C, name="hedft_forceimp")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hedft_update_aimp(ctx) bind(C, name="hedft_update_aimp")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    ! uimp_k(:,:,:,1) of potenimpini (real, (nx/2+1,ny,nz)), m_imp_u(1), mu4
    integer(c_int) function hedft_set_quantum_impurity(ctx, uimp_k, nbytes, m_imp_u, mu4) &
        bind(C, name="hedft_set_quantum_impurity")
      import :: c_int, c_ptr, c_double, c_size_t
      type(c_ptr), value :: ctx
      type(*), intent(in) :: uimp_k(*)
      integer(c_size_t), value :: nbytes
      real(c_double), value :: m_imp_u, mu4
    end function
    integer(c_int) function hedft_set_quantum_impurities(ctx, nq, uimp_k, uimp_kk, pair, nbytes, m_imp_u, mu4) &
        bind(C, name="hedft_set_quantum_impurities")
      import :: c_int, c_int32_t, c_ptr, c_double, c_size_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nq
      type(*), intent(in) :: uimp_k(*), uimp_kk(*)
      integer(c_int32_t), intent(in) :: pair(*)
      integer(c_size_t), value :: nbytes
      real(c_double), intent(in) :: m_imp_u(*)
      real(c_double), value :: mu4
    end function
    integer(c_int) function hedft_select_quantum_impurity(ctx, k) bind(C, name="hedft_select_quantum_impurity")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: k
    end function
    integer(c_int) function hedft_quantum_centroid(ctx, norm, rcm) bind(C, name="hedft_quantum_centroid")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: norm, rcm(3)
    end function
    integer(c_int) function hedft_statics_init(ctx, n4, mu4) bind(C, name="hedft_statics_init")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: n4
      real(c_double), intent(out) :: mu4
    end function
    integer(c_int) function hedft_statics_iteration(ctx, deltat, iterrk, n4, mu4, mu4err) &
         bind(C, name="hedft_statics_iteration")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat, n4
      integer(c_int32_t), value :: iterrk
      real(c_double), intent(inout) :: mu4
      real(c_double), intent(out) :: mu4err
    end function
    integer(c_int) function hedft_tddft_init(ctx) bind(C, name="hedft_tddft_init")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hedft_tddft_iteration(ctx, deltat, iter, nrk, errHe, errimp, errvimp) &
         bind(C, name="hedft_tddft_iteration")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat
      integer(c_int32_t), value :: iter, nrk
      real(c_double), intent(out) :: errHe, errimp, errvimp
    end function
    integer(c_int) function hedft_tddft_run_pc(ctx, deltat, nsteps, use_graph, elapsed_ms) &
         bind(C, name="hedft_tddft_run_pc")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat
      integer(c_int32_t), value :: nsteps, use_graph
      real(c_double), intent(out) :: elapsed_ms
    end function
    integer(c_int) function hedft_statics_iteration_log(ctx, deltat, iterrk, n4, mu4, mu4err, errpc) &
         bind(C, name="hedft_statics_iteration_log")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat, n4
      integer(c_int32_t), value :: iterrk
      real(c_double), intent(inout) :: mu4
      real(c_double), intent(out) :: mu4err, errpc
    end function
    integer(c_int) function hedft_get_tuning(ctx, t) bind(C, name="hedft_get_tuning")
      import :: c_int, c_ptr, hedft_tuning_t
      type(c_ptr), value :: ctx
      type(hedft_tuning_t), intent(out) :: t
    end function
    integer(c_int) function hedft_last_elapsed(ctx, ms) bind(C, name="hedft_last_elapsed")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ms
    end function
    integer(c_int) function hedft_batch_run_pc(ctxs, nctx, deltat, nsteps, chunk, elapsed_ms) &
         bind(C, name="hedft_batch_run_pc")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), intent(in) :: ctxs(*)
      integer(c_int32_t), value :: nctx, nsteps, chunk
      real(c_double), intent(in) :: deltat(*)
      real(c_double), intent(out) :: elapsed_ms
    end function
    integer(c_int) function hedft_tddft_prepare_pc(ctx, deltat, skip, nsteps) bind(C, name="hedft_tddft_prepare_pc")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat
      integer(c_int32_t), value :: skip, nsteps
    end function
    integer(c_int) function hedft_statics_run(ctx, deltat, iter0, niter, n4, use_graph, mu4, mu4err, errpc, &
                                              elapsed_ms) bind(C, name="hedft_statics_run")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat, n4
      integer(c_int32_t), value :: iter0, niter, use_graph
      real(c_double), intent(out) :: mu4, mu4err, errpc, elapsed_ms
    end function
    integer(c_int) function hedft_evolo(ctx, deltat, vdt1, vdt2, iron, n4, mu4, mu4err) bind(C, name="hedft_evolo")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat, vdt1, vdt2, n4
      integer(c_int32_t), value :: iron
      real(c_double), intent(inout) :: mu4
      real(c_double), intent(out) :: mu4err
    end function
    integer(c_int) function hedft_statics_iteration_sd(ctx, deltat, vdt1, vdt2, iron, n4, mu4, mu4err) &
         bind(C, name="hedft_statics_iteration_sd")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat, vdt1, vdt2, n4
      integer(c_int32_t), value :: iron
      real(c_double), intent(inout) :: mu4
      real(c_double), intent(out) :: mu4err
    end function
    integer(c_int) function hedft_statics_log_start(ctx) bind(C, name="hedft_statics_log_start")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hedft_steprkrl(ctx, deltat, mu4, mu4err, n4) bind(C, name="hedft_steprkrl")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat, n4
      real(c_double), intent(inout) :: mu4
      real(c_double), intent(out) :: mu4err
    end function
    integer(c_int) function hedft_steppcrl(ctx, deltat, mu4, mu4err, n4, errpc) bind(C, name="hedft_steppcrl")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat, n4
      real(c_double), intent(inout) :: mu4
      real(c_double), intent(out) :: mu4err, errpc
    end function
    integer(c_int) function hedft_evolox(ctx, deltatx, vdt1, vdt2, ironx, epsx, epsxerr) bind(C, name="hedft_evolox")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deltatx, vdt1, vdt2
      integer(c_int32_t), value :: ironx
      real(c_double), intent(inout) :: epsx
      real(c_double), intent(out) :: epsxerr
    end function
    integer(c_int) function hedft_statics_refresh(ctx) bind(C, name="hedft_statics_refresh")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hedft_set_impurity_pairs(ctx, npair, delpair, vpair, dvpair) &
         bind(C, name="hedft_set_impurity_pairs")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: npair
      real(c_double), value :: delpair
      real(c_double), intent(in) :: vpair(*), dvpair(*)
    end function
    integer(c_int) function hedft_autotune(ctx, deltat, trial_steps, best) bind(C, name="hedft_autotune")
      import :: c_int, c_int32_t, c_ptr, c_double, hedft_tuning_t
      type(c_ptr), value :: ctx
      real(c_double), value :: deltat
      integer(c_int32_t), value :: trial_steps
      type(hedft_tuning_t), intent(out) :: best
    end function
    integer(c_int) function hedft_set_tuning(ctx, t) bind(C, name="hedft_set_tuning")
      import :: c_int, c_ptr, hedft_tuning_t
      type(c_ptr), value :: ctx
      type(hedft_tuning_t), intent(in) :: t
    end function
    integer(c_int) function hedft_read_density(path, cchar, hdr, den, n) bind(C, name="hedft_read_density")
      import :: c_int, c_char, c_size_t, c_double, hedft_denfile_t
      character(kind=c_char), intent(in) :: path(*)
      character(kind=c_char), value :: cchar
      type(hedft_denfile_t), intent(out) :: hdr
      real(c_double), intent(out) :: den(*)
      integer(c_size_t), value :: n
    end function
    integer(c_int) function hedft_write_density(path, hdr, den, n, comment) bind(C, name="hedft_write_density")
      import :: c_int, c_char, c_size_t, c_double, hedft_denfile_t
      character(kind=c_char), intent(in) :: path(*), comment(*)
      type(hedft_denfile_t), intent(in) :: hdr
      real(c_double), intent(in) :: den(*)
      integer(c_size_t), value :: n
    end function
    integer(c_int) function hedft_read_wavefunction(path, cchar, hdr, nimp, rimp, vimp, max_imp, psi, n) &
         bind(C, name="hedft_read_wavefunction")
      import :: c_int, c_int32_t, c_char, c_size_t, c_double, c_double_complex, hedft_denfile_t
      character(kind=c_char), intent(in) :: path(*)
      character(kind=c_char), value :: cchar
      type(hedft_denfile_t), intent(out) :: hdr
      integer(c_int32_t), intent(out) :: nimp
      real(c_double), intent(out) :: rimp(*), vimp(*)
      integer(c_int32_t), value :: max_imp
      complex(c_double_complex), intent(out) :: psi(*)
      integer(c_size_t), value :: n
    end function
    integer(c_int) function hedft_write_wavefunction(path, hdr, nimp, rimp, vimp, psi, n, comment) &
         bind(C, name="hedft_write_wavefunction")
      import :: c_int, c_int32_t, c_char, c_size_t, c_double, c_double_complex, hedft_denfile_t
      character(kind=c_char), intent(in) :: path(*), comment(*)
      type(hedft_denfile_t), intent(in) :: hdr
      integer(c_int32_t), value :: nimp
      real(c_double), intent(in) :: rimp(*), vimp(*)
      complex(c_double_complex), intent(in) :: psi(*)
      integer(c_size_t), value :: n
    end function
  end interface

contains

  !> Stops with the library's message when rc is not HEDFT_OK.
  subroutine hedft_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= HEDFT_OK) then
      write(6, '(A,A,A,I0,A,A)') 'hedft: ', what, ' failed (', rc, '): ', hedft_strerror(rc)
      stop 'hedft error'
    end if
  end subroutine hedft_check

  function hedft_strerror(rc) result(msg)
    integer(c_int), intent(in) :: rc
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: n, i
    cp = hedft_strerror_c(rc)
    n = c_strlen(cp)
    call c_f_pointer(cp, p, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = p(i)
    end do
  end function hedft_strerror

  integer function c_strlen(cp)
    type(c_ptr), intent(in) :: cp
    character(kind=c_char), pointer :: p(:)
    integer :: i
    call c_f_pointer(cp, p, [4096])
    c_strlen = 0
    do i = 1, 4096
      if (p(i) == c_null_char) exit
      c_strlen = i
    end do
  end function c_strlen

end module hedft_gpu
