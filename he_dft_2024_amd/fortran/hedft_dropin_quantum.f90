!> Drop-in replacements for the time-step subroutines of the quantum-impurity
!> program of BCN-4He-DFT (Dynamics_quantum_impurities): STEPRK(deltat),
!> STEPPC(deltat,errHe,errpsix), poten, potenimp and energy keep their names and
!> signatures and run on the GPU through module hedft_gpu, for N_imp <= 4
!> impurities with their uimp_k_k pair kernels.  The module arrays (den, psi,
!> psix, pot4, potx4, upotx, ...) stay
!> on the device between calls.  Two calls are added to the driver:
!>   call hedft_dropin_start(Lambdah, txmean, tymean, tzmean, txsurf, tysurf, tzsurf)
!> after timec has been set (before the time loop, DFT4He3d.f90:661), and
!>   call hedft_dropin_pull()
!> before each output point to refresh den, psi, pot4, psix, potx4 and upotx.
!> The driver always calls potenimp() right before poten() (DFT4He3d.f90:540-541,
!> 676-677; steprk.f90:108-109; steppc.f90:96-97), and hedft_poten computes
!> potx4 and upotx within the same mean-field pass, so potenimp() here only
!> checks the state.  Compiled against the reference's own module files
!> (Makefile target dropin).
module hedft_dropin_qstate
  use iso_c_binding
  implicit none
  type(c_ptr), save :: hctx = c_null_ptr
end module hedft_dropin_qstate

subroutine hedft_dropin_start(Lambdah, txmean, tymean, tzmean, txsurf, tysurf, tzsurf)
  use iso_c_binding
  use hedft_gpu
  use hedft_dropin_qstate
  use grid, only: nx, ny, nz, xmax, ymax, zmax
  use he4, only: h2o2m4, cp4, cpp4, alphas, l, den0s, lsolid, C, beta, den_m
  use lenard4, only: h4, eps4, sigma4, core4
  use deriva, only: npd, icon
  use rho, only: den, psi
  use impur, only: psix, mu4
  use quantal_imp, only: N_imp, uimp_k, uimp_k_k, m_imp_u, selec_gs_k_k
  implicit none
  real(8), intent(in) :: Lambdah, txmean, tymean, tzmean, txsurf, tysurf, tzsurf
  type(hedft_grid_t) :: g
  type(hedft_params_t) :: p
  integer(c_size_t) :: n, nh
  integer(c_int32_t) :: pair(N_imp, N_imp)
  integer :: i, j

  if (N_imp < 1 .or. N_imp > HEDFT_MAX_QIMP) stop 'hedft_dropin_start: 1 <= N_imp <= 4'
  g%nx = nx; g%ny = ny; g%nz = nz; g%pad_ = 0
  g%xmax = xmax; g%ymax = ymax; g%zmax = zmax
  call hedft_default_params(p)
  p%h2o2m4 = h2o2m4; p%cp4 = cp4; p%cpp4 = cpp4; p%alphas = alphas; p%l = l; p%den0s = den0s
  p%csolid = C; p%beta_solid = beta; p%den_m = den_m
  p%h4 = h4; p%eps4 = eps4; p%sigma4 = sigma4
  ! 'OTE' (alpha_s in the energy only, from the start-up term_alfa) has no
  ! counterpart in the library: stop rather than run it as 'OT '
  if (core4 == 'OTE') stop 'hedft drop-in: core4 = OTE is not supported (OT, OP or OTC)'
  p%core_otc = merge(1, 0, core4 == 'OTC')
  p%lsolid = merge(1, 0, lsolid)
  p%npd = npd; p%icon = icon
  p%lambdah = Lambdah
  p%tmean = [txmean, tymean, tzmean]
  p%tsurf = [txsurf, tysurf, tzsurf]
  call hedft_check(hedft_create(g, p, HEDFT_MODE_TDDFT, 0_c_int32_t, hctx), 'create')
  n = int(nx, c_size_t) * ny * nz
  nh = int(nx / 2 + 1, c_size_t) * ny * nz
  ! the pairs potenimp.f90:32 and energy.f90:123 skip: i = j and 'null' selectors
  do j = 1, N_imp
    do i = 1, N_imp
      pair(i, j) = 0
      if (i /= j .and. trim(adjustl(selec_gs_k_k(i, j))) /= 'null') pair(i, j) = 1
    end do
  end do
  call hedft_check(hedft_set_quantum_impurities(hctx, int(N_imp, c_int32_t), uimp_k, uimp_k_k, pair, &
                                                8_c_size_t * nh, m_imp_u, mu4), 'set_quantum_impurities')
  call hedft_check(hedft_upload(hctx, HEDFT_DEN, den, 8_c_size_t * n), 'upload den')
  call hedft_check(hedft_upload(hctx, HEDFT_PSI, psi, 16_c_size_t * n), 'upload psi')
  do i = 1, N_imp
    call hedft_check(hedft_select_quantum_impurity(hctx, int(i - 1, c_int32_t)), 'select')
    call hedft_check(hedft_upload(hctx, HEDFT_PSIX, psix(:, :, :, i), 16_c_size_t * n), 'upload psix')
  end do
  call hedft_check(hedft_select_quantum_impurity(hctx, 0_c_int32_t), 'select')
end subroutine hedft_dropin_start

subroutine hedft_dropin_pull()
  use iso_c_binding
  use hedft_gpu
  use hedft_dropin_qstate
  use grid, only: nx, ny, nz
  use rho, only: den, psi
  use field, only: pot4
  use impur, only: psix, potx4, upotx
  use quantal_imp, only: N_imp
  implicit none
  integer(c_size_t) :: n
  integer :: i
  n = int(nx, c_size_t) * ny * nz
  call hedft_check(hedft_download(hctx, HEDFT_DEN, den, 8_c_size_t * n), 'download den')
  call hedft_check(hedft_download(hctx, HEDFT_PSI, psi, 16_c_size_t * n), 'download psi')
  call hedft_check(hedft_download(hctx, HEDFT_POT4, pot4, 8_c_size_t * n), 'download pot4')
  call hedft_check(hedft_download(hctx, HEDFT_POTX4, potx4, 8_c_size_t * n), 'download potx4')
  do i = 1, N_imp
    call hedft_check(hedft_select_quantum_impurity(hctx, int(i - 1, c_int32_t)), 'select')
    call hedft_check(hedft_download(hctx, HEDFT_PSIX, psix(:, :, :, i), 16_c_size_t * n), 'download psix')
    call hedft_check(hedft_download(hctx, HEDFT_UPOTX, upotx(:, :, :, i), 8_c_size_t * n), 'download upotx')
  end do
  call hedft_check(hedft_select_quantum_impurity(hctx, 0_c_int32_t), 'select')
end subroutine hedft_dropin_pull

SUBROUTINE STEPRK(deltat)
  use hedft_gpu
  use hedft_dropin_qstate
  implicit none
  real(8) :: deltat
  call hedft_check(hedft_steprk(hctx, deltat), 'STEPRK')
end SUBROUTINE STEPRK

SUBROUTINE STEPPC(deltat, errHe, errpsix)
  use iso_c_binding
  use hedft_gpu
  use hedft_dropin_qstate
  implicit none
  real(8) :: deltat, errHe, errpsix
  real(c_double) :: errv
  call hedft_check(hedft_steppc(hctx, deltat, errHe, errpsix, errv), 'STEPPC')
end SUBROUTINE STEPPC

subroutine potenimp()
  use hedft_gpu
  use hedft_dropin_qstate
  implicit none
  ! potx4 and upotx are computed by the poten() that follows (see the header)
  if (.not. c_associated(hctx)) stop 'potenimp: hedft_dropin_start not called'
end subroutine potenimp

subroutine poten()
  use hedft_gpu
  use hedft_dropin_qstate
  implicit none
  call hedft_check(hedft_poten(hctx), 'poten')
end subroutine poten

subroutine energy()
  use hedft_gpu
  use hedft_dropin_qstate
  use energies
  implicit none
  type(hedft_energies_t) :: e
  call hedft_check(hedft_energy(hctx, e), 'energy')
  ekin4 = e%ekin4; elj4 = e%elj4; ecor4 = e%ecor4; ealphas = e%ealphas; esolid = e%esolid
  etot4 = e%etot4; ekinx = e%ekinx; eHeX = e%eHeX; eimpu_impu = e%eimpu_impu; eimpu = e%eimpu
  etot = e%etot
end subroutine energy
