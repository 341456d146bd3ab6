!> Drop-in replacements for the time-step subroutines of the quantum-impurity
!> program of BCN-4He-DFT (Dynamics_quantum_impurities): STEPRK(deltat),
!> STEPPC(deltat,errHe,errpsix), poten, potenimp and energy keep their names and
!> signatures and run on the GPU through module hedft_gpu, for one impurity
!> (N_imp = 1).  The module arrays (den, psi, psix, pot4, potx4, upotx, ...) stay
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
  use quantal_imp, only: N_imp, uimp_k, m_imp_u
  implicit none
  real(8), intent(in) :: Lambdah, txmean, tymean, tzmean, txsurf, tysurf, tzsurf
  type(hedft_grid_t) :: g
  type(hedft_params_t) :: p
  integer(c_size_t) :: n, nh

  if (N_imp /= 1) stop 'hedft_dropin_start: one quantum impurity (N_imp = 1)'
  g%nx = nx; g%ny = ny; g%nz = nz; g%pad_ = 0
  g%xmax = xmax; g%ymax = ymax; g%zmax = zmax
  call hedft_default_params(p)
  p%h2o2m4 = h2o2m4; p%cp4 = cp4; p%cpp4 = cpp4; p%alphas = alphas; p%l = l; p%den0s = den0s
  p%csolid = C; p%beta_solid = beta; p%den_m = den_m
  p%h4 = h4; p%eps4 = eps4; p%sigma4 = sigma4
  p%core_otc = merge(1, 0, core4 == 'OTC')
  p%lsolid = merge(1, 0, lsolid)
  p%npd = npd; p%icon = icon
  p%lambdah = Lambdah
  p%tmean = [txmean, tymean, tzmean]
  p%tsurf = [txsurf, tysurf, tzsurf]
  call hedft_check(hedft_create(g, p, HEDFT_MODE_TDDFT, 0_c_int32_t, hctx), 'create')
  n = int(nx, c_size_t) * ny * nz
  nh = int(nx / 2 + 1, c_size_t) * ny * nz
  call hedft_check(hedft_set_quantum_impurity(hctx, uimp_k(:, :, :, 1), 8_c_size_t * nh, m_imp_u(1), mu4), &
                   'set_quantum_impurity')
  call hedft_check(hedft_upload(hctx, HEDFT_DEN, den, 8_c_size_t * n), 'upload den')
  call hedft_check(hedft_upload(hctx, HEDFT_PSI, psi, 16_c_size_t * n), 'upload psi')
  call hedft_check(hedft_upload(hctx, HEDFT_PSIX, psix(:, :, :, 1), 16_c_size_t * n), 'upload psix')
end subroutine hedft_dropin_start

subroutine hedft_dropin_pull()
  use iso_c_binding
  use hedft_gpu
  use hedft_dropin_qstate
  use grid, only: nx, ny, nz
  use rho, only: den, psi
  use field, only: pot4
  use impur, only: psix, potx4, upotx
  implicit none
  integer(c_size_t) :: n
  n = int(nx, c_size_t) * ny * nz
  call hedft_check(hedft_download(hctx, HEDFT_DEN, den, 8_c_size_t * n), 'download den')
  call hedft_check(hedft_download(hctx, HEDFT_PSI, psi, 16_c_size_t * n), 'download psi')
  call hedft_check(hedft_download(hctx, HEDFT_POT4, pot4, 8_c_size_t * n), 'download pot4')
  call hedft_check(hedft_download(hctx, HEDFT_PSIX, psix(:, :, :, 1), 16_c_size_t * n), 'download psix')
  call hedft_check(hedft_download(hctx, HEDFT_POTX4, potx4, 8_c_size_t * n), 'download potx4')
  call hedft_check(hedft_download(hctx, HEDFT_UPOTX, upotx(:, :, :, 1), 8_c_size_t * n), 'download upotx')
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
