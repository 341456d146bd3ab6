!> Drop-in replacements for the TDDFT hot-path subroutines of BCN-4He-DFT
!> (Dynamics_TDDFT): STEPRK, STEPPC, poten, potenimp, forceimp, derden,
!> term_alfa and energy keep their names and signatures and run on the GPU
!> through module hedft_gpu.  The module arrays (den, psi, pot4, ...) stay on
!> the device between calls; the small impurity state (rimp, vimp, aimp, F)
!> is mirrored every step because the driver itself forms aimp = F/m
!> (DFT4He3d.f90:692-696).  Two calls are added to the driver:
!>   call hedft_dropin_start(Lambdah, txmean, tymean, tzmean, txsurf, tysurf, tzsurf)
!> after its set-up (before the time loop, DFT4He3d.f90:670), and
!>   call hedft_dropin_pull()
!> before each output point (pener/pdenpar/pcurr) to refresh den, psi, pot4.
!> Compiled against the reference's own module files (Makefile target dropin).
module hedft_dropin_state
  use iso_c_binding
  implicit none
  type(c_ptr), save :: hctx = c_null_ptr
end module hedft_dropin_state

subroutine hedft_dropin_start(Lambdah, txmean, tymean, tzmean, txsurf, tysurf, tzsurf)
  use iso_c_binding
  use hedft_gpu
  use hedft_dropin_state
  use grid, only: nx, ny, nz, xmax, ymax, zmax
  use he4, only: h2o2m4, cp4, cpp4, alphas, l, den0s, lsolid, C, beta, den_m
  use lenard4, only: h4, eps4, sigma4, core4
  use deriva, only: npd, icon
  use rho, only: den, psi
  use classicimp, only: N_imp, rimp, vimp, aimp, m_imp, selec_gs_k_k, r_cutoff_gs_k_k, umax_gs_k_k, &
                        drselec_gs_k_k, drr_cutoff_gs_k_k, drumax_gs_k_k
  use interpol, only: DelInter, potion, rmaxinterpol, npot
  implicit none
  real(8), intent(in) :: Lambdah, txmean, tymean, tzmean, txsurf, tysurf, tzsurf
  type(hedft_grid_t) :: g
  type(hedft_params_t) :: p
  real(c_double), allocatable :: tab(:,:), vp(:,:), dvp(:,:)
  integer(c_size_t) :: n
  integer :: k, m, i, ip, npair
  real(8) :: r, delpair, Select_pot

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
  call hedft_check(hedft_upload(hctx, HEDFT_DEN, den, 8_c_size_t * n), 'upload den')
  call hedft_check(hedft_upload(hctx, HEDFT_PSI, psi, 16_c_size_t * n), 'upload psi')
  ! potion(k,i) -> rows contiguous per impurity
  allocate(tab(npot, N_imp))
  tab = transpose(potion)
  call hedft_check(hedft_set_impurities(hctx, int(N_imp, c_int32_t), int(npot, c_int32_t), DelInter, &
                   rmaxinterpol, tab, m_imp), 'set_impurities')
  ! impurity-impurity terms (potenimp.f90:40-58, energy.f90:168-177): the pair
  ! potentials the driver selected, tabulated 20 times finer than potion
  if (N_imp >= 2) then
    npair = 20 * (npot - 1) + 1
    delpair = rmaxinterpol / dble(npair - 1)
    allocate(vp(npair, N_imp * (N_imp - 1) / 2), dvp(npair, N_imp * (N_imp - 1) / 2))
    ip = 0
    do k = 1, N_imp
      do m = k + 1, N_imp
        ip = ip + 1
        do i = 1, npair
          r = dble(i - 1) * delpair
          vp(i, ip) = Select_pot(selec_gs_k_k(k, m), r, r_cutoff_gs_k_k(k, m), umax_gs_k_k(k, m))
          dvp(i, ip) = Select_pot(drselec_gs_k_k(k, m), r, drr_cutoff_gs_k_k(k, m), drumax_gs_k_k(k, m))
        end do
      end do
    end do
    call hedft_check(hedft_set_impurity_pairs(hctx, int(npair, c_int32_t), delpair, vp, dvp), &
                     'set_impurity_pairs')
  end if
  call hedft_check(hedft_set_impurity_state(hctx, rimp, vimp, aimp), 'set_impurity_state')
end subroutine hedft_dropin_start

subroutine hedft_dropin_pull()
  use iso_c_binding
  use hedft_gpu
  use hedft_dropin_state
  use grid, only: nx, ny, nz
  use rho, only: den, psi
  use field, only: pot4
  implicit none
  integer(c_size_t) :: n
  n = int(nx, c_size_t) * ny * nz
  call hedft_check(hedft_download(hctx, HEDFT_DEN, den, 8_c_size_t * n), 'download den')
  call hedft_check(hedft_download(hctx, HEDFT_PSI, psi, 16_c_size_t * n), 'download psi')
  call hedft_check(hedft_download(hctx, HEDFT_POT4, pot4, 8_c_size_t * n), 'download pot4')
end subroutine hedft_dropin_pull

! the driver's aimp (formed on the host) goes to the device before each step,
! with rimp and vimp as forceimp last mirrored them (equal to the device's)
subroutine hedft_dropin_push_aimp()
  use hedft_gpu
  use hedft_dropin_state
  use classicimp, only: rimp, vimp, aimp
  implicit none
  call hedft_check(hedft_set_impurity_state(hctx, rimp, vimp, aimp), 'set aimp')
end subroutine hedft_dropin_push_aimp

SUBROUTINE STEPRK(deltat)
  use hedft_gpu
  use hedft_dropin_state
  implicit none
  real(8) :: deltat
  call hedft_dropin_push_aimp()
  call hedft_check(hedft_steprk(hctx, deltat), 'STEPRK')
end SUBROUTINE STEPRK

SUBROUTINE STEPPC(deltat, errHe, errimp, errvimp)
  use hedft_gpu
  use hedft_dropin_state
  implicit none
  real(8) :: deltat, errHe, errimp, errvimp
  call hedft_dropin_push_aimp()
  call hedft_check(hedft_steppc(hctx, deltat, errHe, errimp, errvimp), 'STEPPC')
end SUBROUTINE STEPPC

subroutine poten()
  use hedft_gpu
  use hedft_dropin_state
  implicit none
  call hedft_check(hedft_poten(hctx), 'poten')
end subroutine poten

subroutine potenimp()
  use hedft_gpu
  use hedft_dropin_state
  implicit none
  call hedft_check(hedft_potenimp(hctx), 'potenimp')
end subroutine potenimp

! forceimp(): F on the device, then rimp, vimp and F back into classicimp
! for the driver's aimp = F/m and its output
subroutine forceimp()
  use iso_c_binding
  use hedft_gpu
  use hedft_dropin_state
  use classicimp, only: N_imp, rimp, vimp, F
  use interpol, only: lstopimp
  implicit none
  integer(c_int32_t) :: ls
  real(c_double) :: adev(N_imp, 3)   ! the device's aimp: the driver forms its own
  call hedft_check(hedft_forceimp(hctx), 'forceimp')
  call hedft_check(hedft_get_impurity_state(hctx, rimp, vimp, adev, F, ls), 'get_impurity_state')
  lstopimp = (ls /= 0)
end subroutine forceimp

subroutine derden()
  use hedft_gpu
  use hedft_dropin_state
  implicit none
  call hedft_check(hedft_derden(hctx), 'derden')
end subroutine derden

subroutine term_alfa()
  use hedft_gpu
  use hedft_dropin_state
  implicit none
  call hedft_check(hedft_term_alfa(hctx), 'term_alfa')
end subroutine term_alfa

subroutine energy()
  use hedft_gpu
  use hedft_dropin_state
  use energies
  implicit none
  type(hedft_energies_t) :: e
  call hedft_check(hedft_energy(hctx, e), 'energy')
  ekin4 = e%ekin4; elj4 = e%elj4; ecor4 = e%ecor4; ealphas = e%ealphas; esolid = e%esolid
  etot4 = e%etot4; ekinx = e%ekinx; eHeX = e%eHeX; eimpu_impu = e%eimpu_impu; eimpu = e%eimpu
  etot = e%etot
end subroutine energy
