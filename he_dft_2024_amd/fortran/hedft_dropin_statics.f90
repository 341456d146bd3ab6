!> Drop-in replacements for the imaginary-time hot path of BCN-4He-DFT
!> (He_statics): STEPRKR, STEPPCR, STEPRKRL, STEPPCRL, evolo, poten, derden,
!> term_alfa and energy keep their names and signatures and run on the GPU
!> through module hedft_gpu.  The module arrays stay on the device between
!> calls.  The statics poten reads the fden, dencg, denalf and grad(den) that the
!> driver prepared after the previous step (BCN4HeDFT.f90:1125-1149), so that
!> host-side preparation is replaced as a whole:
!>   call hedft_dropin_statics_start(n4, mu4)
!> after the density has been read (in place of BCN4HeDFT.f90:724-874: the
!> normalisation to n4, psi = sqrt(den), fden, dencg, denalf, derden,
!> term_alfa, poten and the first mu4),
!>   call hedft_dropin_statics_refresh()
!> in place of the driver's derden / fftfw / dencg / denalf / poten lines after
!> each step (BCN4HeDFT.f90:1125-1160),
!>   call hedft_dropin_statics_log_start()
!> in place of psil = Log(Abs(psi)) (BCN4HeDFT.f90:1081-1083), and
!>   call hedft_dropin_statics_pull()
!> before each output point to refresh den, psi and pot4 on the host.
!> Compiled against the reference's own He_statics/modules.f90 (Makefile
!> target dropin).
module hedft_dropin_statics_state
  use iso_c_binding
  implicit none
  type(c_ptr), save :: sctx = c_null_ptr
end module hedft_dropin_statics_state

subroutine hedft_dropin_statics_start(n4, mu4)
  use iso_c_binding
  use hedft_gpu
  use hedft_dropin_statics_state
  use grid, only: nx, ny, nz, xmax, ymax, zmax
  use he4, only: h2o2m4, cp4, cpp4, alphas, l, den0s, lsolid, C, beta, den_m
  use lenard4, only: h4, eps4, sigma4, core4
  use deriva, only: npd, icon
  use rho, only: den, psimin
  use rkpc, only: Dtmin, Sdmu, Dmumax
  implicit none
  integer(4), intent(in) :: n4
  real(8), intent(out) :: mu4
  type(hedft_grid_t) :: g
  type(hedft_params_t) :: p
  integer(c_size_t) :: n

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
  p%dtmin = Dtmin; p%sdmu = Sdmu; p%dmumax = Dmumax; p%psimin = psimin
  call hedft_check(hedft_create(g, p, HEDFT_MODE_STATICS, 0_c_int32_t, sctx), 'create')
  n = int(nx, c_size_t) * ny * nz
  call hedft_check(hedft_upload(sctx, HEDFT_DEN, den, 8_c_size_t * n), 'upload den')
  call hedft_check(hedft_statics_init(sctx, dble(n4), mu4), 'statics_init')
end subroutine hedft_dropin_statics_start

subroutine hedft_dropin_statics_refresh()
  use hedft_gpu
  use hedft_dropin_statics_state
  implicit none
  call hedft_check(hedft_statics_refresh(sctx), 'statics_refresh')
end subroutine hedft_dropin_statics_refresh

subroutine hedft_dropin_statics_log_start()
  use hedft_gpu
  use hedft_dropin_statics_state
  implicit none
  call hedft_check(hedft_statics_log_start(sctx), 'statics_log_start')
end subroutine hedft_dropin_statics_log_start

subroutine hedft_dropin_statics_pull()
  use iso_c_binding
  use hedft_gpu
  use hedft_dropin_statics_state
  use grid, only: nx, ny, nz
  use rho, only: den, psi
  use field, only: pot4
  implicit none
  integer(c_size_t) :: n
  n = int(nx, c_size_t) * ny * nz
  call hedft_check(hedft_download(sctx, HEDFT_DEN, den, 8_c_size_t * n), 'download den')
  call hedft_check(hedft_download(sctx, HEDFT_PSI, psi, 8_c_size_t * n), 'download psi')
  call hedft_check(hedft_download(sctx, HEDFT_POT4, pot4, 8_c_size_t * n), 'download pot4')
end subroutine hedft_dropin_statics_pull

SUBROUTINE STEPRKR(deltat, mu4, mu4err, n4)
  use hedft_gpu
  use hedft_dropin_statics_state
  implicit none
  real(8) :: deltat, mu4, mu4err
  integer(4) :: n4
  call hedft_check(hedft_steprkr(sctx, deltat, mu4, mu4err, dble(n4)), 'STEPRKR')
end SUBROUTINE STEPRKR

SUBROUTINE STEPPCR(deltat, mu4, mu4err, n4)
  use hedft_gpu
  use hedft_dropin_statics_state
  use rkpc, only: ErrPC
  implicit none
  real(8) :: deltat, mu4, mu4err
  integer(4) :: n4
  call hedft_check(hedft_steppcr(sctx, deltat, mu4, mu4err, dble(n4), ErrPC), 'STEPPCR')
end SUBROUTINE STEPPCR

SUBROUTINE STEPRKRL(deltat, mu4, mu4err, n4)
  use hedft_gpu
  use hedft_dropin_statics_state
  implicit none
  real(8) :: deltat, mu4, mu4err
  integer(4) :: n4
  call hedft_check(hedft_steprkrl(sctx, deltat, mu4, mu4err, dble(n4)), 'STEPRKRL')
end SUBROUTINE STEPRKRL

SUBROUTINE STEPPCRL(deltat, mu4, mu4err, n4)
  use hedft_gpu
  use hedft_dropin_statics_state
  use rkpc, only: ErrPC
  implicit none
  real(8) :: deltat, mu4, mu4err
  integer(4) :: n4
  call hedft_check(hedft_steppcrl(sctx, deltat, mu4, mu4err, dble(n4), ErrPC), 'STEPPCRL')
end SUBROUTINE STEPPCRL

subroutine evolo(deltat, mu4, mu4err, n4)
  use iso_c_binding
  use hedft_gpu
  use hedft_dropin_statics_state
  use util1, only: vdt
  use field, only: iron
  implicit none
  real(8) :: deltat, mu4, mu4err
  integer(4) :: n4
  call hedft_check(hedft_evolo(sctx, deltat, vdt(1), vdt(2), int(iron, c_int32_t), dble(n4), mu4, mu4err), &
                   'evolo')
end subroutine evolo

subroutine poten()
  use hedft_gpu
  use hedft_dropin_statics_state
  implicit none
  call hedft_check(hedft_poten(sctx), 'poten')
end subroutine poten

subroutine derden()
  use hedft_gpu
  use hedft_dropin_statics_state
  implicit none
  call hedft_check(hedft_derden(sctx), 'derden')
end subroutine derden

subroutine term_alfa()
  use hedft_gpu
  use hedft_dropin_statics_state
  implicit none
  call hedft_check(hedft_term_alfa(sctx), 'term_alfa')
end subroutine term_alfa

subroutine energy()
  use hedft_gpu
  use hedft_dropin_statics_state
  use energies
  implicit none
  type(hedft_energies_t) :: e
  call hedft_check(hedft_energy(sctx, e), 'energy')
  ekin4 = e%ekin4; elj4 = e%elj4; ecor4 = e%ecor4; ealphas = e%ealphas; esolid = e%esolid
  etot4 = e%etot4; ekinx = e%ekinx; eimpu = e%eimpu; etot = e%etot
end subroutine energy
