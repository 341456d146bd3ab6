! refq_shim.f90 -- TEST INFRASTRUCTURE ONLY.
! C-callable (bind(C)) wrappers around two routines of the reference's
! quantum-impurity code, compiled from its own sources into oracle/_ref by
! oracle/Makefile: Select_Pot (Dynamics_quantum_impurities/V_impur.f90:90) and
! FT_V_spline (Dynamics_quantum_impurities/FT_V_spline.f90:14), the radial
! Fourier transform potenimpini.f90:25-26 fills uimp_k with.  Tests compare the
! package's own transform of a radial potential against them.

function refq_select_pot(sel, nsel, r, rc, umax) bind(C, name="refq_select_pot") result(v)
  use iso_c_binding
  implicit none
  character(kind=c_char), intent(in) :: sel(*)
  integer(c_int), value :: nsel
  real(c_double), value :: r, rc, umax
  real(c_double) :: v
  character(len=80) :: selec
  real(8), external :: Select_Pot
  integer :: i
  selec = ' '
  do i = 1, nsel
    selec(i:i) = sel(i)
  end do
  v = Select_Pot(selec, r, rc, umax)
end function refq_select_pot

! parabolic /= 0 sets Lparabolic_fit_FT_V_spline (the namelist switch of
! DFT4He3d.f90); a change of the switch drops the cached p-table.
function refq_ft_v_spline(p, pmax, hq, sel, nsel, rc, fc, parabolic) bind(C, name="refq_ft_v_spline") result(v)
  use iso_c_binding
  use For_FT_V_spline
  implicit none
  real(c_double), value :: p, pmax, hq, rc, fc
  character(kind=c_char), intent(in) :: sel(*)
  integer(c_int), value :: nsel, parabolic
  real(c_double) :: v
  character(len=80) :: selec
  real(8), external :: FT_V_spline
  logical :: lp
  integer :: i
  selec = ' '
  do i = 1, nsel
    selec(i:i) = sel(i)
  end do
  lp = parabolic /= 0
  if (lp .neqv. Lparabolic_fit_FT_V_spline) N_p = 0
  Lparabolic_fit_FT_V_spline = lp
  v = FT_V_spline(p, pmax, hq, selec, rc, fc)
end function refq_ft_v_spline
