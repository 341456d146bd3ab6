! ref_shim.f90 -- C-callable bindings (bind(C)) to routines of the reference
! compiled from their own sources into oracle/_ref (test infrastructure only).
! It adds no arithmetic: each entry forwards to the reference routine named.
module ref_shim
use iso_c_binding
implicit none
contains

! Init_deriv_p(npd, ndmax, nthreads): init_deriv_parallel.f90:51-83
subroutine ref_init_deriv(k, kmax, nthr) bind(C, name="ref_init_deriv")
  integer(c_int), value :: k, kmax, nthr
  integer(kind=4) :: k4, kmax4, n4
  k4 = k; kmax4 = kmax; n4 = nthr
  call Init_deriv_p(k4, kmax4, n4)
end subroutine

! Deriv3D_p (real) / Deriv3D_pc (complex): init_deriv_parallel.f90:373-648, 933-1209
subroutine ref_deriv3d(id, nx, ny, nz, h, iv, f, df, icon, iscomplex) bind(C, name="ref_deriv3d")
  use Para_DerivnD
  integer(c_int), value :: id, nx, ny, nz, iv, icon, iscomplex
  real(c_double), value :: h
  type(c_ptr), value :: f, df
  real(kind=8), pointer :: fr(:,:,:), dfr(:,:,:)
  complex(kind=8), pointer :: fc(:,:,:), dfc(:,:,:)
  integer(kind=4) :: nn(3), id4, iv4, icon4
  real(kind=8) :: h8
  nn = (/ nx, ny, nz /); id4 = id; iv4 = iv; icon4 = icon; h8 = h
  if (iscomplex /= 0) then
    call c_f_pointer(f, fc, nn); call c_f_pointer(df, dfc, nn)
    call derivnD(id4, nn, h8, iv4, fc, dfc, icon4)
  else
    call c_f_pointer(f, fr, nn); call c_f_pointer(df, dfr, nn)
    call derivnD(id4, nn, h8, iv4, fr, dfr, icon4)
  end if
end subroutine

! fforma(core,b,eps,sigma,hc,nx,ny,nz,pmod,fvlj): fforma.f90:11-61
subroutine ref_fforma(nx, ny, nz, pmod, fvlj, b, eps, sigma, hc) bind(C, name="ref_fforma")
  integer(c_int), value :: nx, ny, nz
  real(c_double), value :: b, eps, sigma, hc
  real(c_double) :: pmod(nx/2+1, ny, nz), fvlj(nx/2+1, ny, nz)
  integer(kind=4) :: nx4, ny4, nz4
  character(len=2) :: core
  nx4 = nx; ny4 = ny; nz4 = nz; core = 'OT'
  call fforma(core, b, eps, sigma, hc, nx4, ny4, nz4, pmod, fvlj)
end subroutine

! initcg(h,wcgk): initcg.f90:32-49 (reads nx,ny,nz,pmod,twopi from modules)
subroutine ref_initcg(nxa, nya, nza, pm, wcgk, h) bind(C, name="ref_initcg")
  use grid, only: nx, ny, nz
  use gridk, only: pmod
  use util1, only: twopi
  integer(c_int), value :: nxa, nya, nza
  real(c_double) :: pm(nxa/2+1, nya, nza), wcgk(nxa/2+1, nya, nza)
  real(c_double), value :: h
  real(kind=8) :: h8
  nx = nxa; ny = nya; nz = nza
  twopi = 8.0d0*datan(1.0d0)
  if (allocated(pmod)) deallocate(pmod)
  allocate(pmod(nx/2+1, ny, nz))
  pmod = pm
  h8 = h
  call initcg(h8, wcgk)
  deallocate(pmod)
end subroutine

! s13adf(x,ifail): s13adf.f90
function ref_si(x) bind(C, name="ref_si") result(r)
  real(c_double), value :: x
  real(c_double) :: r
  real(kind=8) :: x8
  integer(kind=4) :: ifail
  interface
    double precision function s13adf(x, ifail)
      real(kind=8) :: x
      integer(kind=4) :: ifail
    end function
  end interface
  x8 = x; ifail = 0
  r = s13adf(x8, ifail)
end function

end module ref_shim
