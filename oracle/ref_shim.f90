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


! readen.f90:74-84 records (mode 0): titols, the header, den -- list-directed
! READs executed by the Fortran runtime on the reference's own titols
subroutine ref_read_density(path, plen, cchar, hdr, den, n, ierr) bind(C, name="ref_read_density")
  integer(c_int), value :: plen, n
  character(kind=c_char), intent(in) :: path(plen)
  character(kind=c_char), value :: cchar
  real(c_double), intent(out) :: hdr(13)
  real(c_double), intent(out) :: den(n)
  integer(c_int), intent(out) :: ierr
  character(len=plen) :: fname
  character(len=1) :: cc
  integer :: i, isalto, nxp, nyp, nzp
  logical :: limp
  real(kind=8) :: xmaxp, ymaxp, zmaxp, hxp, hyp, hzp, ximp, yimp, zimp
  do i = 1, plen
    fname(i:i) = path(i)
  end do
  cc = cchar
  ierr = 0
  open(unit=91, file=fname, status='old', iostat=ierr)
  if (ierr /= 0) return
  call titols(91, cc, isalto)
  read(91, *, iostat=ierr) xmaxp, ymaxp, zmaxp, hxp, hyp, hzp, nxp, nyp, nzp, limp, ximp, yimp, zimp
  if (ierr == 0) read(91, *, iostat=ierr) den
  close(91)
  hdr = (/ xmaxp, ymaxp, zmaxp, hxp, hyp, hzp, dble(nxp), dble(nyp), dble(nzp), merge(1d0, 0d0, limp), &
           ximp, yimp, zimp /)
end subroutine

! readenc.f90:94-110 records (mode 2): titols, the header, N_imp, rimp, vimp, psi
subroutine ref_read_wavefunction(path, plen, cchar, hdr, nimp, rimp, vimp, maximp, psi, n, ierr) &
    bind(C, name="ref_read_wavefunction")
  integer(c_int), value :: plen, n, maximp
  character(kind=c_char), intent(in) :: path(plen)
  character(kind=c_char), value :: cchar
  real(c_double), intent(out) :: hdr(9)
  integer(c_int), intent(out) :: nimp, ierr
  real(c_double), intent(out) :: rimp(maximp, 3), vimp(maximp, 3)
  complex(c_double_complex), intent(out) :: psi(n)
  character(len=plen) :: fname
  character(len=1) :: cc
  integer :: i, isalto, nxp, nyp, nzp, ni
  real(kind=8) :: xmaxp, ymaxp, zmaxp, hxp, hyp, hzp
  real(kind=8), allocatable :: r(:,:), v(:,:)
  do i = 1, plen
    fname(i:i) = path(i)
  end do
  cc = cchar
  open(unit=92, file=fname, status='old', iostat=ierr)
  if (ierr /= 0) return
  call titols(92, cc, isalto)
  read(92, *, iostat=ierr) xmaxp, ymaxp, zmaxp, hxp, hyp, hzp, nxp, nyp, nzp
  if (ierr == 0) read(92, *, iostat=ierr) ni
  if (ierr == 0 .and. ni >= 1 .and. ni <= maximp) then
    allocate(r(ni, 3), v(ni, 3))
    read(92, *, iostat=ierr) r
    if (ierr == 0) read(92, *, iostat=ierr) v
    if (ierr == 0) then
      rimp(1:ni, :) = r
      vimp(1:ni, :) = v
    end if
  end if
  if (ierr == 0) read(92, *, iostat=ierr) psi
  close(92)
  nimp = ni
  hdr = (/ xmaxp, ymaxp, zmaxp, hxp, hyp, hzp, dble(nxp), dble(nyp), dble(nzp) /)
end subroutine


! evolo(deltat,mu4,mu4err,n4) (He_statics/evolo.f90:5-96, steepest descent with
! the vdt momentum terms and ironing.f90 smoothing) on module arrays set up
! here from the caller's psi, pot4, psi1, psi2; results copied back
subroutine ref_evolo(nx_, ny_, nz_, hx_, hy_, hz_, icon_, psi_, pot4_, psi1_, psi2_, deltat, vdt1, vdt2, iron_, &
                     n4r, mu4, mu4err, den_, hpsi_) bind(C, name="ref_evolo")
  use deriva, only: icon
  use field, only: hpsi, pot4, iron
  use grid, only: nx, ny, nz, hx, hy, hz, nxyz, dxyz
  use he4, only: lbulk
  use rho, only: den, psi, psinw, psi1, psi2
  use util1, only: vdt, nn
  use work1, only: sto1, sto2, sto3, sto4
  integer(c_int), value :: nx_, ny_, nz_, icon_, iron_
  real(c_double), value :: hx_, hy_, hz_, deltat, vdt1, vdt2, n4r
  real(c_double), intent(inout) :: psi_(nx_, ny_, nz_), psi1_(nx_, ny_, nz_), psi2_(nx_, ny_, nz_), mu4
  real(c_double), intent(in) :: pot4_(nx_, ny_, nz_)
  real(c_double), intent(out) :: mu4err, den_(nx_, ny_, nz_), hpsi_(nx_, ny_, nz_)
  integer(kind=4) :: n4
  real(kind=8) :: d8, m8, e8
  nx = nx_; ny = ny_; nz = nz_; hx = hx_; hy = hy_; hz = hz_
  nxyz = nx * ny * nz; dxyz = hx * hy * hz
  nn = (/ nx, ny, nz /)
  icon = icon_; iron = iron_; vdt = (/ vdt1, vdt2 /); lbulk = .false.
  if (allocated(psi)) deallocate(psi, psinw, psi1, psi2, den, hpsi, pot4, sto1, sto2, sto3, sto4)
  allocate(psi(nx,ny,nz), psinw(nx,ny,nz), psi1(nx,ny,nz), psi2(nx,ny,nz), den(nx,ny,nz), hpsi(nx,ny,nz), &
           pot4(nx,ny,nz), sto1(nx,ny,nz), sto2(nx,ny,nz), sto3(nx,ny,nz), sto4(nx,ny,nz))
  psi = psi_; psi1 = psi1_; psi2 = psi2_; pot4 = pot4_; den = psi**2
  n4 = nint(n4r)
  d8 = deltat; m8 = mu4
  call evolo(d8, m8, e8, n4)
  mu4 = m8; mu4err = e8
  psi_ = psi; psi1_ = psi1; psi2_ = psi2; den_ = den; hpsi_ = hpsi
end subroutine

end module ref_shim
