! Test-infrastructure harness around the reference's AMR decision routines
! checkconv and cal_conv (src/adapt.f:1260-1332, 1122-1192).
!
! This file is ours; the routines it calls are compiled from the reference
! sources where they lie (oracle/Makefile, target ref).  The library it goes
! into, oracle/_ref/libspoc_ref_adapt.so, is the checker for the C
! restatement in oracle/spoc_oracle.c; nothing in the product links or
! loads it.
!
! Arrays follow the reference's layouts: modal arrays (45, ne) column-major
! (nnmod = 45, include/sp.inc:2), pico (9, 45, 3) (ngf_max = 9), tseg and
! ssit (2, nf) 1-based.  LOGICAL results are returned as 0/1 ints.
module spoc_ref_adapt_harness
  use iso_c_binding
  implicit none
  integer, parameter :: LDM = 45, LDF = 9
contains

  subroutine set_const(nc_, ng_, ngf_)
    integer(c_int), value :: nc_, ng_, ngf_
    integer nc, nm, ng, ngf
    common /const/ nc, nm, ng, ngf
    nc = nc_; ng = ng_; ngf = ngf_
    nm = (nc + 1) * (nc + 2) / 2
  end subroutine

  ! checkconv of one field: amax(ne) out.
  subroutine ref_checkconv(nc_, ng_, ngf_, ne, nf, a0, pico, tseg, ssit, amax) &
      bind(C, name="ref_checkconv")
    integer(c_int), value :: nc_, ng_, ngf_, ne, nf
    real(c_double) :: a0(LDM, *), pico(LDF, LDM, 3), amax(*)
    integer(c_int) :: tseg(2, *), ssit(2, *)
    call set_const(nc_, ng_, ngf_)
    call checkconv(nf, ne, a0, pico, tseg, ssit, amax)
  end subroutine

  ! cal_conv: per-element flags and the three global decisions.  (w0, z0, hb
  ! are arguments of cal_conv that it does not read; u0 stands in for them.)
  subroutine ref_cal_conv(nc_, ng_, ngf_, ne, nf, pico, tseg, ssit, u0, v0, h0, fac1, fac2, &
      fac3, fconv, fderaff, conv, deraff, timedir) bind(C, name="ref_cal_conv")
    integer(c_int), value :: nc_, ng_, ngf_, ne, nf
    real(c_double) :: pico(LDF, LDM, 3), u0(LDM, *), v0(LDM, *), h0(LDM, *)
    integer(c_int) :: tseg(2, *), ssit(2, *)
    real(c_double), value :: fac1, fac2, fac3
    integer(c_int) :: fconv(*), fderaff(*), conv, deraff, timedir
    logical, allocatable :: lfconv(:), lfderaff(:)
    logical lconv, lderaff, ltimedir
    double precision f1, f2, f3
    integer i, nn, nc
    call set_const(nc_, ng_, ngf_)
    allocate(lfconv(ne), lfderaff(ne))
    nn = 0
    nc = nc_
    f1 = fac1; f2 = fac2; f3 = fac3
    call cal_conv(nn, ne, nf, nc, pico, ssit, lconv, lfconv, lfderaff, ltimedir, tseg, &
                  u0, v0, h0, u0, u0, u0, lderaff, f1, f2, f3)
    do i = 1, ne
      fconv(i) = merge(1, 0, lfconv(i))
      fderaff(i) = merge(1, 0, lfderaff(i))
    end do
    conv = merge(1, 0, lconv)
    deraff = merge(1, 0, lderaff)
    timedir = merge(1, 0, ltimedir)
    deallocate(lfconv, lfderaff)
  end subroutine

  ! interpol (src/adapt.f:1662-1739) of one field; the COMMON tables must be
  ! installed first (ref_setup of the spoc_ref_harness module, linked into
  ! the same library).
  subroutine ref_interpol(ne, nn, xgr, ygr, in, tseg, pseg, itm, tinter, u, unew, nenew, nnnew, &
      xgrnew, ygrnew, innew) bind(C, name="ref_interpol")
    integer(c_int), value :: ne, nn, nenew, nnnew
    real(c_double) :: xgr(*), ygr(*), u(LDM, *), unew(LDM, *), xgrnew(*), ygrnew(*)
    integer(c_int) :: in(3, *), tseg(2, *), pseg(2, *), itm(3, *), tinter(*), innew(3, *)
    integer nen, nnn, nel, nnl
    nel = ne; nnl = nn; nen = nenew; nnn = nnnew
    call interpol(nel, nnl, xgr, ygr, in, tseg, pseg, itm, tinter, u, unew, nen, nnn, xgrnew, &
                  ygrnew, innew)
  end subroutine

  ! calh2 (src/adapt.f:303-373) of the test function harness_ff below.
  subroutine ref_calh2(ne, xgr, ygr, in, tinter, h0, h1) bind(C, name="ref_calh2")
    integer(c_int), value :: ne
    real(c_double) :: xgr(*), ygr(*), h0(LDM, *), h1(LDM, *)
    integer(c_int) :: in(3, *), tinter(*)
    integer nel
    nel = ne
    call calh2(harness_ff, nel, h0, h1, in, xgr, ygr, tinter)
  end subroutine

  ! A smooth test field: tests/test_interp.py evaluates the same expression.
  double precision function harness_ff(x, y)
    double precision x, y
    harness_ff = x * 1.0d-6 * y * 2.0d-6 + 3.0d0 * x * 1.0d-6 - y * y * 1.0d-12 + 7.0d0
  end function
end module
