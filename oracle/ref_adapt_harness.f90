! Test-infrastructure harness around the reference's AMR routines: the
! decisions checkconv and cal_conv (src/adapt.f:1260-1332, 1122-1192), the
! interpolations interpol and calh2 (1662-1824, 303-373) and the refinement
! hierarchy (init_adap, deraffiner, adapmaille, protridouble, elimpptri,
! inter_only_new; 380-1258).
!
! This file is ours; the routines it calls are compiled from the reference
! sources where they lie (oracle/Makefile, target ref).  The library it goes
! into, oracle/_ref/libspoc_ref_adapt.so, is the checker for the C
! restatement in oracle/spoc_oracle.c and writes the golden fixtures the
! product's host code is tested against; nothing in the product links or
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

  ! The refinement hierarchy (COMMON blocks of src/adapt.f, sized by
  ! include/maille.inc:6-7: nndim = 800, nedim = 1600).
  subroutine ref_adap_init(nn, ne, xgr, ygr, in, tseg, itm) bind(C, name="ref_adap_init")
    integer(c_int), value :: nn, ne
    real(c_double) :: xgr(*), ygr(*)
    integer(c_int) :: in(3, *), tseg(2, *), itm(3, *)
    integer nnl, nel
    nnl = nn; nel = ne
    call init_adap(nnl, nel, xgr, ygr, in, tseg, itm)
  end subroutine

  ! One adaptation as global_refinment runs it between cal_conv and newinter
  ! (src/adapt.f:89-119, with conv true): the per-element flags onto the
  ! hierarchy, deraffiner, adapmaille, protridouble, elimpptri.  fderaff is
  ! handed to deraffiner as global_refinment hands it (read by hierarchy
  ! number); here its entries past ne are .false. (in the reference they are
  ! whatever the local array held).  Outputs: the new mesh and tabsol.
  subroutine ref_adap_cycle(nn, ne, fconv, fderaff, nnnew, nenew, xgrnew, ygrnew, innew, tabs) &
      bind(C, name="ref_adap_cycle")
    integer, parameter :: nndim = 800, nedim = 2 * nndim
    integer(c_int), value :: nn, ne
    integer(c_int) :: fconv(*), fderaff(*), nnnew, nenew, innew(3, *), tabs(*)
    real(c_double) :: xgrnew(*), ygrnew(*)
    integer nn_ab, ne_ab, ndoub
    integer in_ab(3, 2 * nedim), tlevel(2 * nedim), torigi(2 * nedim), tabsol(nedim), &
            tmt(3, 2 * nedim), tenfan(4, 2 * nedim)
    double precision xgr_ab(2 * nndim), ygr_ab(2 * nndim)
    logical tdoubl(2 * nedim), tutili(2 * nedim)
    common /adap_const/ nn_ab, ne_ab, ndoub
    common /adap_int/ tlevel, torigi, tabsol, in_ab, tmt, tenfan
    common /adap_vec/ xgr_ab, ygr_ab
    common /adap_log/ tdoubl, tutili
    logical, allocatable :: fconv_ab(:), fder(:)
    integer i, tk, tk1, nnl, nel, nnn, nen, nemaxnew
    allocate(fconv_ab(2 * nedim), fder(2 * nedim))
    fconv_ab = .false.
    fder = .false.
    do i = 1, ne
      tk = tabsol(i)
      if (tk .gt. ne_ab) then
        tk1 = torigi(tk)
        tutili(tk1) = .true.
      else
        tk1 = tk
      end if
      if (fconv(i) /= 0) fconv_ab(tk1) = .true.
      fder(i) = fderaff(i) /= 0
    end do
    nnl = nn; nel = ne
    call deraffiner(nnl, nel, fconv_ab, fder)
    call adapmaille(fconv_ab, nnn, nen, nemaxnew)
    call protridouble(nnn, nen, nemaxnew)
    call elimpptri(nnn, nen, xgrnew, ygrnew, innew, nemaxnew)
    nnnew = nnn
    nenew = nen
    do i = 1, nen
      tabs(i) = tabsol(i)
    end do
    deallocate(fconv_ab, fder)
  end subroutine

  ! inter_only_new (src/adapt.f:1233-1258).  innew needs a column past
  ! nenew (the routine can read one): the caller passes zeros there.
  subroutine ref_inter_only_new(ne, nenew, in, innew, tinter) bind(C, name="ref_inter_only_new")
    integer(c_int), value :: ne, nenew
    integer(c_int) :: in(3, *), innew(3, *), tinter(*)
    integer nel, nen
    nel = ne; nen = nenew
    call inter_only_new(nel, nen, in, innew, tinter)
  end subroutine

  ! A smooth test field: tests/test_interp.py evaluates the same expression.
  double precision function harness_ff(x, y)
    double precision x, y
    harness_ff = x * 1.0d-6 * y * 2.0d-6 + 3.0d0 * x * 1.0d-6 - y * y * 1.0d-12 + 7.0d0
  end function
end module
