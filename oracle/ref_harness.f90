! Test-infrastructure harness around the reference SPOC Fortran routines.
!
! This file is ours; the routines it calls are compiled from the reference
! sources where they lie (see oracle/Makefile).  The resulting shared
! libraries under oracle/_ref/ are the *checker* for the C restatement in
! oracle/spoc_oracle.c and a CPU timing baseline.  Nothing in the product
! (spoc_amd/) links or loads them.
!
! Every entry point is bind(C) so pytest / the golden-vector script can call
! it through ctypes.  Arrays follow the reference's own layout: modal arrays
! are (45, ne) column-major (leading dimension nnmod = 45, include/sp.inc:2),
! face arrays (9, nf) (ngf_max = 9, include/sp.inc:4).
!
! COMMON blocks are declared exactly as in src/shwwnl.f:61-71 so that the
! reference's cal_var sees the tables this harness installs.
module spoc_ref_harness
  use iso_c_binding
  implicit none
  integer, parameter :: LDM = 45, LDF = 9, LDG = 33
  integer, save :: cipi(2, 0:LDM)
contains

  ! Fill /const/ and the quadrature COMMONs by calling the reference's own
  ! file reader (src/readgauss.f:1-65).  The process cwd must hold
  ! include/tri2_<ng>.dat and include/legen<ngf>.dat.
  subroutine ref_read_gauss(nc_, ng_, ngf_, w, x, y, wf, tf) bind(C, name="ref_read_gauss")
    integer(c_int), value :: nc_, ng_, ngf_
    real(c_double) :: w(*), x(*), y(*), wf(*), tf(*)
    integer nc, nm, ng, ngf
    common /const/ nc, nm, ng, ngf
    double precision wi(LDG), xi(LDG), yi(LDG), wif(LDF), tif(LDF)
    common /gauss/ wi, xi, yi
    common /gaussf/ wif, tif
    integer i
    nc = nc_; ng = ng_; ngf = ngf_
    call read_gauss(ng, xi, yi, wi, ngf, wif, tif, nc, nm)
    do i = 1, ng
      w(i) = wi(i); x(i) = xi(i); y(i) = yi(i)
    end do
    do i = 1, ngf
      wf(i) = wif(i); tf(i) = tif(i)
    end do
  end subroutine

  ! Install given quadrature data and build every operator table exactly as
  ! the driver does (src/shwwnl.f:127-156).  info1/info2 are DPOTRF's info
  ! for amm and amml.
  subroutine ref_setup(nc_, ng_, ngf_, w, x, y, wf, tf, info1, info2) bind(C, name="ref_setup")
    integer(c_int), value :: nc_, ng_, ngf_
    real(c_double) :: w(*), x(*), y(*), wf(*), tf(*)
    integer(c_int) :: info1, info2
    integer nc, nm, ng, ngf
    common /const/ nc, nm, ng, ngf
    double precision wi(LDG), xi(LDG), yi(LDG), wif(LDF), tif(LDF)
    common /gauss/ wi, xi, yi
    common /gaussf/ wif, tif
    double precision pi2(LDG, LDM), pico(LDF, LDM, 3)
    common /gauss2/ pi2
    common /gauss3/ pico
    double precision amm(LDM,LDM), amml(LDM,LDM), agx(LDM,LDM), agy(LDM,LDM), &
         ammx(LDM,LDM), ammy(LDM,LDM), ammo(LDM,LDM)
    common /matrices/ amm, amml, agx, agy, ammx, ammy, ammo
    double precision, external :: f2, dfx, dfy, fx, fy
    integer i, j, info
    nc = nc_; ng = ng_; ngf = ngf_
    nm = (nc + 1) * (nc + 2) / 2
    do i = 1, ng
      wi(i) = w(i); xi(i) = x(i); yi(i) = y(i)
    end do
    do i = 1, ngf
      wif(i) = wf(i); tif(i) = tf(i)
    end do
    call cal_p2(xi, yi, pi2)
    call cal_pico(pico, tif)
    call calip(nc, cipi)
    call cal_mat(LDM, ng, nc, amm, f2, xi, yi, wi)
    call cal_mat(LDM, ng, nc, ammx, fx, xi, yi, wi)
    call cal_mat(LDM, ng, nc, ammy, fy, xi, yi, wi)
    call cal_mat(LDM, ng, nc, agx, dfx, xi, yi, wi)
    call cal_mat(LDM, ng, nc, agy, dfy, xi, yi, wi)
    call cal_matl(LDM, nc, amm, amml)
    do j = 1, nm
      do i = 1, nm
        ammo(i, j) = amm(i, j)
      end do
    end do
    call dpotrf('U', nm, amm, LDM, info)
    info1 = info
    call dpotrf('U', nm - nc - 1, amml, LDM, info)
    info2 = info
  end subroutine

  ! Copy the installed tables out (full leading dimensions).
  subroutine ref_get_tables(pi2o, picoo, ammo_, amml_, agxo, agyo, ammoo, cipo) bind(C, name="ref_get_tables")
    real(c_double) :: pi2o(LDG, LDM), picoo(LDF, LDM, 3), ammo_(LDM, LDM), amml_(LDM, LDM), &
         agxo(LDM, LDM), agyo(LDM, LDM), ammoo(LDM, LDM)
    integer(c_int) :: cipo(2, 0:LDM)
    double precision pi2(LDG, LDM), pico(LDF, LDM, 3)
    common /gauss2/ pi2
    common /gauss3/ pico
    double precision amm(LDM,LDM), amml(LDM,LDM), agx(LDM,LDM), agy(LDM,LDM), &
         ammx(LDM,LDM), ammy(LDM,LDM), ammo(LDM,LDM)
    common /matrices/ amm, amml, agx, agy, ammx, ammy, ammo
    pi2o = pi2; picoo = pico; ammo_ = amm; amml_ = amml
    agxo = agx; agyo = agy; ammoo = ammo; cipo = cipi
  end subroutine

  ! Mesh pointers and geometry (src/shwwnl.f:163-174): pointer_mesh,
  ! check_mesh, cal_rotnf, cal_its, calpre, coriolis.  Sizes: in(3,ne),
  ! x/y(nn), itm/its(3,ne), tseg/pseg/ssit(2,nfcap), norm(2,nfcap),
  ! vect(5,ne), normt(2,ne), bdl(ne), fc(45,ne) (zeroed here), nfcap >= 2400.
  subroutine ref_mesh(nn, ne, in, x, y, nc, f0, beta, nf, itm, its, tseg, pseg, ssit, &
       norm, vect, normt, bdl, fc, nbdl) bind(C, name="ref_mesh")
    integer(c_int), value :: nn, ne, nc
    real(c_double), value :: f0, beta
    integer(c_int) :: in(3, *), nf, itm(3, *), its(3, *), tseg(2, *), pseg(2, *), ssit(2, *), nbdl
    real(c_double) :: x(*), y(*), norm(2, *), vect(5, *), normt(2, *), fc(LDM, *)
    integer(c_int) :: bdl(*)
    integer, allocatable :: nmit(:), tmit(:), imit(:), pbdl(:), pint(:), sbd(:)
    logical, allocatable :: lbdl(:), nobd(:)
    integer nint, nebd, i
    allocate(nmit(8 * 800 + 8), tmit(8 * 800 + 64), imit(8 * 800 + 64), pbdl(ne), pint(ne), &
         sbd(4096), lbdl(ne), nobd(nn))
    call pointer_mesh(nn, ne, nf, in, itm, nmit, tmit, imit, pseg, tseg)
    call check_mesh(ne, nf, in, itm, tseg)
    call cal_rotnf(nf, x, y, norm, pseg)
    call cal_its(nf, itm, its, tseg, ssit)
    call calpre(nn, ne, nf, in, x, y, vect, itm, norm, normt, lbdl, nint, pint, nbdl, pbdl, &
         nobd, nebd, sbd, tseg, pseg)
    do i = 1, ne
      fc(:, i) = 0.d0
      bdl(i) = 0
      if (lbdl(i)) bdl(i) = 1
    end do
    call coriolis(nc, ne, in, x, y, fc, f0, beta)
  end subroutine

  ! One reference cal_var call (the link-time variant this library was built
  ! with).  bdl arrives as int32 and is converted to LOGICAL.
  subroutine ref_cal_var(nf, ne, u0, v0, h0, u1, v1, h1, tseg, ssit, itm, its, norm, bdl, &
       vect, normt, mu0, mb0, fc, beta, w0, z0, g0, hb, time) bind(C, name="ref_cal_var")
    integer(c_int), value :: nf, ne
    real(c_double) :: u0(LDM, *), v0(LDM, *), h0(LDM, *), u1(LDM, *), v1(LDM, *), h1(LDM, *)
    integer(c_int) :: tseg(2, *), ssit(2, *), itm(3, *), its(3, *), bdl(*)
    real(c_double) :: norm(2, *), vect(5, *), normt(2, *), fc(LDM, *), w0(LDM, *), z0(LDM, *), &
         hb(LDM, *)
    real(c_double), value :: mu0, mb0, beta, g0, time
    logical, allocatable :: lbdl(:)
    integer :: pseg(2, 1), nbdl, pbdl(1), nint, pint(1), i
    allocate(lbdl(ne))
    do i = 1, ne
      lbdl(i) = bdl(i) /= 0
    end do
    nbdl = 0; nint = 0; pseg = 0; pbdl = 0; pint = 0
    call cal_var(nf, ne, u0, v0, h0, u1, v1, h1, tseg, pseg, ssit, itm, its, norm, nbdl, pbdl, &
         nint, pint, lbdl, cipi, vect, normt, mu0, mb0, fc, beta, w0, z0, g0, hb, time)
  end subroutine

  ! nsteps of the driver's RK4 loop (src/shwwnl.f:233-286), same expression
  ! order, on (45, ne) arrays updated in place.
  subroutine ref_rk4(nf, ne, u0, v0, h0, tseg, ssit, itm, its, norm, bdl, vect, normt, mu0, &
       mb0, fc, beta, w0, z0, g0, hb, dt, nsteps) bind(C, name="ref_rk4")
    integer(c_int), value :: nf, ne, nsteps
    real(c_double) :: u0(LDM, *), v0(LDM, *), h0(LDM, *)
    integer(c_int) :: tseg(2, *), ssit(2, *), itm(3, *), its(3, *), bdl(*)
    real(c_double) :: norm(2, *), vect(5, *), normt(2, *), fc(LDM, *), w0(LDM, *), z0(LDM, *), &
         hb(LDM, *)
    real(c_double), value :: mu0, mb0, beta, g0, dt
    integer nc, nm, ng, ngf
    common /const/ nc, nm, ng, ngf
    double precision, allocatable :: u1(:,:), v1(:,:), h1(:,:), k(:,:,:,:)
    logical, allocatable :: lbdl(:)
    integer :: pseg(2, 1), nbdl, pbdl(1), nint, pint(1), i, j, ite, s
    double precision time
    allocate(u1(LDM, ne), v1(LDM, ne), h1(LDM, ne), k(LDM, ne, 3, 4), lbdl(ne))
    u1 = 0.d0; v1 = 0.d0; h1 = 0.d0; k = 0.d0
    do i = 1, ne
      lbdl(i) = bdl(i) /= 0
    end do
    nbdl = 0; nint = 0; pseg = 0; pbdl = 0; pint = 0
    time = 0.d0
    do ite = 1, nsteps
      do s = 1, 4
        if (s == 1) then
          call cal_var(nf, ne, u0, v0, h0, k(1,1,1,s), k(1,1,2,s), k(1,1,3,s), tseg, pseg, ssit, &
               itm, its, norm, nbdl, pbdl, nint, pint, lbdl, cipi, vect, normt, mu0, mb0, fc, &
               beta, w0, z0, g0, hb, time)
        else
          call cal_var(nf, ne, u1, v1, h1, k(1,1,1,s), k(1,1,2,s), k(1,1,3,s), tseg, pseg, ssit, &
               itm, its, norm, nbdl, pbdl, nint, pint, lbdl, cipi, vect, normt, mu0, mb0, fc, &
               beta, w0, z0, g0, hb, time)
        end if
        if (s <= 2) then
          do i = 1, ne
            do j = 1, nm
              u1(j, i) = u0(j, i) + dt * k(j, i, 1, s) * 0.5d0
              v1(j, i) = v0(j, i) + dt * k(j, i, 2, s) * 0.5d0
              h1(j, i) = h0(j, i) + dt * k(j, i, 3, s) * 0.5d0
            end do
          end do
        else if (s == 3) then
          do i = 1, ne
            do j = 1, nm
              u1(j, i) = u0(j, i) + dt * k(j, i, 1, s)
              v1(j, i) = v0(j, i) + dt * k(j, i, 2, s)
              h1(j, i) = h0(j, i) + dt * k(j, i, 3, s)
            end do
          end do
        end if
      end do
      do i = 1, ne
        do j = 1, nm
          u0(j, i) = u0(j, i) + dt / 6.d0 * (k(j,i,1,1) + 2.d0 * k(j,i,1,2) + 2.d0 * k(j,i,1,3) &
               + k(j,i,1,4))
          v0(j, i) = v0(j, i) + dt / 6.d0 * (k(j,i,2,1) + 2.d0 * k(j,i,2,2) + 2.d0 * k(j,i,2,3) &
               + k(j,i,2,4))
          h0(j, i) = h0(j, i) + dt / 6.d0 * (k(j,i,3,1) + 2.d0 * k(j,i,3,2) + 2.d0 * k(j,i,3,3) &
               + k(j,i,3,4))
        end do
      end do
    end do
  end subroutine

  ! Global integrals (src/energy.f:3-96).
  subroutine ref_energy(ne, u0, v0, h0, g0, hb, ke, pe, mass, in, x, y) bind(C, name="ref_energy")
    integer(c_int), value :: ne
    real(c_double) :: u0(LDM, *), v0(LDM, *), h0(LDM, *), hb(LDM, *), x(*), y(*)
    real(c_double), value :: g0
    real(c_double) :: ke, pe, mass
    integer(c_int) :: in(3, *)
    double precision amm(LDM,LDM), amml(LDM,LDM), agx(LDM,LDM), agy(LDM,LDM), &
         ammx(LDM,LDM), ammy(LDM,LDM), ammo(LDM,LDM)
    common /matrices/ amm, amml, agx, agy, ammx, ammy, ammo
    call energy(ne, u0, v0, h0, g0, hb, ke, pe, mass, in, x, y, ammo)
  end subroutine

  ! Single-kernel entry points, for unit-level pins.
  subroutine ref_cal_flu(nf, tseg, fbdy, a0, ssit) bind(C, name="ref_cal_flu")
    integer(c_int), value :: nf
    integer(c_int) :: tseg(2, *), ssit(2, *)
    real(c_double) :: fbdy(LDF, *), a0(LDM, *)
    call cal_flu(nf, tseg, fbdy, a0, ssit)
  end subroutine

  subroutine ref_dtrsm4(m, which, b) bind(C, name="ref_dtrsm4")
    integer(c_int), value :: m, which
    real(c_double) :: b(*)
    double precision amm(LDM,LDM), amml(LDM,LDM), agx(LDM,LDM), agy(LDM,LDM), &
         ammx(LDM,LDM), ammy(LDM,LDM), ammo(LDM,LDM)
    common /matrices/ amm, amml, agx, agy, ammx, ammy, ammo
    double precision dia(LDM)
    integer i
    if (which == 0) then
      do i = 1, m
        dia(i) = 1.d0 / amm(i, i)
      end do
      call dtrsm4(m, amm, b, dia)
    else
      do i = 1, m
        dia(i) = 1.d0 / amml(i, i)
      end do
      call dtrsm4(m, amml, b, dia)
    end if
  end subroutine

  subroutine ref_reduc(nc, gn, gnr, inverse) bind(C, name="ref_reduc")
    integer(c_int), value :: nc, inverse
    real(c_double) :: gn(*), gnr(*)
    if (inverse == 0) then
      call reduc_gn(nc, gn, gnr, cipi)
    else
      call reduc_inv_gn(nc, gn, gnr, cipi)
    end if
  end subroutine

  real(c_double) function ref_po(n, x) bind(C, name="ref_po")
    integer(c_int), value :: n
    real(c_double), value :: x
    double precision, external :: po
    ref_po = po(n, x)
  end function

  real(c_double) function ref_dpo(n, x) bind(C, name="ref_dpo")
    integer(c_int), value :: n
    real(c_double), value :: x
    double precision, external :: dpo
    ref_dpo = dpo(n, x)
  end function
end module
