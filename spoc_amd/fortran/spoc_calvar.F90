! Drop-in replacement for the reference's link-time subroutine cal_var
! (src/calvar_lin.f:7-10, src/calvarnl_fs.f:9-12, src/calvarnl_ns.f:9-12,
! src/calvarnl_inv.f:6-9): same name, same argument list, same COMMON
! blocks read (src/shwwnl.f:61-71).  It forwards to the C-ABI of
! include/spoc_abi.h through ISO_C_BINDING.  Compiled once per variant with
! -DSPOC_VARIANT=0 (lin), 1 (fs), 2 (ns), 3 (inv); the makefile picks the
! object exactly as the reference's makefiles pick calvar_*.o.
!
! Environment: SPOC_PRECISION=fp32 selects the FP32-MFMA path (default: the
! plain FP64 path, which reproduces the reference bit for bit);
! SPOC_DEVICE=n selects the HIP device.
!
! The mesh and the per-run fields are uploaded on the first call and again
! whenever any of them changes: the shim keeps copies of every mesh operand
! (itm, its, tseg, ssit, bdl, norm, vect, normt) and of rows 1..nm of the
! fields (fc, w0, z0, hb) and compares them exactly on each call, since the
! adaptive driver remeshes between calls (src/shwwnl_adap.f:420-432), possibly
! to a mesh with the same ne and nf.

#ifndef SPOC_VARIANT
#define SPOC_VARIANT 1
#endif

module spoc_abi
  use iso_c_binding
  implicit none
  ! the device context, shared by the cal_var and cal_conv shims, and the
  ! element count of the mesh it holds
  type(c_ptr), save :: spoc_ctx_h = c_null_ptr
  integer, save :: spoc_ctx_ne = -1
  ! status codes of the C-ABI (include/spoc_abi.h)
  integer(c_int), parameter :: SPOC_OK = 0, SPOC_EINVAL = -1, SPOC_EHIP = -2, &
       SPOC_ESTATE = -3, SPOC_ESINGULAR = -4, SPOC_EMESH = -5, SPOC_EFILE = -6, &
       SPOC_ENAN = -7, SPOC_ENOMEM = -8
  interface
    integer(c_int) function spoc_ctx_create(ctx, device, nc, ng, ngf, wi, wif, pi2, ld_g, pico, &
         ld_f, ld_m, amm, amml, agx, agy, ammo) bind(C, name="spoc_ctx_create")
      import
      type(c_ptr) :: ctx
      integer(c_int), value :: device, nc, ng, ngf, ld_g, ld_f, ld_m
      real(c_double) :: wi(*), wif(*), pi2(*), pico(*), amm(*), amml(*), agx(*), agy(*), ammo(*)
    end function
    integer(c_int) function spoc_mesh_set(ctx, ne, nf, itm, its, tseg, ssit, norm, vect, normt, &
         bdl, nn, in, xgr, ygr) bind(C, name="spoc_mesh_set")
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: ne, nf, nn
      integer(c_int) :: itm(*), its(*), tseg(*), ssit(*), bdl(*)
      real(c_double) :: norm(*), vect(*), normt(*)
      type(c_ptr), value :: in, xgr, ygr
    end function
    integer(c_int) function spoc_fields_set(ctx, ld, fc, w0, z0, hb, mu0, mb0, g0) &
         bind(C, name="spoc_fields_set")
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: ld
      real(c_double) :: fc(*), w0(*), z0(*), hb(*)
      real(c_double), value :: mu0, mb0, g0
    end function
    integer(c_int) function spoc_rhs(ctx, variant, precision, ld, u0, v0, h0, u1, v1, h1) &
         bind(C, name="spoc_rhs")
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: variant, precision, ld
      real(c_double) :: u0(*), v0(*), h0(*), u1(*), v1(*), h1(*)
    end function
    integer(c_int) function spoc_state_upload(ctx, ld, u0, v0, h0) bind(C, name="spoc_state_upload")
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: ld
      real(c_double) :: u0(*), v0(*), h0(*)
    end function
    integer(c_int) function spoc_state_download(ctx, ld, u0, v0, h0) &
         bind(C, name="spoc_state_download")
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: ld
      real(c_double) :: u0(*), v0(*), h0(*)
    end function
    integer(c_int) function spoc_rk4_steps(ctx, variant, precision, dt, nsteps) &
         bind(C, name="spoc_rk4_steps")
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: variant, precision, nsteps
      real(c_double), value :: dt
    end function
    integer(c_int) function spoc_energy(ctx, ke, pe, mass) bind(C, name="spoc_energy")
      import
      type(c_ptr), value :: ctx
      real(c_double) :: ke, pe, mass
    end function
    integer(c_int) function spoc_cal_conv(ctx, fac1, fac2, fac3, fconv, fderaff, conv, deraff, &
         timedir, supmax, amax3) bind(C, name="spoc_cal_conv")
      import
      type(c_ptr), value :: ctx
      real(c_double), value :: fac1, fac2, fac3
      integer(c_int) :: fconv(*), fderaff(*), conv(*), deraff(*), timedir(*)
      type(c_ptr), value :: supmax, amax3
    end function
    type(c_ptr) function spoc_last_error() bind(C, name="spoc_last_error")
      import
    end function
  end interface
contains
  subroutine spoc_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(spoc_last_error(), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write(*, *) 'spoc: ', where, ' failed (', rc, '): ', msg(1:n)
    stop 1
  end subroutine
end module

subroutine cal_var(nf, ne, u0, v0, h0, u1, v1, h1, tseg, pseg, ssit, itm, its, norm, nbdl, pbdl, &
     nint, pint, bdl, cip, vect, normt, mu0, mb0, fc, beta, w0, z0, g0, hb, time)
  use iso_c_binding
  use spoc_abi
  implicit none
  integer, parameter :: nnmod = 45, ngf_max = 9, ng_max = 33
  integer nc, nm, ng, ngf
  common /const/ nc, nm, ng, ngf
  double precision wi(ng_max), xi(ng_max), yi(ng_max), wif(ngf_max), tif(ngf_max)
  common /gauss/ wi, xi, yi
  common /gaussf/ wif, tif
  double precision pi2(ng_max, nnmod), pico(ngf_max, nnmod, 3)
  common /gauss2/ pi2
  common /gauss3/ pico
  double precision amm(nnmod, nnmod), amml(nnmod, nnmod), agx(nnmod, nnmod), agy(nnmod, nnmod), &
       ammx(nnmod, nnmod), ammy(nnmod, nnmod), ammo(nnmod, nnmod)
  common /matrices/ amm, amml, agx, agy, ammx, ammy, ammo
  integer nf, ne, nbdl, nint
  integer tseg(2, *), pseg(2, *), ssit(2, *), itm(3, *), its(3, *), pbdl(*), pint(*), cip(*)
  logical bdl(*)
  double precision u0(nnmod, *), v0(nnmod, *), h0(nnmod, *), u1(nnmod, *), v1(nnmod, *), &
       h1(nnmod, *), norm(2, *), vect(5, *), normt(2, *), fc(nnmod, *), w0(nnmod, *), &
       z0(nnmod, *), hb(nnmod, *), mu0, mb0, beta, g0, time
  integer, save :: last_ne = -1, last_nf = -1, prec = 0
  double precision, save :: last_par(3) = -1.d300
  ! copies of every operand the device holds, compared exactly on each call:
  ! mesh (itm, its, tseg, ssit, bdl, norm, vect, normt) and fields (rows
  ! 1..nm of fc, w0, z0, hb)
  integer(c_int), allocatable, save :: s_itm(:), s_its(:), s_tseg(:), s_ssit(:), ibdl(:)
  double precision, allocatable, save :: s_geo(:), s_fld(:, :, :)
  logical :: remesh, refield
  integer :: i, dev, stat, l
  character(len=16) :: env
  if (.not. c_associated(spoc_ctx_h)) then
    dev = 0
    call get_environment_variable('SPOC_DEVICE', env, l, stat)
    if (stat == 0 .and. l > 0) read(env, *) dev
    call get_environment_variable('SPOC_PRECISION', env, l, stat)
    if (stat == 0 .and. l > 0 .and. env(1:4) == 'fp32') prec = 1
    call spoc_check(spoc_ctx_create(spoc_ctx_h, dev, nc, ng, ngf, wi, wif, pi2, ng_max, pico, ngf_max, &
         nnmod, amm, amml, agx, agy, ammo), 'spoc_ctx_create')
  end if
  ! remeshing between calls (src/shwwnl_adap.f:420-432) may keep ne and nf
  remesh = ne /= last_ne .or. nf /= last_nf .or. .not. allocated(s_itm)
  if (.not. remesh) remesh = int_differs(3 * ne, itm, s_itm) .or. &
       int_differs(3 * ne, its, s_its) .or. int_differs(2 * nf, tseg, s_tseg) .or. &
       int_differs(2 * nf, ssit, s_ssit) .or. bdl_differs() .or. geo_differs()
  if (remesh) then
    if (allocated(s_itm)) deallocate(s_itm, s_its, s_tseg, s_ssit, ibdl, s_geo)
    allocate(s_itm(3 * ne), s_its(3 * ne), s_tseg(2 * nf), s_ssit(2 * nf), ibdl(ne), &
         s_geo(2 * nf + 7 * ne))
    call int_copy(3 * ne, itm, s_itm)
    call int_copy(3 * ne, its, s_its)
    call int_copy(2 * nf, tseg, s_tseg)
    call int_copy(2 * nf, ssit, s_ssit)
    do i = 1, ne
      ibdl(i) = 0
      if (bdl(i)) ibdl(i) = 1
    end do
    call real_copy(2 * nf, norm, s_geo(1))
    call real_copy(5 * ne, vect, s_geo(2 * nf + 1))
    call real_copy(2 * ne, normt, s_geo(2 * nf + 5 * ne + 1))
    call spoc_check(spoc_mesh_set(spoc_ctx_h, ne, nf, itm, its, tseg, ssit, norm, vect, normt, &
         ibdl, 0, c_null_ptr, c_null_ptr, c_null_ptr), 'spoc_mesh_set')
    spoc_ctx_ne = ne
    last_ne = ne
    last_nf = nf
  end if
  ! a new mesh drops the device fields; otherwise compare them
  refield = remesh .or. .not. allocated(s_fld)
  if (.not. refield) refield = last_par(1) /= mu0 .or. last_par(2) /= mb0 .or. &
       last_par(3) /= g0
  if (.not. refield) refield = fld_differs(fc, 1) .or. fld_differs(w0, 2) .or. &
       fld_differs(z0, 3) .or. fld_differs(hb, 4)
  if (refield) then
    if (allocated(s_fld)) deallocate(s_fld)
    allocate(s_fld(nm, ne, 4))
    do i = 1, ne
      s_fld(:, i, 1) = fc(1:nm, i)
      s_fld(:, i, 2) = w0(1:nm, i)
      s_fld(:, i, 3) = z0(1:nm, i)
      s_fld(:, i, 4) = hb(1:nm, i)
    end do
    call spoc_check(spoc_fields_set(spoc_ctx_h, nnmod, fc, w0, z0, hb, mu0, mb0, g0), &
         'spoc_fields_set')
    last_par = [mu0, mb0, g0]
  end if
  call spoc_check(spoc_rhs(spoc_ctx_h, SPOC_VARIANT, prec, nnmod, u0, v0, h0, u1, v1, h1), &
       'spoc_rhs')
contains
  logical function int_differs(n, a, b)
    integer, intent(in) :: n
    integer, intent(in) :: a(n)
    integer(c_int), intent(in) :: b(n)
    integer :: k
    int_differs = .true.
    do k = 1, n
      if (a(k) /= b(k)) return
    end do
    int_differs = .false.
  end function
  subroutine int_copy(n, a, b)
    integer, intent(in) :: n
    integer, intent(in) :: a(n)
    integer(c_int), intent(out) :: b(n)
    b = a
  end subroutine
  subroutine real_copy(n, a, b)
    integer, intent(in) :: n
    double precision, intent(in) :: a(n)
    double precision, intent(out) :: b(n)
    b = a
  end subroutine
  logical function real_differs(n, a, b)
    integer, intent(in) :: n
    double precision, intent(in) :: a(n), b(n)
    integer :: k
    real_differs = .true.
    do k = 1, n
      if (a(k) /= b(k)) return
    end do
    real_differs = .false.
  end function
  logical function bdl_differs()
    integer :: k
    bdl_differs = .true.
    do k = 1, ne
      if (bdl(k) .neqv. (ibdl(k) /= 0)) return
    end do
    bdl_differs = .false.
  end function
  logical function geo_differs()
    geo_differs = real_differs(2 * nf, norm, s_geo(1)) .or. &
         real_differs(5 * ne, vect, s_geo(2 * nf + 1)) .or. &
         real_differs(2 * ne, normt, s_geo(2 * nf + 5 * ne + 1))
  end function
  logical function fld_differs(a, k)
    double precision, intent(in) :: a(nnmod, *)
    integer, intent(in) :: k
    integer :: e
    fld_differs = .true.
    do e = 1, ne
      if (real_differs(nm, a(1, e), s_fld(1, e, k))) return
    end do
    fld_differs = .false.
  end function
end subroutine
