! Drop-in replacement for the adaptive driver's refinement decision routine
! cal_conv (src/adapt.f:1122-1192): same name, same argument list, called by
! src/shwwnl_adap.f:385-386 every perraff steps.  The three jump
! indicators (checkconv, src/adapt.f:1260-1332) and the decisions come from
! spoc_cal_conv on the device, bit-identical to the reference's.
!
! It works on the context that the cal_var shim (spoc_calvar.F90) created
! and filled with the current mesh: the driver calls cal_var on that mesh
! before every cal_conv.  pico, ssit and tseg describe that same mesh and
! are not read again.  Like the reference it prints conv, deraff, timedir.
!
! The reference defines cal_conv inside adapt.f, next to routines the
! driver still needs; linking this object therefore needs adapt.f's own
! cal_conv renamed (oracle/Makefile does it with objcopy on adapt.o).

subroutine cal_conv(nn, ne, nf, nc, pico, ssit, conv, fconv, fderaff, timedir, tseg, u0, v0, &
     h0, w0, z0, hb, deraff, fac1, fac2, fac3)
  use iso_c_binding
  use spoc_abi
  implicit none
  integer, parameter :: nnmod = 45, ngf_max = 9
  integer nn, ne, nf, nc
  integer tseg(2, *), ssit(2, *)
  double precision pico(ngf_max, nnmod, 3), u0(nnmod, *), v0(nnmod, *), h0(nnmod, *), &
       w0(nnmod, *), z0(nnmod, *), hb(nnmod, *), fac1, fac2, fac3
  logical conv, fconv(*), fderaff(*), timedir, deraff
  integer(c_int), allocatable :: ic(:), id(:)
  integer(c_int) :: iconv(1), ideraff(1), itimedir(1)
  integer i
  if (.not. c_associated(spoc_ctx_h) .or. spoc_ctx_ne /= ne) then
    write(*, *) 'spoc: cal_conv needs cal_var to have run on this mesh first'
    stop 1
  end if
  call spoc_check(spoc_state_upload(spoc_ctx_h, nnmod, u0, v0, h0), 'spoc_state_upload')
  allocate(ic(ne), id(ne))
  call spoc_check(spoc_cal_conv(spoc_ctx_h, fac1, fac2, fac3, ic, id, iconv, ideraff, itimedir, &
       c_null_ptr, c_null_ptr), 'spoc_cal_conv')
  do i = 1, ne
    fconv(i) = ic(i) /= 0
    fderaff(i) = id(i) /= 0
  end do
  conv = iconv(1) /= 0
  deraff = ideraff(1) /= 0
  timedir = itimedir(1) /= 0
  deallocate(ic, id)
  write(*, *) conv, deraff, timedir
end subroutine
