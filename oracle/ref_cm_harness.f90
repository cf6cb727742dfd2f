! Test-infrastructure harness: the reference's mesh preparation (the CM tool,
! src/chmal.f:2-45) callable through ctypes.  Checker only; the product never
! loads it.
module spoc_ref_cm_harness
  use iso_c_binding
  implicit none
contains
  ! pointer_mesh + changemaille on in(3,ne) (modified in place), exactly the
  ! sequence of program chmaille (src/chmal.f:24-34) minus the file I/O.
  subroutine ref_changemaille(nn, ne, in) bind(C, name="ref_changemaille")
    integer(c_int), value :: nn, ne
    integer(c_int) :: in(3, *)
    integer, allocatable :: itm(:,:), nmit(:), tmit(:), imit(:), pseg(:,:), tseg(:,:), sbd(:)
    integer nf, nebd, i
    allocate(itm(3, ne), nmit(8 * 800 + 8), tmit(8 * 800 + 64), imit(8 * 800 + 64), &
         pseg(2, 4 * ne + 2400), tseg(2, 4 * ne + 2400), sbd(4 * ne + 16))
    call pointer_mesh(nn, ne, nf, in, itm, nmit, tmit, imit, pseg, tseg)
    nebd = 0
    do i = 1, nf
      if (tseg(2, i) == 0) then
        nebd = nebd + 1
        sbd(nebd) = i
      end if
    end do
    call changemaille(nn, ne, nf, nebd, sbd, in, itm, tseg, pseg)
  end subroutine
end module
