! rrtmgsw_lead.f90 -- the leading blocks of the drop-in's array dummies, as contiguous memory.
!
! The reference reads its array dummies by their leading block only: x(1:ncol, 1:nlay) of a
! profile (src/rrtmg_sw_rad.f90:1534 play(iplon,l)), cldfmcl(1:ngptsw, 1:ncol, 1:nlay)
! (src/rrtmg_sw_rad.f90:1614), cldfrac(1:ncol, 1:nlay) (src/mcica_subcol_gen_sw.f90:458). A GCM may
! therefore pass arrays whose extents exceed the counts: pcols-sized arrays with ncol columns in
! use, as CAM does. The C-ABI takes exactly the counts' column-major layout. These helpers return
! the address of the actual argument when its extents equal the counts, otherwise that of a
! packed copy of the leading block (inputs) or of a temporary that back2 / back3 unpack after the
! call (outputs). An extent below its count stops with a message. An input the call does not
! read (used = .false.) is passed as it is.
module rrtmgsw_lead
  use iso_c_binding
  implicit none
  private
  public :: buf1, buf2, buf3, lead1, lead2, lead3, outb2, outb3, back2, back3
  integer, parameter :: rb = selected_real_kind(12)
  type :: buf1
    real(rb), allocatable :: a(:)
  end type
  type :: buf2
    real(rb), allocatable :: a(:,:)
  end type
  type :: buf3
    real(rb), allocatable :: a(:,:,:)
  end type
contains
  subroutine too_small()
    error stop "RRTMG_SW: AN ARRAY ARGUMENT IS SMALLER THAN NCOL x NLAY"
  end subroutine too_small

  function lead1(x, n1, t, used) result(p)
    real(rb), intent(in), target, contiguous :: x(:)
    integer, intent(in) :: n1
    type(buf1), target, intent(inout) :: t
    logical, intent(in) :: used
    type(c_ptr) :: p
    if (.not. used .or. size(x) == n1) then
      p = c_loc(x)
      return
    end if
    if (size(x) < n1) call too_small()
    allocate(t%a(n1))
    t%a = x(1:n1)
    p = c_loc(t%a)
  end function lead1

  function lead2(x, n1, n2, t, used) result(p)
    real(rb), intent(in), target, contiguous :: x(:,:)
    integer, intent(in) :: n1, n2
    type(buf2), target, intent(inout) :: t
    logical, intent(in) :: used
    type(c_ptr) :: p
    if (.not. used .or. (size(x, 1) == n1 .and. size(x, 2) == n2)) then
      p = c_loc(x)
      return
    end if
    if (size(x, 1) < n1 .or. size(x, 2) < n2) call too_small()
    allocate(t%a(n1, n2))
    t%a = x(1:n1, 1:n2)
    p = c_loc(t%a)
  end function lead2

  function lead3(x, n1, n2, n3, t, used) result(p)
    real(rb), intent(in), target, contiguous :: x(:,:,:)
    integer, intent(in) :: n1, n2, n3
    type(buf3), target, intent(inout) :: t
    logical, intent(in) :: used
    type(c_ptr) :: p
    if (.not. used .or. (size(x, 1) == n1 .and. size(x, 2) == n2 .and. size(x, 3) == n3)) then
      p = c_loc(x)
      return
    end if
    if (size(x, 1) < n1 .or. size(x, 2) < n2 .or. size(x, 3) < n3) call too_small()
    allocate(t%a(n1, n2, n3))
    t%a = x(1:n1, 1:n2, 1:n3)
    p = c_loc(t%a)
  end function lead3

  function outb2(x, n1, n2, t) result(p)
    real(rb), intent(inout), target, contiguous :: x(:,:)
    integer, intent(in) :: n1, n2
    type(buf2), target, intent(inout) :: t
    type(c_ptr) :: p
    if (size(x, 1) == n1 .and. size(x, 2) == n2) then
      p = c_loc(x)
      return
    end if
    if (size(x, 1) < n1 .or. size(x, 2) < n2) call too_small()
    allocate(t%a(n1, n2))
    p = c_loc(t%a)
  end function outb2

  function outb3(x, n1, n2, n3, t) result(p)
    real(rb), intent(inout), target, contiguous :: x(:,:,:)
    integer, intent(in) :: n1, n2, n3
    type(buf3), target, intent(inout) :: t
    type(c_ptr) :: p
    if (size(x, 1) == n1 .and. size(x, 2) == n2 .and. size(x, 3) == n3) then
      p = c_loc(x)
      return
    end if
    if (size(x, 1) < n1 .or. size(x, 2) < n2 .or. size(x, 3) < n3) call too_small()
    allocate(t%a(n1, n2, n3))
    p = c_loc(t%a)
  end function outb3

  subroutine back2(x, t)
    real(rb), intent(inout), contiguous :: x(:,:)
    type(buf2), intent(in) :: t
    if (allocated(t%a)) x(1:size(t%a, 1), 1:size(t%a, 2)) = t%a
  end subroutine back2

  subroutine back3(x, t)
    real(rb), intent(inout), contiguous :: x(:,:,:)
    type(buf3), intent(in) :: t
    if (allocated(t%a)) x(1:size(t%a, 1), 1:size(t%a, 2), 1:size(t%a, 3)) = t%a
  end subroutine back3
end module rrtmgsw_lead
