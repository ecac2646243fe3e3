! example_pad.f90 -- for the example host programs: re-home an array in one whose column extent is
! pcols > ncol, as a GCM's pcols-sized arrays are, with the columns beyond ncol set to a value no
! atmosphere has, so that a call that reads them gives results far off.
module example_pad
  implicit none
  private
  public :: pad1, pad2, pad3a, pad3b
  integer, parameter :: rb = selected_real_kind(12)
  real(rb), parameter :: junk = -1.0e30_rb
contains
  ! x(ncol) -> x(pcols)
  subroutine pad1(x, pcols)
    real(rb), allocatable, intent(inout) :: x(:)
    integer, intent(in) :: pcols
    real(rb), allocatable :: t(:)
    allocate(t(pcols))
    t = junk
    t(1:size(x)) = x
    call move_alloc(t, x)
  end subroutine pad1

  ! x(ncol, n2) -> x(pcols, n2)
  subroutine pad2(x, pcols)
    real(rb), allocatable, intent(inout) :: x(:,:)
    integer, intent(in) :: pcols
    real(rb), allocatable :: t(:,:)
    allocate(t(pcols, size(x, 2)))
    t = junk
    t(1:size(x, 1), :) = x
    call move_alloc(t, x)
  end subroutine pad2

  ! x(ncol, n2, n3) -> x(pcols, n2, n3)
  subroutine pad3a(x, pcols)
    real(rb), allocatable, intent(inout) :: x(:,:,:)
    integer, intent(in) :: pcols
    real(rb), allocatable :: t(:,:,:)
    allocate(t(pcols, size(x, 2), size(x, 3)))
    t = junk
    t(1:size(x, 1), :, :) = x
    call move_alloc(t, x)
  end subroutine pad3a

  ! x(n1, ncol, n3) -> x(n1, pcols, n3)
  subroutine pad3b(x, pcols)
    real(rb), allocatable, intent(inout) :: x(:,:,:)
    integer, intent(in) :: pcols
    real(rb), allocatable :: t(:,:,:)
    allocate(t(size(x, 1), pcols, size(x, 3)))
    t = junk
    t(:, 1:size(x, 2), :) = x
    call move_alloc(t, x)
  end subroutine pad3b
end module example_pad
