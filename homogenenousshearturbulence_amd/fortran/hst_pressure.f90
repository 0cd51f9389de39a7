! Pressure at snapshot times (the reference's module hst_pressure,
! src/hst_pressure.f90): the fluctuation pressure of the current velocity
! (compute_pressure, :35-112) solved on the GPU, and its file (write_pressure).
module hst_pressure

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api
  use hst_setup, only: state_in, to_device, from_device, S_V
  use hst_io, only: field_write

  implicit none
  private
  public :: compute_pressure, write_pressure

contains

  ! p (an array of the field shape: its interior rows; the ghost rows are
  ! left as they are) from the current velocity, whose ghost rows must be
  ! filled.  The right-hand side goes to rhs(:, :, :, 1), as in the reference.
  subroutine compute_pressure(p)
    complex(C_DOUBLE_COMPLEX), intent(inout) :: p(-2:ny + 1, -nz:nz, 0:nx)
    call state_in(S_V)
    call to_device(HST_RHS2, p)
    call gpu_check(hst_gpu_pressure(time, HST_RHS2), 'compute_pressure')
    call from_device(HST_RHS2, p)
    if (.not. state_resident) call from_device(HST_RHS1, rhs(:, :, :, 1))
  end subroutine compute_pressure

  ! the pressure of the current velocity to a headerless one-component file
  subroutine write_pressure(filename)
    character(len=*), intent(in) :: filename
    complex(C_DOUBLE_COMPLEX), allocatable :: p(:, :, :)
    if (has_terminal) print '(A,F12.5)', '   writing '//trim(filename)//' at time', time
    if (state_resident) then
      allocate (p(-2:ny + 1, -nz:nz, 0:nx))
      p = 0
      call compute_pressure(p)
      call field_write(filename, p)
      deallocate (p)
    else
      call compute_pressure(rhs(:, :, :, 2))
      call field_write(filename, rhs(:, :, :, 2))
    end if
  end subroutine write_pressure

end module hst_pressure
