! Pressure at snapshot times (the reference's module hst_pressure): the
! fluctuation pressure of the current velocity, solved on the GPU into the
! spare right-hand side (compute_pressure, src/hst_pressure.f90:35-112), then
! written as a headerless one-component CPL file (write_pressure, :114-122).
module hst_pressure

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api
  use hst_io, only: field_write

  implicit none
  private
  public :: compute_pressure, write_pressure

contains

  ! p into the device field dst (HST_RHS2 or HST_SCRATCH; the ghost rows of
  ! the velocity must be filled, as after every step)
  subroutine compute_pressure(dst)
    integer(C_INT), intent(in) :: dst
    call gpu_check(hst_gpu_pressure(time, dst), 'pressure')
  end subroutine compute_pressure

  subroutine write_pressure(filename)
    character(len=*), intent(in) :: filename
    if (has_terminal) print '(A,F12.5)', '   writing '//trim(filename)//' at time', time
    call compute_pressure(HST_RHS2)
    call field_write(filename, HST_RHS2)
  end subroutine write_pressure

end module hst_pressure
