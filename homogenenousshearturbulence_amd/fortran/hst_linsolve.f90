! The wall-normal cyclic pentadiagonal systems (the reference's module
! hst_linsolve, src/hst_linsolve.f90): one system per Fourier column, every
! column at once on the GPU (bordering with one slab, libhstgpu).
module hst_linsolve

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api
  use hst_setup, only: to_device, from_device
  use hst_derivatives, only: shear_shifts

  implicit none
  private
  public :: init_linsolve, free_linsolve, line_solve
  public :: KIND_D2V, KIND_ETA, KIND_POISSON, KIND_D0, KIND_DY

  integer(C_INT), parameter :: KIND_D2V = 1, KIND_ETA = 2, KIND_POISSON = 3, KIND_D0 = 4, KIND_DY = 5

contains

  ! The solver's workspace is part of the device state (allocate_fields).
  subroutine init_linsolve()
  end subroutine init_linsolve

  subroutine free_linsolve()
  end subroutine free_linsolve

  !   KIND_D2V     [lambda (D2 - k2 D0) - ni (D4 - 2 k2 D2 + k2^2 D0)] dst = src
  !   KIND_ETA     [lambda D0 - ni (D2 - k2 D0)] dst = src
  !   KIND_POISSON (D2 - k2 D0) dst = src
  !   KIND_D0      D0 dst = src
  !   KIND_DY      D0 dst = D1 src      (the ghost rows of src filled)
  ! for the interior rows of dst; its ghost rows are left as they are, and so
  ! is the (0,0) column for KIND_POISSON (singular; zero for KIND_D2V).  The
  ! wrap phase is that of shift_x, shift_z, or of the current time without
  ! them.  src and dst have the layout of V(:, :, :, 1) and may be any
  ! arrays of that shape (sections of V, rhs, oldrhs or the caller's own).
  ! On the GPU src goes to the spare field and the solution to rhs(:, :, :, 2)'s
  ! (which dst fills first, so that what the solve does not write comes back
  ! unchanged); with a resident state that field is free between substeps.
  subroutine line_solve(sys, lambda, src, dst, shift_x, shift_z)
    integer(C_INT), intent(in) :: sys
    real(C_DOUBLE), intent(in) :: lambda
    complex(C_DOUBLE_COMPLEX), intent(in) :: src(-2:ny + 1, -nz:nz, 0:nx)
    complex(C_DOUBLE_COMPLEX), intent(inout) :: dst(-2:ny + 1, -nz:nz, 0:nx)
    real(C_DOUBLE), intent(in), optional :: shift_x, shift_z
    real(C_DOUBLE) :: sx, sz
    if (present(shift_x) .neqv. present(shift_z)) error stop 'line_solve: give both shift_x and shift_z or neither'
    if (present(shift_x)) then
      sx = shift_x
      sz = shift_z
    else
      call shear_shifts(time, sx, sz)
    end if
    call to_device(HST_SCRATCH, src)
    call to_device(HST_RHS2, dst)
    call gpu_check(hst_gpu_line_solve(sys, lambda, HST_SCRATCH, HST_RHS2, sx, sz), 'line_solve')
    call from_device(HST_RHS2, dst)
  end subroutine line_solve

end module hst_linsolve
