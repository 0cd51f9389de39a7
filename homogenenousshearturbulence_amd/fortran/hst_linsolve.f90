! The wall-normal cyclic pentadiagonal systems (the reference's module
! hst_linsolve): one system per Fourier column, all columns at once on the
! GPU, bordering with one slab.  src and dst name device fields.
module hst_linsolve

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api
  use hst_derivatives, only: shear_shifts

  implicit none
  private
  public :: init_linsolve, free_linsolve, line_solve
  public :: KIND_D2V, KIND_ETA, KIND_POISSON, KIND_D0, KIND_DY

  integer(C_INT), parameter :: KIND_D2V = 1, KIND_ETA = 2, KIND_POISSON = 3, KIND_D0 = 4, KIND_DY = 5

contains

  ! The solver workspace is allocated with the fields (hst_gpu_init).
  subroutine init_linsolve()
  end subroutine init_linsolve

  subroutine free_linsolve()
  end subroutine free_linsolve

  !   KIND_D2V     [lambda (D2 - k2 D0) - ni (D4 - 2 k2 D2 + k2^2 D0)] x = src
  !   KIND_ETA     [lambda D0 - ni (D2 - k2 D0)] x = src
  !   KIND_POISSON (D2 - k2 D0) x = src
  !   KIND_D0      D0 x = src
  !   KIND_DY      D0 x = D1 src      (ghost rows of src filled)
  ! into the interior rows of dst; the wrap phase of shift_x, shift_z, or of
  ! the current time without them.
  subroutine line_solve(sys, lambda, src, dst, shift_x, shift_z)
    integer(C_INT), intent(in) :: sys
    real(C_DOUBLE), intent(in) :: lambda
    integer(C_INT), intent(in) :: src, dst
    real(C_DOUBLE), intent(in), optional :: shift_x, shift_z
    real(C_DOUBLE) :: sx, sz
    if (present(shift_x)) then
      sx = shift_x; sz = shift_z
    else
      call shear_shifts(time, sx, sz)
    end if
    call gpu_check(hst_gpu_line_solve(sys, lambda, src, dst, sx, sz), 'line_solve')
  end subroutine line_solve

end module hst_linsolve
