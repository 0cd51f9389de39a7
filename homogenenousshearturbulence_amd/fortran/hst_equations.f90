! The equations (the reference's module hst_equations): one RK3 step as the
! fused device call, and its pieces with the reference's names for callers
! that run them one at a time.
module hst_equations

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api
  use hst_transforms, only: invalidate

  implicit none
  private
  public :: timestep, buildrhs_prepare, buildrhs, shear_shift, linsolve

contains

  ! One full step of length deltat: three RK substeps (transform, products,
  ! right-hand sides, shear shift, clock, implicit solves, u and w, ghosts).
  ! The clock advances by 2/RK_rai(1,i)*deltat per substep in FP64; the CFL
  ! number of the third substep is left in cfl.
  subroutine timestep()
    call gpu_check(hst_gpu_timestep(deltat, time, cfl), 'timestep')
    call invalidate()
  end subroutine timestep

  subroutine buildrhs_prepare(ODE)
    real(C_DOUBLE), intent(in) :: ODE(3)
    call gpu_check(hst_gpu_buildrhs_prepare(ODE, deltat, time), 'buildrhs_prepare')
  end subroutine buildrhs_prepare

  subroutine buildrhs(ODE, g)
    real(C_DOUBLE), intent(in) :: ODE(3)
    integer(C_INT), intent(in) :: g
    call gpu_check(hst_gpu_buildrhs(ODE, g), 'buildrhs')
  end subroutine buildrhs

  subroutine shear_shift(dt_sub)
    real(C_DOUBLE), intent(in) :: dt_sub
    call gpu_check(hst_gpu_shear_shift(dt_sub, time), 'shear_shift')
  end subroutine shear_shift

  subroutine linsolve(lambda)
    real(C_DOUBLE), intent(in) :: lambda
    call gpu_check(hst_gpu_linsolve(lambda, time), 'linsolve')
    call invalidate()
  end subroutine linsolve

end module hst_equations
