! The equations in velocity-vorticity form (the reference's module
! hst_equations, src/hst_equations.f90): one RK3 step as one fused device
! call, and the pieces of a substep under the reference's names for callers
! that run them one at a time:
!
!   do i = 1, 3
!     call transform_to_physical(); call buildrhs_prepare(RK_rai(:, i))
!     do g = 1, 2
!       call build_products(g); call products_to_spectral(); call buildrhs(RK_rai(:, i), g)
!     end do
!     call shear_shift(2/RK_rai(1, i)*deltat); time = time + ...; call linsolve(RK_rai(1, i)/deltat)
!   end do
!
! Each piece takes the host arrays it reads to the GPU and brings back those
! it writes (nothing moves when the state is resident; see hst_params).
module hst_equations

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api
  use hst_fft, only: VVdz
  use hst_transforms, only: product_group
  use hst_setup, only: state_in, state_out, S_V, S_RHS, S_OLD

  implicit none
  private
  public :: timestep, buildrhs_prepare, buildrhs, shear_shift, linsolve

contains

  ! One full step of length deltat: three substeps (transforms, products,
  ! right-hand sides, shear shift, clock, implicit solves, u and w, ghosts).
  ! The clock advances by 2/RK_rai(1,i)*deltat per substep in FP64; cfl
  ! receives the CFL number of the third substep's velocity.
  subroutine timestep()
    call state_in(S_V + S_OLD)
    call gpu_check(hst_gpu_timestep(deltat, time, cfl), 'timestep')
    call state_out(S_V + S_RHS + S_OLD)
  end subroutine timestep

  ! the implicit part and the carried explicit terms of a substep into rhs;
  ! oldrhs restarts from the new explicit terms
  subroutine buildrhs_prepare(ODE)
    real(C_DOUBLE), intent(in) :: ODE(3)
    call state_in(S_V + S_OLD)
    call gpu_check(hst_gpu_buildrhs_prepare(ODE, deltat, time), 'buildrhs_prepare')
    call state_out(S_RHS + S_OLD)
  end subroutine buildrhs_prepare

  ! the nonlinear terms of product group g, whose spectra the last
  ! products_to_spectral left in VVdz
  subroutine buildrhs(ODE, g)
    real(C_DOUBLE), intent(in) :: ODE(3)
    integer(C_INT), intent(in) :: g
    if (product_group() /= g) then
      print '(A,I0,A,I0)', ' ERROR: buildrhs of group ', g, ' but VVdz holds the spectra of group ', product_group()
      error stop 1
    end if
    call state_in(S_RHS + S_OLD)
    call gpu_check(hst_gpu_set_products(g, VVdz), 'buildrhs (product spectra)')
    call gpu_check(hst_gpu_buildrhs(ODE, g), 'buildrhs')
    call state_out(S_RHS + S_OLD)
  end subroutine buildrhs

  ! the mean-shear advection of rhs and oldrhs over a substep of length dt_sub
  ! (default node phase, or the exact shift)
  subroutine shear_shift(dt_sub)
    real(C_DOUBLE), intent(in) :: dt_sub
    call state_in(S_RHS + S_OLD)
    call gpu_check(hst_gpu_shear_shift(dt_sub, time), 'shear_shift')
    call state_out(S_RHS + S_OLD)
  end subroutine shear_shift

  ! the implicit solves of a substep at the current time: v, eta, dv/dy,
  ! u and w, their ghost rows
  subroutine linsolve(lambda)
    real(C_DOUBLE), intent(in) :: lambda
    call state_in(S_RHS)
    call gpu_check(hst_gpu_linsolve(lambda, time), 'linsolve')
    call state_out(S_V)
  end subroutine linsolve

end module hst_equations
