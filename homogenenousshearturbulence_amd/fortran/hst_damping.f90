! Large-scale damping (the reference's module hst_damping,
! src/hst_damping.f90): the rates of the columns inside the damping region of
! the shear frame, built on the GPU (and rebuilt when the shear direction
! turns, inside the substep's buildrhs_prepare), applied by the substep as
! the explicit term -g u (soft) or as the hard mask of the right-hand sides.
module hst_damping

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api
  use hst_setup, only: state_in, state_out, S_RHS, S_OLD

  implicit none
  private
  public :: damping_active, damping_setup, damping_update, damping_soft_factor, damping_theta, damping_mask

  real(C_DOUBLE), parameter :: PI = 3.141592653589793d0

contains

  logical function damping_active()
    damping_active = (damp_kpar > 0.0d0 .and. damp_kperp > 0.0d0 .and. (damp_hard .or. damp_rate > 0.0d0))
  end function damping_active

  ! the rate table is part of the device state (allocate_fields)
  subroutine damping_setup()
  end subroutine damping_setup

  ! the table follows the shear direction inside buildrhs_prepare on the GPU
  subroutine damping_update(t)
    real(C_DOUBLE), intent(in) :: t
  end subroutine damping_update

  real(C_DOUBLE) function damping_soft_factor()
    damping_soft_factor = 0.0d0
    if (damping_active() .and. .not. damp_hard) damping_soft_factor = 1.0d0
  end function damping_soft_factor

  ! direction of e_par (radians from x towards z) at time t, s2t = S2(t)
  real(C_DOUBLE) function damping_theta(t, s2t)
    real(C_DOUBLE), intent(in) :: t, s2t
    if (damp_follow) then
      damping_theta = atan2(s2t, S)
    else
      damping_theta = damp_angle*PI/180.0d0
    end if
  end function damping_theta

  ! the hard mask: rhs and oldrhs of the masked columns set to zero (for a
  ! host that runs the substep's pieces; timestep() does it inside)
  subroutine damping_mask()
    call state_in(S_RHS + S_OLD)
    call gpu_check(hst_gpu_damping_mask(), 'damping_mask')
    call state_out(S_RHS + S_OLD)
  end subroutine damping_mask

end module hst_damping
