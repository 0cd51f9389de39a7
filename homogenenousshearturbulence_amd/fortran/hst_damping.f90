! Large-scale damping (the reference's module hst_damping): the rates of the
! columns inside the damping region of the shear frame are built on the GPU
! (rebuilt when the shear direction turns) and applied by the substep: the
! explicit term -g u (soft) or the hard mask of the right-hand sides.
module hst_damping

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api

  implicit none
  private
  public :: damping_active, damping_soft_factor, damping_theta, damping_mask

  real(C_DOUBLE), parameter :: PI = 3.141592653589793d0

contains

  logical function damping_active()
    damping_active = (damp_kpar > 0.0d0 .and. damp_kperp > 0.0d0 .and. (damp_hard .or. damp_rate > 0.0d0))
  end function damping_active

  real(C_DOUBLE) function damping_soft_factor()
    damping_soft_factor = 0.0d0
    if (damping_active() .and. .not. damp_hard) damping_soft_factor = 1.0d0
  end function damping_soft_factor

  ! direction of e_par (radians from x towards z) at time t
  real(C_DOUBLE) function damping_theta(t, s2t)
    real(C_DOUBLE), intent(in) :: t, s2t     ! s2t = S2(t)
    if (damp_follow) then
      damping_theta = atan2(s2t, S)
    else
      damping_theta = damp_angle*PI/180.0d0
    end if
  end function damping_theta

  ! the hard mask: rhs and oldrhs of the masked columns set to zero (for a host
  ! that runs the substep's pieces; timestep() does it inside)
  subroutine damping_mask()
    call gpu_check(hst_gpu_damping_mask(), 'damping_mask')
  end subroutine damping_mask

end module hst_damping
