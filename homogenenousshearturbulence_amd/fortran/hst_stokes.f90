! The Stokes layer (the reference's module hst_stokes): the oscillating
! spanwise mean profile at mid-box and its body force.  The profile, the force
! rows and the region mask live on the GPU; a substep imposes the profile after
! its solve (hst_gpu_timestep), and a host imposes it once after loading a
! field.
module hst_stokes

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api

  implicit none
  private
  public :: stokes_active, stokes_apply

contains

  logical function stokes_active()
    stokes_active = (sl_amplitude /= 0.0d0)
  end function stokes_active

  ! the (0,0) mode of w set to the layer's profile at the current time, its
  ! ghost rows refilled (stokes_apply + fill_ghosts(3), src/hst_stokes.f90:103-118)
  subroutine stokes_apply()
    call gpu_check(hst_gpu_stokes_apply(time), 'stokes_apply')
  end subroutine stokes_apply

end module hst_stokes
