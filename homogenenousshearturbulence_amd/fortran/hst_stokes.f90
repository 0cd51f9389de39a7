! The Stokes layer (the reference's module hst_stokes, src/hst_stokes.f90): an
! oscillating spanwise mean profile at mid-box and the body force that drives
! it.  The profile, the force rows and the region mask live in the library,
! which sets them up with the device state; a substep imposes the profile
! after its solve, and a host imposes it once after loading a field.
module hst_stokes

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api
  use hst_setup, only: state_in, state_out, S_V

  implicit none
  private
  public :: stokes_active, stokes_setup, stokes_apply

contains

  logical function stokes_active()
    stokes_active = (sl_amplitude /= 0.0d0)
  end function stokes_active

  ! the layer's tables are made by allocate_fields (hst_gpu_init)
  subroutine stokes_setup()
  end subroutine stokes_setup

  ! the (0,0) mode of w set to the layer's profile at the current time, its
  ! ghost rows refilled (stokes_apply + fill_ghosts(3), src/hst_stokes.f90:103-118)
  subroutine stokes_apply()
    call state_in(S_V)
    call gpu_check(hst_gpu_stokes_apply(time), 'stokes_apply')
    call state_out(S_V)
  end subroutine stokes_apply

end module hst_stokes
