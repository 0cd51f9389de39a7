! The field a run starts from when there is no restart file (the reference's
! module hst_initial): the seeded divergence-free random field or the
! Taylor-Green vortex, built on the host by the library into V and uploaded.
module hst_initial

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api
  use hst_setup, only: deck_struct, upload_velocity

  implicit none
  private
  public :: generate_initial_field, uniform_from_key

contains

  subroutine generate_initial_field()
    type(hst_deck_t) :: d
    if (has_terminal) write (*, '(A,I0,A,F6.2,A,ES9.2)') &
      '   generating initial field: seed = ', seed, '  kpeak = ', kpeak, '  amplitude = ', amplitude
    d = deck_struct()
    call gpu_check(hst_initial_field(d, V), 'generate_initial_field')
    call upload_velocity()
  end subroutine generate_initial_field

  ! A uniform deviate in [0, 1) from the seed and the mode indices (an
  ! fmix32-shaped avalanche masked to 32 bits).
  pure function uniform_from_key(seed, component, iy, iz, ix) result(u)
    integer(C_INT), intent(in) :: seed, component, iy, iz, ix
    real(C_DOUBLE) :: u
    integer(C_INT64_T) :: h
    integer(C_INT64_T), parameter :: MASK32 = 4294967295_C_INT64_T
    integer(C_INT64_T), parameter :: MIX_A = 2146121005_C_INT64_T
    integer(C_INT64_T), parameter :: MIX_B = 2032289749_C_INT64_T
    h = int(seed, C_INT64_T)
    h = h + 2654435761_C_INT64_T*int(component, C_INT64_T)
    h = h + 2246822519_C_INT64_T*int(iy, C_INT64_T)
    h = h + 3266489917_C_INT64_T*int(iz, C_INT64_T)
    h = h + 668265263_C_INT64_T*int(ix, C_INT64_T)
    h = iand(h, MASK32)
    h = ieor(h, ishft(h, -16))
    h = iand(h*MIX_A, MASK32)
    h = ieor(h, ishft(h, -13))
    h = iand(h*MIX_B, MASK32)
    h = ieor(h, ishft(h, -16))
    u = real(h, C_DOUBLE)*(1.0d0/4294967296.0d0)
  end function uniform_from_key

end module hst_initial
