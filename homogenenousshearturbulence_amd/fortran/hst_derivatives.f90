! Compact finite differences in y and the shear-periodic ghost rows (the
! reference's module hst_derivatives, src/hst_derivatives.f90).  The weights
! der(iy, n, j) come from the library at set-up (the same construction and
! bits); the clock helpers S2(t), its integral and the wrap shifts are the
! closed forms the reference evaluates on the host.
module hst_derivatives

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api
  use hst_setup, only: to_device, from_device

  implicit none
  private
  public :: setup_derivatives, fill_ghosts, fill_ghosts_field, shear_shifts
  public :: s2_of, s2_integral, gamma_y_of

  real(C_DOUBLE), parameter :: TWOPI = 6.283185307179586d0

contains

  ! der, y, dyl were filled by allocate_fields from the library's tables.
  subroutine setup_derivatives()
    if (ny < 8) error stop 'setup_derivatives: ny must be at least 8'
    if (.not. allocated(der)) error stop 'setup_derivatives: call allocate_fields first'
  end subroutine setup_derivatives

  ! S2(t) = s2_mean + s2_amplitude sin(2 pi (t - s2_start)/s2_period) from s2_start on
  real(C_DOUBLE) function s2_of(t)
    real(C_DOUBLE), intent(in) :: t
    s2_of = 0.0d0
    if ((s2_amplitude == 0.0d0 .and. s2_mean == 0.0d0) .or. t < s2_start) return
    if (s2_period > 0.0d0) then
      s2_of = s2_mean + s2_amplitude*sin(TWOPI/s2_period*(t - s2_start))
    else
      s2_of = s2_mean + s2_amplitude
    end if
  end function s2_of

  ! the integral of S2 from t1 to t2 (the part after s2_start)
  real(C_DOUBLE) function s2_integral(t1, t2)
    real(C_DOUBLE), intent(in) :: t1, t2
    real(C_DOUBLE) :: a, b, omega
    s2_integral = 0.0d0
    if ((s2_amplitude == 0.0d0 .and. s2_mean == 0.0d0) .or. t2 <= s2_start) return
    a = max(t1, s2_start); b = t2
    if (s2_period > 0.0d0) then
      omega = TWOPI/s2_period
      s2_integral = s2_mean*(b - a) + s2_amplitude/omega*(cos(omega*(a - s2_start)) - cos(omega*(b - s2_start)))
    else
      s2_integral = (s2_mean + s2_amplitude)*(b - a)
    end if
  end function s2_integral

  real(C_DOUBLE) function gamma_y_of(t)
    real(C_DOUBLE), intent(in) :: t
    gamma_y_of = s2_integral(s2_start, t)
  end function gamma_y_of

  ! displacements of the upper image at time t
  subroutine shear_shifts(t, shift_x, shift_z)
    real(C_DOUBLE), intent(in) :: t
    real(C_DOUBLE), intent(out) :: shift_x, shift_z
    shift_x = modulo(S*t*ly, lx)
    shift_z = modulo(gamma_y_of(t)*ly, lz)
  end subroutine shear_shifts

  ! Ghost rows of velocity component c (and c2) at the current time.
  subroutine fill_ghosts(c, c2)
    integer(C_INT), intent(in) :: c
    integer(C_INT), intent(in), optional :: c2
    real(C_DOUBLE) :: sx, sz
    call shear_shifts(time, sx, sz)
    call ghosts_of_component(c, sx, sz)
    if (present(c2)) call ghosts_of_component(c2, sx, sz)
  end subroutine fill_ghosts

  subroutine ghosts_of_component(c, sx, sz)
    integer(C_INT), intent(in) :: c
    real(C_DOUBLE), intent(in) :: sx, sz
    if (.not. state_resident) call to_device(HST_V1 + c - 1, V(:, :, :, c))
    call gpu_check(hst_gpu_fill_ghosts(HST_V1 + c - 1, sx, sz), 'fill_ghosts')
    if (.not. state_resident) call from_device(HST_V1 + c - 1, V(:, :, :, c))
  end subroutine ghosts_of_component

  ! Ghost rows of any array of the field shape for the displacements
  ! shift_x, shift_z (the work runs in the library's spare field).
  subroutine fill_ghosts_field(field, shift_x, shift_z)
    complex(C_DOUBLE_COMPLEX), intent(inout) :: field(-2:ny + 1, -nz:nz, 0:nx)
    real(C_DOUBLE), intent(in) :: shift_x, shift_z
    call to_device(HST_SCRATCH, field)
    call gpu_check(hst_gpu_fill_ghosts(HST_SCRATCH, shift_x, shift_z), 'fill_ghosts_field')
    call from_device(HST_SCRATCH, field)
  end subroutine fill_ghosts_field

end module hst_derivatives
