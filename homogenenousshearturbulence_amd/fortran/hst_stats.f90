! Runtime statistics in the layout of the CPL code (the reference's module
! hst_stats): statistics/Runtimedata and statistics/variances_runtime.dat,
! opened for append, and the column-energy spectra statistics/spectra/exz<n>.dat.
! The reductions and the column energies' running sum stay on the GPU
! (hst_gpu_outstats, hst_gpu_spectrum).
module hst_stats

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api
  use hst_derivatives, only: s2_of, gamma_y_of
  use hst_setup, only: state_in, S_V

  implicit none
  private
  public :: outstats, open_runtimedata, close_runtimedata, spectrum_write

  integer, parameter :: unit_rt = 121, unit_var = 122, unit_sl = 123, unit_dp = 124
  real(C_DOUBLE), save :: t_acc0 = 0.0d0     ! start of the running spectrum average

contains

  subroutine open_runtimedata()
    if (has_terminal) open (unit=unit_rt, file='statistics/Runtimedata', action='write', position='append')
    if (has_terminal) open (unit=unit_var, file='statistics/variances_runtime.dat', action='write', position='append')
    if (has_terminal .and. sl_amplitude /= 0.0d0) &
      open (unit=unit_sl, file='statistics/stokes_runtime.dat', action='write', position='append')
    if (has_terminal .and. damping_on()) &
      open (unit=unit_dp, file='statistics/damping_runtime.dat', action='write', position='append')
    t_acc0 = time
  end subroutine open_runtimedata

  ! damping_active (src/hst_damping.f90:45-47)
  logical function damping_on()
    damping_on = (damp_kpar > 0.0d0 .and. damp_kperp > 0.0d0 .and. (damp_hard .or. damp_rate > 0.0d0))
  end function damping_on

  ! The column energies averaged since the previous call (or the start), with
  ! the (0,0) column about its box mean, then a new average (spectrum_write,
  ! src/hst_stats.f90:110-134; the text layout is the reference's).
  subroutine spectrum_write(filename)
    character(len=*), intent(in) :: filename
    real(C_DOUBLE), allocatable :: E(:, :)
    integer(C_INT32_T) :: nacc
    integer(C_INT) :: ix, iz
    integer :: unit
    allocate (E(-nz:nz, 0:nx))
    call gpu_check(hst_gpu_spectrum(E, nacc), 'spectrum')
    if (has_terminal) then
      open (newunit=unit, file=filename, action='write')
      write (unit, '(A,F12.5,A,F12.5,A,I0,A)') '# column energies ly/2 sum_y |u|^2 averaged over t = ', t_acc0, &
        ' .. ', time, ' (', nacc, ' samples);  ix  iz  kx  kz  E'
      do ix = 0, nx
        do iz = -nz, nz
          write (unit, '(I5,1X,I5,1X,ES15.7,1X,ES15.7,1X,ES15.7)') ix, iz, alfa0*ix, beta0*iz, E(iz, ix)
        end do
      end do
      close (unit)
    end if
    deallocate (E)
    t_acc0 = time
  end subroutine spectrum_write

  subroutine close_runtimedata()
    if (has_terminal) close (unit_rt)
    if (has_terminal) close (unit_var)
    if (has_terminal .and. sl_amplitude /= 0.0d0) close (unit_sl)
    if (has_terminal .and. damping_on()) close (unit_dp)
  end subroutine close_runtimedata

  ! energy, diss, uv, vw, uu, vv, ww, uw (box means) and the profile integrals mfx, mfz
  subroutine outstats()
    real(C_DOUBLE) :: o(10), e(8)
    call state_in(S_V)
    call gpu_check(hst_gpu_outstats(time, o), 'outstats')
    if (has_terminal) then
      write (*, '(F12.5,2X,ES11.4,2X,F8.4,4(2X,ES13.6))') time, deltat, cfl*deltat, o(1), o(2), o(3), o(4)
      write (unit_rt, '(13(ES23.15,1X))') time, o(9), o(10), S, s2_of(time), S*time, gamma_y_of(time), deltat, &
        cfl*deltat, o(1), o(2), o(3), o(4)
      write (unit_var, '(5(ES23.15,1X))') time, o(5), o(6), o(7), o(8)
      flush (unit_rt); flush (unit_var)
    end if
    ! the region line of the layer and the damping line (src/hst_stats.f90:270-284)
    if (sl_amplitude /= 0.0d0 .or. damping_on()) then
      call gpu_check(hst_gpu_outstats_extra(e), 'outstats_extra')
      if (has_terminal .and. sl_amplitude /= 0.0d0) then
        write (unit_sl, '(5(ES23.15,1X))') time, e(1), e(2), e(3), e(4)
        flush (unit_sl)
      end if
      if (has_terminal .and. damping_on()) then
        write (unit_dp, '(2(ES23.15,1X),I8,2(1X,ES23.15))') time, e(5), nint(e(6)), e(7), e(8)
        flush (unit_dp)
      end if
    end if
  end subroutine outstats

end module hst_stats
