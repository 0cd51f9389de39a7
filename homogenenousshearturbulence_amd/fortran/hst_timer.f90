! Wall-clock time of the steps (timing = .true. in &time_control); each toc
! waits for the device first.
module hst_timer

  use, intrinsic :: iso_c_binding
  use hst_params, only: timing, has_terminal
  use hst_fft, only: device_sync

  implicit none
  private
  public :: tic, toc, timer_report, T_STEP, T_OTHER

  integer, parameter :: T_STEP = 1, T_OTHER = 2, NPHASE = 2
  character(len=32), parameter :: names(NPHASE) = [character(len=32) :: 'RK3 steps (GPU)', 'statistics, I/O, rest']
  real(C_DOUBLE), save :: acc(NPHASE) = 0.0d0
  integer(8), save :: t_last = 0

contains

  subroutine tic()
    if (.not. timing) return
    call device_sync()
    call system_clock(t_last)
  end subroutine tic

  subroutine toc(phase)
    integer, intent(in) :: phase
    integer(8) :: t, rate
    if (.not. timing) return
    call device_sync()
    call system_clock(t, rate)
    acc(phase) = acc(phase) + real(t - t_last, C_DOUBLE)/real(rate, C_DOUBLE)
    t_last = t
  end subroutine toc

  subroutine timer_report(nsteps)
    integer(C_SIZE_T), intent(in) :: nsteps
    integer :: i
    real(C_DOUBLE) :: total
    if (.not. (timing .and. has_terminal)) return
    total = sum(acc)
    write (*, '(A)') ' '
    write (*, '(A,I0,A)') '   time per phase over ', nsteps, ' steps:'
    do i = 1, NPHASE
      write (*, '(A,A,F10.3,A,F9.5,A,F6.1,A)') '     ', names(i), acc(i), ' s', acc(i)/max(nsteps, 1_C_SIZE_T), &
        ' s/step', 100.0d0*acc(i)/max(total, tiny(total)), ' %'
    end do
  end subroutine timer_report

end module hst_timer
