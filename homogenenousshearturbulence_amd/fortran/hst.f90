! hst on one MI355X: the run driver.   Usage:  hst [deck]   (default hst.in)
!
! The state stays on the GPU for the whole run (allocate_fields with
! resident = .true.); the host sees the velocity only when it writes a file.
! Output follows the reference's schedule (src/hst.f90:93-125): a line of
! statistics at the start and whenever the step crosses a multiple of
! dt_stat, a snapshot (velocity, pressure, column-energy spectrum) at each
! multiple of dt_field, the restart file Dati.cart.out at each multiple of
! dt_save and at the end; the step length follows the CFL number (cflmax)
! or stays at the deck's deltat.  The run stops at t_max, after nstep steps
! or after wall_max seconds in the loop.
program hst

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api
  use hst_input, only: read_input, print_input
  use hst_setup, only: select_device, allocate_fields, free_fields
  use hst_fft, only: init_fft, free_fft, device_sync
  use hst_derivatives, only: setup_derivatives, fill_ghosts
  use hst_linsolve, only: init_linsolve, free_linsolve
  use hst_equations, only: timestep
  use hst_initial, only: generate_initial_field
  use hst_io, only: restart_read, restart_write, make_output_dirs
  use hst_stats, only: outstats, open_runtimedata, close_runtimedata, spectrum_write
  use hst_pressure, only: write_pressure
  use hst_timer, only: timer_start, tic, toc, timer_report, T_OTHER
  use hst_stokes, only: stokes_setup, stokes_apply
  use hst_damping, only: damping_setup, damping_active

  implicit none

  real(C_DOUBLE) :: loop_seconds = 0.0d0
  logical :: out_of_time = .false.

  call start()
  do while (keep_going())
    call one_step()
  end do
  call finish()

contains

  subroutine start()
    character(len=256) :: deck
    integer(C_INT) :: c
    call select_device()
    deck = 'hst.in'
    if (command_argument_count() > 0) call get_command_argument(1, deck)
    call read_input(trim(deck))
    if (has_terminal) call print_input()
    call allocate_fields(resident=.true.)
    call init_fft()
    call init_linsolve()
    call setup_derivatives()
    call damping_setup()
    call make_output_dirs()
    if (restart_from_file) then
      call restart_read(trim(restart_file))
    else
      call generate_initial_field()
    end if
    call stokes_setup()
    call stokes_apply()
    do c = 1, 3
      call fill_ghosts(c)
    end do
    call open_runtimedata()
    ! the first step from the CFL number of the start field
    if (deltat == 0.0d0) deltat = 1.0d0
    call gpu_check(hst_gpu_compute_cfl(cfl), 'compute_cfl')
    deltat = step_length(deltat, cfl)
    if (has_terminal .and. damping_active() .and. .not. damp_hard .and. damp_rate*deltat > 1.0d0) &
      print '(A,ES10.3,A)', '   note: the explicit damping takes damp_rate*deltat = ', damp_rate*deltat, ' > 1'
    ifield = floor(time/dt_field)
    if (has_terminal) print '(/,A,/,A)', '        time       deltat       cfl         energy           diss' // &
      '             uv             vw', '   (y-averaged; diss = nu <grad u : grad u>)'
    call outstats()
    call timer_start()
  end subroutine start

  logical function keep_going()
    keep_going = time < t_max - 0.5d0*deltat .and. istep < nstep .and. .not. out_of_time
  end function keep_going

  subroutine one_step()
    integer(8) :: c0, c1, rate
    call system_clock(c0, rate)
    istep = istep + 1
    call timestep()
    call tic()
    call write_what_is_due()
    deltat = step_length(deltat, cfl)
    call toc(T_OTHER)
    call device_sync()
    call system_clock(c1)
    loop_seconds = loop_seconds + real(c1 - c0, C_DOUBLE)/real(rate, C_DOUBLE)
    if (wall_max > 0.0d0) out_of_time = loop_seconds > wall_max
    if (has_terminal .and. (istep <= 5 .or. mod(istep, 50_C_SIZE_T) == 0)) &
      print '(A,I0,A,F9.5,A,F11.2,A)', '   step ', istep, ' took ', real(c1 - c0, C_DOUBLE)/real(rate, C_DOUBLE), &
      ' s (', loop_seconds, ' s in the loop)'
  end subroutine one_step

  ! statistics, snapshot and restart file whose interval ends at this step
  subroutine write_what_is_due()
    character(len=48) :: name
    if (due(dt_stat)) call outstats()
    if (due(dt_field)) then
      ifield = ifield + 1
      write (name, '(A,I0,A)') 'fields/field', ifield, '.fld'
      if (has_terminal) print '(A,F12.5)', '   snapshot '//trim(name)//' at t =', time
      call restart_write(trim(name))
      write (name, '(A,I0,A)') 'p_fields/pField', ifield, '.fld'
      call write_pressure(trim(name))
      write (name, '(A,I0,A)') 'statistics/spectra/exz', ifield, '.dat'
      call spectrum_write(trim(name))
    end if
    if (due(dt_save)) then
      if (has_terminal) print '(A,F12.5)', '   restart file Dati.cart.out at t =', time
      call restart_write('Dati.cart.out')
    end if
  end subroutine write_what_is_due

  ! .true. when a multiple of period lies within half a step of the clock
  logical function due(period)
    real(C_DOUBLE), intent(in) :: period
    due = .false.
    if (period <= 0.0d0) return
    due = floor((time + 0.5d0*deltat)/period) > floor((time - 0.5d0*deltat)/period)
  end function due

  ! The next step length (src/hst.f90:163-169): cflmax/cfl when cflmax is
  ! set, at most the deck's deltat when that is set too; the deck's deltat
  ! alone without cflmax; otherwise the step stays.
  pure real(C_DOUBLE) function step_length(dt, c) result(d)
    real(C_DOUBLE), intent(in) :: dt, c
    d = dt
    if (cflmax > 0.0d0 .and. c > 0.0d0) d = cflmax/c
    if (dt_fixed > 0.0d0) then
      if (cflmax > 0.0d0) then
        d = min(d, dt_fixed)
      else
        d = dt_fixed
      end if
    end if
  end function step_length

  subroutine finish()
    if (has_terminal) then
      print '(A,F12.5,A,I0,A,F10.2,A,F9.5,A)', '   run ends at t =', time, ' after ', istep, ' steps, ', &
        loop_seconds, ' s in the loop (', loop_seconds/max(istep, 1_C_SIZE_T), ' s/step)'
      if (out_of_time) print '(A,F10.0,A)', '   the loop stopped at wall_max =', wall_max, ' s'
    end if
    call timer_report(istep)
    call restart_write('Dati.cart.out')
    call close_runtimedata()
    call free_linsolve()
    call free_fft()
    call free_fields()
  end subroutine finish

end program hst
