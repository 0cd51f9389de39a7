! Direct numerical simulation of homogeneous shear turbulence: the driver of
! the reference (src/hst.f90) on one MI355X.  Usage:  hst [hst.in]
! Set-up, then per step: timestep(), statistics, snapshots (velocity,
! pressure and the averaged column-energy spectrum), restart file, new
! time step from the CFL number.  The run ends at t_max, after nstep steps or
! after wall_max seconds, whichever comes first.
program hst

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_input
  use hst_setup
  use hst_fft
  use hst_derivatives
  use hst_linsolve
  use hst_transforms
  use hst_equations
  use hst_initial, only: generate_initial_field
  use hst_io
  use hst_stats
  use hst_pressure
  use hst_timer
  use hst_stokes, only: stokes_apply

  implicit none

  character(len=256) :: deck
  character(len=40) :: fname
  integer(8) :: c0, c1, rate
  real(C_DOUBLE) :: elapsed
  logical :: wall_reached = .false.

  call select_device()
  deck = 'hst.in'
  if (command_argument_count() >= 1) call get_command_argument(1, deck)
  call read_input(trim(deck))
  if (has_terminal) call print_input()
  call allocate_fields()
  call init_fft()
  call init_linsolve()
  call setup_derivatives()
  call make_output_dirs()
  if (restart_from_file) then
    call restart_read(trim(restart_file))
  else
    if (has_terminal) print *, '   restart_from_file = .false.: generating initial field'
    call generate_initial_field()
  end if
  ! the Stokes layer's profile imposed on the loaded field (stokes_setup, stokes_apply)
  call stokes_apply()
  call fill_ghosts(1)
  call fill_ghosts(2)
  call fill_ghosts(3)
  call open_runtimedata()

  ! first time step from the CFL of the initial field
  if (deltat == 0.0d0) deltat = 1.0d0
  call compute_cfl()
  call new_timestep()
  ifield = floor(time/dt_field)
  if (has_terminal) write (*, '(A)') '        time       deltat       cfl         energy           diss' // &
    '             uv             vw      (y-averaged; diss = nu <grad u : grad u>)'
  call outstats()

  elapsed = 0.0d0
  do while (time < t_max - 0.5d0*deltat .and. istep < nstep .and. .not. wall_reached)
    call system_clock(c0, rate)
    istep = istep + 1
    call tic()
    call timestep()
    call toc(T_STEP)
    if (crossed(dt_stat)) call outstats()
    if (crossed(dt_field)) then
      ifield = ifield + 1
      write (fname, '(A,I0,A)') 'fields/field', ifield, '.fld'
      if (has_terminal) print '(A,F12.5)', '   writing '//trim(fname)//' at time', time
      call restart_write(trim(fname))
      write (fname, '(A,I0,A)') 'p_fields/pField', ifield, '.fld'
      call write_pressure(trim(fname))
      write (fname, '(A,I0,A)') 'statistics/spectra/exz', ifield, '.dat'
      call spectrum_write(trim(fname))
    end if
    if (crossed(dt_save)) then
      if (has_terminal) print '(A,F12.5)', '   writing Dati.cart.out at time', time
      call restart_write('Dati.cart.out')
    end if
    call new_timestep()
    call toc(T_OTHER)
    call device_sync()
    call system_clock(c1)
    elapsed = elapsed + real(c1 - c0, C_DOUBLE)/real(rate, C_DOUBLE)
    if (wall_max > 0.0d0) wall_reached = elapsed > wall_max
    if (has_terminal .and. (mod(istep, 50_C_SIZE_T) == 0 .or. istep <= 5)) &
      write (*, '(A,I0,A,F9.5,A,F12.2,A)') '   step ', istep, ': ', real(c1 - c0, C_DOUBLE)/real(rate, C_DOUBLE), &
      ' s/step, ', elapsed, ' s elapsed'
  end do

  if (has_terminal) print '(A,F12.5,A,I0,A,F10.2,A,F9.5,A)', '   end of run at time', time, ' after ', istep, &
    ' steps; ', elapsed, ' s, ', elapsed/max(istep, 1_C_SIZE_T), ' s/step; writing Dati.cart.out'
  if (has_terminal .and. wall_reached) print '(A,F10.0,A)', '   (wall-clock limit wall_max =', wall_max, ' s reached)'
  call timer_report(istep)
  call restart_write('Dati.cart.out')
  call close_runtimedata()
  call free_linsolve()
  call free_fft()
  call free_fields()

contains

  logical function crossed(dt)
    real(C_DOUBLE), intent(in) :: dt
    crossed = .false.
    if (dt > 0.0d0) crossed = floor((time + 0.5d0*deltat)/dt) > floor((time - 0.5d0*deltat)/dt)
  end function crossed

  ! next time step: from cflmax when the deck gives deltat = 0, else the fixed
  ! value (capped by cflmax if that is set)
  subroutine new_timestep()
    if (cflmax > 0.0d0 .and. cfl > 0.0d0) deltat = cflmax/cfl
    if (dt_fixed > 0.0d0) deltat = min(deltat, dt_fixed)
    if (dt_fixed > 0.0d0 .and. cflmax <= 0.0d0) deltat = dt_fixed
  end subroutine new_timestep

end program hst
