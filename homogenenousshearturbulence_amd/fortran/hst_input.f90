! The input deck hst.in (module hst_input of the reference): one namelist file
! with the groups &mesh, &physics, &time_control and &init and the optional
! &damping; their keys and defaults are the reference's
! (src/hst_input.f90:19-42), so a deck written for it runs here unchanged.
! The arithmetic of the transforms is chosen outside the deck, by the
! environment: HST_PRECISION = fp64 (default) or mixed.
module hst_input

  use, intrinsic :: iso_c_binding
  use hst_params

  implicit none
  private
  public :: read_input, print_input, fft_fit

  real(C_DOUBLE), parameter :: PI = 3.141592653589793d0

contains

  subroutine read_input(filename)
    character(len=*), intent(in) :: filename
    namelist /mesh/ nx, ny, nz, alfa0, beta0, ly, ystretch, line_chunk, transport, npy
    namelist /physics/ re, S, linear, exact_shift, s2_mean, s2_amplitude, s2_period, s2_start, &
      sl_amplitude, sl_period, sl_delta, sl_start, sl_bodyforce, sl_ramp
    namelist /time_control/ deltat, cflmax, t_max, nstep, wall_max, dt_stat, dt_field, dt_save, time, &
      time_from_restart, restart_from_file, restart_remap, restart_file, timing
    namelist /init/ initial_condition, amplitude, seed, kpeak, tg_amplitude, tg_mode
    namelist /damping/ damp_kpar, damp_kperp, damp_rate, damp_hard, damp_edge, damp_norm, damp_mean, damp_follow, &
      damp_angle
    integer :: unit, ios
    character(len=200) :: why

    call defaults()
    open (newunit=unit, file=filename, status='old', action='read', iostat=ios)
    if (ios /= 0) call deck_error('the deck '//trim(filename)//' cannot be opened')
    ! each group from the top of the file; only &damping may be missing
    rewind (unit); read (unit, nml=mesh, iostat=ios, iomsg=why)
    if (ios /= 0) call deck_error('&mesh: '//trim(why))
    rewind (unit); read (unit, nml=physics, iostat=ios, iomsg=why)
    if (ios /= 0) call deck_error('&physics: '//trim(why))
    rewind (unit); read (unit, nml=time_control, iostat=ios, iomsg=why)
    if (ios /= 0) call deck_error('&time_control: '//trim(why))
    rewind (unit); read (unit, nml=init, iostat=ios, iomsg=why)
    if (ios /= 0) call deck_error('&init: '//trim(why))
    rewind (unit); read (unit, nml=damping, iostat=ios, iomsg=why)
    if (ios > 0) call deck_error('&damping: '//trim(why))
    close (unit)
    call validate()
    call derive()
    call choose_precision()
  end subroutine read_input

  subroutine defaults()
    nx = 63; ny = 128; nz = 63
    alfa0 = 2.0d0*PI/3.0d0; beta0 = 2.0d0*PI; ly = 2.0d0; ystretch = 0.0d0
    line_chunk = 0; transport = 'auto'; npy = 0
    re = 1000.0d0; S = 1.0d0; linear = .false.; exact_shift = .false.
    s2_mean = 0.0d0; s2_amplitude = 0.0d0; s2_period = 0.0d0; s2_start = 0.0d0
    sl_amplitude = 0.0d0; sl_period = 1.0d0; sl_delta = 0.02d0; sl_start = 0.0d0
    sl_bodyforce = .true.; sl_ramp = .false.
    deltat = 0.0d0; cflmax = 1.0d0; t_max = 100.0d0; nstep = 1000000; wall_max = 0.0d0
    dt_stat = 0.01d0; dt_field = 10.0d0; dt_save = 10.0d0
    time = 0.0d0; time_from_restart = .false.; restart_from_file = .true.
    restart_remap = 'strict'; restart_file = 'Dati.cart.out'; timing = .false.
    initial_condition = 'random'; amplitude = 1.0d-3; seed = 1; kpeak = 4.0d0
    tg_amplitude = 1.0d0; tg_mode = 1
    damp_kpar = 0.0d0; damp_kperp = 0.0d0; damp_rate = 0.0d0; damp_hard = .false.; damp_edge = 0.0d0
    damp_norm = 2.0d0; damp_mean = .false.; damp_follow = .true.; damp_angle = 0.0d0
  end subroutine defaults

  subroutine deck_error(what)
    character(len=*), intent(in) :: what
    print '(A)', ' hst.in: '//what
    error stop 1
  end subroutine deck_error

  ! The combinations the solver cannot run.
  subroutine validate()
    if (npy > 1) call deck_error('npy > 1: this build runs the whole grid on one GPU (npy = 0 or 1)')
    if (all(transport /= [character(len=8) :: 'auto', 'mpi', 'nccl'])) &
      call deck_error('transport is one of auto, mpi, nccl (and unused with one GPU)')
    if (all(restart_remap /= [character(len=16) :: 'strict', 'mode_index'])) &
      call deck_error('restart_remap is strict or mode_index')
    if (all(initial_condition /= [character(len=16) :: 'random', 'taylor_green'])) &
      call deck_error('initial_condition is random or taylor_green')
    if (tg_mode < 1) call deck_error('tg_mode counts modes: 1 or more')
    if (sl_amplitude /= 0.0d0) then
      if (s2_amplitude /= 0.0d0 .or. s2_mean /= 0.0d0) &
        call deck_error('a Stokes layer (sl_amplitude) and a spanwise shear (s2_mean, s2_amplitude) exclude each other')
      if (sl_period <= 0.0d0) call deck_error('sl_period of the Stokes layer must be > 0')
    end if
    if ((damp_kpar > 0.0d0) .neqv. (damp_kperp > 0.0d0)) &
      call deck_error('damp_kpar and damp_kperp switch the damping on together (both > 0) or off (both 0)')
    if (damp_norm < 1.0d0) call deck_error('damp_norm (the p of the region''s p-norm) must be >= 1')
    if (damp_edge < 0.0d0 .or. damp_edge >= 1.0d0) call deck_error('damp_edge lies in [0, 1)')
    if (damp_rate < 0.0d0) call deck_error('damp_rate must be >= 0')
  end subroutine validate

  ! The sizes and constants that follow from the deck (src/hst_input.f90:92-103).
  subroutine derive()
    ni = 1.0d0/re
    dt_fixed = deltat
    lx = 2.0d0*PI/alfa0
    lz = 2.0d0*PI/beta0
    nxd = padded_size(3*(nx + 1)/2)
    nzd = padded_size(3*nz)
    dx = PI/(alfa0*nxd)
    dz = 2.0d0*PI/(beta0*nzd)
    dy = ly/ny
    factor = 1.0d0/(2.0d0*nxd*nzd)
  end subroutine derive

  ! the smallest n' >= n with fft_fit(n')
  integer(C_INT) function padded_size(n) result(m)
    integer(C_INT), intent(in) :: n
    m = n
    do while (.not. fft_fit(m))
      m = m + 1
    end do
  end function padded_size

  ! n = 2^k or 3 2^k: the transform lengths the FFTs take
  logical function fft_fit(n)
    integer(C_INT), intent(in) :: n
    integer(C_INT) :: odd
    odd = n
    if (odd < 1) then
      fft_fit = .false.
      return
    end if
    odd = ishft(odd, -trailz(odd))
    fft_fit = (odd == 1 .or. odd == 3)
  end function fft_fit

  subroutine choose_precision()
    character(len=16) :: v
    integer :: n, st
    precision = 0
    call get_environment_variable('HST_PRECISION', v, n, st)
    if (st /= 0 .or. n == 0) return
    select case (trim(v))
    case ('fp64')
      precision = 0
    case ('mixed')
      precision = 1
    case default
      call deck_error('HST_PRECISION (environment) is fp64 or mixed, not '//trim(v))
    end select
  end subroutine choose_precision

  ! The run's parameters, one group per line.
  subroutine print_input()
    write (*, '(/,A)') ' hst: homogeneous shear turbulence, RK3 substep on one MI355X'
    write (*, '(2X,A12,":",3(1X,A,"=",I0),3X,A,"=",I0,1X,A,"=",I0)') 'grid', 'nx', nx, 'ny', ny, 'nz', nz, &
      'nxd', nxd, 'nzd', nzd
    write (*, '(2X,A12,":",3(1X,A,"=",ES12.5),1X,A,"=",F6.3)') 'box', 'lx', lx, 'ly', ly, 'lz', lz, 'ystretch', ystretch
    write (*, '(2X,A12,":",2(1X,A,"=",ES12.5),2(1X,A,"=",L1))') 'flow', 'Re', re, 'S', S, 'linear', linear, &
      'exact_shift', exact_shift
    if (s2_amplitude /= 0.0d0 .or. s2_mean /= 0.0d0) &
      write (*, '(2X,A12,":",4(1X,A,"=",ES12.5))') 'S2(t)', 'mean', s2_mean, 'amplitude', s2_amplitude, &
      'period', s2_period, 'start', s2_start
    if (sl_amplitude /= 0.0d0) &
      write (*, '(2X,A12,":",4(1X,A,"=",ES12.5),2(1X,A,"=",L1))') 'Stokes layer', 'A', sl_amplitude, &
      'T', sl_period, 'delta', sl_delta, 't0', sl_start, 'bodyforce', sl_bodyforce, 'ramp', sl_ramp
    if (damp_kpar > 0.0d0) &
      write (*, '(2X,A12,":",3(1X,A,"=",ES12.5),1X,A,"=",L1,1X,A,"=",ES12.5,2(1X,A,"=",L1),1X,A,"=",F7.2)') &
      'damping', 'k_par', damp_kpar, 'k_perp', damp_kperp, 'p', damp_norm, 'hard', damp_hard, 'rate', damp_rate, &
      'mean', damp_mean, 'follow', damp_follow, 'angle', damp_angle
    write (*, '(2X,A12,":",3(1X,A,"=",ES12.5),1X,A,"=",I0)') 'time step', 'deltat', deltat, 'cflmax', cflmax, &
      't_max', t_max, 'nstep', nstep
    write (*, '(2X,A12,":",4(1X,A,"=",ES12.5))') 'output', 'dt_stat', dt_stat, 'dt_field', dt_field, &
      'dt_save', dt_save, 'wall_max', wall_max
    if (restart_from_file) then
      write (*, '(2X,A12,":",1X,A,1X,A,"=",A,1X,A,"=",L1)') 'start', 'file '//trim(restart_file), 'remap', &
        trim(restart_remap), 'time_from_restart', time_from_restart
    else
      write (*, '(2X,A12,":",1X,A,1X,A,"=",I0,1X,A,"=",ES12.5,1X,A,"=",ES12.5)') 'start', trim(initial_condition), &
        'seed', seed, 'kpeak', kpeak, 'amplitude', amplitude
    end if
    write (*, '(2X,A12,":",1X,A,/)') 'arithmetic', merge('mixed (FP32 transforms, FP64 solves)', &
                                                        'FP64                                ', precision == 1)
  end subroutine print_input

end module hst_input
