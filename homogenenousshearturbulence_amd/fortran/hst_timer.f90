! Time per phase of the run (timing = .true. in &time_control), reported at
! the end under the seven phases of the reference's timer
! (src/hst_timer.f90:26-30).  The GPU phases come from the library's own
! table of kernel times (HIP events around every launch on its stream), so
! they are what the device spent; each kernel is filed under the phase of the
! reference routine it implements.  Kernels fuse routines here, so a fused
! kernel is filed under the first routine it does: the x pass (inverse x
! transforms, products, forward x transforms) under the products, the
! right-hand sides with the default node-phase shift under the products too,
! the exact shift fused with the D2V and ETA solves under the solves.  The
! host's own time (statistics output, snapshots, files) is measured with the
! wall clock around it (tic / toc), after a device synchronisation.
module hst_timer

  use, intrinsic :: iso_c_binding
  use hst_params, only: timing, has_terminal
  use hst_gpu_api
  use hst_fft, only: device_sync

  implicit none
  private
  public :: timer_start, tic, toc, timer_report
  public :: T_TRANSFORM, T_PREPARE, T_PRODUCTS, T_SHIFT, T_SOLVE, T_RECOVER, T_OTHER

  integer, parameter :: T_TRANSFORM = 1, T_PREPARE = 2, T_PRODUCTS = 3, T_SHIFT = 4, T_SOLVE = 5, &
                        T_RECOVER = 6, T_OTHER = 7, NPHASE = 7
  character(len=34), parameter :: phase_name(NPHASE) = [character(len=34) :: &
    'to physical: FFTs, transposes, CFL', 'buildrhs_prepare', 'products: FFTs, transposes, rhs', &
    'shear_shift', 'implicit solves', 'ghosts, dv/dy, u and w', 'statistics, I/O, rest']
  real(C_DOUBLE), save :: host_s(NPHASE) = 0.0d0
  integer(8), save :: mark = 0

contains

  ! start of the timed loop: the kernel table starts empty
  subroutine timer_start()
    if (.not. timing) return
    call gpu_check(hst_gpu_profile(1_C_INT), 'profile on')
  end subroutine timer_start

  subroutine tic()
    if (.not. timing) return
    call device_sync()
    call system_clock(mark)
  end subroutine tic

  ! host time since the last tic/toc into phase p
  subroutine toc(p)
    integer, intent(in) :: p
    integer(8) :: now, rate
    if (.not. timing) return
    call device_sync()
    call system_clock(now, rate)
    host_s(p) = host_s(p) + real(now - mark, C_DOUBLE)/real(rate, C_DOUBLE)
    mark = now
  end subroutine toc

  ! the reference phase of a kernel of the library's table
  integer function phase_of(kernel)
    character(len=*), intent(in) :: kernel
    select case (kernel)
    case ('zift', 'xpass_cfl', 'uw_from_eta')
      phase_of = T_TRANSFORM
    case ('buildrhs_prepare', 'damping_table', 'stokes_force')
      phase_of = T_PREPARE
    case ('xpass', 'zfft', 'buildrhs', 'buildrhs_group', 'closure')
      phase_of = T_PRODUCTS
    case ('shift_phases', 'exact_shift', 'exact_shift_old', 'phase_shift', 'damping_mask')
      phase_of = T_SHIFT
    case ('shift_penta', 'penta_d2v_eta', 'penta', 'penta_factor', 'd0_solve', 'dy_solve')
      phase_of = T_SOLVE
    case ('ghosts', 'dvdy', 'dy_recover', 'stokes_apply')
      phase_of = T_RECOVER
    case default
      phase_of = T_OTHER
    end select
  end function phase_of

  subroutine timer_report(nsteps)
    integer(C_SIZE_T), intent(in) :: nsteps
    integer(C_INT) :: n, i, k
    integer :: p
    integer(C_INT64_T) :: launches
    real(C_DOUBLE) :: ms, bytes, total, per(NPHASE)
    character(kind=c_char) :: cname(64)
    character(len=64) :: kname
    character(len=400) :: members(NPHASE)
    if (.not. (timing .and. has_terminal)) return
    call device_sync()
    per = host_s
    members = ''
    n = hst_gpu_profile_count()
    do i = 0, n - 1
      call gpu_check(hst_gpu_profile_entry(i, cname, 64_C_INT, launches, ms, bytes), 'profile entry')
      kname = ''
      do k = 1, 64
        if (cname(k) == c_null_char) exit
        kname(k:k) = cname(k)
      end do
      p = phase_of(trim(kname))
      per(p) = per(p) + 1.0d-3*ms
      members(p) = trim(members(p))//' '//trim(kname)
    end do
    total = sum(per)
    write (*, '(A)') ' '
    write (*, '(A,I0,A)') '   time per phase over ', nsteps, ' steps (GPU kernels by HIP events; host work by the clock):'
    do p = 1, NPHASE
      write (*, '(A,A,F10.3,A,F9.5,A,F6.1,A,A)') '     ', phase_name(p), per(p), ' s', &
        per(p)/max(nsteps, 1_C_SIZE_T), ' s/step', 100.0d0*per(p)/max(total, tiny(total)), ' %  ', trim(members(p))
    end do
    write (*, '(A,A,F10.3,A,F9.5,A)') '     ', 'total                             ', total, ' s', &
      total/max(nsteps, 1_C_SIZE_T), ' s/step'
    call gpu_check(hst_gpu_profile(0_C_INT), 'profile off')
  end subroutine timer_report

end module hst_timer
