! Runtime statistics in the layout of the CPL code (the reference's module
! hst_stats): statistics/Runtimedata and statistics/variances_runtime.dat,
! opened for append.  The reductions run on the GPU (hst_gpu_outstats).
module hst_stats

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api
  use hst_derivatives, only: s2_of, gamma_y_of

  implicit none
  private
  public :: outstats, open_runtimedata, close_runtimedata

  integer, parameter :: unit_rt = 121, unit_var = 122

contains

  subroutine open_runtimedata()
    if (has_terminal) open (unit=unit_rt, file='statistics/Runtimedata', action='write', position='append')
    if (has_terminal) open (unit=unit_var, file='statistics/variances_runtime.dat', action='write', position='append')
  end subroutine open_runtimedata

  subroutine close_runtimedata()
    if (has_terminal) close (unit_rt)
    if (has_terminal) close (unit_var)
  end subroutine close_runtimedata

  ! energy, diss, uv, vw, uu, vv, ww, uw (box means) and the profile integrals mfx, mfz
  subroutine outstats()
    real(C_DOUBLE) :: o(10)
    call gpu_check(hst_gpu_outstats(time, o), 'outstats')
    if (has_terminal) then
      write (*, '(F12.5,2X,ES11.4,2X,F8.4,4(2X,ES13.6))') time, deltat, cfl*deltat, o(1), o(2), o(3), o(4)
      write (unit_rt, '(13(ES23.15,1X))') time, o(9), o(10), S, s2_of(time), S*time, gamma_y_of(time), deltat, &
        cfl*deltat, o(1), o(2), o(3), o(4)
      write (unit_var, '(5(ES23.15,1X))') time, o(5), o(6), o(7), o(8)
      flush (unit_rt); flush (unit_var)
    end if
  end subroutine outstats

end module hst_stats
