! Everything the run knows on the host: mesh, physics, clock, derived grid
! data and the host copy of the velocity (for I/O).  The fields themselves
! live on the GPU (libhstgpu); a field is named by its id (hst_gpu_api):
!   HST_V1..HST_V3 = V(:, :, :, 1:3), HST_RHS1/2 = rhs(:, :, :, 1:2),
!   HST_OLD1/2 = oldrhs(:, :, :, 1:2), HST_SCRATCH = a spare field.
! The names and meanings follow the reference's module hst_params.
module hst_params

  use, intrinsic :: iso_c_binding

  implicit none

  !---------------------------------------------------------------- mesh ----
  integer(C_INT), save :: nx, ny, nz, nxd, nzd
  real(C_DOUBLE), save :: alfa0, beta0, lx, ly, lz, ystretch
  integer(C_INT), save :: line_chunk, npy
  character(len=8), save :: transport
  !------------------------------------------------------------- physics ----
  real(C_DOUBLE), save :: re, ni, S
  logical, save :: linear, exact_shift
  real(C_DOUBLE), save :: s2_mean, s2_amplitude, s2_period, s2_start
  real(C_DOUBLE), save :: sl_amplitude, sl_period, sl_delta, sl_start
  logical, save :: sl_bodyforce, sl_ramp
  real(C_DOUBLE), save :: damp_kpar, damp_kperp, damp_rate, damp_edge, damp_norm, damp_angle
  logical, save :: damp_hard, damp_mean, damp_follow
  !------------------------------------------------------- clock and I/O ----
  real(C_DOUBLE), save :: deltat, dt_fixed, cflmax, cfl = 0.0d0
  real(C_DOUBLE), save :: t_max, dt_stat, dt_field, dt_save, wall_max
  real(C_DOUBLE), save :: time, time0 = 0.0d0
  integer(C_SIZE_T), save :: nstep, istep = 0, ifield = 0
  logical, save :: time_from_restart, restart_from_file, timing
  character(len=16), save :: restart_remap
  character(len=256), save :: restart_file
  !------------------------------------------------------ initial field ----
  real(C_DOUBLE), save :: amplitude, kpeak, tg_amplitude
  integer(C_INT), save :: seed, tg_mode
  character(len=16), save :: initial_condition
  !--------------------------------------------------- derived grid data ----
  real(C_DOUBLE), allocatable, save :: y(:), dyl(:), der(:, :, :)   ! y(-2:ny+1), dyl(0:ny-1), der(0:ny-1, 0:3, -2:2)
  real(C_DOUBLE), save :: dx, dy, dz, factor
  ! host copy of the velocity, reference layout V(iy, iz, ix, c)
  complex(C_DOUBLE_COMPLEX), allocatable, target, save :: V(:, :, :, :)
  ! arithmetic of the transforms (environment HST_PRECISION = fp64 | mixed)
  integer(C_INT), save :: precision = 0
  logical, save :: has_terminal = .true.

  !--------------------------------------- Runge-Kutta (Rai-Moin) table ----
  real(C_DOUBLE), dimension(3, 3), save :: RK_rai = reshape( &
      (/120.0d0/32.0d0, 2.0d0, 0.0d0, &
        120.0d0/8.0d0, 50.0d0/8.0d0, 34.0d0/8.0d0, &
        120.0d0/20.0d0, 90.0d0/20.0d0, 50.0d0/20.0d0/), shape=(/3, 3/))

end module hst_params
