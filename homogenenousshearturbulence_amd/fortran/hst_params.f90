! Names of the run, as the reference's module hst_params declares them
! (src/hst_params.f90): mesh, physics, clock, the one-rank layout, derived
! grid arrays, the solution arrays and the RK3 table, so that code written
! against the reference's modules compiles against these.
!
! Where the state lives.  libhstgpu holds the fields on the GPU in its own
! layout.  The host arrays V, rhs and oldrhs below, in the reference's layout
! V(iy, iz, ix, c), are the state a caller sees:
!   * by default (allocate_fields()) every module routine takes what it reads
!     from these arrays to the GPU and brings back what it writes, so a
!     caller may read and write them between any two calls, as the reference
!     allows between its `target update` directives (which compile to nothing
!     here: callers are built without OpenMP offload);
!   * with allocate_fields(resident=.true.) (the driver hst) the state stays
!     on the GPU between calls and V is brought to the host only where the
!     driver asks for it (I/O); rhs and oldrhs are not allocated.
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
  !---------------------------------------------------------- rank layout ----
  ! One process, one GPU, the whole grid: x modes 0..nx, z points 1..nzd,
  ! rows 0..ny-1 (setup_decomposition fills these in).
  integer(C_INT), save :: nproc = 1, iproc = 0, npxz = 1, ipxz = 0, ipy = 0
  integer(C_INT), save :: nx0, nxN, nxB, nz0, nzN, nzB, ny0, nyN, nyB
  logical, save :: has_terminal = .true., has_average = .true.
  !--------------------------------------------------- derived grid data ----
  real(C_DOUBLE), allocatable, save :: y(:)            ! y(-2:ny+1)
  real(C_DOUBLE), allocatable, save :: dyl(:)          ! dyl(0:ny-1)
  real(C_DOUBLE), allocatable, save :: der(:, :, :)    ! der(0:ny-1, 0:3, -2:2), the library's weights
  real(C_DOUBLE), allocatable, save :: k2(:, :)        ! k2(iz, ix) = alfa^2 + beta^2
  complex(C_DOUBLE_COMPLEX), allocatable, save :: ialfa(:), ibeta(:)   ! ialfa(ix), ibeta(iz)
  integer(C_INT), allocatable, save :: izd(:)          ! izd(iz): z mode -> index of the padded z line, from 0
  real(C_DOUBLE), save :: dx, dy, dz, factor
  !---------------------------------------------------- solution arrays ----
  complex(C_DOUBLE_COMPLEX), allocatable, target, save :: V(:, :, :, :)        ! V(-2:ny+1, -nz:nz, 0:nx, 3)
  complex(C_DOUBLE_COMPLEX), allocatable, target, save :: rhs(:, :, :, :)      ! rhs(..., 2): eta, d2v
  complex(C_DOUBLE_COMPLEX), allocatable, target, save :: oldrhs(:, :, :, :)   ! oldrhs(..., 2)
  logical, save :: state_resident = .false.   ! .true.: the state stays on the GPU between calls
  ! arithmetic of the transforms (environment HST_PRECISION = fp64 | mixed)
  integer(C_INT), save :: precision = 0

  !--------------------------------------- Runge-Kutta (Rai-Moin) table ----
  ! column i = substep i: (unknown/deltat, new explicit, old explicit)
  real(C_DOUBLE), dimension(3, 3), save :: RK_rai = reshape( &
      (/120.0d0/32.0d0, 2.0d0, 0.0d0, &
        120.0d0/8.0d0, 50.0d0/8.0d0, 34.0d0/8.0d0, &
        120.0d0/20.0d0, 90.0d0/20.0d0, 50.0d0/20.0d0/), shape=(/3, 3/))

end module hst_params
