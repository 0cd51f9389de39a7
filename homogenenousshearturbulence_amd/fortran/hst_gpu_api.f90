! The iso_c_binding interface of libhstgpu.so (include/hst_gpu.h): the deck
! and grid structs and every entry point.  The wrappers of the other modules
! stop the run on a non-zero status with the library's message, as every error
! of the reference is fatal.
module hst_gpu_api

  use, intrinsic :: iso_c_binding

  implicit none

  type, bind(c) :: hst_deck_t
    integer(c_int32_t) :: nx, ny, nz
    integer(c_int32_t) :: linear, exact_shift
    integer(c_int32_t) :: seed, tg_mode
    integer(c_int32_t) :: initial_condition
    real(c_double) :: alfa0, beta0, ly, ystretch
    real(c_double) :: re, S, s2_mean, s2_amplitude, s2_period, s2_start
    real(c_double) :: amplitude, kpeak, tg_amplitude
    real(c_double) :: sl_amplitude
    real(c_double) :: damp_kpar, damp_kperp
    real(c_double) :: sl_period, sl_delta, sl_start
    real(c_double) :: damp_rate, damp_edge, damp_norm, damp_angle
    integer(c_int32_t) :: sl_bodyforce, sl_ramp
    integer(c_int32_t) :: damp_hard, damp_mean, damp_follow
  end type hst_deck_t

  type, bind(c) :: hst_grid_info_t
    integer(c_int32_t) :: nxd, nzd, nxm, nzm, nyg, ncol
    real(c_double) :: lx, lz, dx, dy, dz, factor, ni
  end type hst_grid_info_t

  integer(c_int), parameter :: HST_V1 = 1, HST_V2 = 2, HST_V3 = 3, HST_RHS1 = 4, HST_RHS2 = 5, &
                               HST_OLD1 = 6, HST_OLD2 = 7, HST_SCRATCH = 8
  integer(c_int), parameter :: HST_FP64 = 0, HST_MIXED = 1

  interface
    integer(c_int) function hst_grid_setup(deck, info, y, dyl, der) bind(c, name='hst_grid_setup')
      import :: c_int, c_double, hst_deck_t, hst_grid_info_t
      type(hst_deck_t), intent(in) :: deck
      type(hst_grid_info_t), intent(out) :: info
      real(c_double), intent(out) :: y(*), dyl(*), der(*)
    end function
    integer(c_int) function hst_initial_field(deck, V) bind(c, name='hst_initial_field')
      import :: c_int, c_double_complex, hst_deck_t
      type(hst_deck_t), intent(in) :: deck
      complex(c_double_complex), intent(out) :: V(*)
    end function
    integer(c_int) function hst_initial_energy(deck, q2) bind(c, name='hst_initial_energy')
      import :: c_int, c_double, hst_deck_t
      type(hst_deck_t), intent(in) :: deck
      real(c_double), intent(out) :: q2
    end function
    integer(c_int) function hst_gpu_init(deck, precision) bind(c, name='hst_gpu_init')
      import :: c_int, hst_deck_t
      type(hst_deck_t), intent(in) :: deck
      integer(c_int), value :: precision
    end function
    integer(c_int) function hst_gpu_free() bind(c, name='hst_gpu_free')
      import :: c_int
    end function
    integer(c_int) function hst_gpu_sync() bind(c, name='hst_gpu_sync')
      import :: c_int
    end function
    integer(c_int) function hst_gpu_last_error(buf, len) bind(c, name='hst_gpu_last_error')
      import :: c_int, c_char
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_int), value :: len
    end function
    integer(c_int) function hst_gpu_grid(info, y, dyl, der) bind(c, name='hst_gpu_grid')
      import :: c_int, c_double, hst_grid_info_t
      type(hst_grid_info_t), intent(out) :: info
      real(c_double), intent(out) :: y(*), dyl(*), der(*)
    end function
    integer(c_int) function hst_gpu_upload_field(field, host) bind(c, name='hst_gpu_upload_field')
      import :: c_int, c_double_complex
      integer(c_int), value :: field
      complex(c_double_complex), intent(in) :: host(*)
    end function
    integer(c_int) function hst_gpu_download_field(field, host) bind(c, name='hst_gpu_download_field')
      import :: c_int, c_double_complex
      integer(c_int), value :: field
      complex(c_double_complex), intent(out) :: host(*)
    end function
    integer(c_int) function hst_gpu_zero_field(field) bind(c, name='hst_gpu_zero_field')
      import :: c_int
      integer(c_int), value :: field
    end function
    integer(c_int) function hst_gpu_fill_ghosts(field, sx, sz) bind(c, name='hst_gpu_fill_ghosts')
      import :: c_int, c_double
      integer(c_int), value :: field
      real(c_double), value :: sx, sz
    end function
    integer(c_int) function hst_gpu_nonlinear(cfl) bind(c, name='hst_gpu_nonlinear')
      import :: c_int, c_double
      real(c_double), intent(out) :: cfl
    end function
    integer(c_int) function hst_gpu_compute_cfl(cfl) bind(c, name='hst_gpu_compute_cfl')
      import :: c_int, c_double
      real(c_double), intent(out) :: cfl
    end function
    integer(c_int) function hst_gpu_product_to_field(p, field) bind(c, name='hst_gpu_product_to_field')
      import :: c_int
      integer(c_int), value :: p, field
    end function
    integer(c_int) function hst_gpu_roundtrip() bind(c, name='hst_gpu_roundtrip')
      import :: c_int
    end function
    integer(c_int) function hst_gpu_buildrhs_prepare(ode, deltat, time) bind(c, name='hst_gpu_buildrhs_prepare')
      import :: c_int, c_double
      real(c_double), intent(in) :: ode(3)
      real(c_double), value :: deltat, time
    end function
    integer(c_int) function hst_gpu_buildrhs(ode, g) bind(c, name='hst_gpu_buildrhs')
      import :: c_int, c_double
      real(c_double), intent(in) :: ode(3)
      integer(c_int), value :: g
    end function
    integer(c_int) function hst_gpu_shear_shift(dt_sub, time) bind(c, name='hst_gpu_shear_shift')
      import :: c_int, c_double
      real(c_double), value :: dt_sub, time
    end function
    integer(c_int) function hst_gpu_line_solve(kind, lambda, src, dst, sx, sz) bind(c, name='hst_gpu_line_solve')
      import :: c_int, c_double
      integer(c_int), value :: kind, src, dst
      real(c_double), value :: lambda, sx, sz
    end function
    integer(c_int) function hst_gpu_linsolve(lambda, time) bind(c, name='hst_gpu_linsolve')
      import :: c_int, c_double
      real(c_double), value :: lambda, time
    end function
    integer(c_int) function hst_gpu_timestep(deltat, time, cfl) bind(c, name='hst_gpu_timestep')
      import :: c_int, c_double
      real(c_double), value :: deltat
      real(c_double), intent(inout) :: time
      real(c_double), intent(out) :: cfl
    end function
    integer(c_int) function hst_gpu_outstats(time, out) bind(c, name='hst_gpu_outstats')
      import :: c_int, c_double
      real(c_double), value :: time
      real(c_double), intent(out) :: out(10)
    end function
    integer(c_int) function hst_gpu_stokes_apply(time) bind(c, name='hst_gpu_stokes_apply')
      import :: c_int, c_double
      real(c_double), value :: time
    end function
    integer(c_int) function hst_gpu_project_divergence(time) bind(c, name='hst_gpu_project_divergence')
      import :: c_int, c_double
      real(c_double), value :: time
    end function
    integer(c_int) function hst_gpu_damping_mask() bind(c, name='hst_gpu_damping_mask')
      import :: c_int
    end function
    integer(c_int) function hst_gpu_outstats_extra(out) bind(c, name='hst_gpu_outstats_extra')
      import :: c_int, c_double
      real(c_double), intent(out) :: out(8)
    end function
    integer(c_int) function hst_gpu_pressure(time, dst) bind(c, name='hst_gpu_pressure')
      import :: c_int, c_double
      real(c_double), value :: time
      integer(c_int), value :: dst
    end function
    integer(c_int) function hst_gpu_spectrum(E, nacc) bind(c, name='hst_gpu_spectrum')
      import :: c_int, c_int32_t, c_double
      real(c_double), intent(out) :: E(*)
      integer(c_int32_t), intent(out) :: nacc
    end function
    integer(c_int) function hst_gpu_set_chunk_rows(rows) bind(c, name='hst_gpu_set_chunk_rows')
      import :: c_int
      integer(c_int), value :: rows
    end function
    integer(c_int) function hst_gpu_memory(bytes, chunk) bind(c, name='hst_gpu_memory')
      import :: c_int, c_int64_t
      integer(c_int64_t), intent(out) :: bytes
      integer(c_int), intent(out) :: chunk
    end function
    ! per-field transforms of the reference's buffers (host arrays)
    integer(c_int) function hst_gpu_fft_z(lines, nlines, sign) bind(c, name='hst_gpu_fft_z')
      import :: c_int, c_int64_t, c_double_complex
      complex(c_double_complex), intent(inout) :: lines(*)
      integer(c_int64_t), value :: nlines
      integer(c_int), value :: sign
    end function
    integer(c_int) function hst_gpu_fft_x(lines, nlines, dir) bind(c, name='hst_gpu_fft_x')
      import :: c_int, c_int64_t, c_double_complex
      complex(c_double_complex), intent(inout) :: lines(*)
      integer(c_int64_t), value :: nlines
      integer(c_int), value :: dir
    end function
    integer(c_int) function hst_gpu_products(g, nlines, vel, prod) bind(c, name='hst_gpu_products')
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: g
      integer(c_int64_t), value :: nlines
      real(c_double), intent(in) :: vel(*)
      real(c_double), intent(inout) :: prod(*)
    end function
    integer(c_int) function hst_gpu_set_products(g, vvdz) bind(c, name='hst_gpu_set_products')
      import :: c_int, c_double_complex
      integer(c_int), value :: g
      complex(c_double_complex), intent(in) :: vvdz(*)
    end function
    ! per-kernel timing (HIP events on the library's stream)
    integer(c_int) function hst_gpu_profile(enable) bind(c, name='hst_gpu_profile')
      import :: c_int
      integer(c_int), value :: enable
    end function
    integer(c_int) function hst_gpu_profile_count() bind(c, name='hst_gpu_profile_count')
      import :: c_int
    end function
    integer(c_int) function hst_gpu_profile_entry(i, name, len, launches, ms, bytes) &
      bind(c, name='hst_gpu_profile_entry')
      import :: c_int, c_char, c_int64_t, c_double
      integer(c_int), value :: i
      character(kind=c_char), intent(out) :: name(*)
      integer(c_int), value :: len
      integer(c_int64_t), intent(out) :: launches
      real(c_double), intent(out) :: ms, bytes
    end function
  end interface

contains

  ! stop with the library's message on a non-zero status (the reference's `check`)
  subroutine gpu_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char) :: buf(512)
    character(len=512) :: msg
    integer :: i, ierr
    if (rc == 0) return
    ierr = hst_gpu_last_error(buf, 512_c_int)
    msg = ' '
    do i = 1, 512
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    end do
    print '(A,A,A,I0,A)', ' ERROR: ', where, ' failed (status ', rc, '): '//trim(msg)
    error stop 1
  end subroutine gpu_check

end module hst_gpu_api
