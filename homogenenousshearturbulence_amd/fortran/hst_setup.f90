! Allocation of the device state and the host copy of the velocity; the deck
! handed to the library (allocate_fields / free_fields of the reference's
! hst_setup, where the fields are mapped to the device once).
module hst_setup

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api

  implicit none
  private
  public :: allocate_fields, free_fields, select_device, deck_struct

contains

  ! One process, one GPU: HIP's current device (HIP_VISIBLE_DEVICES picks it).
  subroutine select_device()
  end subroutine select_device

  function deck_struct() result(d)
    type(hst_deck_t) :: d
    d%nx = nx; d%ny = ny; d%nz = nz
    d%linear = merge(1, 0, linear); d%exact_shift = merge(1, 0, exact_shift)
    d%seed = seed; d%tg_mode = tg_mode
    d%initial_condition = merge(1, 0, trim(initial_condition) == 'taylor_green')
    d%alfa0 = alfa0; d%beta0 = beta0; d%ly = ly; d%ystretch = ystretch
    d%re = re; d%S = S; d%s2_mean = s2_mean; d%s2_amplitude = s2_amplitude; d%s2_period = s2_period
    d%s2_start = s2_start
    d%amplitude = amplitude; d%kpeak = kpeak; d%tg_amplitude = tg_amplitude
    d%sl_amplitude = sl_amplitude; d%damp_kpar = damp_kpar; d%damp_kperp = damp_kperp
    d%sl_period = sl_period; d%sl_delta = sl_delta; d%sl_start = sl_start
    d%sl_bodyforce = merge(1, 0, sl_bodyforce); d%sl_ramp = merge(1, 0, sl_ramp)
    d%damp_rate = damp_rate; d%damp_edge = damp_edge; d%damp_norm = damp_norm; d%damp_angle = damp_angle
    d%damp_hard = merge(1, 0, damp_hard); d%damp_mean = merge(1, 0, damp_mean)
    d%damp_follow = merge(1, 0, damp_follow)
  end function deck_struct

  ! The device state (fields, transform and solver buffers, tables), zeroed,
  ! and the host grid y, dyl, der as the library computed them.
  subroutine allocate_fields()
    type(hst_deck_t) :: d
    type(hst_grid_info_t) :: info
    d = deck_struct()
    call gpu_check(hst_gpu_init(d, precision), 'hst_gpu_init')
    allocate (y(-2:ny + 1), dyl(0:ny - 1), der(0:ny - 1, 0:3, -2:2))
    call gpu_check(hst_gpu_grid(info, y, dyl, der), 'hst_gpu_grid')
    allocate (V(-2:ny + 1, -nz:nz, 0:nx, 1:3)); V = 0
  end subroutine allocate_fields

  subroutine free_fields()
    integer(c_int) :: rc
    rc = hst_gpu_free()
    if (allocated(V)) deallocate (V)
    if (allocated(y)) deallocate (y, dyl, der)
  end subroutine free_fields

end module hst_setup
