! Set-up of the device state and of the host arrays (the reference's module
! hst_setup: setup_decomposition, allocate_fields, free_fields), and the
! moves of the state between the host arrays and the GPU that every other
! module makes around its device calls.
module hst_setup

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api

  implicit none
  private
  public :: select_device, setup_decomposition, allocate_fields, free_fields, deck_struct
  public :: state_in, state_out, download_velocity, upload_velocity, to_device, from_device
  public :: S_V, S_RHS, S_OLD

  ! which of the state arrays a call reads (state_in) or writes (state_out)
  integer, parameter :: S_V = 1, S_RHS = 2, S_OLD = 4

contains

  ! One process, one GPU: the device HIP makes current (HIP_VISIBLE_DEVICES
  ! chooses it); the reference's per-rank device choice has nothing to do.
  subroutine select_device()
  end subroutine select_device

  ! The one-rank layout: every x mode, z point and row on this process.
  subroutine setup_decomposition()
    nx0 = 0; nxN = nx; nxB = nx + 1
    nz0 = 1; nzN = nzd; nzB = nzd
    ny0 = 0; nyN = ny - 1; nyB = ny
    nproc = 1; iproc = 0; npxz = 1; ipxz = 0; npy = 1; ipy = 0
    has_terminal = .true.; has_average = .true.
  end subroutine setup_decomposition

  function deck_struct() result(d)
    type(hst_deck_t) :: d
    d%nx = nx; d%ny = ny; d%nz = nz
    d%linear = merge(1, 0, linear); d%exact_shift = merge(1, 0, exact_shift)
    d%seed = seed; d%tg_mode = tg_mode
    d%initial_condition = merge(1, 0, trim(initial_condition) == 'taylor_green')
    d%alfa0 = alfa0; d%beta0 = beta0; d%ly = ly; d%ystretch = ystretch
    ! the library takes the viscosity as 1/re: a caller that set ni itself
    ! (the reference's tests do, after read_input) gets that ni
    d%re = re
    if (ni /= 1.0d0/re) d%re = 1.0d0/ni
    d%S = S; d%s2_mean = s2_mean; d%s2_amplitude = s2_amplitude; d%s2_period = s2_period
    d%s2_start = s2_start
    d%amplitude = amplitude; d%kpeak = kpeak; d%tg_amplitude = tg_amplitude
    d%sl_amplitude = sl_amplitude; d%damp_kpar = damp_kpar; d%damp_kperp = damp_kperp
    d%sl_period = sl_period; d%sl_delta = sl_delta; d%sl_start = sl_start
    d%sl_bodyforce = merge(1, 0, sl_bodyforce); d%sl_ramp = merge(1, 0, sl_ramp)
    d%damp_rate = damp_rate; d%damp_edge = damp_edge; d%damp_norm = damp_norm; d%damp_angle = damp_angle
    d%damp_hard = merge(1, 0, damp_hard); d%damp_mean = merge(1, 0, damp_mean)
    d%damp_follow = merge(1, 0, damp_follow)
  end function deck_struct

  ! The device state (zeroed), the grid tables as the library computed them
  ! (y, dyl, der: the same bits), the wavenumber tables, and the host arrays:
  ! V always; rhs and oldrhs unless the state is resident on the GPU.
  subroutine allocate_fields(resident)
    logical, intent(in), optional :: resident
    type(hst_deck_t) :: d
    type(hst_grid_info_t) :: info
    integer(C_INT) :: ix, iz
    state_resident = .false.
    if (present(resident)) state_resident = resident
    call setup_decomposition()
    d = deck_struct()
    call gpu_check(hst_gpu_init(d, precision), 'hst_gpu_init')
    allocate (y(-2:ny + 1), dyl(0:ny - 1), der(0:ny - 1, 0:3, -2:2))
    call gpu_check(hst_gpu_grid(info, y, dyl, der), 'hst_gpu_grid')
    allocate (k2(-nz:nz, 0:nx), ialfa(0:nx), ibeta(-nz:nz), izd(-nz:nz))
    do ix = 0, nx
      ialfa(ix) = cmplx(0.0d0, alfa0*ix, C_DOUBLE)
    end do
    do iz = -nz, nz
      ibeta(iz) = cmplx(0.0d0, beta0*iz, C_DOUBLE)
      izd(iz) = modulo(iz, nzd)
    end do
    do ix = 0, nx
      do iz = -nz, nz
        k2(iz, ix) = (alfa0*ix)**2 + (beta0*iz)**2
      end do
    end do
    allocate (V(-2:ny + 1, -nz:nz, 0:nx, 1:3))
    V = 0
    if (.not. state_resident) then
      allocate (rhs(-2:ny + 1, -nz:nz, 0:nx, 1:2), oldrhs(-2:ny + 1, -nz:nz, 0:nx, 1:2))
      rhs = 0
      oldrhs = 0
    end if
  end subroutine allocate_fields

  subroutine free_fields()
    integer(C_INT) :: rc
    rc = hst_gpu_free()
    if (allocated(V)) deallocate (V)
    if (allocated(rhs)) deallocate (rhs)
    if (allocated(oldrhs)) deallocate (oldrhs)
    if (allocated(y)) deallocate (y, dyl, der)
    if (allocated(k2)) deallocate (k2, ialfa, ibeta, izd)
  end subroutine free_fields

  ! One host array of the field shape to device field fid, and back.
  subroutine to_device(fid, a)
    integer(C_INT), intent(in) :: fid
    complex(C_DOUBLE_COMPLEX), intent(in) :: a(-2:ny + 1, -nz:nz, 0:nx)
    call gpu_check(hst_gpu_upload_field(fid, a), 'upload field')
  end subroutine to_device

  subroutine from_device(fid, a)
    integer(C_INT), intent(in) :: fid
    complex(C_DOUBLE_COMPLEX), intent(inout) :: a(-2:ny + 1, -nz:nz, 0:nx)
    call gpu_check(hst_gpu_download_field(fid, a), 'download field')
  end subroutine from_device

  subroutine upload_velocity()
    integer(C_INT) :: c
    do c = 1, 3
      call to_device(HST_V1 + c - 1, V(:, :, :, c))
    end do
  end subroutine upload_velocity

  subroutine download_velocity()
    integer(C_INT) :: c
    do c = 1, 3
      call from_device(HST_V1 + c - 1, V(:, :, :, c))
    end do
  end subroutine download_velocity

  ! Before a device call: the state arrays it reads (a sum of S_V, S_RHS,
  ! S_OLD) to the GPU.  Nothing moves when the state is resident.
  subroutine state_in(which)
    integer, intent(in) :: which
    integer(C_INT) :: k
    if (state_resident) return
    if (iand(which, S_V) /= 0) call upload_velocity()
    do k = 1, 2
      if (iand(which, S_RHS) /= 0) call to_device(HST_RHS1 + k - 1, rhs(:, :, :, k))
      if (iand(which, S_OLD) /= 0) call to_device(HST_OLD1 + k - 1, oldrhs(:, :, :, k))
    end do
  end subroutine state_in

  ! After a device call: the state arrays it wrote back to the host arrays.
  subroutine state_out(which)
    integer, intent(in) :: which
    integer(C_INT) :: k
    if (state_resident) return
    if (iand(which, S_V) /= 0) call download_velocity()
    do k = 1, 2
      if (iand(which, S_RHS) /= 0) call from_device(HST_RHS1 + k - 1, rhs(:, :, :, k))
      if (iand(which, S_OLD) /= 0) call from_device(HST_OLD1 + k - 1, oldrhs(:, :, :, k))
    end do
  end subroutine state_out

end module hst_setup
