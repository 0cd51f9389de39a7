! The transform buffers and the per-field transforms of the reference's
! module hst_fft (src/hst_fft.f90): complex z transforms of VVdz, real x
! transforms between VVdx and its real view rVVdx, and between the real
! products and VVdp, all three fields of a buffer at a time or one by one.
!
!   VVdz(nzd, nx+1, -2:ny+1, 3)          z lines, complex
!   VVdx(nxd+1, nzd, -2:ny+1, 3)         x modes ...  rVVdx(2(nxd+1), nzd, -2:ny+1, 3) the same bytes as reals
!   VVdp, products                       the same pair for three products
!
! The buffers are host arrays; FFT(m), IFT(m), RFT(m) and HFT(m) run the
! transform of field m on the GPU (hst_gpu_fft_z, hst_gpu_fft_x: FP64,
! unnormalised, in place) and bring the lines back.  The time step does not
! use these: its transforms are stages of the fused kernels behind
! hst_gpu_timestep.  With the state resident on the GPU (the driver) the
! buffers are not allocated.
module hst_fft

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api

  implicit none
  private
  public :: init_fft, free_fft, FFT, IFT, RFT, HFT, device_sync, fields_by_field
  public :: VVdz, VVdx, rVVdx, VVdp, products

  complex(C_DOUBLE_COMPLEX), allocatable, target, save :: VVdz(:, :, :, :), VVdx(:, :, :, :), VVdp(:, :, :, :)
  real(C_DOUBLE), pointer, save :: rVVdx(:, :, :, :) => null(), products(:, :, :, :) => null()

contains

  subroutine init_fft()
    real(C_DOUBLE), pointer :: flat(:, :, :, :)
    if (state_resident) return
    allocate (VVdz(nzd, nx + 1, -2:ny + 1, 3))
    allocate (VVdx(nxd + 1, nzd, -2:ny + 1, 3), VVdp(nxd + 1, nzd, -2:ny + 1, 3))
    VVdz = 0
    VVdx = 0
    VVdp = 0
    ! the real views: 2(nxd+1) reals per x line, rows -2..ny+1
    call c_f_pointer(c_loc(VVdx), flat, [2*(nxd + 1), nzd, ny + 4, 3])
    rVVdx(1:, 1:, -2:, 1:) => flat
    call c_f_pointer(c_loc(VVdp), flat, [2*(nxd + 1), nzd, ny + 4, 3])
    products(1:, 1:, -2:, 1:) => flat
  end subroutine init_fft

  subroutine free_fft()
    nullify (rVVdx, products)
    if (allocated(VVdz)) deallocate (VVdz, VVdx, VVdp)
  end subroutine free_fft

  ! .true. when the per-field buffers exist (the state is not resident)
  logical function fields_by_field()
    fields_by_field = allocated(VVdz)
  end function fields_by_field

  subroutine need_buffers(name)
    character(len=*), intent(in) :: name
    if (.not. allocated(VVdz)) then
      print *, 'ERROR: ', name, ': the transform buffers are not allocated (init_fft after allocate_fields(), ', &
        'not with a resident state)'
      error stop 1
    end if
  end subroutine need_buffers

  ! Wait for everything queued on the device.
  subroutine device_sync()
    call gpu_check(hst_gpu_sync(), 'hst_gpu_sync')
  end subroutine device_sync

  ! z transforms of field m of VVdz: forward (FFT, exp(-i ...)) and backward (IFT)
  subroutine FFT(m)
    integer(C_INT), intent(in) :: m
    call need_buffers('FFT')
    call gpu_check(hst_gpu_fft_z(VVdz(:, :, :, m), int((nx + 1)*(ny + 4), C_INT64_T), -1_C_INT), 'FFT')
  end subroutine FFT

  subroutine IFT(m)
    integer(C_INT), intent(in) :: m
    call need_buffers('IFT')
    call gpu_check(hst_gpu_fft_z(VVdz(:, :, :, m), int((nx + 1)*(ny + 4), C_INT64_T), 1_C_INT), 'IFT')
  end subroutine IFT

  ! x modes of field m of VVdx -> its real points rVVdx (RFT, c2r), and the
  ! real products of field m -> their modes VVdp (HFT, r2c)
  subroutine RFT(m)
    integer(C_INT), intent(in) :: m
    call need_buffers('RFT')
    call gpu_check(hst_gpu_fft_x(VVdx(:, :, :, m), int(nzd*(ny + 4), C_INT64_T), 1_C_INT), 'RFT')
  end subroutine RFT

  subroutine HFT(m)
    integer(C_INT), intent(in) :: m
    call need_buffers('HFT')
    call gpu_check(hst_gpu_fft_x(VVdp(:, :, :, m), int(nzd*(ny + 4), C_INT64_T), -1_C_INT), 'HFT')
  end subroutine HFT

end module hst_fft
