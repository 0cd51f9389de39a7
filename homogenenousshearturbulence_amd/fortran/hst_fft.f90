! The transform layer of the reference (module hst_fft): on the GPU the z and
! x transforms of a field are not separate library calls but stages of the
! fused kernels behind hst_transforms (z inverse with the padding, x inverse
! with the products and x forward, z forward with the truncation), so the
! per-field FFT, IFT, RFT and HFT entry points stop with that explanation.
module hst_fft

  use, intrinsic :: iso_c_binding
  use hst_gpu_api

  implicit none
  private
  public :: init_fft, free_fft, FFT, IFT, RFT, HFT, device_sync

contains

  ! The transform buffers are allocated with the fields (hst_gpu_init).
  subroutine init_fft()
  end subroutine init_fft

  subroutine free_fft()
  end subroutine free_fft

  ! Wait for everything queued on the device.
  subroutine device_sync()
    call gpu_check(hst_gpu_sync(), 'hst_gpu_sync')
  end subroutine device_sync

  subroutine fused(name)
    character(len=*), intent(in) :: name
    print *, 'ERROR: ', name, ': the per-field transforms are fused on the GPU; call transform_to_physical, ', &
      'build_products and products_to_spectral (hst_transforms)'
    error stop 1
  end subroutine fused

  subroutine FFT(m)
    integer(C_INT), intent(in) :: m
    call fused('FFT')
  end subroutine FFT

  subroutine IFT(m)
    integer(C_INT), intent(in) :: m
    call fused('IFT')
  end subroutine IFT

  subroutine RFT(m)
    integer(C_INT), intent(in) :: m
    call fused('RFT')
  end subroutine RFT

  subroutine HFT(m)
    integer(C_INT), intent(in) :: m
    call fused('HFT')
  end subroutine HFT

end module hst_fft
