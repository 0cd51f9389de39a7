! The spectral <-> physical round trip (the reference's module
! hst_transforms).  On the GPU one fused pass takes u, v, w to physical space,
! forms the six products of both groups (and the CFL number) and brings them
! back; transform_to_physical runs it, build_products(g) selects the group
! whose three product spectra the following calls see.
module hst_transforms

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api

  implicit none
  private
  public :: transform_to_physical, build_products, products_to_spectral, vvdz_to_field, compute_cfl, invalidate

  logical, save :: fresh = .false.
  integer(C_INT), save :: group = 1
  real(C_DOUBLE), save :: cfl_last = 0.0d0

contains

  subroutine transform_to_physical()
    call gpu_check(hst_gpu_nonlinear(cfl_last), 'transform_to_physical')
    fresh = .true.
  end subroutine transform_to_physical

  ! products of group g: 1 uu, vv, ww; 2 uv, vw, uw (times 1/(2 nxd nzd))
  subroutine build_products(g)
    integer(C_INT), intent(in) :: g
    if (.not. fresh) call transform_to_physical()
    group = g
  end subroutine build_products

  ! the spectra of the selected group are ready after the fused pass
  subroutine products_to_spectral()
  end subroutine products_to_spectral

  ! product m of the selected group into a device field
  subroutine vvdz_to_field(field, m)
    integer(C_INT), intent(in) :: field, m
    call gpu_check(hst_gpu_product_to_field(3*(group - 1) + m, field), 'vvdz_to_field')
  end subroutine vvdz_to_field

  ! CFL number of the velocity last transformed (or transformed now)
  subroutine compute_cfl()
    if (fresh) then
      cfl = cfl_last
    else
      call gpu_check(hst_gpu_compute_cfl(cfl), 'compute_cfl')
    end if
  end subroutine compute_cfl

  ! the velocity changed: the next transform_to_physical is needed again
  subroutine invalidate()
    fresh = .false.
  end subroutine invalidate

end module hst_transforms
