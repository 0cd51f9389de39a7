! The spectral <-> physical round trip of the reference's module
! hst_transforms (src/hst_transforms.f90), field by field on the buffers of
! hst_fft:
!
!   transform_to_physical   V -> rVVdx(:, :, :, 1:3): u, v, w at the 2 nxd x
!                           nzd points of every row -2..ny+1
!   build_products(g)       rVVdx -> products(:, :, :, 1:3): group 1 uu, vv,
!                           ww; group 2 uv, vw, uw; times factor
!   products_to_spectral    products -> VVdz(:, :, :, 1:3); mode (iz, ix) of
!                           product m at VVdz(izd(iz) + 1, ix + 1, iy, m)
!   vvdz_to_field(f, m)     that product into an array of the field shape
!   compute_cfl             the CFL number of the velocity
!
! Every transform and product runs on the GPU (hst_fft; hst_gpu_products);
! the host moves the lines between the z and x layouts (with one process
! the reference's transposes are local copies too).  The time step does not
! come here: its transforms are fused on the GPU (hst_gpu_timestep).
module hst_transforms

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api
  use hst_fft
  use hst_setup, only: state_in, S_V

  implicit none
  private
  public :: transform_to_physical, build_products, products_to_spectral, vvdz_to_field, compute_cfl
  public :: product_group

  integer(C_INT), save :: group = 0          ! the group of the last build_products
  integer(C_INT), save :: spectra_group = 0  ! the group whose spectra VVdz holds

contains

  ! the group of the product spectra in VVdz (0: none yet)
  integer(C_INT) function product_group()
    product_group = spectra_group
  end function product_group

  subroutine need_buffers(name)
    character(len=*), intent(in) :: name
    if (.not. fields_by_field()) then
      print *, 'ERROR: ', name, ' works on the host buffers of hst_fft, which a resident state does not have'
      error stop 1
    end if
  end subroutine need_buffers

  subroutine transform_to_physical()
    integer(C_INT) :: m, iy, ix, iz, k
    call need_buffers('transform_to_physical')
    ! V into the padded z lines: modes 0..nz first, -nz..-1 at the end
    VVdz = 0
    do m = 1, 3
      do iy = -2, ny + 1
        do ix = 0, nx
          do iz = -nz, nz
            VVdz(izd(iz) + 1, ix + 1, iy, m) = V(iy, iz, ix, m)
          end do
        end do
      end do
    end do
    do m = 1, 3
      call IFT(m)
    end do
    ! z lines -> x lines, the x modes above nx zero (the 3/2 padding)
    do m = 1, 3
      do iy = -2, ny + 1
        do k = 1, nzd
          VVdx(1:nx + 1, k, iy, m) = VVdz(k, 1:nx + 1, iy, m)
          VVdx(nx + 2:nxd + 1, k, iy, m) = 0
        end do
      end do
    end do
    do m = 1, 3
      call RFT(m)
    end do
  end subroutine transform_to_physical

  subroutine build_products(g)
    integer(C_INT), intent(in) :: g
    if (g /= 1 .and. g /= 2) error stop 'build_products: the group is 1 or 2'
    call need_buffers('build_products')
    call gpu_check(hst_gpu_products(g, int(nzd*(ny + 4), C_INT64_T), rVVdx, products), 'build_products')
    group = g
  end subroutine build_products

  subroutine products_to_spectral()
    integer(C_INT) :: m, iy, k
    call need_buffers('products_to_spectral')
    do m = 1, 3
      call HFT(m)
    end do
    do m = 1, 3
      do iy = -2, ny + 1
        do k = 1, nzd
          VVdz(k, 1:nx + 1, iy, m) = VVdp(1:nx + 1, k, iy, m)
        end do
      end do
    end do
    do m = 1, 3
      call FFT(m)
    end do
    spectra_group = group
  end subroutine products_to_spectral

  ! Product m of the last products_to_spectral into an array of the field
  ! shape, every row.
  subroutine vvdz_to_field(field, m)
    complex(C_DOUBLE_COMPLEX), intent(out) :: field(-2:ny + 1, -nz:nz, 0:nx)
    integer(C_INT), intent(in) :: m
    integer(C_INT) :: ix, iy, iz
    call need_buffers('vvdz_to_field')
    do ix = 0, nx
      do iz = -nz, nz
        do iy = -2, ny + 1
          field(iy, iz, ix) = VVdz(izd(iz) + 1, ix + 1, iy, m)
        end do
      end do
    end do
  end subroutine vvdz_to_field

  ! max over the interior rows of |u|/dx + |v|/dyl + |w|/dz at the physical
  ! points, of the current velocity (transformed once more on the GPU)
  subroutine compute_cfl()
    call state_in(S_V)
    call gpu_check(hst_gpu_compute_cfl(cfl), 'compute_cfl')
  end subroutine compute_cfl

end module hst_transforms
