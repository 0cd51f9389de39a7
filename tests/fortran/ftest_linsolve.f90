! line_solve through the module interface with array arguments, as the
! reference's tests/test_linsolve.f90 calls it: for every system kind the
! operator is applied on the host to a seeded field through its ghost rows
! (filled at a nonzero time, so the wrap phase takes part), the system is
! solved on the GPU and must return the field; KIND_DY must satisfy
! D0 dst = D1 src through the ghost rows of dst (fill_ghosts_field).
program ftest_linsolve

  use, intrinsic :: iso_c_binding
  use hst_params
  use test_common
  use hst_linsolve
  use hst_derivatives, only: shear_shifts, fill_ghosts, fill_ghosts_field
  use hst_initial, only: uniform_from_key

  implicit none

  real(C_DOUBLE), parameter :: lambda = 123.4d0
  character(len=7), parameter :: label(5) = ['D2V    ', 'ETA    ', 'POISSON', 'D0     ', 'DY     ']
  integer(C_INT) :: kind, ix, iz, iy
  real(C_DOUBLE) :: sx, sz, rel, worst

  call test_start()
  time = 0.37d0
  call test_setup()
  call shear_shifts(time, sx, sz)

  ! a seeded field in V(:, :, :, 1), its (0,0) column zero (singular for two kinds)
  V = 0
  do ix = 0, nx
    do iz = -nz, nz
      if (ix == 0 .and. iz == 0) cycle
      do iy = 0, ny - 1
        V(iy, iz, ix, 1) = cmplx(uniform_from_key(seed + 11, 1, iy, iz, ix) - 0.5d0, &
                                 uniform_from_key(seed + 12, 1, iy, iz, ix) - 0.5d0, C_DOUBLE)
      end do
    end do
  end do
  call fill_ghosts(1)

  worst = 0.0d0
  do kind = KIND_D2V, KIND_DY
    if (kind == KIND_DY) then
      rhs(:, :, :, 2) = 0
      call line_solve(kind, lambda, V(:, :, :, 1), rhs(:, :, :, 2))
      call fill_ghosts_field(rhs(:, :, :, 2), sx, sz)
      rel = dy_identity_error()
    else
      call apply_operator(kind, V(:, :, :, 1), rhs(:, :, :, 1))
      rhs(:, :, :, 2) = 0
      call line_solve(kind, lambda, rhs(:, :, :, 1), rhs(:, :, :, 2))
      rel = maxval(abs(rhs(0:ny - 1, :, :, 2) - V(0:ny - 1, :, :, 1)))/maxval(abs(V(0:ny - 1, :, :, 1)))
    end if
    print '(3X,A7,A,ES10.2)', label(kind), ': relative error ', rel
    worst = max(worst, rel)
  end do
  call test_finish(worst <= 1.0d-11)

contains

  ! the coefficient of row iy, band entry j, of system kind (before the wrap phase)
  real(C_DOUBLE) function band(kind, iy, j, kk)
    integer(C_INT), intent(in) :: kind, iy, j
    real(C_DOUBLE), intent(in) :: kk
    select case (kind)
    case (KIND_D2V)
      band = lambda*(der(iy, 2, j) - kk*der(iy, 0, j)) - ni*(der(iy, 3, j) - 2.0d0*kk*der(iy, 2, j) + kk*kk*der(iy, 0, j))
    case (KIND_ETA)
      band = lambda*der(iy, 0, j) - ni*(der(iy, 2, j) - kk*der(iy, 0, j))
    case (KIND_POISSON)
      band = der(iy, 2, j) - kk*der(iy, 0, j)
    case default
      band = der(iy, 0, j)
    end select
  end function band

  ! out = A(kind) f on the interior rows, the stencils reaching the ghost rows
  subroutine apply_operator(kind, f, out)
    integer(C_INT), intent(in) :: kind
    complex(C_DOUBLE_COMPLEX), intent(in) :: f(-2:ny + 1, -nz:nz, 0:nx)
    complex(C_DOUBLE_COMPLEX), intent(out) :: out(-2:ny + 1, -nz:nz, 0:nx)
    integer(C_INT) :: ix, iz, iy, j
    out = 0
    do ix = 0, nx
      do iz = -nz, nz
        do iy = 0, ny - 1
          do j = -2, 2
            out(iy, iz, ix) = out(iy, iz, ix) + band(kind, iy, j, k2(iz, ix))*f(iy + j, iz, ix)
          end do
        end do
      end do
    end do
  end subroutine apply_operator

  ! max |D0 dst - D1 src| relative to max |D1 src|
  real(C_DOUBLE) function dy_identity_error() result(rel)
    complex(C_DOUBLE_COMPLEX) :: lhs, d1
    real(C_DOUBLE) :: err, ref
    integer(C_INT) :: ix, iz, iy, j
    err = 0.0d0
    ref = 0.0d0
    do ix = 0, nx
      do iz = -nz, nz
        do iy = 0, ny - 1
          lhs = 0
          d1 = 0
          do j = -2, 2
            lhs = lhs + der(iy, 0, j)*rhs(iy + j, iz, ix, 2)
            d1 = d1 + der(iy, 1, j)*V(iy + j, iz, ix, 1)
          end do
          err = max(err, abs(lhs - d1))
          ref = max(ref, abs(d1))
        end do
      end do
    end do
    rel = err/ref
  end function dy_identity_error

end program ftest_linsolve
