! The transform layer and the field files through the module interface, with
! the calls of the reference's tests/test_roundtrip.f90: a seeded field goes
! to physical space (transform_to_physical: z then x transforms of the
! buffers of hst_fft), is scaled by factor into the products buffer, and goes
! back (products_to_spectral); vvdz_to_field must return it to 1e-12 of its
! largest value.  Then restart_write and restart_read must return it too.
program ftest_roundtrip

  use, intrinsic :: iso_c_binding
  use hst_params
  use test_common
  use hst_fft
  use hst_transforms
  use hst_initial, only: generate_initial_field
  use hst_io, only: restart_write, restart_read
  use hst_derivatives, only: fill_ghosts

  implicit none

  complex(C_DOUBLE_COMPLEX), allocatable :: V0(:, :, :, :), W(:, :, :)
  integer(C_INT) :: m
  real(C_DOUBLE) :: vmax, err, worst

  call test_start()
  time = 0.0d0
  call test_setup()
  call generate_initial_field()
  do m = 1, 3
    call fill_ghosts(m)
  end do
  V0 = V
  allocate (W(-2:ny + 1, -nz:nz, 0:nx))
  vmax = maxval(abs(V0))
  worst = 0.0d0

  call transform_to_physical()
  call compute_cfl()
  products(1:2*nxd, :, :, :) = rVVdx(1:2*nxd, :, :, :)*factor
  call products_to_spectral()
  do m = 1, 3
    call vvdz_to_field(W, m)
    err = maxval(abs(W - V0(:, :, :, m)))
    print '(A,I1,A,ES10.2,A,ES10.2,A)', '   round trip of component ', m, ': max error ', err, ' (max |V| ', vmax, ')'
    worst = max(worst, err)
  end do
  print '(A,F10.4)', '   cfl ', cfl

  call restart_write('ftest_roundtrip.fld')
  V = 0
  call restart_read('ftest_roundtrip.fld')
  err = maxval(abs(V(0:ny - 1, :, :, :) - V0(0:ny - 1, :, :, :)))
  print '(A,ES10.2)', '   restart_write / restart_read: max error ', err
  worst = max(worst, err)
  call test_finish(worst <= 1.0d-12*vmax .and. cfl > 0.0d0)

end program ftest_roundtrip
