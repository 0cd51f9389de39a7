! The energy input of each nonlinear product alone, through the pieces of a
! substep (the procedure of the reference's tests/test_conservation.f90):
! S = 0, nearly inviscid, one step of dt = 1e-4 from the seeded field after
! one unmasked step, with all products but one zeroed in physical space
! between build_products and products_to_spectral.  In the continuum uu, vv,
! ww alone inject no energy and uv + vw + uw together inject none; the
! discrete rates (dq2/dt)/q2 must stay below 1e-2 (anything of order one is a
! wrong term).  All products kept must conserve too.
program ftest_conservation

  use, intrinsic :: iso_c_binding
  use hst_params
  use test_common
  use hst_fft, only: products
  use hst_transforms
  use hst_equations
  use hst_initial, only: generate_initial_field
  use hst_derivatives, only: fill_ghosts

  implicit none

  character(len=3), parameter :: pname(0:6) = ['all', 'uu ', 'vv ', 'ww ', 'uv ', 'vw ', 'uw ']
  real(C_DOUBLE) :: rate(0:6), q_before
  integer :: keep

  call test_start()
  S = 0.0d0
  linear = .false.
  ni = 1.0d-12
  call test_setup()

  do keep = 0, 6
    time = 0.0d0
    deltat = 1.0d-4
    oldrhs = 0
    call generate_initial_field()
    call fill_ghosts(1)
    call fill_ghosts(2)
    call fill_ghosts(3)
    ! the solver projects the seeded field onto its discretely divergence-free
    ! space in its first step, which moves the energy by more than a step of
    ! the nonlinear terms: measure from the projected field
    call timestep()
    q_before = box_energy()
    call masked_step(keep)
    rate(keep) = (box_energy() - q_before)/deltat/q_before
    print '(3X,A3,A,ES11.3)', pname(keep), ': (dq2/dt)/q2 = ', rate(keep)
  end do
  print '(3X,A,ES11.3)', 'uv + vw + uw: (dq2/dt)/q2 = ', rate(4) + rate(5) + rate(6)
  call test_finish(maxval(abs(rate(0:3))) <= 1.0d-2 .and. abs(rate(4) + rate(5) + rate(6)) <= 1.0d-2)

contains

  ! one RK3 step from the pieces, only product `keep` (1..6) kept, or all (0)
  subroutine masked_step(keep)
    integer, intent(in) :: keep
    integer :: i, g, p
    real(C_DOUBLE) :: dt_sub
    do i = 1, 3
      call transform_to_physical()
      call buildrhs_prepare(RK_rai(:, i))
      do g = 1, 2
        call build_products(g)
        if (keep /= 0) then
          do p = 1, 3
            if (3*(g - 1) + p /= keep) products(:, :, :, p) = 0.0d0
          end do
        end if
        call products_to_spectral()
        call buildrhs(RK_rai(:, i), g)
      end do
      dt_sub = 2.0d0/RK_rai(1, i)*deltat
      call shear_shift(dt_sub)
      time = time + dt_sub
      call linsolve(RK_rai(1, i)/deltat)
    end do
  end subroutine masked_step

  ! q2 = <u_i u_i> over the box (the interior rows)
  real(C_DOUBLE) function box_energy() result(q)
    integer(C_INT) :: ix
    real(C_DOUBLE) :: w
    q = 0.0d0
    do ix = 0, nx
      w = merge(1.0d0, 2.0d0, ix == 0)
      q = q + w*sum(abs(V(0:ny - 1, :, ix, :))**2)
    end do
    q = q/ny
  end function box_energy

end program ftest_conservation
