! Start and end of the Fortran test programs (the names of the reference's
! tests/test_common.f90, one process): test_start reads the deck (first
! argument, default hst.in); the caller may change parameters before
! test_setup, which makes the device state and the host arrays; test_finish
! frees them and stops with status 1 unless passed.
module test_common

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_input, only: read_input
  use hst_setup, only: select_device, setup_decomposition, allocate_fields, free_fields
  use hst_fft, only: init_fft, free_fft
  use hst_linsolve, only: init_linsolve, free_linsolve
  use hst_derivatives, only: setup_derivatives
  use hst_damping, only: damping_setup

  implicit none
  private
  public :: test_start, test_setup, test_finish

contains

  subroutine test_start()
    character(len=256) :: deck
    call select_device()
    deck = 'hst.in'
    if (command_argument_count() > 0) call get_command_argument(1, deck)
    call read_input(trim(deck))
  end subroutine test_start

  subroutine test_setup()
    call setup_decomposition()
    call allocate_fields()
    call init_fft()
    call init_linsolve()
    call setup_derivatives()
    call damping_setup()
  end subroutine test_setup

  subroutine test_finish(passed)
    logical, intent(in) :: passed
    call free_linsolve()
    call free_fft()
    call free_fields()
    if (.not. passed) then
      print '(A)', ' FAILED'
      error stop 1
    end if
    print '(A)', ' PASSED'
  end subroutine test_finish

end module test_common
