! Field files in the format of the CPL code (module hst_io of the reference,
! src/hst_io.f90:5-33 for the format), read and written by this one process
! with stream I/O.
!
!   file    = header, "Vfield=" LF, array       (velocity: Dati.cart.out, fields/field<n>.fld)
!           = array                             (pressure: p_fields/pField<n>.fld)
!   header  = text lines of "key=value" tokens separated by " " TAB, then five
!             "name=" LF <8 raw bytes> LF records (time, S, S2, gamma_x, gamma_y)
!   array   = A(c, r, iz, ix), c fastest: components (u, w, v) (one for the
!             pressure), rows r = -2..ny+1 (the four ghost rows as plain
!             periodic copies), iz = -nz..nz, ix = 0..nx, complex128
module hst_io

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api
  use hst_setup, only: upload_velocity, download_velocity
  use hst_derivatives, only: s2_of, gamma_y_of
  use hst_initial, only: generate_initial_field

  implicit none
  private
  public :: restart_read, restart_write, field_write, make_output_dirs

  character, parameter :: LF = achar(10), TAB = achar(9)
  ! file component k holds our component FILE_COMP(k): the file's (u, v, w)
  ! are its spanwise-last convention, (u, w, v) in ours
  integer, parameter :: FILE_COMP(3) = [1, 3, 2]
  integer, parameter :: HEAD_MAX = 4096

contains

  subroutine make_output_dirs()
    if (has_terminal) call execute_command_line('mkdir -p fields p_fields statistics statistics/spectra vtks')
  end subroutine make_output_dirs

  ! ------------------------------------------------------------------ reading

  ! Loads a velocity file into V and the device (the caller fills the ghost
  ! rows), or, when there is no such file, the deck's initial field.  Another
  ! resolution is accepted: the Fourier modes both grids have are taken (x and
  ! z are mode indices), every other mode is zero, and each row of our y grid
  ! takes the file row nearest to it on the file's uniform grid (the periodic
  ! end wrapped).  Another alpha0 or beta0 needs restart_remap = 'mode_index',
  ! and the discrete divergence is then repaired on the GPU.
  subroutine restart_read(filename)
    character(len=*), intent(in) :: filename
    character(len=HEAD_MAX) :: head
    integer :: unit, ios, body
    integer(C_INT) :: fnx, fnz, fny
    real(C_DOUBLE) :: f_alfa0, f_beta0, f_re, f_time, f_S
    complex(C_DOUBLE_COMPLEX), allocatable :: A(:, :, :, :)
    logical :: other_box

    open (newunit=unit, file=trim(filename), access='stream', form='unformatted', status='old', action='read', &
          iostat=ios)
    if (ios /= 0) then
      if (has_terminal) print '(A)', '   no field file '//trim(filename)//': starting from the initial condition'
      call generate_initial_field()
      return
    end if
    head = ''
    read (unit, pos=1, iostat=ios) head
    body = index(head, 'Vfield='//LF)
    if (body == 0) then
      close (unit)
      print '(A)', ' ERROR: '//trim(filename)//': no "Vfield=" record, so not a velocity file of the CPL format'
      error stop 1
    end if
    body = body + len('Vfield=') + 1            ! first byte of the array
    ! the file's sizes: its nx, its spanwise modes ("ny"), its rows + 1 ("nz")
    fnx = nint(header_number(head, 'nx'))
    fnz = nint(header_number(head, 'ny'))
    fny = nint(header_number(head, 'nz')) - 1
    f_alfa0 = header_number(head, 'alpha0')
    f_beta0 = header_number(head, 'beta0')
    f_re = header_number(head, 'Re')
    f_time = header_raw(unit, head, 'time')
    f_S = header_raw(unit, head, 'S')
    if (fnx < 0 .or. fnz < 0 .or. fny < 1) then
      close (unit)
      print '(A,3(1X,I0))', ' ERROR: '//trim(filename)//' declares impossible sizes (nx, nz, ny):', fnx, fnz, fny
      error stop 1
    end if
    if (has_terminal) then
      print '(A)', '   loading '//trim(filename)
      if (fnx /= nx .or. fnz /= nz .or. fny /= ny) &
        print '(A,3(1X,I0),A,3(1X,I0),A)', '   resolution of the file (nx, nz, ny) =', fnx, fnz, fny, &
        ', of the deck =', nx, nz, ny, ': modes and rows remapped'
      if (abs(f_re - re) > 1.0d-9*re .or. f_S /= S) &
        print '(A,2(1X,ES12.5))', '   (the file was written at another Re or S:', f_re, f_S
    end if
    other_box = abs(f_alfa0 - alfa0) > 1.0d-4*max(abs(alfa0), 1.0d0) .or. &
                abs(f_beta0 - beta0) > 1.0d-4*max(abs(beta0), 1.0d0)
    if (other_box) then
      if (trim(restart_remap) /= 'mode_index') then
        close (unit)
        print '(A,4(1X,ES14.7))', ' ERROR: the box of the file differs (alpha0, beta0 of file and deck):', &
          f_alfa0, f_beta0, alfa0, beta0
        print '(A)', '        restart_remap = ''mode_index'' maps the mode indices and repairs the divergence'
        error stop 1
      end if
      if (has_terminal) print '(A)', '   another box: mode indices kept, divergence repaired on the new wavenumbers'
    end if
    if (time_from_restart) time = f_time

    allocate (A(3, -2:fny + 1, -fnz:fnz, 0:fnx))
    read (unit, pos=body) A
    close (unit)
    call take_rows(A, fnx, fnz, fny)
    deallocate (A)
    call upload_velocity()
    if (other_box) then
      call gpu_check(hst_gpu_project_divergence(time), 'divergence repair')
      call download_velocity()
    end if
  end subroutine restart_read

  ! V from the file array: for every row of our grid the nearest file row,
  ! then the modes the two grids share.
  subroutine take_rows(A, fnx, fnz, fny)
    integer(C_INT), intent(in) :: fnx, fnz, fny
    complex(C_DOUBLE_COMPLEX), intent(in) :: A(3, -2:fny + 1, -fnz:fnz, 0:fnx)
    integer(C_INT) :: nearest(0:ny - 1), iy, k, mx, mz
    real(C_DOUBLE) :: step
    step = ly/real(fny, C_DOUBLE)
    do iy = 0, ny - 1
      nearest(iy) = modulo(nint(modulo(y(iy), ly)/step), fny)
    end do
    mx = min(nx, fnx)
    mz = min(nz, fnz)
    V = 0
    do k = 1, 3
      do iy = 0, ny - 1
        V(iy, -mz:mz, 0:mx, FILE_COMP(k)) = A(k, nearest(iy), -mz:mz, 0:mx)
      end do
    end do
  end subroutine take_rows

  ! The number after "key=" in the text part of the header: the key starts
  ! the header or follows a separator, the value runs to the next one.
  real(C_DOUBLE) function header_number(head, key) result(x)
    character(len=*), intent(in) :: head, key
    integer :: p, q
    character :: before
    x = -huge(1.0d0)
    p = 0
    do
      q = index(head(p + 1:), key//'=')
      if (q == 0) return
      p = p + q
      before = ' '
      if (p > 1) before = head(p - 1:p - 1)
      if (before == ' ' .or. before == TAB .or. before == LF) exit
    end do
    p = p + len(key) + 1
    q = scan(head(p:), ' '//TAB//LF)
    if (q <= 1) return
    read (head(p:p + q - 2), *) x
  end function header_number

  ! The 8 raw bytes of the record "name=" LF <double> LF of the header.
  real(C_DOUBLE) function header_raw(unit, head, name) result(x)
    integer, intent(in) :: unit
    character(len=*), intent(in) :: head, name
    integer :: p
    p = index(head, LF//name//'='//LF)
    if (p == 0) then
      print '(A)', ' ERROR: the field header has no '//name//'= record'
      error stop 1
    end if
    read (unit, pos=p + len(name) + 3) x
  end function header_raw

  ! ------------------------------------------------------------------ writing

  ! The velocity of the device state with the CPL header.
  subroutine restart_write(filename)
    character(len=*), intent(in) :: filename
    call download_velocity()
    call write_file(filename, cpl_header(), V, 3)
  end subroutine restart_write

  ! One array of the field shape as a headerless one-component CPL file.
  subroutine field_write(filename, field)
    character(len=*), intent(in) :: filename
    complex(C_DOUBLE_COMPLEX), intent(in) :: field(-2:ny + 1, -nz:nz, 0:nx)
    call write_file(filename, '', field, 1)
  end subroutine field_write

  ! head, then the array of ncomp components of f (our layout f(iy, iz, ix, c))
  subroutine write_file(filename, head, f, ncomp)
    character(len=*), intent(in) :: filename, head
    integer, intent(in) :: ncomp
    complex(C_DOUBLE_COMPLEX), intent(in) :: f(-2:ny + 1, -nz:nz, 0:nx, ncomp)
    complex(C_DOUBLE_COMPLEX), allocatable :: A(:, :, :, :)
    integer :: unit, k, ix, iz, c
    allocate (A(ncomp, -2:ny + 1, -nz:nz, 0:nx))
    do k = 1, ncomp
      c = k
      if (ncomp == 3) c = FILE_COMP(k)
      do ix = 0, nx
        do iz = -nz, nz
          A(k, 0:ny - 1, iz, ix) = f(0:ny - 1, iz, ix, c)
        end do
      end do
    end do
    ! the ghost rows of the file: plain periodic copies of the interior
    A(:, ny:ny + 1, :, :) = A(:, 0:1, :, :)
    A(:, -2:-1, :, :) = A(:, ny - 2:ny - 1, :, :)
    if (.not. has_terminal) return
    open (newunit=unit, file=trim(filename), access='stream', form='unformatted', status='replace', action='write')
    if (len(head) > 0) write (unit) head
    write (unit) A
    close (unit)
    deallocate (A)
  end subroutine write_file

  ! The header of a velocity file at the current time.
  function cpl_header() result(h)
    character(len=:), allocatable :: h
    real(C_DOUBLE) :: htcoeff
    htcoeff = -1.0d0
    if (ystretch > 0.0d0) htcoeff = ystretch
    h = fields([character(len=40) :: 'nx='//int_text(nx), 'ny='//int_text(nz), 'nz='//int_text(ny + 1), &
                'alpha0='//real_text(alfa0), 'beta0='//real_text(beta0), 'htcoeff='//real_text(htcoeff), &
                'Re='//real_text(re)], 'Pr=0.71')
    h = h//fields([character(len=40) :: 'deltat='//real_text(deltat), 't_max='//real_text(t_max), &
                   'dt_field='//real_text(dt_field)], 'dt_save='//real_text(dt_save))
    h = h//'t_field='//real_text(time)//LF
    h = h//fields([character(len=40) :: 'meanpx=0', 'meanflowx=0', 'meanpy=0'], 'meanflowy=0')
    h = h//raw_record('time', time)//raw_record('S', S)//raw_record('S2', s2_of(time))// &
        raw_record('gamma_x', S*time)//raw_record('gamma_y', gamma_y_of(time))//'Vfield='//LF
  end function cpl_header

  ! one header line: the tokens, each followed by " " TAB, then the last and LF
  function fields(tokens, last) result(line)
    character(len=*), intent(in) :: tokens(:), last
    character(len=:), allocatable :: line
    integer :: i
    line = ''
    do i = 1, size(tokens)
      line = line//trim(tokens(i))//' '//TAB
    end do
    line = line//last//LF
  end function fields

  function raw_record(name, x) result(r)
    character(len=*), intent(in) :: name
    real(C_DOUBLE), intent(in) :: x
    character(len=:), allocatable :: r
    character(len=8) :: bytes
    bytes = transfer(x, bytes)
    r = name//'='//LF//bytes//LF
  end function raw_record

  function int_text(i) result(s)
    integer(C_INT), intent(in) :: i
    character(len=:), allocatable :: s
    character(len=16) :: t
    write (t, '(I0)') i
    s = trim(t)
  end function int_text

  ! the list-directed G0.17 text of x that the CPL tools expect, with the
  ! leading zero of a number below one written out (0.1..., -0.1...)
  function real_text(x) result(s)
    real(C_DOUBLE), intent(in) :: x
    character(len=:), allocatable :: s
    character(len=32) :: t
    write (t, '(G0.17)') x
    s = trim(adjustl(t))
    if (s(1:1) == '.') then
      s = '0'//s
    else if (len(s) > 1) then
      if (s(1:2) == '-.') s = '-0'//s(2:)
    end if
  end function real_text

end module hst_io
