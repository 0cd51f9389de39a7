! Velocity files in the layout of the CPL code (the reference's module
! hst_io): restart file Dati.cart.out and snapshots fields/field<n>.fld.  A
! text header as io.cpl writes it, five raw doubles (time, S, S2, gamma_x,
! gamma_y), the line "Vfield=", then the array (0..nx, -nz..nz, -2..ny+1) of
! (u, w, v) complex, C order with the components innermost; the four ghost
! rows are written as plain periodic copies.  One process writes and reads the
! whole file with stream I/O.
module hst_io

  use, intrinsic :: iso_c_binding
  use hst_params
  use hst_gpu_api
  use hst_derivatives, only: s2_of, gamma_y_of
  use hst_initial, only: generate_initial_field, upload_velocity

  implicit none
  private
  public :: restart_read, restart_write, make_output_dirs, download_velocity, field_write

  character, parameter :: LF = achar(10), TAB = achar(9)
  integer, parameter :: CPL_ORDER(3) = [1, 3, 2]     ! our component behind CPL's (u, v, w)

contains

  subroutine make_output_dirs()
    if (has_terminal) call execute_command_line('mkdir -p fields p_fields statistics statistics/spectra vtks')
  end subroutine make_output_dirs

  ! the device velocity into the host copy V (`target update from(V)`)
  subroutine download_velocity()
    integer(C_INT) :: c
    do c = 1, 3
      call gpu_check(hst_gpu_download_field(c, V(:, :, :, c)), 'download V')
    end do
  end subroutine download_velocity

  ! Reads filename into V, or generates the initial field when the file is
  ! absent; the device copy is updated, the caller fills the ghosts
  ! (restart_read, src/hst_io.f90:63-190).  A file of another resolution is
  ! remapped: the Fourier modes present in both fields are copied (x and z are
  ! mode indices), y is resampled by the nearest row of the file's uniform
  ! grid.  Other alpha0/beta0 need restart_remap = 'mode_index', which then
  ! repairs the discrete divergence on the device (project_restart).
  subroutine restart_read(filename)
    character(len=*), intent(in) :: filename
    integer :: io, unit, p
    character(len=4096) :: head
    integer(C_INT) :: r_nx, r_ny, r_nz, source_ny, ix, iy, iz, src_iy
    real(C_DOUBLE) :: r_alfa0, r_beta0, r_re, r_time, r_S, source_dy
    complex(C_DOUBLE_COMPLEX), allocatable :: buf(:, :, :, :)

    open (newunit=unit, file=trim(filename), access='stream', form='unformatted', status='old', action='read', &
          iostat=io)
    if (io /= 0) then
      if (has_terminal) print *, '   restart file '//trim(filename)//' not found'
      call generate_initial_field()
      return
    end if
    if (has_terminal) print *, '   reading '//trim(filename)
    head = ''
    read (unit, pos=1, iostat=io) head
    p = index(head, 'Vfield='//LF)
    if (p == 0) then
      print *, 'ERROR: '//trim(filename)//' has no Vfield= line (not a CPL field file)'
      error stop 1
    end if
    r_nx = int(header_value(head, 'nx='))
    r_ny = int(header_value(head, 'ny='))
    r_nz = int(header_value(head, 'nz='))
    r_alfa0 = header_value(head, 'alpha0=')
    r_beta0 = header_value(head, 'beta0=')
    r_re = header_value(head, 'Re=')
    source_ny = r_nz - 1
    if (r_nx < 0 .or. r_ny < 0 .or. source_ny < 1) then
      print *, 'ERROR: invalid mesh dimensions in '//trim(filename)
      error stop 1
    end if
    if (has_terminal .and. (r_nx /= nx .or. r_ny /= nz .or. source_ny /= ny)) then
      print *, '   interpolating restart mesh: file (nx, nz, ny) = ', r_nx, r_ny, source_ny
      print *, '                              deck (nx, nz, ny) = ', nx, nz, ny
    end if
    if (abs(r_alfa0 - alfa0) > 1.0d-4*max(abs(alfa0), 1.0d0) .or. abs(r_beta0 - beta0) > 1.0d-4*max(abs(beta0), 1.0d0)) &
      then
      if (trim(restart_remap) == 'mode_index') then
        if (has_terminal) print *, '   restart_remap=mode_index: remapping signed Fourier indices and projecting' // &
          ' on target wavenumbers'
      else
        print *, 'ERROR: restart alpha0/beta0 do not match the deck:', r_alfa0, r_beta0, alfa0, beta0
        error stop 1
      end if
    end if
    p = index(head, 'time='//LF)
    read (unit, pos=p + 6) r_time
    p = index(head, LF//'S='//LF)
    read (unit, pos=p + 4) r_S
    if (has_terminal .and. (abs(r_re - re) > 1.0d-9*re .or. r_S /= S)) print *, '   note: Re or S differ from the file'
    if (time_from_restart) time = r_time
    ! the file holds x = 0..r_nx, z = -r_ny..r_ny and the rows -2..source_ny+1
    allocate (buf(3, -2:source_ny + 1, -r_ny:r_ny, 0:r_nx))
    p = index(head, 'Vfield='//LF)
    read (unit, pos=p + 8) buf
    close (unit)
    ! the CPL shear grid is uniform whatever the target's stretching: the
    ! nearest row by coordinate, the end point wrapped (src/hst_io.f90:155-167)
    source_dy = ly/real(source_ny, C_DOUBLE)
    V = 0
    do ix = 0, min(nx, r_nx)
      do iz = max(-nz, -r_ny), min(nz, r_ny)
        do iy = 0, ny - 1
          src_iy = nint(modulo(y(iy), ly)/source_dy)
          if (src_iy == source_ny) src_iy = 0
          V(iy, iz, ix, CPL_ORDER) = buf(:, src_iy, iz, ix)
        end do
      end do
    end do
    deallocate (buf)
    call upload_velocity()
    if (trim(restart_remap) == 'mode_index') then
      call gpu_check(hst_gpu_project_divergence(time), 'project_restart')
      call download_velocity()
    end if
  end subroutine restart_read

  ! Writes the device velocity with the CPL header.
  subroutine restart_write(filename)
    character(len=*), intent(in) :: filename
    character(len=:), allocatable :: head
    complex(C_DOUBLE_COMPLEX), allocatable :: buf(:, :, :, :)
    integer(C_INT) :: ix, iz
    integer :: unit
    call download_velocity()
    head = 'nx='//str_i(nx)//' '//TAB//'ny='//str_i(nz)//' '//TAB//'nz='//str_i(ny + 1)// &
           ' '//TAB//'alpha0='//str_r(alfa0)//' '//TAB//'beta0='//str_r(beta0)// &
           ' '//TAB//'htcoeff='//str_r(merge(ystretch, -1.0d0, ystretch > 0.0d0))//' '//TAB//'Re='//str_r(re)// &
           ' '//TAB//'Pr=0.71'//LF// &
           'deltat='//str_r(deltat)//' '//TAB//'t_max='//str_r(t_max)//' '//TAB// &
           'dt_field='//str_r(dt_field)//' '//TAB//'dt_save='//str_r(dt_save)//LF// &
           't_field='//str_r(time)//LF// &
           'meanpx=0 '//TAB//'meanflowx=0 '//TAB//'meanpy=0 '//TAB//'meanflowy=0'//LF// &
           'time='//LF//raw8(time)//LF//'S='//LF//raw8(S)//LF//'S2='//LF//raw8(s2_of(time))//LF// &
           'gamma_x='//LF//raw8(S*time)//LF//'gamma_y='//LF//raw8(gamma_y_of(time))//LF//'Vfield='//LF
    allocate (buf(3, -2:ny + 1, -nz:nz, 0:nx))
    do ix = 0, nx
      do iz = -nz, nz
        buf(:, 0:ny - 1, iz, ix) = transpose(V(0:ny - 1, iz, ix, CPL_ORDER))
        buf(:, ny:ny + 1, iz, ix) = buf(:, 0:1, iz, ix)
        buf(:, -2:-1, iz, ix) = buf(:, ny - 2:ny - 1, iz, ix)
      end do
    end do
    open (newunit=unit, file=trim(filename), access='stream', form='unformatted', status='replace', action='write')
    write (unit) head
    write (unit) buf
    close (unit)
    deallocate (buf)
  end subroutine restart_write

  ! One device field as a headerless CPL file of one component (field_write,
  ! src/hst_io.f90:273-289): rows 0..ny-1, then ny, ny+1 as copies of 0, 1
  ! and -2, -1 as copies of ny-2, ny-1, in the file order (-2..ny+1, -nz..nz,
  ! 0..nx) with y fastest.
  subroutine field_write(filename, fid)
    character(len=*), intent(in) :: filename
    integer(C_INT), intent(in) :: fid
    complex(C_DOUBLE_COMPLEX), allocatable :: f(:, :, :), buf(:, :, :)
    integer(C_INT) :: ix, iz
    integer :: unit
    allocate (f(-2:ny + 1, -nz:nz, 0:nx), buf(-2:ny + 1, -nz:nz, 0:nx))
    call gpu_check(hst_gpu_download_field(fid, f), 'download field')
    do ix = 0, nx
      do iz = -nz, nz
        buf(0:ny - 1, iz, ix) = f(0:ny - 1, iz, ix)
        buf(ny:ny + 1, iz, ix) = f(0:1, iz, ix)
        buf(-2:-1, iz, ix) = f(ny - 2:ny - 1, iz, ix)
      end do
    end do
    open (newunit=unit, file=trim(filename), access='stream', form='unformatted', status='replace', action='write')
    write (unit) buf
    close (unit)
    deallocate (f, buf)
  end subroutine field_write

  real(C_DOUBLE) function header_value(head, key)
    character(len=*), intent(in) :: head, key
    integer :: p, q
    p = index(head, key)
    if (p == 0) then
      header_value = -huge(1.0d0)
      return
    end if
    p = p + len(key)
    q = p
    do while (q <= len(head) .and. head(q:q) /= ' ' .and. head(q:q) /= TAB .and. head(q:q) /= LF)
      q = q + 1
    end do
    read (head(p:q - 1), *) header_value
  end function header_value

  function str_i(i) result(s)
    integer(C_INT), intent(in) :: i
    character(len=:), allocatable :: s
    character(len=32) :: t
    write (t, '(I0)') i
    s = trim(t)
  end function str_i

  ! G0.17 with the leading zero of a fraction written out (0.1..., -0.1...)
  function str_r(x) result(s)
    real(C_DOUBLE), intent(in) :: x
    character(len=:), allocatable :: s
    character(len=32) :: t
    write (t, '(G0.17)') x
    s = trim(adjustl(t))
    if (s(1:1) == '.') s = '0'//s
    if (len(s) > 1) then
      if (s(1:2) == '-.') s = '-0'//s(2:)
    end if
  end function str_r

  function raw8(x) result(s)
    real(C_DOUBLE), intent(in) :: x
    character(len=8) :: s
    s = transfer(x, s)
  end function raw8

end module hst_io
