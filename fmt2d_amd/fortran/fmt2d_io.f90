! Readers of the reference's text inputs (read_tomo2d.f90) and the interface
! interpolation the host programs share.  Errors stop the program with the
! reference's messages.
module fmt2d_io
  use iso_c_binding, only: c_double
  implicit none
  private
  public :: read_smesh, read_refl, read_geom, iface_z

  integer, parameter :: dp = c_double

contains

  ! .smesh: nx nz vwater vair, xpos, topo, zpos_rel, then the velocities of
  ! each x node (read_tomo2d.f90:8-79); kept as vel(iz, ix), the layout the
  ! library takes.
  subroutine read_smesh(filename, nx, nz, vwater, vair, xpos, topo, zrel, vel)
    character(len=*), intent(in) :: filename
    integer, intent(out) :: nx, nz
    real(dp), intent(out) :: vwater, vair
    real(dp), allocatable, intent(out) :: xpos(:), topo(:), zrel(:), vel(:,:)
    integer :: ios, ix
    open(10, file=trim(filename), status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(*,'(A,A)') 'Error: cannot open smesh file: ', trim(filename)
      stop 1
    end if
    read(10, *, iostat=ios) nx, nz, vwater, vair
    if (ios /= 0) then
      write(*,'(A)') 'Error: reading smesh header (nx, nz, vwater, vair)'
      stop 1
    end if
    if (nx < 2 .or. nz < 2) then
      write(*,'(A,I0,A,I0)') 'Error: smesh grid too small: nx=', nx, ' nz=', nz
      stop 1
    end if
    allocate(xpos(nx), topo(nx), zrel(nz), vel(nz, nx))
    read(10, *, iostat=ios) xpos(1:nx)
    if (ios == 0) read(10, *, iostat=ios) topo(1:nx)
    if (ios == 0) read(10, *, iostat=ios) zrel(1:nz)
    do ix = 1, nx
      if (ios == 0) read(10, *, iostat=ios) vel(1:nz, ix)
    end do
    if (ios /= 0) then
      write(*,'(A,A)') 'Error: reading smesh file ', trim(filename)
      stop 1
    end if
    close(10)
    write(*,'(A,A,A,I0,A,I0)') '  Read smesh: ', trim(filename), '  nx = ', nx, '  nz = ', nz
  end subroutine read_smesh

  ! Interface file: "x z" per non-empty line, malformed lines skipped
  ! (read_tomo2d.f90:83-137).
  subroutine read_refl(filename, n, x, z)
    character(len=*), intent(in) :: filename
    integer, intent(out) :: n
    real(dp), allocatable, intent(out) :: x(:), z(:)
    integer, parameter :: max_lines = 100000
    real(dp), allocatable :: tx(:), tz(:)
    character(len=512) :: line
    integer :: ios
    open(11, file=trim(filename), status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(*,'(A,A)') 'Error: cannot open refl file: ', trim(filename)
      stop 1
    end if
    allocate(tx(max_lines), tz(max_lines))
    n = 0
    do
      read(11, '(A)', iostat=ios) line
      if (ios /= 0) exit
      if (len_trim(line) == 0) cycle
      n = n + 1
      if (n > max_lines) then
        write(*,'(A)') 'Error: refl file exceeds max_lines'
        stop 1
      end if
      read(line, *, iostat=ios) tx(n), tz(n)
      if (ios /= 0) n = n - 1
    end do
    close(11)
    if (n < 1) then
      write(*,'(A,A)') 'Error: refl file has no valid data: ', trim(filename)
      stop 1
    end if
    allocate(x(n), z(n))
    x = tx(1:n)
    z = tz(1:n)
    write(*,'(A,A,A,I0)') '  Read refl: ', trim(filename), '  nx_refl = ', n
  end subroutine read_refl

  ! Geometry: nobs, then per OBS "S x z nrec" and nrec lines
  ! "R x z ray_type t_obs unc" (read_tomo2d.f90:141-248), read in two passes.
  subroutine read_geom(filename, nobs, obs_x, obs_z, nrec, rec_x, rec_z, obs_id, rec_phase, rec_t)
    character(len=*), intent(in) :: filename
    integer, intent(out) :: nobs, nrec
    real(dp), allocatable, intent(out) :: obs_x(:), obs_z(:), rec_x(:), rec_z(:), rec_t(:)
    integer, allocatable, intent(out) :: obs_id(:), rec_phase(:)
    character(len=1) :: tag
    integer :: ios, iobs, irec, pass, nr, idx, rt
    real(dp) :: rx, rz, tt, unc
    open(12, file=trim(filename), status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(*,'(A,A)') 'Error: cannot open geom file: ', trim(filename)
      stop 1
    end if
    nrec = 0
    do pass = 1, 2
      rewind(12)
      read(12, *, iostat=ios) nobs
      if (ios /= 0 .or. nobs < 1) then
        write(*,'(A)') 'Error: reading nobs from geom file'
        stop 1
      end if
      if (pass == 1) allocate(obs_x(nobs), obs_z(nobs))
      idx = 0
      do iobs = 1, nobs
        read(12, *, iostat=ios) tag, obs_x(iobs), obs_z(iobs), nr
        if (ios /= 0) then
          write(*,'(A,I0)') 'Error: reading OBS header for iobs = ', iobs
          stop 1
        end if
        do irec = 1, nr
          read(12, *, iostat=ios) tag, rx, rz, rt, tt, unc
          if (ios /= 0) then
            write(*,'(A,I0,A,I0)') 'Error: reading rec for iobs = ', iobs, ' irec = ', irec
            stop 1
          end if
          idx = idx + 1
          if (pass == 2) then
            if (idx > nrec) then
              write(*,'(A)') 'Error: geom file inconsistent between passes (idx > nrec_total)'
              stop 1
            end if
            rec_x(idx) = rx
            rec_z(idx) = rz
            obs_id(idx) = iobs
            rec_phase(idx) = rt
            rec_t(idx) = tt
          end if
        end do
      end do
      if (pass == 1) then
        nrec = idx
        allocate(rec_x(nrec), rec_z(nrec), obs_id(nrec), rec_phase(nrec), rec_t(nrec))
      end if
    end do
    close(12)
    write(*,'(A,A,A,I0,A,I0)') '  Read geom: ', trim(filename), '  nobs = ', nobs, '  nrec_total = ', nrec
  end subroutine read_geom

  ! interpolate_iface_z (invert2d_main.f90:906-930, fmt2d_main.f90:1389-1415)
  ! on nodes x(:), z(:) spaced x(2) - x(1)
  real(dp) function iface_z(x, z, px)
    real(dp), intent(in) :: x(:), z(:), px
    integer :: n, ixx
    real(dp) :: d, w
    n = size(x)
    if (n < 2) then
      iface_z = z(1)
      return
    end if
    d = x(2) - x(1)
    ixx = floor((px - x(1)) / d) + 1
    if (ixx < 1) then
      iface_z = z(1)
      return
    end if
    if (ixx >= n) then
      iface_z = z(n)
      return
    end if
    w = (px - x(ixx)) / d
    if (w < 0.0_dp) w = 0.0_dp
    if (w > 1.0_dp) w = 1.0_dp
    iface_z = (1.0_dp - w) * z(ixx) + w * z(ixx + 1)
  end function iface_z

end module fmt2d_io
