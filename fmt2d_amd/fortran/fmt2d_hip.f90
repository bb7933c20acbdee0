! Forward modelling on the MI355X library: the host program of the
! reference's fmt2d with the twelve time fields of every OBS and all rays
! computed on the GPU through the C ABI (module fmt2d_hip,
! include/fmt2d_hip.h), in the forward program's variant
! (FMT2D_FWD_PROGRAM).
!
! Control flow of src/fmt2d_main.f90: fmt2d.in (:59-68), the model and
! geometry files, receivers in the water moved to the seafloor (:463-479),
! the rays of each receiver line (an explicit phase 1..5, or Pg plus PsedP,
! PmP and Pg_ss by offset windows, :486-735), the forward of all OBS and
! their rays (:757-1363), then write_arrivals2d (arrivals, observed times
! and entry points, :1655-1719) and, with write_rays, write_rays_for_source
! (:1722-1800).  The files carry the reference's names and edit
! descriptors, so they compare byte for byte with the reference's on the
! same input.  Run in a directory holding fmt2d.in and the files it names.
program fmt2d_hip_forward
  use iso_c_binding
  use fmt2d_hip
  use fmt2d_io
  implicit none

  integer, parameter :: dp = c_double
  integer(c_int), parameter :: FWD_PROGRAM = 2
  integer, parameter :: nsec_of_path(5) = [1, 2, 2, 4, 4]

  character(len=256) :: smesh_file, refl_file, sed_file, geom_file
  character(len=256) :: arrivals_file, obs_file, rays_prefix
  logical :: write_rays

  integer :: nx, nz, nsed, nmoho, nobs, nrec, nray
  real(dp) :: vwater, vair
  real(dp), allocatable :: xpos(:), topo(:), zrel(:), vel(:,:)
  real(dp), allocatable :: sed_x(:), sed_z(:), moho_x(:), moho_z(:)
  real(dp), allocatable :: obs_x(:), obs_z(:), src_z(:), rec_x(:), rec_z(:), rec_t(:)
  integer, allocatable :: obs_id(:), rec_phase(:)
  logical, allocatable :: in_water(:)

  ! rays in receiver order, each receiver's rays in the reference's order
  integer(c_int), allocatable :: ray_rec(:), ray_src(:), ray_path(:), valid(:)
  real(dp), allocatable :: ray_x(:), ray_z(:), t_calc(:), wtime(:), entry_x(:), entry_z(:)
  integer, allocatable :: direct(:)

  type(c_ptr) :: ctx
  integer :: count0, count1, count_rate

  call system_clock(count0, count_rate)
  write(*,'(A)') '============================================'
  write(*,'(A)') '  FMT2D-HIP - forward modelling on the MI355X library'
  write(*,'(A)') '============================================'

  open(1, file='fmt2d.in', status='old', action='read')
  read(1, '(A)') smesh_file
  read(1, '(A)') refl_file
  read(1, '(A)') sed_file
  read(1, '(A)') geom_file
  read(1, *) write_rays
  read(1, '(A)') arrivals_file
  read(1, '(A)') obs_file
  read(1, '(A)') rays_prefix
  close(1)

  call read_smesh(smesh_file, nx, nz, vwater, vair, xpos, topo, zrel, vel)
  call read_refl(refl_file, nmoho, moho_x, moho_z)
  call read_refl(sed_file, nsed, sed_x, sed_z)
  call read_geom(geom_file, nobs, obs_x, obs_z, nrec, rec_x, rec_z, obs_id, rec_phase, rec_t)

  call fmt2d_check(fmt2d_ctx_create(0_c_int, ctx), 'fmt2d_ctx_create')
  call fmt2d_check(fmt2d_set_model(ctx, int(nx, c_int), int(nz, c_int), xpos, topo, zrel, vel, vwater, &
       int(nsed, c_int), sed_x, sed_z, int(nmoho, c_int), moho_x, moho_z), 'fmt2d_set_model')

  call setup_rays()

  write(*,'(A)') '  Starting forward modeling...'
  call fmt2d_check(fmt2d_forward(ctx, int(nobs, c_int), obs_x, FWD_PROGRAM), 'fmt2d_forward')
  call fmt2d_check(fmt2d_trace(ctx, int(nray, c_int), ray_src, ray_path, ray_x, ray_z, t_calc, valid), &
       'fmt2d_trace')
  call collect_rays()
  write(*,'(A,I0,A,I0)') '  Rays: ', nray, '  valid: ', count(valid /= 0)

  call write_arrivals()
  if (write_rays) call write_ray_files()

  call fmt2d_check(fmt2d_ctx_destroy(ctx), 'fmt2d_ctx_destroy')
  call system_clock(count1)
  write(*,'(A,F12.2,A)') '  Total time: ', dble(count1 - count0) / dble(count_rate), ' sec'
  write(*,'(A)') '  FMT2D finished'

contains

  ! Receivers above the seafloor are in the water (:463-479).  Rays per
  ! receiver line (:669-735): the phase given, or by source-receiver offset
  ! Pg always, PsedP within 3 km .. 6 x the mean sediment thickness, PmP
  ! within 3-30 km and Pg_ss within 50 km.
  subroutine setup_rays()
    real(dp) :: avg_sed_thick, psedp_max_offset, off
    integer :: n, r, pass, k
    allocate(in_water(nrec), src_z(nobs))
    do n = 1, nrec
      in_water(n) = rec_z(n) < iface_z(xpos, topo, rec_x(n))
    end do
    do n = 1, nobs
      src_z(n) = iface_z(xpos, topo, obs_x(n))
    end do
    avg_sed_thick = 0.0_dp
    do n = 1, nsed
      avg_sed_thick = avg_sed_thick + (sed_z(n) - iface_z(xpos, topo, sed_x(n)))
    end do
    avg_sed_thick = avg_sed_thick / dble(nsed)
    psedp_max_offset = 6.0_dp * avg_sed_thick
    do pass = 1, 2
      r = 0
      do n = 1, nrec
        if (rec_phase(n) >= 1 .and. rec_phase(n) <= 5) then
          call add_ray(pass, r, n, rec_phase(n))
        else
          off = abs(rec_x(n) - obs_x(obs_id(n)))
          call add_ray(pass, r, n, 1)
          if (off >= 3.0_dp .and. off <= psedp_max_offset) call add_ray(pass, r, n, 2)
          if (off >= 3.0_dp .and. off <= 30.0_dp) call add_ray(pass, r, n, 3)
          if (off <= 50.0_dp) call add_ray(pass, r, n, 4)
        end if
      end do
      if (pass == 1) then
        nray = r
        allocate(ray_rec(nray), ray_src(nray), ray_path(nray), valid(nray), ray_x(nray), ray_z(nray))
        allocate(t_calc(nray), wtime(nray), entry_x(nray), entry_z(nray), direct(nray))
      end if
    end do
    k = 0
    write(*,'(A,F8.2,A)') '  PsedP max offset:       ', psedp_max_offset, ' km'
    write(*,'(A,I0,A,I0)') '  Receivers: ', nrec, '  rays: ', nray
  end subroutine setup_rays

  subroutine add_ray(pass, r, n, path)
    integer, intent(in) :: pass, n, path
    integer, intent(inout) :: r
    r = r + 1
    if (pass == 1) return
    ray_rec(r) = n
    ray_src(r) = obs_id(n) - 1
    ray_path(r) = path
    ray_x(r) = rec_x(n)
    ray_z(r) = rec_z(n)
  end subroutine add_ray

  ! Water leg, entry point and direct-water flag of every ray.  The entry
  ! point is reported where the reference records one: receivers in the
  ! water, rays of Pg, PsedP and PmP (:1244-1247, :1318-1322).
  subroutine collect_rays()
    integer :: r
    integer(c_int) :: npts(4), flags(3)
    real(dp) :: exz(2), wt
    do r = 1, nray
      call fmt2d_check(fmt2d_ray_info(ctx, int(r - 1, c_int), exz, wt, npts, flags), 'fmt2d_ray_info')
      wtime(r) = wt
      direct(r) = flags(3)
      entry_x(r) = 0.0_dp
      entry_z(r) = 0.0_dp
      if (in_water(ray_rec(r)) .and. ray_path(r) <= 3) then
        entry_x(r) = exz(1)
        entry_z(r) = exz(2)
      end if
    end do
  end subroutine collect_rays

  ! write_arrivals2d (:1655-1719)
  subroutine write_arrivals()
    integer :: r, nvalid
    real(dp) :: t_arr, t_solid
    nvalid = count(valid /= 0)
    open(40, file=trim(arrivals_file), status='replace', action='write')
    open(41, file=trim(obs_file), status='replace', action='write')
    write(41, '(I0)') nvalid
    do r = 1, nray
      if (valid(r) /= 0) then
        t_arr = t_calc(r)
        t_solid = t_arr - wtime(r)
      else
        t_arr = -1.0_dp
        t_solid = -1.0_dp
      end if
      write(40, '(3I8,2F15.6,F12.6)') ray_rec(r), ray_src(r) + 1, ray_path(r), t_arr, t_solid, wtime(r)
      if (valid(r) /= 0) write(41, '(3I8,F15.6)') ray_rec(r), ray_src(r) + 1, ray_path(r), t_arr
    end do
    close(40)
    close(41)
    open(42, file='entry_points.dat', status='replace', action='write')
    write(42, '(I0)') nvalid
    do r = 1, nray
      if (valid(r) /= 0) write(42, '(3I8,2F15.6,F12.6)') ray_rec(r), ray_src(r) + 1, ray_path(r), &
           entry_x(r), entry_z(r), wtime(r)
    end do
    close(42)
    write(*,'(A,A,A,I0,A)') '  Observed times written to ', trim(obs_file), ' (', nvalid, ' valid)'
    write(*,'(A,A)') '  Arrivals written to ', trim(arrivals_file)
  end subroutine write_arrivals

  ! write_rays_for_source (:1722-1800): per OBS, each valid ray from the
  ! receiver: a direct water wave as a straight line, else the water leg down
  ! to the ray's entry point, then the sections' points from the receiver
  ! end to the source.
  subroutine write_ray_files()
    character(len=64) :: fname
    real(dp), allocatable :: px(:), pz(:)
    real(dp) :: x_rcv, z_rcv, x_entry, z_sf, dist_direct, sx, sz
    integer :: isrc, r, n, k, isec, npts_water
    integer(c_int) :: npts(4), flags(3), np
    real(dp) :: exz(2), wt
    do isrc = 1, nobs
      write(fname, '(A,A,I0,A)') trim(rays_prefix), '_src', isrc, '.dat'
      open(50, file=trim(fname), status='replace', action='write')
      sx = obs_x(isrc)
      sz = src_z(isrc)
      do r = 1, nray
        if (ray_src(r) + 1 /= isrc) cycle
        if (valid(r) == 0) cycle
        n = ray_rec(r)
        write(50, '(A,I0,1X,I0,1X,I0,1X,F12.6)') '> ', isrc, n, ray_path(r), t_calc(r)
        x_rcv = rec_x(n)
        z_rcv = rec_z(n)
        if (direct(r) /= 0) then
          dist_direct = sqrt((x_rcv - sx)**2 + (z_rcv - sz)**2)
          npts_water = max(3, nint(dist_direct / 0.1_dp) + 1)
          do k = 1, npts_water
            write(50, '(F12.6,1X,F12.6)') x_rcv + (sx - x_rcv) * dble(k - 1) / dble(npts_water - 1), &
                 z_rcv + (sz - z_rcv) * dble(k - 1) / dble(npts_water - 1)
          end do
          write(50, *)
          cycle
        end if
        call fmt2d_check(fmt2d_ray_info(ctx, int(r - 1, c_int), exz, wt, npts, flags), 'fmt2d_ray_info')
        x_entry = x_rcv
        z_sf = iface_z(xpos, topo, x_rcv)
        if (in_water(n)) then
          isec = nsec_of_path(ray_path(r))
          if (npts(isec) > 0) then
            call section_points(r, isec, npts(isec), px, pz)
            x_entry = px(1)
            z_sf = pz(1)
          else
            x_entry = entry_x(r)
            z_sf = entry_z(r)
          end if
        end if
        if (z_sf > z_rcv + 0.001_dp) then
          write(50, '(F12.6,1X,F12.6)') x_rcv, z_rcv
          npts_water = max(2, nint(abs(z_sf - z_rcv) / 0.05_dp) + 1)
          do k = 1, npts_water
            write(50, '(F12.6,1X,F12.6)') x_rcv + (x_entry - x_rcv) * dble(k) / dble(npts_water), &
                 z_rcv + (z_sf - z_rcv) * dble(k) / dble(npts_water)
          end do
        end if
        do isec = nsec_of_path(ray_path(r)), 1, -1
          if (npts(isec) <= 0) cycle
          call section_points(r, isec, npts(isec), px, pz)
          do k = 1, npts(isec)
            write(50, '(F12.6,1X,F12.6)') px(k), pz(k)
          end do
        end do
        write(50, *)
      end do
      close(50)
    end do
  end subroutine write_ray_files

  subroutine section_points(r, isec, n, px, pz)
    integer, intent(in) :: r, isec
    integer(c_int), intent(in) :: n
    real(dp), allocatable, intent(inout) :: px(:), pz(:)
    integer(c_int) :: np
    if (allocated(px)) deallocate(px, pz)
    allocate(px(n), pz(n))
    call fmt2d_check(fmt2d_ray_points(ctx, int(r - 1, c_int), int(isec, c_int), px, pz, n, np), 'fmt2d_ray_points')
  end subroutine section_points

end program fmt2d_hip_forward
