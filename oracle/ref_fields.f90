! Field-dump driver for the reference FMT2D modules (oracle test infrastructure).
!
! Builds the propagation model exactly as the reference's executables do
! (src/fmt2d_main.f90:91-418, src/invert2d_main.f90:139-480: five interfaces,
! z-resampled velocity grid, intersections, regions 1-6, regular-node tags,
! velocities) by calling the reference's own routines, then runs the
! per-source stage schedule of run_forward2d (src/inversion2d.f90:4184-4416)
! for every listed source and writes the raw region fields after each stage.
!
! It is linked against the reference modules compiled from the reference sources
! by oracle/build_ref.sh and never ships with the product.
!
! Input (ref_fields.in, one item per line):
!   smesh file, Moho file, sediment-base file, output file ("none": no dump),
!   number of sources, then one source x per line,
!   number of stages to run per source (1 = tf1 only, 12 = all).
! Output (stream binary, little endian):
!   header ints: nx, nz ; doubles: dx, x0, z0
!   fully_regular(nx,nz) int32, node_region(nx,nz) int32
!   for region 1..6: nnode, ngnode, i1(nnode), i2(nnode), i3(nnode),
!                    x(nnode), z(nnode), velocity(nnode), region_rnode_id(nx,nz)
!   for intersection 1..5: nnode, x, z, normal(2,nnode), vel_top, vel_bot
!   then per source, per stage: itf, region id, T(nnode), grad(2,nnode)
program ref_fields
  use type_definitions2d
  use global_variables2d
  use read_tomo2d
  use mod_fmt2d
  implicit none

  character(len=512) :: f_smesh, f_moho, f_sed, f_out
  integer :: nsrc, nstage, is, n, m, ix, iz, ireg, ipath, k, itf, itf_prev
  real(dp), allocatable :: src_x(:)
  integer :: nx_s, nz_s, nx_m, nx_sd, u
  real(dp) :: vwater, vair, sx, sz
  real(dp), allocatable :: xpos(:), topo(:), zrel(:), vel(:,:)
  real(dp), allocatable :: mx(:), mz(:), sdx(:), sdz(:)
  type(Tregion2d), pointer :: reg
  type(Ttime_field2d), pointer :: tf, tfp
  ! stage table of run_forward2d: (path, position) -> tf, region, istart
  integer, parameter :: npath = 5
  integer, parameter :: path_ntf(npath) = (/1, 2, 2, 4, 4/)
  integer, parameter :: path_tf(4, npath) = reshape((/ 1,0,0,0,  2,3,0,0,  4,5,0,0, &
                                                       1,6,7,8,  9,10,11,12 /), (/4, npath/))
  integer, parameter :: tf_region(12) = (/5, 2, 2, 6, 6, 1, 1, 1, 1, 1, 5, 1/)
  integer, parameter :: tf_istart(12) = (/2, 2, 3, 2, 4, 2, 1, 2, 2, 1, 2, 2/)
  type(Ttime_field2d), target :: tfs(12)
  logical :: reflect
  integer :: m_top, m_below
  integer(8) :: c0, c1, crate
  logical :: dump

  open(10, file='ref_fields.in', status='old', action='read')
  read(10, '(A)') f_smesh
  read(10, '(A)') f_moho
  read(10, '(A)') f_sed
  read(10, '(A)') f_out
  read(10, *) nsrc
  allocate(src_x(nsrc))
  do is = 1, nsrc
    read(10, *) src_x(is)
  end do
  read(10, *) nstage
  close(10)

  call read_smesh(f_smesh, nx_s, nz_s, vwater, xpos, topo, zrel, vel, vair)
  vwater_global = vwater
  vair_global = vair
  call read_refl(f_moho, nx_m, mx, mz)
  call read_refl(f_sed, nx_sd, sdx, sdz)

  ! ---- interfaces: top, seafloor, sediment base, Moho, bottom
  n_interfaces = 5
  allocate(intrface(5))
  do n = 1, 5
    call interface2d_defaults(intrface(n))
    intrface(n)%id = n
  end do
  call set_iface(intrface(1), xpos, spread(0.0_dp, 1, nx_s))
  call set_iface(intrface(2), xpos, topo)
  call set_iface(intrface(3), sdx, sdz)
  call set_iface(intrface(4), mx, mz)
  call set_iface(intrface(5), xpos, spread(zrel(nz_s) + maxval(topo), 1, nx_s))

  ! ---- propagation grid (square cells of the smesh x spacing)
  n_vgrids = 1
  n_vtypes = 1
  allocate(vgrid(1, NVTYPE))
  do n = 1, NVTYPE
    call vgrid2d_defaults(vgrid(1, n))
  end do
  allocate(pgrid)
  call pgrid2d_defaults(pgrid)
  pgrid%nx = nx_s
  pgrid%dz = xpos(2) - xpos(1)
  pgrid%dx = pgrid%dz
  pgrid%z0 = 0.0_dp
  pgrid%zmax = zrel(nz_s) + maxval(topo) + 5.0_dp
  pgrid%nz = nint((pgrid%zmax - pgrid%z0) / pgrid%dz) + 1
  pgrid%x0 = xpos(1)
  pgrid%xmax = xpos(nx_s)
  pgrid%tolerance = min(pgrid%dx, pgrid%dz) * 0.01_dp
  pgrid%is_main_grid = .true.
  pgrid%index_x0 = 1
  allocate(pgrid%x(pgrid%nx), pgrid%z(pgrid%nz))
  do n = 1, pgrid%nx
    pgrid%x(n) = pgrid%x0 + (n-1)*pgrid%dx
  end do
  do n = 1, pgrid%nz
    pgrid%z(n) = pgrid%z0 + (n-1)*pgrid%dz
  end do

  call resample_velocity()

  ! ---- intersections and regions
  n_intersections = 5
  allocate(intersection(5))
  do n = 1, 5
    call intersection2d_defaults(intersection(n))
    intersection(n)%id = n
  end do
  do n = 1, 5
    call find_intersection2d(intersection(n), intrface(n), pgrid)
  end do
  n_regions = 6
  allocate(region(6))
  do n = 1, 4
    call region2d_defaults(region(n))
    region(n)%id = n
    region(n)%ivgrid = 1
    region(n)%itop => intersection(n)
    region(n)%ibot => intersection(n+1)
    call define_region2d(region(n), region(n)%itop, region(n)%ibot, pgrid)
  end do
  if (associated(intersection(2)%rright_node_id)) deallocate(intersection(2)%rright_node_id)
  if (associated(intersection(5)%rleft_node_id)) deallocate(intersection(5)%rleft_node_id)
  call region2d_defaults(region(5))
  region(5)%id = 5
  region(5)%ivgrid = 1
  region(5)%itop => intersection(2)
  region(5)%ibot => intersection(5)
  call define_region2d(region(5), region(5)%itop, region(5)%ibot, pgrid)
  intersection(2)%regright => region(5)
  intersection(5)%regleft => region(5)
  call region2d_defaults(region(6))
  region(6)%id = 6
  region(6)%ivgrid = 1
  region(6)%itop => intersection(2)
  region(6)%ibot => intersection(4)
  call define_region2d(region(6), region(6)%itop, region(6)%ibot, pgrid)
  call tag_regular_nodes2d(pgrid)
  call velocities_to_grid2d(pgrid)
  do n = 1, 5
    call velocities_to_intersection2d(intersection(n))
  end do
  do n = 1, 6
    call velocities_to_region2d(region(n), pgrid)
  end do
  intersection(1)%regleft => null()
  intersection(1)%regright => region(1)
  intersection(3)%regleft => region(2)
  intersection(3)%regright => region(3)
  intersection(4)%regleft => region(3)
  intersection(4)%regright => region(4)
  intersection(5)%regright => null()

  ! ---- geometry dump (output file "none": timing run without dumps)
  dump = trim(f_out) /= 'none'
  if (dump) open(newunit=u, file=trim(f_out), access='stream', form='unformatted', status='replace')
  if (dump) then
  write(u) pgrid%nx, pgrid%nz, pgrid%dx, pgrid%x0, pgrid%z0
  write(u) ((merge(1, 0, pgrid%fully_regular(ix, iz)), ix=1,pgrid%nx), iz=1,pgrid%nz)
  write(u) ((pgrid%node_region(ix, iz), ix=1,pgrid%nx), iz=1,pgrid%nz)
  do ireg = 1, 6
    reg => region(ireg)
    write(u) reg%nnode, reg%ngnode
    write(u) (reg%node(n)%i1, n=1,reg%nnode), (reg%node(n)%i2, n=1,reg%nnode), &
             (reg%node(n)%i3, n=1,reg%nnode)
    write(u) reg%x(1:reg%nnode), reg%z(1:reg%nnode), reg%velocity(VTYPE_P, 1:reg%nnode)
    write(u) ((reg%region_rnode_id(ix, iz), ix=1,pgrid%nx), iz=1,pgrid%nz)
  end do
  do n = 1, 5
    write(u) intersection(n)%nnode
    write(u) intersection(n)%x, intersection(n)%z, intersection(n)%normal, &
             intersection(n)%vel_top(VTYPE_P, :), intersection(n)%vel_bot(VTYPE_P, :)
  end do
  endif

  ! ---- per-source stage schedule (run_forward2d, without ray tracing)
  do n = 1, 12
    call time_field2d_defaults(tfs(n))
    tfs(n)%id = n
    tfs(n)%reg => region(tf_region(n))
    tfs(n)%vtype = VTYPE_P
    tfs(n)%istart => intersection(tf_istart(n))
    allocate(tfs(n)%arrivaltime(region(tf_region(n))%nnode))
    allocate(tfs(n)%time_gradient(2, region(tf_region(n))%nnode))
  end do
  call system_clock(c0, crate)
  do is = 1, nsrc
    sx = src_x(is)
    sz = iface_z_at(intrface(2), sx)
    do ipath = 1, npath
      do k = 1, path_ntf(ipath)
        itf = path_tf(k, ipath)
        if (itf > nstage) cycle
        tf => tfs(itf)
        reg => tf%reg
        reg%node_status = -1
        reg%arrivaltime = huge_time
        reg%time_gradient = 0.0_dp
        if (k == 1) then
          call initialize_region_from_source2d(reg, sx, sz, 0.0_dp, VTYPE_P)
        else
          itf_prev = path_tf(k-1, ipath)
          tfp => tfs(itf_prev)
          tfp%reg%arrivaltime = tfp%arrivaltime
          tfp%reg%time_gradient = tfp%time_gradient
          call transfer_times_to_intersection(tf%istart, tfp)
          reflect = (ipath == 2 .and. k == 2) .or. (ipath == 3 .and. k == 2) .or. &
                    (ipath == 4 .and. k >= 2) .or. (ipath == 5 .and. k >= 2)
          if (reflect) call reflect_gradient_at_interface(tf%istart, reg, VTYPE_P)
          reg%node_status = -1
          reg%arrivaltime = huge_time
          reg%time_gradient = 0.0_dp
          call sweep_region_from_interface2d(reg, tf%istart, VTYPE_P, source_dummy())
        endif
        call propagate2d(reg, VTYPE_P)
        if (reg%id == 1) then
          if ((ipath == 4 .and. (k == 2 .or. k == 4)) .or. (ipath == 5 .and. (k == 1 .or. k == 4))) then
            do ix = 1, pgrid%nx
              m_top = reg%region_rnode_id(ix, 1)
              m_below = reg%region_rnode_id(ix, 2)
              if (m_top > 0 .and. m_below > 0) then
                if (reg%time_gradient(2, m_top) > 0.0_dp .or. abs(reg%time_gradient(2, m_top)) < 0.01_dp) &
                  reg%time_gradient(1:2, m_top) = reg%time_gradient(1:2, m_below)
              endif
            end do
          else if ((ipath == 4 .and. k == 3) .or. (ipath == 5 .and. k == 2)) then
            do ix = 1, pgrid%nx
              m_top = reg%region_rnode_id(ix, 1)
              m_below = reg%region_rnode_id(ix, 2)
              if (m_top > 0 .and. m_below > 0) then
                if (reg%time_gradient(2, m_top) < 0.0_dp .or. abs(reg%time_gradient(2, m_top)) < 0.01_dp) &
                  reg%time_gradient(1:2, m_top) = reg%time_gradient(1:2, m_below)
              endif
            end do
          endif
        endif
        tf%arrivaltime = reg%arrivaltime
        tf%time_gradient = reg%time_gradient
        if (dump) then
          write(u) itf, reg%id
          write(u) reg%arrivaltime(1:reg%nnode), reg%time_gradient(1:2, 1:reg%nnode)
        endif
      end do
    end do
  end do
  call system_clock(c1)
  if (dump) close(u)
  write(*,'(A,F12.6)') 'stage loop seconds: ', real(c1 - c0, dp) / real(crate, dp)

contains

  subroutine set_iface(f, x, z)
    type(Tinterface2d) :: f
    real(dp) :: x(:), z(:)
    f%nx = size(x)
    f%nnode = size(x)
    f%dx = x(2) - x(1)
    f%x0 = x(1)
    allocate(f%x(size(x)), f%z_nodes(size(x)))
    f%x = x
    f%z_nodes = z
  end subroutine set_iface

  function iface_z_at(f, px) result(zv)
    ! linear interpolation on the interface nodes (fmt2d_main.f90:1389-1415)
    type(Tinterface2d) :: f
    real(dp) :: px, zv, w
    integer :: i
    i = floor((px - f%x(1)) / f%dx) + 1
    if (i < 1) then
      zv = f%z_nodes(1)
      return
    endif
    if (i >= f%nx) then
      zv = f%z_nodes(f%nx)
      return
    endif
    w = min(1.0_dp, max(0.0_dp, (px - f%x(i)) / f%dx))
    zv = (1.0_dp - w) * f%z_nodes(i) + w * f%z_nodes(i+1)
  end function iface_z_at

  function source_dummy() result(s)
    type(Tsource2d) :: s
    call source2d_defaults(s)
  end function source_dummy

  subroutine resample_velocity()
    ! Velocity on the uniform-dz grid: water above the seafloor, linear in
    ! depth-below-seafloor inside the smesh, last row below it
    ! (fmt2d_main.f90:225-321).
    integer :: i, j, jo, nzn
    real(dp) :: zsf, pr, w
    type(Tvelocity_grid2d), pointer :: vg
    nzn = pgrid%nz
    vg => vgrid(1, VTYPE_P)
    vg%nx = nx_s
    vg%nz = nzn
    vg%dx = xpos(2) - xpos(1)
    vg%dz = pgrid%dz
    vg%x0 = xpos(1)
    vg%z0 = 0.0_dp
    vg%nnode = nx_s * nzn
    allocate(vg%x(nx_s), vg%z(nzn), vg%velocity(NVTYPE, nx_s, nzn), &
             vg%velocity_init(NVTYPE, nx_s, nzn), vg%active(NVTYPE, nx_s, nzn))
    vg%x = xpos
    do j = 1, nzn
      vg%z(j) = (j - 1) * pgrid%dz
    end do
    vg%velocity = 0.0_dp
    vg%active = .false.
    do i = 1, nx_s
      zsf = topo(i)
      do j = 1, nzn
        pr = vg%z(j) - zsf
        if (pr < 0.0_dp) then
          vg%velocity(VTYPE_P, i, j) = vwater
        else if (pr >= zrel(nz_s)) then
          vg%velocity(VTYPE_P, i, j) = vel(i, nz_s)
        else
          do jo = 1, nz_s - 1
            if (pr >= zrel(jo) .and. pr < zrel(jo+1)) exit
          end do
          if (jo >= nz_s) jo = nz_s - 1
          w = (pr - zrel(jo)) / (zrel(jo+1) - zrel(jo))
          vg%velocity(VTYPE_P, i, j) = (1.0_dp - w) * vel(i, jo) + w * vel(i, jo+1)
        endif
        if (pr >= 0.0_dp) vg%active(VTYPE_P, i, j) = .true.
      end do
    end do
    vg%velocity_init = vg%velocity
  end subroutine resample_velocity

end program ref_fields
