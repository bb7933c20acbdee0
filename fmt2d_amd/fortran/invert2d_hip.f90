! Travel-time inversion on the MI355X library: the host program of the
! reference's invert2d with the forward, rays, Frechet rows, LSQR solves and
! line-search forwards running on the GPU through the C ABI (module
! fmt2d_hip, include/fmt2d_hip.h).
!
! Control flow of src/invert2d_main.f90 (invert2d.in :60-96, frechet.in as
! initialize_inversion2d reads it, frechet2d.f90:14-128, receivers in the
! water :517-531) and of the outer loop of lsqr_inversion2d
! (src/inversion2d.f90:112-999): initial forward and residuals, then per
! iteration the alternating mode, a forward, the data selection, the
! system, the damped update, the line search, the log line, the best model
! and the stopping rules; then the best model restored, write_model2d
! (:3842-3962) and the final forward and residuals.  The files carry the
! reference's names and edit descriptors (write_residuals2d :3342-3377,
! write_ray_density2d :3380-3458), so they compare byte for byte with the
! reference's on the same input.  Run in a directory holding invert2d.in,
! frechet.in and the files they name.
program invert2d_hip
  use iso_c_binding
  use fmt2d_hip
  use fmt2d_io
  implicit none

  integer, parameter :: dp = c_double
  real(dp), parameter :: huge_time = 1.0e20_dp
  integer, parameter :: nsec_of_path(5) = [1, 2, 2, 4, 4]

  ! invert2d.in
  character(len=256) :: smesh_file, refl_file, sed_file, geom_file
  integer :: niter_inv, lsqr_max_iter_in, smooth_type, alt_inv_mode_in
  real(dp) :: lam_damp_in, lam_smooth_in, inv_crit_chi_in, target_pert_in
  real(dp) :: lam_damp_iface_in, lam_smooth_iface_in, refl_weight_in
  real(dp) :: phase_wt_pmp_in, phase_wt_pgss_in, phase_wt_psed_in, corr_x_in, corr_z_in
  real(dp) :: target_dd_in, iface_step_ratio_in, max_rel_pert_in, corr_x_iface_in, iface_dx_target_in
  ! inversion mesh spacings (km; 0: the propagation grid), an optional last
  ! line of this build's own
  real(dp) :: mesh_dx_in, mesh_dx_iface_in

  ! model and acquisition
  integer :: nx, nz, nsed, nmoho, nobs, nrec, nray
  real(dp) :: vwater, vair
  real(dp), allocatable :: xpos(:), topo(:), zrel(:), vel(:,:)
  real(dp), allocatable :: sed_x(:), sed_z(:), moho_x(:), moho_z(:)
  real(dp), allocatable :: obs_x(:), obs_z(:), rec_x(:), rec_z(:), rec_t(:)
  integer, allocatable :: obs_id(:), rec_phase(:)
  logical, allocatable :: in_water(:)

  ! one ray per receiver line of phase 1..5, in file order
  integer(c_int), allocatable :: ray_rec(:), ray_src(:), ray_path(:), valid(:), fresh(:)
  real(dp), allocatable :: ray_x(:), ray_z(:), t_obs(:), t_calc(:)

  ! inversion
  logical :: do_inversion, inv_vel
  integer :: n_inv_iface
  integer(c_int) :: iface_list(3)
  integer(c_int) :: n_vel_par, n_par
  type(c_ptr) :: ctx
  type(fmt2d_grid_info) :: ginfo

  integer :: count0, count1, count_rate

  call system_clock(count0, count_rate)
  write(*,'(A)') '============================================'
  write(*,'(A)') '  INVERT2D-HIP - inversion on the MI355X library'
  write(*,'(A)') '============================================'

  call read_invert2d_in()
  call read_smesh(smesh_file, nx, nz, vwater, vair, xpos, topo, zrel, vel)
  call read_refl(refl_file, nmoho, moho_x, moho_z)
  call read_refl(sed_file, nsed, sed_x, sed_z)
  call read_geom(geom_file, nobs, obs_x, obs_z, nrec, rec_x, rec_z, obs_id, rec_phase, rec_t)

  call fmt2d_check(fmt2d_ctx_create(0_c_int, ctx), 'fmt2d_ctx_create')
  call fmt2d_check(fmt2d_set_model(ctx, int(nx, c_int), int(nz, c_int), xpos, topo, zrel, vel, vwater, &
       int(nsed, c_int), sed_x, sed_z, int(nmoho, c_int), moho_x, moho_z), 'fmt2d_set_model')
  call fmt2d_check(fmt2d_grid_info_get(ctx, ginfo), 'fmt2d_grid_info_get')
  write(*,'(A,I6,A,I6)') '  Propagation grid: nx=', ginfo%nx, ' nz=', ginfo%nz

  call setup_rays()
  call read_frechet_in()
  if (do_inversion) then
    call fmt2d_check(fmt2d_set_inversion(ctx, merge(1_c_int, 0_c_int, inv_vel), int(n_inv_iface, c_int), &
         iface_list), 'fmt2d_set_inversion')
    if (mesh_dx_in > 0.0_dp .or. mesh_dx_iface_in > 0.0_dp) then
      call fmt2d_check(fmt2d_set_inversion_mesh(ctx, mesh_stride(mesh_dx_in), mesh_stride(mesh_dx_iface_in)), &
           'fmt2d_set_inversion_mesh')
      write(*,'(A,2I4)') '  Inversion mesh strides (velocity, interfaces): ', &
           mesh_stride(mesh_dx_in), mesh_stride(mesh_dx_iface_in)
    end if
    call fmt2d_check(fmt2d_inversion_info(ctx, n_vel_par, n_par), 'fmt2d_inversion_info')
    write(*,'(A,I8)') '  Inversion parameters: n_inv_parms = ', n_par
    if (n_par > 0) call invert()
  end if

  call fmt2d_check(fmt2d_ctx_destroy(ctx), 'fmt2d_ctx_destroy')
  call system_clock(count1)
  write(*,'(A)') '  Inversion complete'
  write(*,'(A,F12.2,A)') '  Total time: ', dble(count1 - count0) / dble(count_rate), ' sec'

contains

  ! grid nodes per inversion mesh spacing (1: the grid)
  integer(c_int) function mesh_stride(dx_mesh)
    real(dp), intent(in) :: dx_mesh
    mesh_stride = 1
    if (dx_mesh > 0.0_dp) mesh_stride = max(1, nint(dx_mesh / ginfo%dx))
  end function mesh_stride

  ! invert2d.in: four file names, fourteen numbers, then the optional
  ! smooth_type / target_dd / iface_step_ratio / max_rel_pert line, the
  ! alternating mode, corr_x_iface and iface_dx_target (invert2d_main.f90:60-96).
  ! A further optional line, "mesh_dx mesh_dx_iface", sets an inversion mesh
  ! coarser than the propagation grid (multiples of the grid spacing).
  subroutine read_invert2d_in()
    open(1, file='invert2d.in', status='old', action='read')
    read(1, '(A)') smesh_file
    read(1, '(A)') refl_file
    read(1, '(A)') sed_file
    read(1, '(A)') geom_file
    read(1, *) niter_inv
    read(1, *) lam_damp_in
    read(1, *) lam_smooth_in
    read(1, *) inv_crit_chi_in
    read(1, *) lsqr_max_iter_in
    read(1, *) target_pert_in
    read(1, *) lam_damp_iface_in
    read(1, *) lam_smooth_iface_in
    read(1, *) refl_weight_in
    read(1, *) phase_wt_pmp_in
    read(1, *) phase_wt_pgss_in
    read(1, *) phase_wt_psed_in
    read(1, *) corr_x_in
    read(1, *) corr_z_in
    smooth_type = 2
    target_dd_in = 0.0_dp
    iface_step_ratio_in = 0.0_dp
    max_rel_pert_in = 0.0_dp
    alt_inv_mode_in = -1
    corr_x_iface_in = corr_x_in
    iface_dx_target_in = 0.0_dp
    mesh_dx_in = 0.0_dp
    mesh_dx_iface_in = 0.0_dp
    read(1, *, end=10) smooth_type, target_dd_in, iface_step_ratio_in, max_rel_pert_in
    read(1, *, end=10) alt_inv_mode_in
    read(1, *, end=10) corr_x_iface_in
    read(1, *, end=10) iface_dx_target_in
    read(1, *, end=10) mesh_dx_in, mesh_dx_iface_in
10  continue
    close(1)
    if (iface_step_ratio_in <= 0.0_dp) iface_step_ratio_in = 0.5_dp
    if (max_rel_pert_in <= 0.0_dp) max_rel_pert_in = 0.30_dp
    if (alt_inv_mode_in < 0 .or. alt_inv_mode_in > 2) alt_inv_mode_in = -1
    if (corr_x_iface_in <= 0.0_dp) corr_x_iface_in = corr_x_in
    write(*,'(A,A)') '  smesh_vp file:    ', trim(smesh_file)
    write(*,'(A,A)') '  moho file:        ', trim(refl_file)
    write(*,'(A,A)') '  sed_bottom file:  ', trim(sed_file)
    write(*,'(A,A)') '  geom file:        ', trim(geom_file)
    write(*,'(A,I4)') '  niter:            ', niter_inv
    write(*,'(A,I4)') '  smooth_type:      ', smooth_type
    write(*,'(A,I4)') '  alt_inv_mode:     ', alt_inv_mode_in
  end subroutine read_invert2d_in

  ! frechet.in (frechet2d.f90:14-128): a flag, the velocity grids to invert
  ! (this model has one, grid 1 of P velocity), then the interfaces.
  subroutine read_frechet_in()
    integer :: i, ngrids, input_list(1000), type_list(100)
    do_inversion = .false.
    inv_vel = .false.
    n_inv_iface = 0
    open(1, file='frechet.in', status='old', action='read')
    read(1, *, end=100) i
    if (i == 0) then
      close(1)
      return
    end if
    do_inversion = .true.
    read(1, *, end=100) ngrids
    if (ngrids > 0) then
      read(1, *, end=100) input_list(1:ngrids)
      read(1, *, end=100) type_list(1:ngrids)
      if (ngrids /= 1 .or. input_list(1) /= 1 .or. type_list(1) /= 1) then
        write(*,'(A)') 'Error: frechet.in: the model has one velocity grid (1) of P velocity (type 1)'
        stop 1
      end if
      inv_vel = .true.
    end if
    read(1, *, end=100) n_inv_iface
    if (n_inv_iface > 3) then
      write(*,'(A)') 'Error: frechet.in: at most three interfaces (2 seafloor, 3 sediment base, 4 Moho)'
      stop 1
    end if
    if (n_inv_iface > 0) read(1, *, end=100) iface_list(1:n_inv_iface)
    close(1)
    write(*,'(A,I4)') '  Velocity grids to invert: ', merge(1, 0, inv_vel)
    write(*,'(A,I4)') '  Interfaces to invert:     ', n_inv_iface
    return
100 continue
    write(*,'(A)') 'Error reading frechet.in'
    close(1)
    stop
  end subroutine read_frechet_in

  ! Receivers above the seafloor of the set-up are in the water
  ! (invert2d_main.f90:517-531); one ray per receiver line of phase 1..5.
  subroutine setup_rays()
    integer :: n, r
    allocate(in_water(nrec))
    nray = 0
    do n = 1, nrec
      in_water(n) = rec_z(n) < iface_z(xpos, topo, rec_x(n))
      if (rec_phase(n) >= 1 .and. rec_phase(n) <= 5) nray = nray + 1
    end do
    allocate(ray_rec(nray), ray_src(nray), ray_path(nray), valid(nray), fresh(nray))
    allocate(ray_x(nray), ray_z(nray), t_obs(nray), t_calc(nray))
    r = 0
    do n = 1, nrec
      if (rec_phase(n) < 1 .or. rec_phase(n) > 5) cycle
      r = r + 1
      ray_rec(r) = n
      ray_src(r) = obs_id(n) - 1
      ray_path(r) = rec_phase(n)
      ray_x(r) = rec_x(n)
      ray_z(r) = rec_z(n)
      t_obs(r) = rec_t(n)
    end do
    valid = 1
    write(*,'(A,I0,A,I0)') '  Sources: ', nobs, '  rays: ', nray
  end subroutine setup_rays

  ! run_forward2d (inversion2d.f90:4145-4545): all fields of every source,
  ! then the rays.  A ray that fails stays invalid until the caller resets
  ! the flags, as ray%valid is only set at the start of an iteration.
  subroutine run_forward()
    call fmt2d_check(fmt2d_forward(ctx, int(nobs, c_int), obs_x, 0_c_int), 'fmt2d_forward')
    call fmt2d_check(fmt2d_trace(ctx, int(nray, c_int), ray_src, ray_path, ray_x, ray_z, t_calc, fresh), &
         'fmt2d_trace')
    where (fresh == 0) valid = 0
  end subroutine run_forward

  ! Depths of interface k (2..4) as the library holds them.
  subroutine model_depths(k, z)
    integer, intent(in) :: k
    real(dp), allocatable, intent(out) :: z(:)
    integer(c_int) :: n
    real(dp) :: dummy(1)
    call fmt2d_check(fmt2d_model_get(ctx, int(k, c_int), dummy, 0_c_int, n), 'fmt2d_model_get')
    allocate(z(n))
    call fmt2d_check(fmt2d_model_get(ctx, int(k, c_int), z, n, n), 'fmt2d_model_get')
  end subroutine model_depths

  ! The velocity grid, v(ix, iz) on the propagation grid's nodes.
  subroutine model_velocity(v)
    real(dp), allocatable, intent(out) :: v(:,:)
    integer(c_int) :: n
    allocate(v(ginfo%nx, ginfo%nz))
    call fmt2d_check(fmt2d_model_get(ctx, 0_c_int, v, int(size(v), c_int), n), 'fmt2d_model_get')
  end subroutine model_velocity

  ! Node positions of interface k.
  subroutine iface_nodes_x(k, x)
    integer, intent(in) :: k
    real(dp), allocatable, intent(out) :: x(:)
    select case (k)
    case (3)
      allocate(x(nsed))
      x = sed_x
    case (4)
      allocate(x(nmoho))
      x = moho_x
    case default
      allocate(x(nx))
      x = xpos
    end select
  end subroutine iface_nodes_x

  ! write_residuals2d (inversion2d.f90:3342-3377).  Receivers in the water
  ! are listed on the current seafloor, where the forward leaves them.
  subroutine write_residuals(filename)
    character(len=*), intent(in) :: filename
    real(dp), allocatable :: sf(:)
    real(dp) :: rz
    integer :: r, n, nd
    call model_depths(2, sf)
    open(30, file=trim(filename), status='replace', action='write')
    write(30,'(A)') '# src_id  rcv_x  rcv_z  phase_id  obs_time  calc_time  residual'
    nd = 0
    do r = 1, nray
      if (valid(r) == 0) cycle
      if (t_obs(r) < 0.0_dp) cycle
      n = ray_rec(r)
      rz = rec_z(n)
      if (in_water(n)) rz = iface_z(xpos, sf, rec_x(n))
      write(30,'(I6,F12.3,F10.3,I4,F12.6,F12.6,ES14.6)') obs_id(n), rec_x(n), rz, ray_path(r), t_obs(r), &
           t_calc(r), t_obs(r) - t_calc(r)
      nd = nd + 1
    end do
    close(30)
    write(*,'(A,A,A,I6)') '  Written residuals to ', trim(filename), ' ndata=', nd
  end subroutine write_residuals

  ! write_ray_density2d (inversion2d.f90:3380-3458): segment midpoints of
  ! every valid picked ray spread bilinearly onto the velocity grid.
  subroutine write_ray_density(filename)
    character(len=*), intent(in) :: filename
    real(dp), allocatable :: cnt(:,:), vz(:), px_(:), pz_(:)
    real(dp) :: dxv, dzv, px, pz, wx, wz, exz(2), wt
    integer(c_int) :: npts(4), flags(3), np
    integer :: r, isec, ip, ivx, ivz, ix, iz, nxv, nzv
    nxv = nx
    nzv = ginfo%nz
    dxv = xpos(2) - xpos(1)
    dzv = ginfo%dz
    allocate(cnt(nxv, nzv), vz(nzv))
    do iz = 1, nzv
      vz(iz) = (iz - 1) * dzv
    end do
    cnt = 0.0_dp
    do r = 1, nray
      if (valid(r) == 0) cycle
      if (t_obs(r) < 0.0_dp) cycle
      call fmt2d_check(fmt2d_ray_info(ctx, int(r - 1, c_int), exz, wt, npts, flags), 'fmt2d_ray_info')
      do isec = 1, nsec_of_path(ray_path(r))
        if (npts(isec) < 2) cycle
        allocate(px_(npts(isec)), pz_(npts(isec)))
        call fmt2d_check(fmt2d_ray_points(ctx, int(r - 1, c_int), int(isec, c_int), px_, pz_, npts(isec), np), &
             'fmt2d_ray_points')
        do ip = 2, np
          px = 0.5_dp * (px_(ip) + px_(ip - 1))
          pz = 0.5_dp * (pz_(ip) + pz_(ip - 1))
          ivx = floor((px - xpos(1)) / dxv) + 1
          if (ivx < 1) ivx = 1
          if (ivx >= nxv) ivx = nxv - 1
          wx = (px - xpos(ivx)) / dxv
          if (wx < 0.0_dp) wx = 0.0_dp
          if (wx > 1.0_dp) wx = 1.0_dp
          if (pz < vz(1) .or. pz >= vz(nzv)) cycle
          ivz = floor((pz - vz(1)) / dzv) + 1
          if (ivz < 1) ivz = 1
          if (ivz >= nzv) ivz = nzv - 1
          wz = (pz - vz(ivz)) / dzv
          if (wz < 0.0_dp) wz = 0.0_dp
          if (wz > 1.0_dp) wz = 1.0_dp
          cnt(ivx, ivz) = cnt(ivx, ivz) + (1.0_dp - wx) * (1.0_dp - wz)
          cnt(ivx + 1, ivz) = cnt(ivx + 1, ivz) + wx * (1.0_dp - wz)
          cnt(ivx, ivz + 1) = cnt(ivx, ivz + 1) + (1.0_dp - wx) * wz
          cnt(ivx + 1, ivz + 1) = cnt(ivx + 1, ivz + 1) + wx * wz
        end do
        deallocate(px_, pz_)
      end do
    end do
    open(31, file=trim(filename), status='replace', action='write')
    write(31,'(I6,I6)') nxv, nzv
    do ix = 1, nxv
      write(31,'(ES14.6)') xpos(ix)
    end do
    do iz = 1, nzv
      write(31,'(ES14.6)') vz(iz)
    end do
    do ix = 1, nxv
      do iz = 1, nzv
        if (cnt(ix, iz) > 0.0_dp) write(31,'(I6,I6,ES14.6)') ix, iz, cnt(ix, iz)
      end do
    end do
    close(31)
    write(*,'(A,A,A,F10.1)') '  Written ray density to ', trim(filename), ' max_count=', maxval(cnt)
  end subroutine write_ray_density

  ! write_model2d (inversion2d.f90:3842-3962): the velocity grid back on
  ! the .smesh's depths below the seafloor, the inverted interfaces, and
  ! the model of iteration iter on the propagation grid.
  subroutine write_model(iter)
    integer, intent(in) :: iter
    real(dp), allocatable :: v(:,:), vz(:), sf(:), zk(:), xk(:)
    logical, allocatable :: active(:,:)
    real(dp) :: z_abs, wz, v_interp, wx, zsf
    integer :: ix, iz, iz_out, iz_lo, iz_hi, nzv, ivx, k
    character(len=256) :: filename
    nzv = ginfo%nz
    call model_depths(2, sf)
    if (inv_vel) then
      call model_velocity(v)
      allocate(vz(nzv), active(nx, nzv))
      do iz = 1, nzv
        vz(iz) = (iz - 1) * ginfo%dz
      end do
      ! active: below the seafloor of the set-up (invert2d_main.f90:354-363, 459-474)
      do ix = 1, nx
        ivx = ix
        if (ivx >= nx) ivx = nx - 1
        wx = dble(ix - ivx)
        if (wx < 0.0_dp) wx = 0.0_dp
        if (wx > 1.0_dp) wx = 1.0_dp
        zsf = (1.0_dp - wx) * topo(ivx) + wx * topo(ivx + 1)
        do iz = 1, nzv
          active(ix, iz) = vz(iz) - zsf >= 0.0_dp
        end do
      end do
      write(filename,'(A,I0,A,I0,A)') 'vgrid_', 1, '_vtype', 1, '.smesh'
      open(20, file=trim(filename), status='replace', action='write')
      write(20,'(I6,I6,F10.3,F10.3)') nx, nz, vwater, vair
      write(20,'(10F12.4)') (xpos(ix), ix=1,nx)
      write(20,'(10F12.4)') (sf(ix), ix=1,nx)
      write(20,'(10F12.4)') (zrel(iz), iz=1,nz)
      do ix = 1, nx
        do iz = 1, nz
          z_abs = sf(ix) + zrel(iz)
          iz_lo = 1
          iz_hi = nzv
          do iz_out = 1, nzv - 1
            if (vz(iz_out) <= z_abs .and. vz(iz_out + 1) >= z_abs) then
              iz_lo = iz_out
              iz_hi = iz_out + 1
              exit
            end if
          end do
          if (z_abs <= vz(1)) then
            v_interp = v(ix, 1)
          else if (z_abs >= vz(nzv)) then
            v_interp = v(ix, nzv)
          else
            wz = (z_abs - vz(iz_lo)) / (vz(iz_hi) - vz(iz_lo))
            if (active(ix, iz_lo) .and. v(ix, iz_lo) >= 2.0_dp) then
              v_interp = (1.0_dp - wz) * v(ix, iz_lo) + wz * v(ix, iz_hi)
            else
              v_interp = v(ix, iz_hi)
            end if
          end if
          write(20,'(F10.4)', advance='no') v_interp
        end do
        write(20,*)
      end do
      close(20)
    end if
    do k = 1, n_inv_iface
      call iface_nodes_x(iface_list(k), xk)
      call model_depths(iface_list(k), zk)
      write(filename,'(A,I0,A)') 'iface_', iface_list(k), '.refl'
      open(20, file=trim(filename), status='replace', action='write')
      do ix = 1, size(zk)
        write(20,'(F12.3,F12.3)') xk(ix), zk(ix)
      end do
      close(20)
    end do
    write(filename,'(A,I0,A)') 'model_iter', iter, '.dat'
    open(20, file=trim(filename), status='replace', action='write')
    write(20,'(A,I0)') 'Iteration ', iter
    if (inv_vel) then
      write(20,'(A,I0,A,I0)') 'vgrid ', 1, ' vtype ', 1
      do ix = 1, nx
        write(20,'(10F10.4)') (v(ix, iz), iz=1,nzv)
      end do
    end if
    do k = 1, n_inv_iface
      call model_depths(iface_list(k), zk)
      write(20,'(A,I0)') 'iface ', iface_list(k)
      write(20,'(10F12.4)') (zk(ix), ix=1,size(zk))
    end do
    close(20)
  end subroutine write_model

  ! The outer loop of lsqr_inversion2d (inversion2d.f90:112-999).
  subroutine invert()
    type(fmt2d_system_params) :: sp
    type(fmt2d_lsqr_params) :: lp
    type(fmt2d_damping_params) :: dpar
    type(fmt2d_damping_result) :: dres
    type(fmt2d_line_search_params) :: lsp
    type(fmt2d_line_search_result) :: lres
    real(dp), allocatable :: dm(:), dm_sum(:), best_vel(:,:), cur(:,:), zk(:)
    real(dp), allocatable :: best_if(:,:)
    integer, allocatable :: best_nif(:)
    real(dp) :: prev_rms, best_rms, rms_orig, rms_new, eps_min, res_check
    integer :: iter, best_iter, ls_prev_idx, alt_inv_mode, ndata, n_outliers, r, k
    integer(c_int) :: nd_sys, nnz

    allocate(dm(n_par), dm_sum(n_par))
    dm_sum = 0.0_dp
    allocate(best_if(max(nx, nsed, nmoho), max(n_inv_iface, 1)), best_nif(max(n_inv_iface, 1)))
    prev_rms = huge_time
    best_rms = huge_time
    best_iter = 0
    ls_prev_idx = 1

    lp = fmt2d_lsqr_params(0.0_dp, 0.0_dp, lam_smooth_in, lam_smooth_iface_in, int(smooth_type, c_int), &
         int(min(lsqr_max_iter_in, 2000), c_int), corr_x_in, corr_z_in, corr_x_iface_in)
    dpar = fmt2d_damping_params(lam_damp_in, lam_damp_iface_in, target_pert_in, target_dd_in, max_rel_pert_in)

    open(99, file='invert_log_latest.txt', status='replace', action='write')

    valid = 1
    call run_forward()
    call write_residuals('residuals_initial.dat')
    if (inv_vel) call write_ray_density('ray_density_initial.dat')

    do iter = 1, niter_inv
      ! alternating mode: 0 joint, 1 velocity only, 2 interfaces only; by
      ! default three velocity iterations per interface iteration
      if (n_inv_iface > 0) then
        if (alt_inv_mode_in == -1) then
          if (mod(iter, 4) == 0) then
            alt_inv_mode = 2
          else
            alt_inv_mode = 1
          end if
        else
          alt_inv_mode = alt_inv_mode_in
        end if
      else
        alt_inv_mode = 0
      end if
      write(*,'(A,I4,A,I1)') '  Iteration ', iter, ' alt_inv_mode=', alt_inv_mode

      valid = 1
      call run_forward()

      ndata = 0
      n_outliers = 0
      do r = 1, nray
        if (valid(r) /= 0 .and. t_obs(r) >= 0.0_dp) then
          res_check = t_obs(r) - t_calc(r)
          if (abs(res_check) > inv_crit_chi_in) then
            n_outliers = n_outliers + 1
            cycle
          end if
          ndata = ndata + 1
        end if
      end do
      write(*,'(A,I6,A,I4,A,F6.2)') '  ndata=', ndata, ' n_outliers=', n_outliers, ' crit_chi=', inv_crit_chi_in
      if (ndata == 0) then
        write(*,'(A,I4,A)') 'Iteration ', iter, ' no valid data'
        exit
      end if

      sp = fmt2d_system_params(inv_crit_chi_in, phase_wt_psed_in, phase_wt_pmp_in, phase_wt_pgss_in, &
           refl_weight_in, int(alt_inv_mode, c_int))
      call fmt2d_check(fmt2d_build_system(ctx, t_obs, sp, nd_sys, nnz, eps_min), 'fmt2d_build_system')
      if (nd_sys /= ndata) then
        write(*,'(A,I0,A,I0)') 'Error: system rows ', nd_sys, ' /= selected data ', ndata
        stop 2
      end if
      call fmt2d_check(fmt2d_damped_update(ctx, lp, dpar, dm_sum, dm, dres), 'fmt2d_damped_update')
      write(*,'(A,ES10.3,A,ES10.3,A,F8.3,A,A,F8.3,A)') '    Final: wdv=', dres%wd, ' wdd=', dres%wd_iface, &
           ' avg_dmv=', dres%avg_dmv * 100.0_dp, '%', ' avg_dmd=', dres%avg_dmd, 'km'
      lsp = fmt2d_line_search_params(inv_crit_chi_in, iface_step_ratio_in, int(ls_prev_idx, c_int))
      call fmt2d_check(fmt2d_line_search(ctx, lsp, t_obs, dm, valid, lres), 'fmt2d_line_search')
      ls_prev_idx = lres%next_first_step
      dm_sum = dm_sum + dm
      rms_orig = lres%rms_orig
      rms_new = lres%rms_new
      write(*,'(A,F6.3,A,I0,A)') '    step=', lres%step, ' (', lres%forwards, ' trial forwards)'

      write(*,'(A,I4,A,I6,A,F14.8,A,F14.8,A,ES10.3,A,ES10.3,A,ES10.3)') 'Iteration ', iter, &
           ' ndata=', ndata, ' RMS=', rms_orig, ' RMS_new=', rms_new, &
           ' avg_dmv=', dres%avg_dmv, ' wdv=', dres%wd, ' wdd=', dres%wd_iface
      write(99,'(A,I4,A,I6,A,F14.8,A,F14.8,A,ES10.3,A,ES10.3,A,ES10.3)') 'Iteration ', iter, &
           ' ndata=', ndata, ' RMS=', rms_orig, ' RMS_new=', rms_new, &
           ' avg_dmv=', dres%avg_dmv, ' wdv=', dres%wd, ' wdd=', dres%wd_iface
      flush(99)

      if (rms_new < best_rms) then
        best_rms = rms_new
        best_iter = iter
        if (inv_vel) then
          call model_velocity(cur)
          if (.not. allocated(best_vel)) allocate(best_vel(size(cur, 1), size(cur, 2)))
          best_vel = cur
        end if
        do k = 1, n_inv_iface
          call model_depths(iface_list(k), zk)
          best_nif(k) = size(zk)
          best_if(1:size(zk), k) = zk
        end do
        write(*,'(A,I4,A,F14.8)') '    ** New best model at iter ', iter, ' RMS=', rms_new
      end if

      if (rms_orig < 1.0e-4_dp) exit
      if (iter > 1 .and. prev_rms > 0.0_dp) then
        if (prev_rms > rms_orig .and. (prev_rms - rms_orig) / prev_rms < 1.0e-5_dp .and. iter > 5) exit
      end if
      prev_rms = rms_orig
    end do

    if (best_iter > 0 .and. best_iter < iter) then
      write(*,'(A,I4,A,F14.8,A,I4)') '  Restoring best model from iter ', best_iter, ' RMS=', best_rms, &
           ' (final iter=', iter, ')'
      if (inv_vel) call fmt2d_check(fmt2d_model_put(ctx, 0_c_int, best_vel, int(size(best_vel), c_int)), &
           'fmt2d_model_put')
      do k = 1, n_inv_iface
        call fmt2d_check(fmt2d_model_put(ctx, iface_list(k), best_if(1:best_nif(k), k), &
             int(best_nif(k), c_int)), 'fmt2d_model_put')
      end do
    end if
    close(99)

    call write_model(iter)
    call run_forward()
    call write_residuals('residuals_final.dat')
    if (inv_vel) call write_ray_density('ray_density_final.dat')
  end subroutine invert

end program invert2d_hip
