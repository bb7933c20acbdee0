! Fortran binding of the MI355X library (include/fmt2d_hip.h) through
! iso_c_binding: the C structs as bind(C) types and one interface per entry
! point used by the Fortran host programs.  Status codes: 0 ok, < 0 error
! (text from fmt2d_last_error).
module fmt2d_hip
  use iso_c_binding
  implicit none

  type, bind(C) :: fmt2d_grid_info
    integer(c_int) :: nx, nz
    real(c_double) :: dx, dz, x0, z0
    integer(c_int) :: region_nnode(7)
  end type

  type, bind(C) :: fmt2d_system_params
    real(c_double) :: crit_chi, phase_wt_psed, phase_wt_pmp, phase_wt_pgss, refl_weight
    integer(c_int) :: alt_mode
  end type

  type, bind(C) :: fmt2d_lsqr_params
    real(c_double) :: wdv, wdv_iface, w_smooth, w_smooth_iface
    integer(c_int) :: smooth_type, max_iter
    real(c_double) :: corr_x, corr_z, corr_x_iface
  end type

  type, bind(C) :: fmt2d_damping_params
    real(c_double) :: lam_damp, lam_damp_iface, target_dv, target_dd, max_rel_pert
  end type

  type, bind(C) :: fmt2d_damping_result
    real(c_double) :: wd, wd_iface, avg_dmv, avg_dmd
    integer(c_int) :: solves, iterations
  end type

  type, bind(C) :: fmt2d_line_search_params
    real(c_double) :: crit_chi, iface_step_ratio
    integer(c_int) :: first_step
  end type

  type, bind(C) :: fmt2d_line_search_result
    real(c_double) :: step, rms_orig, rms_new
    integer(c_int) :: next_first_step, forwards
  end type

  interface
    integer(c_int) function fmt2d_ctx_create(device, ctx) bind(C, name="fmt2d_ctx_create")
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr) :: ctx
    end function
    integer(c_int) function fmt2d_ctx_destroy(ctx) bind(C, name="fmt2d_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function fmt2d_last_error() bind(C, name="fmt2d_last_error")
      import :: c_ptr
    end function
    integer(c_int) function fmt2d_set_model(ctx, nx_s, nz_s, xpos, topo, zrel, vel, vwater, &
        nsed, sed_x, sed_z, nmoho, moho_x, moho_z) bind(C, name="fmt2d_set_model")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nx_s, nz_s, nsed, nmoho
      real(c_double), intent(in) :: xpos(*), topo(*), zrel(*), vel(*)
      real(c_double), value :: vwater
      real(c_double), intent(in) :: sed_x(*), sed_z(*), moho_x(*), moho_z(*)
    end function
    integer(c_int) function fmt2d_grid_info_get(ctx, info) bind(C, name="fmt2d_grid_info_get")
      import :: c_int, c_ptr, fmt2d_grid_info
      type(c_ptr), value :: ctx
      type(fmt2d_grid_info), intent(out) :: info
    end function
    integer(c_int) function fmt2d_forward(ctx, nsrc, src_x, flags) bind(C, name="fmt2d_forward")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsrc, flags
      real(c_double), intent(in) :: src_x(*)
    end function
    integer(c_int) function fmt2d_set_inversion(ctx, invert_velocity, n_iface, iface_ids) &
        bind(C, name="fmt2d_set_inversion")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: invert_velocity, n_iface
      integer(c_int), intent(in) :: iface_ids(*)
    end function
    integer(c_int) function fmt2d_set_inversion_mesh(ctx, vel_stride, iface_stride) &
        bind(C, name="fmt2d_set_inversion_mesh")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: vel_stride, iface_stride
    end function
    integer(c_int) function fmt2d_inversion_info(ctx, n_vel_par, n_par) bind(C, name="fmt2d_inversion_info")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: n_vel_par, n_par
    end function
    integer(c_int) function fmt2d_trace(ctx, nray, src, path, rx, rz, t_rcv, valid) bind(C, name="fmt2d_trace")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nray
      integer(c_int), intent(in) :: src(*), path(*)
      real(c_double), intent(in) :: rx(*), rz(*)
      real(c_double), intent(out) :: t_rcv(*)
      integer(c_int), intent(out) :: valid(*)
    end function
    integer(c_int) function fmt2d_build_system(ctx, t_obs, p, ndata, nnz, eps_min) bind(C, name="fmt2d_build_system")
      import :: c_int, c_ptr, c_double, fmt2d_system_params
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: t_obs(*)
      type(fmt2d_system_params), intent(in) :: p
      integer(c_int), intent(out) :: ndata, nnz
      real(c_double), intent(out) :: eps_min
    end function
    integer(c_int) function fmt2d_damped_update(ctx, p, d, dm_sum, dm_out, r) bind(C, name="fmt2d_damped_update")
      import :: c_int, c_ptr, c_double, fmt2d_lsqr_params, fmt2d_damping_params, fmt2d_damping_result
      type(c_ptr), value :: ctx
      type(fmt2d_lsqr_params), intent(in) :: p
      type(fmt2d_damping_params), intent(in) :: d
      real(c_double), intent(in) :: dm_sum(*)
      real(c_double), intent(out) :: dm_out(*)
      type(fmt2d_damping_result), intent(out) :: r
    end function
    integer(c_int) function fmt2d_line_search(ctx, p, t_obs, dm, valid, r) bind(C, name="fmt2d_line_search")
      import :: c_int, c_ptr, c_double, fmt2d_line_search_params, fmt2d_line_search_result
      type(c_ptr), value :: ctx
      type(fmt2d_line_search_params), intent(in) :: p
      real(c_double), intent(in) :: t_obs(*)
      real(c_double), intent(inout) :: dm(*)
      integer(c_int), intent(inout) :: valid(*)
      type(fmt2d_line_search_result), intent(out) :: r
    end function
    integer(c_int) function fmt2d_ray_info(ctx, iray, entry_xz, water_time, npts, flags) bind(C, name="fmt2d_ray_info")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: iray
      real(c_double), intent(out) :: entry_xz(2), water_time
      integer(c_int), intent(out) :: npts(4), flags(3)
    end function
    integer(c_int) function fmt2d_ray_points(ctx, iray, section, x, z, cap, npts) bind(C, name="fmt2d_ray_points")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: iray, section, cap
      real(c_double), intent(out) :: x(*), z(*)
      integer(c_int), intent(out) :: npts
    end function
    integer(c_int) function fmt2d_model_get(ctx, what, out, cap, n) bind(C, name="fmt2d_model_get")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: what, cap
      real(c_double), intent(out) :: out(*)
      integer(c_int), intent(out) :: n
    end function
    integer(c_int) function fmt2d_model_put(ctx, what, in, n) bind(C, name="fmt2d_model_put")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: what, n
      real(c_double), intent(in) :: in(*)
    end function
  end interface

contains

  ! Stops with the library's message when a call failed.
  subroutine fmt2d_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(fmt2d_last_error(), msg, [4096])
    n = 0
    do while (n < 4096)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write(*, '(A,A,I0,A,4096A1)') trim(what), ' failed (', rc, '): ', msg(1:n)
    stop 2
  end subroutine fmt2d_check

end module fmt2d_hip
