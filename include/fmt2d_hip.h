/*
 * fmt2d_hip.h -- C ABI of the MI355X implementation of FMT2D's forward hot
 * path.  Plain C types only; every call returns 0 on success or a negative
 * FMT2D_E* code, with a message from fmt2d_last_error().  One context per
 * host thread; calls on one context are serialised on its HIP stream; host
 * arrays are caller-owned and only touched during the call.
 *
 * Reference interfaces replaced (FMT2D, src/...):
 *   fmt2d_set_model   model build of fmt2d_main.f90:91-418 / invert2d_main.f90:139-480
 *                     (find_intersection2d, define_region2d, tag_regular_nodes2d,
 *                     velocities_to_grid2d/intersection2d/region2d,
 *                     propagate2d.f90:1335-1979)
 *   fmt2d_forward     the per-source stage loop of run_forward2d
 *                     (inversion2d.f90:4184-4416): initialize_region_from_source2d
 *                     (propagate2d.f90:2307), transfer_times_to_intersection (:2414),
 *                     reflect_gradient_at_interface (:2632),
 *                     sweep_region_from_interface2d (:1982), propagate2d (:33),
 *                     the water top-row fix and the PmP head-wave post-pass
 *   fmt2d_get_field   read-back of one time field (Ttime_field2d arrivaltime /
 *                     time_gradient, mod_fmt2d.f90:262-279) in the reference's
 *                     region node order
 *   fmt2d_set_inversion / fmt2d_inversion_info
 *                     the parameterisation read from frechet.in
 *                     (initialize_inversion2d, frechet2d.f90:14-118) over the
 *                     resampled velocity grid (invert2d_main.f90:346-475)
 *   fmt2d_trace       the ray loop of run_forward2d (inversion2d.f90:4491-4537):
 *                     find_optimal_entry_inv (:6344-6438) for receivers in the
 *                     water, trace_ray_from_receiver2d / trace_ray_section2d
 *                     (rays2d.f90:7-762), receiver time + water leg
 *   fmt2d_ray_info / fmt2d_ray_points
 *                     Tray2d / Tray_section2d point lists (mod_fmt2d.f90)
 *   fmt2d_frechet_csr compute_frechet_for_ray2d (inversion2d.f90:4548-4787) for
 *                     every valid ray, packed as CSR rows (read_frechet_data2d,
 *                     :3965-4052)
 *   fmt2d_build_system / fmt2d_system_get
 *                     data selection and weighting of lsqr_inversion2d
 *                     (inversion2d.f90:280-486): outlier cut, phase weights,
 *                     mvscale, G*mvscale*w, gtg_diag, eps_min
 *   fmt2d_lsqr        run_lsqr (inversion2d.f90:6578-7167) with the Laplacian
 *                     smoothing block (apply_laplacian_smooth_2d, :1161-1249)
 */
#ifndef FMT2D_HIP_H
#define FMT2D_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FMT2D_OK 0
#define FMT2D_EINVAL (-1)
#define FMT2D_EHIP (-2)
#define FMT2D_ESTATE (-3)
#define FMT2D_ENOCONV (-4)

/* number of time fields per source: tf1..tf12 of the reference */
#define FMT2D_NTF 12

typedef struct fmt2d_ctx fmt2d_ctx;

typedef struct {
  int nx, nz;               /* propagation grid */
  double dx, dz, x0, z0;
  int region_nnode[7];      /* [1..6]: node count of each region (reference numbering) */
} fmt2d_grid_info;

typedef struct {
  int nsrc, nfield;         /* sources and resident fields of the last forward */
  int launches;             /* tile-solver launches over all stage levels */
  long long tiles;          /* tiles processed over all launches */
  long long region_nodes;   /* sum of region node counts over the fields */
  double ms_total;          /* wall time of the forward */
  double ms_fim;            /* summed device time of the tile-solver launches (HIP events) */
  long long exit_nodes;     /* nodes checked exactly by the heap-exhaustion stop */
  long long sweeps;         /* in-tile sweeps over all tile visits */
  long long capped;         /* tile visits that reached the sweep cap */
  int max_sweeps;           /* most sweeps of one tile visit */
} fmt2d_forward_stats;

/* Context on HIP device `device` (owns all device memory). */
int fmt2d_ctx_create(int device, fmt2d_ctx** out);
int fmt2d_ctx_destroy(fmt2d_ctx* ctx);
const char* fmt2d_last_error(void);

/* Model from the .smesh content (vel is nx_s*nz_s with iz fastest, as in the
 * file), the sediment-base and the Moho interface nodes. */
int fmt2d_set_model(fmt2d_ctx* ctx, int nx_s, int nz_s, const double* xpos, const double* topo,
                    const double* zrel, const double* vel, double vwater, int nsed, const double* sed_x,
                    const double* sed_z, int nmoho, const double* moho_x, const double* moho_z);
int fmt2d_grid_info_get(fmt2d_ctx* ctx, fmt2d_grid_info* info);

/* All twelve time fields of every source (sources placed on the seafloor at
 * src_x), solved as four dependency levels of batched stages.  Fields stay
 * resident on the device.  flags: bit0 = skip the PmP head-wave post-pass. */
int fmt2d_forward(fmt2d_ctx* ctx, int nsrc, const double* src_x, int flags);
int fmt2d_forward_stats_get(fmt2d_ctx* ctx, fmt2d_forward_stats* st);
/* Solver tuning: arrival-time window of the tile schedule in tile crossing
 * times (at 4 km/s), the sweep cap of one tile visit, and the in-tile sweep
 * order (4: four-colour Gauss-Seidel, 2: two-colour Gauss-Seidel, 1: Jacobi).
 * Defaults 3.0, 64, 2. */
int fmt2d_set_tuning(fmt2d_ctx* ctx, double window_tiles, int max_sweeps, int colors);
/* Due-tile count and the most sweeps of one tile visit, per solver launch of
 * the last forward (diagnostics); either array may be NULL. */
int fmt2d_forward_trace(fmt2d_ctx* ctx, int* counts, int* max_sweeps, int max, int* n);
/* Tile visits of the last forward by sweep count: hist[k] visits of k + 1
 * sweeps, hist[63] 64 or more (diagnostics). */
int fmt2d_forward_sweep_hist(fmt2d_ctx* ctx, long long* hist);

/* Region id (1..6) of time field itf (1..12). */
int fmt2d_tf_region(int itf);

/* Field itf (1..12) of source isrc (0-based) in the reference's region node
 * order (grid nodes ix-outer/iz-inner, then top, then bottom interface
 * nodes); arrays hold region_nnode[fmt2d_tf_region(itf)] values. */
int fmt2d_get_field(fmt2d_ctx* ctx, int isrc, int itf, double* T, double* dTdx, double* dTdz);

/* Inversion parameterisation (frechet.in): invert the velocity grid and/or
 * the listed interfaces (2 seafloor, 3 sediment base, 4 Moho).  Default:
 * velocity only. */
int fmt2d_set_inversion(fmt2d_ctx* ctx, int invert_velocity, int n_iface, const int* iface_ids);
/* Number of velocity parameters and of all parameters (n_inv_parms). */
int fmt2d_inversion_info(fmt2d_ctx* ctx, int* n_vel_par, int* n_par);
/* Traces nray rays on the fields of the last fmt2d_forward.  Ray r runs
 * from receiver (rx[r], rz[r]) back to source src[r] (0-based index into the
 * last forward's sources) along path path[r] (1 Pg, 2 PsedP, 3 PmP, 4 Pg_ss,
 * 5 Pg_ss2).  t_rcv: receiver time including the water leg; valid: 1/0. */
int fmt2d_trace(fmt2d_ctx* ctx, int nray, const int* src, const int* path, const double* rx, const double* rz,
                double* t_rcv, int* valid);
/* Ray r of the last trace: entry point (x, z) the ray starts from, water-leg
 * time, points per section (4), flags (headwave, diffracted). */
int fmt2d_ray_info(fmt2d_ctx* ctx, int iray, double* entry_xz, double* water_time, int* npts, int* flags);
/* Points of section `section` (1 = at the source) of ray r; with x == NULL
 * only *npts is returned. */
int fmt2d_ray_points(fmt2d_ctx* ctx, int iray, int section, double* x, double* z, int cap, int* npts);
/* Frechet rows of the last trace as CSR (one row per ray, empty for invalid
 * rays; columns are parameter numbers).  Call with col == NULL for *nnz and
 * row_ptr, then with buffers of cap >= nnz.  index_base 1 gives Fortran
 * indexing for row_ptr and col. */
int fmt2d_frechet_csr(fmt2d_ctx* ctx, int index_base, int* row_ptr, int* col, double* val, int cap, int* nnz);

/* Data selection and scaling (lsqr_inversion2d's system assembly). */
typedef struct {
  double crit_chi;        /* inv_crit_chi: rows with |t_obs - t_calc| > crit_chi are outliers */
  double phase_wt_psed;   /* data weights of PsedP, PmP, Pg_ss/Pg_ss2 rows */
  double phase_wt_pmp;
  double phase_wt_pgss;
  double refl_weight;     /* mvscale of interface parameters */
  int alt_mode;           /* 0 joint, 1 velocity only, 2 interface only */
} fmt2d_system_params;

typedef struct {
  double wdv, wdv_iface;           /* damping weights (velocity, interface) */
  double w_smooth, w_smooth_iface; /* smoothing weights */
  int smooth_type;                 /* 1: Laplacian (the supported form) */
  int max_iter;
} fmt2d_lsqr_params;

typedef struct {
  int iterations;
  double rnorm, bnorm, test2;
  double ms;              /* wall time of the solve */
} fmt2d_lsqr_stats;

/* Builds the weighted, scaled system from the Frechet rows of the last trace
 * and the observed times t_obs[nray] (< 0: no pick).  Returns the number of
 * data rows, non-zeros and eps_min = max(1e-6, 1e-4 max gtg_diag). */
int fmt2d_build_system(fmt2d_ctx* ctx, const double* t_obs, const fmt2d_system_params* p, int* ndata, int* nnz,
                       double* eps_min);
/* gtg_diag (n_par) and the weighted residuals (ndata) of the built system;
 * either pointer may be NULL. */
int fmt2d_system_get(fmt2d_ctx* ctx, double* gtg_diag, double* residual);
/* One LSQR solve of [G; w L; wd I] dm = [r; -w L dm_sum; -wd dm_sum] with the
 * reference's column mask and stopping rule (test2 < 1e-4). */
int fmt2d_lsqr(fmt2d_ctx* ctx, const fmt2d_lsqr_params* p, const double* dm_sum, double* dm_out,
               fmt2d_lsqr_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
