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

/* fmt2d_forward flags */
#define FMT2D_FWD_NO_HEADWAVE 1  /* skip the PmP head-wave post-pass (inversion2d.f90:4294-4409) */
#define FMT2D_FWD_PROGRAM 2      /* as the forward program fmt2d (fmt2d_main.f90:757-1363): no
                                  * head-wave post-pass, reflection normals from interface_normal2d
                                  * at each node (:1563), and the rays of fmt2d_trace by its rules:
                                  * entry search skipping seafloor points within 3 dx of the source
                                  * (:1803-1905), Pg of a receiver in the water compared with the
                                  * direct water wave (:1249-1280), the water leg reported for
                                  * PsedP and PmP only (:1334-1337) */

typedef struct fmt2d_ctx fmt2d_ctx;

typedef struct {
  int nx, nz;               /* propagation grid */
  double dx, dz, x0, z0;
  int region_nnode[7];      /* [1..6]: node count of each region (reference numbering) */
} fmt2d_grid_info;

typedef struct {
  int nsrc, nfield;         /* sources and resident fields of the last forward */
  int launches;             /* tile-solver launches over all stage levels (0: the
                               fields left by a line search were the answer) */
  long long tiles;          /* tiles processed over all launches */
  long long region_nodes;   /* sum of region node counts over the fields */
  double ms_total;          /* wall time of the forward */
  double ms_fim;            /* summed device time of the tile-solver launches: first wave start to last
                               wave end of each, on the device's constant clock (the launches run in
                               replayed HIP graphs) */
  long long exit_nodes;     /* nodes checked exactly by the heap-exhaustion stop */
  long long sweeps;         /* in-tile sweeps over all tile visits */
  long long capped;         /* tile visits that reached the sweep cap */
  int max_sweeps;           /* most sweeps of one tile visit */
  int dispatches;           /* tile-kernel dispatches: launches plus empty ones that close a batch
                               (FMT2D_LAUNCH_GRAPH), one per stage level (FMT2D_LAUNCH_LEVEL) */
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
 * resident on the device.  flags: FMT2D_FWD_* (0: as the inversion's
 * run_forward2d, inversion2d.f90:4145-4545). */
int fmt2d_forward(fmt2d_ctx* ctx, int nsrc, const double* src_x, int flags);
int fmt2d_forward_stats_get(fmt2d_ctx* ctx, fmt2d_forward_stats* st);
/* Solver tuning: arrival-time window of the tile schedule in tile crossing
 * times (at 4 km/s), the sweep cap of one tile visit, and the in-tile sweep
 * order (4: four-colour Gauss-Seidel, 2: two-colour Gauss-Seidel, 1: Jacobi).
 * Defaults 3.0, 16, 2. */
int fmt2d_set_tuning(fmt2d_ctx* ctx, double window_tiles, int max_sweeps, int colors);
/* Sweep cap of one visit of a tile cut by an interface (its visits are the
 * slow ones; a capped visit continues in a later launch; results are
 * identical for any cap).  Default 8.  (Solver tuning of this build, no
 * counterpart in the reference's heap-ordered march.) */
int fmt2d_set_heavy_sweeps(fmt2d_ctx* ctx, int max_sweeps_heavy);
/* Storage precision of the tile solver's grid-node fields: 64 (default;
 * every result bit-exact against the reference) or 32: T stored in 32-bit
 * fixed point (units of 2^-24 s = 6.0e-8 s, times up to 255 s; one grid of
 * values for every node, so the result does not depend on the order of
 * the updates), raw gradients in FP32, every update still computed in FP64
 * and rounded to the stored values; 8x8 tiles.  Interface nodes, seeds and
 * everything after the solve (post-passes, rays, Frechet rows) stay FP64;
 * the adjoint gradient needs an FP64 forward.  Halves the solver's field
 * traffic; fmt2d_field_snapshot / fmt2d_field_compare measure its effect
 * per field.  No counterpart in the reference. */
int fmt2d_set_field_precision(fmt2d_ctx* ctx, int bits);
/* Keep a copy of T of every field of the last forward (device memory), and
 * compare a later forward of the same sources with it: per field (order
 * isrc * 12 + itf - 1) the largest |dT| over slots both reach, the count of
 * slots only one reaches and the count of slots with |dT| > tol.  With an
 * output NULL or cap 0 only *n.  (Diagnostics of this build.) */
int fmt2d_field_snapshot(fmt2d_ctx* ctx);
int fmt2d_field_compare(fmt2d_ctx* ctx, double tol, double* max_dt, long long* n_reach_diff, long long* n_over,
                        int cap, int* n);
/* Tile shape of the solver: 8x8 (default; one node per lane of a
 * wavefront), 16x8 or 8x16 (two nodes per lane, one of each checkerboard
 * colour).  Results are identical for every shape.  (Solver tuning of this
 * build, no counterpart in the reference.) */
int fmt2d_set_tile(fmt2d_ctx* ctx, int tile_x, int tile_z);
typedef struct {
  int tile_x, tile_z;       /* chosen tile shape */
  int max_sweeps, max_sweeps_heavy, colors;
  double window_tiles;
  double ms_best;           /* forward wall time with the chosen schedule */
  double ms_start;          /* forward wall time with the schedule set on entry */
  int trials;               /* forwards run */
} fmt2d_tune_result;
/* Autotuner of the tile schedule on the model and sources at hand (no
 * counterpart in the reference): times one forward per candidate by
 * coordinate descent over the tile shape (8x8, 16x8, 8x16), the sweep caps
 * of light and heavy tiles and, with level 1, the arrival window and the
 * in-tile order; keeps the fastest in the context.  Every schedule gives
 * the same fields bit for bit, so the context ends solved for these
 * sources.  About 11 forwards at level 0, 16 at level 1. */
int fmt2d_autotune(fmt2d_ctx* ctx, int nsrc, const double* src_x, int flags, int level, fmt2d_tune_result* out);
/* Solver launches allowed per stage level before fmt2d_forward gives up with
 * FMT2D_ENOCONV (0, the default: 64 x (tile columns + tile rows) + 4000).  A
 * forward that gave up leaves the context usable: the next forward starts
 * from clean tile state.  (Solver control of this build, no counterpart in
 * the reference's heap-ordered march.) */
int fmt2d_set_launch_cap(fmt2d_ctx* ctx, int max_launches);
/* How the tile solver's launches reach the device.  FMT2D_LAUNCH_GRAPH (the
 * default): each launch is a split and a tile dispatch; batches of 8
 * launches are captured once per stage level in a HIP graph and replayed
 * until the active set is empty.  FMT2D_LAUNCH_LEVEL: one cooperative
 * dispatch per stage level runs all its launches, as phases separated by
 * grid-wide barriers (FMT2D_EHIP where the device has no cooperative
 * launch).  Both run the same schedule (the same split of the active set
 * at each launch) and reach the same fields bit for bit.  (Solver control
 * of this build, no counterpart in the reference.) */
#define FMT2D_LAUNCH_GRAPH 0
#define FMT2D_LAUNCH_LEVEL 1
int fmt2d_set_solver_launch(fmt2d_ctx* ctx, int mode);
/* Due-tile count and the most sweeps of one tile visit, per solver launch of
 * the last forward (diagnostics); either array may be NULL. */
int fmt2d_forward_trace(fmt2d_ctx* ctx, int* counts, int* max_sweeps, int max, int* n);
/* Tile visits of the last forward by sweep count: hist[k] visits of k + 1
 * sweeps, hist[63] 64 or more (diagnostics). */
int fmt2d_forward_sweep_hist(fmt2d_ctx* ctx, long long* hist);

typedef struct {
  long long reached;            /* slots of the field with a time (< 1e19) */
  unsigned long long hash;      /* order-free hash of the bits of T, dT/dx, dT/dz over all slots */
  double max_resid, mean_resid; /* | |grad T| v - 1 | over fully regular reached grid nodes
                                   whose four neighbours are reached */
  long long n_resid;            /* nodes in the residual */
  long long n_resid_big;        /* of them, residuals above 0.05 */
} fmt2d_field_summary;
/* Summary of every field of the last forward, computed on the device, in
 * the order isrc * 12 + itf - 1 (diagnostics of this build: checks over all
 * fields of a large batch without reading them back; no counterpart in the
 * reference).  With out == NULL or cap == 0 only *n. */
int fmt2d_forward_summary(fmt2d_ctx* ctx, fmt2d_field_summary* out, int cap, int* n);

/* Region id (1..6) of time field itf (1..12). */
int fmt2d_tf_region(int itf);

/* Field itf (1..12) of source isrc (0-based) in the reference's region node
 * order (grid nodes ix-outer/iz-inner, then top, then bottom interface
 * nodes); arrays hold region_nnode[fmt2d_tf_region(itf)] values. */
int fmt2d_get_field(fmt2d_ctx* ctx, int isrc, int itf, double* T, double* dTdx, double* dTdz);

/* Inversion parameterisation (frechet.in): invert the velocity grid and/or
 * the listed interfaces (2 seafloor, 3 sediment base, 4 Moho).  Default:
 * velocity only.  Resets the inversion mesh to the grid. */
int fmt2d_set_inversion(fmt2d_ctx* ctx, int invert_velocity, int n_iface, const int* iface_ids);
/* Device memory bound of the Frechet-row and G P assembly: at most
 * max_summands (ray, node) contributions are sorted at once (about 40 bytes
 * each; default 2^28); rows are assembled in chunks of whole rays, so the
 * result does not depend on it.  The assembly's scratch (up to the chunk's
 * ~40 bytes per summand) stays allocated in the context between assemblies.
 * The rows of the last trace are assembled again on next use.  (No counterpart in the reference, which accumulates
 * each ray into a dense row, inversion2d.f90:4573-4577.) */
int fmt2d_set_frechet_chunk(fmt2d_ctx* ctx, long long max_summands);
/* Inversion mesh coarser than the propagation grid (no counterpart in the
 * reference, whose inversion grid is the propagation grid,
 * fmt2d_main.f90:200-205): velocity parameters at every vel_stride-th grid
 * column and row, interface depth parameters at every iface_stride-th node
 * of each inverted interface (the reference reads an interface spacing,
 * iface_dx_target, invert2d_main.f90:89-93, and leaves it unused).  Mesh
 * nodes below the set-up seafloor are parameters, numbered as the grid's.
 * The prolongation P (bilinear over a velocity mesh cell, linear along an
 * interface; nodes that are not parameters contribute zero) maps mesh
 * parameters to the Frechet columns: fmt2d_build_system, the LSQR solves,
 * fmt2d_damped_update and fmt2d_line_search work on G P, and a model step
 * moves the grid by P dm.  Strides 1, 1 (the default) are the reference's
 * parameterisation, with G P = G bit for bit.  The strides must divide the
 * grid's cell counts and each inverted interface's node spacings. */
int fmt2d_set_inversion_mesh(fmt2d_ctx* ctx, int vel_stride, int iface_stride);
/* Number of velocity parameters and of all parameters (n_inv_parms) of the
 * inversion mesh: the length of the model updates (dm) of
 * fmt2d_damped_update, fmt2d_lsqr and fmt2d_line_search.  Equal to the
 * Frechet column counts when the mesh is the grid. */
int fmt2d_inversion_info(fmt2d_ctx* ctx, int* n_vel_par, int* n_par);
/* Velocity nodes of the inversion mesh (vnx x vnz) and the number of
 * Frechet columns (grid velocity parameters + interface input nodes). */
int fmt2d_mesh_info(fmt2d_ctx* ctx, int* vnx, int* vnz, int* n_frechet);
/* Traces nray rays on the fields of the last fmt2d_forward.  Ray r runs
 * from receiver (rx[r], rz[r]) back to source src[r] (0-based index into the
 * last forward's sources) along path path[r] (1 Pg, 2 PsedP, 3 PmP, 4 Pg_ss,
 * 5 Pg_ss2).  t_rcv: receiver time including the water leg; valid: 1/0. */
int fmt2d_trace(fmt2d_ctx* ctx, int nray, const int* src, const int* path, const double* rx, const double* rz,
                double* t_rcv, int* valid);
/* Nearest-node search of the ray interpolators' fallbacks
 * (interpolate_arrivaltime_reg, find_position_in_region2d,
 * rays2d.f90:960-976, 1076-1131): 0 (default) a column index of each
 * region, visited outward from the point; 1 the reference's scan over the
 * region's whole node list.  Both return the same node (first minimum in
 * node order).  fmt2d_ray_nearest_calls: searches made by the last trace.
 * (Diagnostics of this build.) */
int fmt2d_set_ray_search(fmt2d_ctx* ctx, int mode);
int fmt2d_ray_nearest_calls(fmt2d_ctx* ctx, long long* n);
/* The nearest node (field slot) of region `region` (1..6) for n points by
 * both searches: idx (column index; -2 when the region has none) and scan
 * (diagnostics: they must agree). */
int fmt2d_ray_nearest_check(fmt2d_ctx* ctx, int region, int n, const double* px, const double* pz, int* idx,
                            int* scan);
/* Ray r of the last trace: entry point (x, z) the ray starts from, water-leg
 * time, points per section (4), flags[3] (headwave, diffracted, direct water
 * wave: a Pg of FMT2D_FWD_PROGRAM taken as the straight water path, no
 * sections). */
int fmt2d_ray_info(fmt2d_ctx* ctx, int iray, double* entry_xz, double* water_time, int* npts, int* flags);
/* Points of section `section` (1 = at the source) of ray r; with x == NULL
 * only *npts is returned. */
int fmt2d_ray_points(fmt2d_ctx* ctx, int iray, int section, double* x, double* z, int cap, int* npts);
/* Frechet rows of the last trace as CSR (one row per ray, empty for invalid
 * rays; columns are the grid's parameter numbers, fmt2d_param_index).  Call with col == NULL for *nnz and
 * row_ptr, then with buffers of cap >= nnz.  index_base 1 gives Fortran
 * indexing for row_ptr and col. */
int fmt2d_frechet_csr(fmt2d_ctx* ctx, int index_base, int* row_ptr, int* col, double* val, int cap, int* nnz);
/* y = G x (transpose 0: x has n_frechet entries, y nray) or y = G^T x
 * (transpose 1: x nray, y n_frechet) with the Frechet rows of the last trace as
 * fmt2d_frechet_csr returns them (unscaled: d t / d slowness of a velocity
 * node, d t / d depth of an interface node); each sum in the order
 * sparse_Ax / sparse_ATx accumulate it (inversion2d.f90:2997-3042).  The
 * adjoint of the receiver times, e.g. for a torch.autograd backward. */
int fmt2d_frechet_apply(fmt2d_ctx* ctx, int transpose, const double* x, double* y);
/* fmt2d_frechet_apply without any device allocation (no counterpart in the
 * reference; for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops).  x and y are staged in a fixed 16 MiB
 * device array of the library (one per device, held until the process
 * ends): x and y together must fit in 2^21 doubles (FMT2D_EINVAL beyond).
 * The rows (and, for transpose 1, their column copy) must already be built
 * by one fmt2d_frechet_apply call on the same trace (FMT2D_ESTATE
 * otherwise).  Results equal fmt2d_frechet_apply's bit for bit.  Calls that
 * overlap on one device share that array and race. */
int fmt2d_frechet_apply_noalloc(fmt2d_ctx* ctx, int transpose, const double* x, double* y);
typedef struct {
  int fields;                 /* stage fields swept (those some receiver time depends on) */
  int rounds;                 /* rounds of the reverse sweep over all stage levels */
  long long leaves_unmatched; /* reached nodes whose recomputed winning update, or head-wave chain, did
                                 not reproduce the stored time (0 expected): kept as leaves */
  long long dropped_partials; /* partial derivatives dropped (a value held 8 already; 0 expected) */
  double ms_record, ms_reverse, ms_total;
  long long peak_bytes;       /* device memory of the sweep (adjoint fields, edges, counters) */
} fmt2d_adjoint_stats;
/* Adjoint-state gradient of a misfit of the receiver times of the last
 * trace (no counterpart in the reference, which inverts with the ray-based
 * rows of compute_frechet_for_ray2d, inversion2d.f90:4548-4787): dJdt[r] is
 * d J / d t_r for every ray (0 or NaN: none).  Every reached node of the
 * solved fields is one local update (regular_update2d, time_from_1_node /
 * time_from_2_nodes, propagate2d.f90:605-946, or its seed,
 * :1982-2411); the adjoint of the receiver times (interpolation weights at
 * the entry point) is swept back over those updates in reverse dependency
 * order, stage by stage through the interface re-seeding and the PmP
 * head-wave pass, onto the region velocities and interface-node depths.
 * The updates' partial derivatives are recomputed from the resident fields
 * one stage level at a time (the forward stores no adjoint data; the sweep
 * holds one level's edges).  grad_vel (nz x nx, x fastest, the layout of
 * fmt2d_model_get 0): d J / d velocity of each propagation-grid node (the
 * region and interface velocities are linear in them,
 * velocities_to_region2d / velocities_to_intersection2d,
 * propagate2d.f90:1698-1929).  grad_iface (5 x nx, interface 1..5 in turn,
 * may be NULL): d J / d depth of each interface node through the node
 * positions the updates and seeds use and the interface velocities sampled
 * at them, grid velocities and the cut-cell topology held fixed.  At nodes
 * where two updates tie (coincident grid and interface nodes on the sea
 * surface) the times are not differentiable; the gradient follows the
 * update the forward kept.  Needs a forward without FMT2D_FWD_PROGRAM. */
int fmt2d_adjoint_gradient(fmt2d_ctx* ctx, const double* dJdt, double* grad_vel, double* grad_iface,
                           fmt2d_adjoint_stats* stats);
/* Device memory the adjoint sweep may give to one batch of edge records
 * (about 120 bytes per field slot; the fields of a stage level are swept in
 * batches that fit).  0: the default, 4 GiB.  No counterpart in the
 * reference. */
int fmt2d_set_adjoint_memory(fmt2d_ctx* ctx, long long bytes);
/* Frechet column (1-based; 0: none) of each velocity-grid node (nz x nx, x
 * fastest) as initialize_inversion2d numbers the parameters
 * (frechet2d.f90:44-60); the interfaces' input nodes follow.  These are the
 * inversion parameters when the inversion mesh is the grid (the default;
 * see fmt2d_set_inversion_mesh).  With node_param == NULL or cap == 0 only
 * *n. */
int fmt2d_param_index(fmt2d_ctx* ctx, int* node_param, int cap, int* n);

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
  int smooth_type;                 /* 1 Laplacian, 2 Gaussian (R_h, R_v, interface R_if), 3 Laplacian + R_if */
  int max_iter;
  double corr_x, corr_z;           /* Gaussian correlation lengths (km), smooth_type 2 */
  double corr_x_iface;             /* interface correlation length (km), smooth_type 2 and 3; <= 0: none */
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
/* Reductions of the LSQR solves (fmt2d_lsqr, fmt2d_damped_update, the
 * solves of the damping search): 0 (default) every norm and sparse product in the
 * reference's order, so the model update and iteration count are the
 * reference's bit for bit; 1 double-double tree-reduced norms and
 * wave-reduced products, about 2x faster per iteration, with rounding of their
 * own that the Krylov recurrences amplify: on the test systems the model
 * update moves by up to 7e-3 relative and the iteration count by up to 2.
 * No counterpart in the reference. */
int fmt2d_set_lsqr_mode(fmt2d_ctx* ctx, int mode);
/* *out = x[0]^2 + x[1]^2 + ... added front to back in FP64 (the ordered
 * mode's norm kernel on a host vector of n entries), bit for bit the plain
 * loop's sum.  (Diagnostics of this build, no counterpart in the
 * reference.) */
int fmt2d_ordered_sumsq(fmt2d_ctx* ctx, const double* x, long long n, double* out);
/* One LSQR solve of [G; w L; wd I] dm = [r; -w L dm_sum; -wd dm_sum] with the
 * reference's column mask and stopping rule (test2 < 1e-4); smooth_type 2
 * solves [G; w R_h; w R_v; wd I] (run_lsqr, inversion2d.f90:6578-7167; the
 * Gaussian operators of :1252-1636 are built on first use for the current
 * parameterisation and correlation lengths, then reused). */
int fmt2d_lsqr(fmt2d_ctx* ctx, const fmt2d_lsqr_params* p, const double* dm_sum, double* dm_out,
               fmt2d_lsqr_stats* stats);

/* Damping of the model update (lsqr_inversion2d). */
typedef struct {
  double lam_damp;        /* > 0: fixed velocity damping; <= 0: automatic search */
  double lam_damp_iface;  /* fixed interface damping (> 0, interfaces inverted); else lam_damp */
  double target_dv;       /* search target of the mean velocity update (<= 0: 0.03) */
  double target_dd;       /* search target of the mean interface depth update, km (<= 0: 0.5) */
  double max_rel_pert;    /* cap of a velocity parameter's update */
} fmt2d_damping_params;

typedef struct {
  double wd, wd_iface;    /* damping weights of the final solve */
  double avg_dmv, avg_dmd; /* mean velocity / interface-depth update of the final solve (with dm_sum) */
  int solves;             /* LSQR solves run */
  int iterations;         /* iterations of the final solve */
} fmt2d_damping_result;

/* The iteration's model update: fixed damping, or the automatic search (no
 * damping first, then secant searches of the interface and the velocity
 * damping weights for the target mean updates, one LSQR solve per trial),
 * then inactive and fixed parameters zeroed and the velocity update capped
 * (lsqr_inversion2d, inversion2d.f90:533-773).  p carries the smoothing
 * settings; its damping weights are ignored. */
int fmt2d_damped_update(fmt2d_ctx* ctx, const fmt2d_lsqr_params* p, const fmt2d_damping_params* d,
                        const double* dm_sum, double* dm_out, fmt2d_damping_result* res);
/* The model as the context holds it: what 0 the velocity grid (nz x nx,
 * x fastest; the propagation grid nodes), 1..5 an interface's node depths
 * (2 seafloor, 3 sediment base, 4 Moho).  With out == NULL or cap == 0
 * only *n. */
int fmt2d_model_get(fmt2d_ctx* ctx, int what, double* out, int cap, int* n);
/* Replaces the velocity grid (what 0) or an interface's node depths (2..4)
 * and rebuilds the geometry, as restore_model_state does
 * (inversion2d.f90:7242-7276); n must match the model's size. */
int fmt2d_model_put(fmt2d_ctx* ctx, int what, const double* in, int n);
/* apply_model_step (inversion2d.f90:7173-7240): slowness update of the
 * velocity grid with the 2-10 km/s and +-30 % of the initial model clamps,
 * interface depths moved by at most 0.5 km, geometry rebuilt. */
int fmt2d_apply_model_step(fmt2d_ctx* ctx, double step, double step_iface, const double* dm);

typedef struct {
  double crit_chi;          /* data selection of the RMS (inv_crit_chi) */
  double iface_step_ratio;  /* interface step = ratio x velocity step */
  int first_step;           /* ladder index to start at, 1..6 (1.0, 0.75, 0.5, 0.25, 0.125, 0.0625) */
} fmt2d_line_search_params;

typedef struct {
  double step;              /* accepted step (0: every step failed, model restored) */
  double rms_orig, rms_new;
  int next_first_step;      /* ladder index for the next iteration */
  int forwards;             /* trial forwards run */
} fmt2d_line_search_result;

/* Line search of lsqr_inversion2d (inversion2d.f90:775-889): for each step
 * of the ladder from first_step, the model moved by step x dm (interfaces by
 * step x iface_step_ratio), a forward and the rays of the last trace (no
 * Frechet rows), the RMS of the selected residuals (compute_rms_after_forward,
 * :3298-3339); stops at the first step below the RMS before the update,
 * applies the best one and scales dm in place by the accepted steps.  Needs
 * the system of this iteration (fmt2d_build_system).  A ray that one trial
 * forward loses stays out of the RMS of the later trials, as ray%valid is
 * only set again at the start of the next iteration (:283-288).  valid
 * (nray, may be NULL): in, the validity of the iteration's rays (NULL: that
 * of the last trace); out, and-ed with every trial forward's.  When a trial's
 * model step, forward or trace fails, the model the caller passed in is put
 * back (restore_model_state, :7242-7276) before the error returns. */
int fmt2d_line_search(fmt2d_ctx* ctx, const fmt2d_line_search_params* p, const double* t_obs, double* dm,
                      int* valid, fmt2d_line_search_result* res);

#ifdef __cplusplus
}
#endif
#endif
