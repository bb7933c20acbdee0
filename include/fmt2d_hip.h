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
 * order (4: four-colour Gauss-Seidel, 1: Jacobi).  Defaults 2.0, 64, 4. */
int fmt2d_set_tuning(fmt2d_ctx* ctx, double window_tiles, int max_sweeps, int colors);
/* Due-tile count of every solver launch of the last forward (diagnostics). */
int fmt2d_forward_trace(fmt2d_ctx* ctx, int* counts, int max, int* n);

/* Region id (1..6) of time field itf (1..12). */
int fmt2d_tf_region(int itf);

/* Field itf (1..12) of source isrc (0-based) in the reference's region node
 * order (grid nodes ix-outer/iz-inner, then top, then bottom interface
 * nodes); arrays hold region_nnode[fmt2d_tf_region(itf)] values. */
int fmt2d_get_field(fmt2d_ctx* ctx, int isrc, int itf, double* T, double* dTdx, double* dTdz);

#ifdef __cplusplus
}
#endif
#endif
