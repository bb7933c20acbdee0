// Device-side descriptors and kernel launchers of the eikonal path.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "eik_common.h"
#include "heap_exit.h"

namespace fmt2d {

enum { SEED_POINT = 0, SEED_IFACE = 1 };

// One traveltime field (one source x one stage) resident in HBM.
struct FieldDesc {
  int region;        // 1..6
  int kind;          // SEED_POINT / SEED_IFACE
  double sx, sz, t0; // point source
  int ix0, iz0;      // lower-left node of the source cell (0-based)
  int fixed_key;     // grid slot of an on-node source, -1 none
  int side;          // start interface: 0 = region top, 1 = region bottom
  const double* isT; // start-interface times after transfer/reflection (nx)
  const double* isGX;
  const double* isGZ;
  double* T;         // (nz+2)*nx traveltime
  double* GX;        // raw gradient (fixed-point iteration)
  double* GZ;
  double* OGX;       // gradient after the refit / fix post-passes (output)
  double* OGZ;
  uint8_t* TIE;      // equal-time support mask of the grid nodes (TIE_* bits)
  // 32-bit storage of the tile solver (T32 == nullptr: FP64 storage): the
  // grid nodes' T in unsigned 32-bit fixed point (T32_UNIT s, one grid of values for
  // every node, so the stored field does not depend on the order of the
  // updates), the raw gradients as FP32; compute stays FP64 and the
  // interface nodes stay FP64.  Copied to T / GX / GZ after the solve.
  unsigned* T32;
  float* GX32;
  float* GZ32;
};
constexpr double T32_SCALE = 16777216.0;         // 2^24 units per second
constexpr double T32_UNIT = 1.0 / T32_SCALE;     // 6.0e-8 s
constexpr double T32_MAX = 255.0;                // times up to here are stored (the grid spans 256 s)
constexpr unsigned T32_NONE = 0xFFFFFFFFu;       // unreached

// Interface stage: where the start-interface values come from.
struct StageLink {
  int prev;           // field index of the previous stage
  int prev_side;      // row of the start interface in the previous region (0 top, 1 bottom)
  int isec;           // intersection id 1..5 (index into IsecDev array)
  int reflect;        // apply Snell re-projection
  int target_side;    // +1 start interface is the target's top, -1 bottom, 0 by gradient sign
  int target_has_iface;
  const double* target_vel; // target region's velocity row of that interface (nx)
};

struct IsecDev {
  const double* nx;
  const double* nz;
  const double* vel_top;
  const double* vel_bot;
};

// Moho head-wave post-pass map (src/inversion2d.f90:4296-4345)
struct HeadwaveMap {
  int n;
  double dxif;
  const int* slot;     // field slot of the node the reference addresses, -1 none
  const double* vhw;   // head-wave velocity per Moho node
};

hipError_t launch_fill(const FieldDesc* descs, int nfield, int nslot, hipStream_t st);
hipError_t launch_transfer_reflect(const FieldDesc* descs_all, int f0, const StageLink* links_all, int nfield, int nx,
                                   int nz, const IsecDev* isecs, hipStream_t st);
hipError_t launch_seed_tiles(const FieldDesc* descs, const RegionDev* regs, int nfield, int nx, int ntx, int ntz,
                             int tx, int tz, unsigned int* in_set, unsigned long long* tkey, int2* list_next,
                             int* count_next, unsigned long long* front_next, hipStream_t st);
// per launch: clear the due / next counters and the next list's field
// fronts, then split the active list (its count on the device) into due
// tiles and carry-over (front: the active list's fronts)
hipError_t launch_schedule(const int2* list, const int* count, int nfield, int ntx, int ntiles,
                           unsigned long long* tkey, const unsigned long long* front, unsigned long long* front_next,
                           double window, int color, unsigned int* in_set, int2* due, int* count_due, int2* carry,
                           int* count_carry, int* launch_no, int* trace, int trace_cap, const FieldDesc* descs,
                           const uint8_t* heavy, int due_cap, int grid, hipStream_t st);
// Per launch, span holds SPAN_SLOTS + 1 stamps of the device's constant
// clock: the start, then end stamps (the launch ends at the largest).
constexpr int SPAN_SLOTS = 64;
// tile shapes tx x tz: 8x8, 16x8, 8x16 (hipErrorInvalidValue otherwise)
hipError_t launch_fim(const FieldDesc* descs, const RegionDev* regs, const int2* list, int grid, const int* count_due,
                      int ntx, int ntz, int tx, int tz, unsigned int* in_set, unsigned long long* tkey,
                      int2* list_next, int* count_next, unsigned long long* front_next, int max_sweeps,
                      int max_sweeps_heavy, int colors, unsigned long long* acc, const int* launch_no, int* sweep_max,
                      int trace_cap, unsigned long long* span, int due_cap, int f32, int* flag, hipStream_t st);
// All launches of a dependency level in one cooperative dispatch.  count
// (16 ints): [1] / [2] sizes of active lists 0 / 1 (list 0 seeded, list 1
// empty), [4] launches run (written at the end), [0] / [3] and [8] / [11]
// due counters and [12] / [13] tile hand-out counters of even / odd
// launches (all zero at the start); bar: level_bar_words(grid) zeros (the
// grid barrier's counters); front[0]
// holds list 0's fronts, front[1] no key (0x7F bytes).  trace[b]: due tiles of
// launch b.
struct LevelSched {
  const FieldDesc* descs;
  const RegionDev* regs;
  const uint8_t* heavy;
  int2* list[2];
  int2* due;
  int* count;
  unsigned int* in_set;
  unsigned long long* tkey;
  unsigned long long* front[2];
  unsigned long long* acc;
  int* trace;
  int* sweep_max;
  unsigned long long* span;
  int* flag;
  unsigned* bar;
  double window;
  int nfield, ntx, ntz, ntiles, due_cap;
  int max_sweeps, max_sweeps_heavy, colors, max_launch, trace_cap;
};
// blocks of the level kernel's cooperative grid (all resident on the device)
hipError_t level_grid(int device, int tx, int tz, int f32, int* grid);
size_t level_bar_words(int grid);
hipError_t launch_fim_level(const LevelSched& a, int tx, int tz, int f32, int grid, hipStream_t st);
// per field: max |T - snap| over slots both reach, count of slots one
// reaches, count of slots both reach with |T - snap| > tol
hipError_t launch_field_compare(const FieldDesc* descs, int nfield, const double* snap, long long nslot, double tol,
                                double* max_dt, long long* n_diff, long long* n_over, hipStream_t st);
// 32-bit storage: reset the level's arrays (nslotg grid slots per field),
// and copy them out to T / GX / GZ
hipError_t launch_f32_reset(const FieldDesc* descs, int nfield, int nslotg, hipStream_t st);
hipError_t launch_f32_export(const FieldDesc* descs, int nfield, int nslotg, hipStream_t st);
hipError_t launch_member_time(const FieldDesc* descs, const RegionDev* regs, int nfield, int nx, int nz,
                              hipStream_t st);
// device work arrays of the heap-exhaustion stop (heap_exit.h); bin arrays
// hold nfield*EXIT_NBIN entries, `ent` holds `cap` entries
struct ExitWork {
  ExitField* ef;
  int* bcnt;
  int* bdiff;
  int* boff;
  unsigned long long* bmax;
  unsigned long long* bm;
  int* total;
  ExitEntry* ent;
  long long cap;
};
// init + stats + bins + scan; afterwards *total holds the number of nodes
// in open bins
hipError_t launch_exit_bins(const FieldDesc* descs, const RegionDev* regs, int nfield, int nslot, ExitWork& w,
                            hipStream_t st);
// gather + exact check (needs cap >= *total) + truncation of the fields
hipError_t launch_exit_check(const FieldDesc* descs, const RegionDev* regs, int nfield, int nslot, ExitWork& w,
                             long long ncand_bound, hipStream_t st);
// also sets *stop_flag when a grid node's time is NaN (the reference's stop
// on a negative discriminant too large, propagate2d.f90:745-756)
hipError_t launch_refit(const FieldDesc* descs, const RegionDev* regs, int nfield, int nx, int nz, int* stop_flag,
                        hipStream_t st);
struct FieldSummary {
  long long reached;         // slots with a time
  unsigned long long hash;   // order-free hash of the bits of T and the output gradient
  double max_resid, sum_resid;
  long long n_resid, n_resid_big;  // nodes in the eikonal residual, residuals above 0.05
};
hipError_t launch_field_summary(const FieldDesc* descs, const RegionDev* regs, int nfield, int nx, int nz,
                                FieldSummary* out, hipStream_t st);
hipError_t launch_water_fix(const FieldDesc* descs, const RegionDev* regs, const int* which, const int* up, int nsel,
                            int nx, hipStream_t st);
// save (nsel x hm->n, may be NULL): the Moho-node times before the pass
hipError_t launch_headwave(const FieldDesc* descs, const RegionDev* regs, const int* which, int nsel,
                           const HeadwaveMap* hm, double* scratch, double* save, hipStream_t st);
// exchanges the Moho-node times of fields which[0..nsel) with save
hipError_t launch_headwave_swap(const FieldDesc* descs, const int* which, int nsel, const HeadwaveMap* hm, int n,
                                double* save, hipStream_t st);

}  // namespace fmt2d
