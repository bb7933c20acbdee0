// Context of the C ABI (private to the library): device buffers of the
// model, the resident stage fields of the last forward and the rays of the
// last trace.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/fmt2d_hip.h"
#include "eik_kernels.h"
#include "heap_exit.h"
#include "model.h"
#include "rays.h"
#include "lsqr.h"

using namespace fmt2d;

int fmt2d_fail(int code, const std::string& msg);
#define fail fmt2d_fail

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess) return fail(FMT2D_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

namespace {

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;    // elements in use
  size_t cap = 0;  // elements allocated: a request within it keeps the allocation
  hipError_t alloc(size_t count) {
    if (p && count <= cap) {
      n = count;
      return hipSuccess;
    }
    release();
    const hipError_t e = hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) { p = nullptr; return e; }
    n = count;
    cap = std::max<size_t>(count, 1);
    return hipSuccess;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    cap = 0;
  }
  ~DBuf() { release(); }
};

// stage table of run_forward2d (src/invert2d_main.f90:564-811,
// src/inversion2d.f90:4184-4416): region, start interface, previous stage,
// water-layer top-row fix (+1 up-going, -1 down-going)
const int TF_REGION[13] = {0, 5, 2, 2, 6, 6, 1, 1, 1, 1, 1, 5, 1};
const int TF_ISTART[13] = {0, 2, 2, 3, 2, 4, 2, 1, 2, 2, 1, 2, 2};
const int TF_PREV[13] = {0, 0, 0, 2, 0, 4, 1, 6, 7, 0, 9, 10, 11};
const int TF_WATERFIX[13] = {0, 0, 0, 0, 0, 0, 1, -1, 1, 1, -1, 0, 1};
const int LEVELS[4][4] = {{1, 2, 4, 9}, {3, 5, 6, 10}, {7, 11, 0, 0}, {8, 12, 0, 0}};
const int LEVEL_N[4] = {4, 4, 2, 2};

}  // namespace

// The inversion mesh: every vstride-th grid column and row carries a
// velocity parameter, every istride-th node of an inverted interface a depth
// parameter (strides 1: the grid itself, the reference's parameterisation).
// P prolongs mesh parameters to Frechet columns: bilinear in the velocity
// cell, linear along an interface; mesh nodes that are not parameters (in
// the water at set-up) contribute nothing.  The system matrix is G P and a
// model step moves the grid by P dm.
struct InvMesh {
  int vstride = 1, istride = 1;
  int vnx = 0, vnz = 0;                // mesh velocity nodes
  std::vector<double> vx, vz;          // their x and depth below the grid top
  std::vector<int> inv;                // (vnz, vnx) parameter number, 0: none
  int n_vel = 0, n_par = 0;
  int if_start[NIF + 1] = {}, if_nnode[NIF + 1] = {};
  std::vector<double> ifx[NIF + 1];    // mesh interface node x
  std::vector<int> p_rp, p_col;        // P: row per Frechet column, 0-based mesh parameters
  std::vector<double> p_w;
};

struct fmt2d_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool have_model = false;
  HostModel m;
  // model on the device
  DBuf<double> xg, zg, rvel[NREG + 1], ifx[NREG + 1][2], ifz[NREG + 1][2];
  DBuf<uint16_t> rflags[NREG + 1];
  DBuf<int> ifcc[NREG + 1][2];
  DBuf<uint8_t> ifok[NREG + 1][2];
  DBuf<RegionDev> regs;
  DBuf<double> isn[NIF + 1][4];  // nx, nz, vel_top, vel_bot
  DBuf<IsecDev> isecs;
  DBuf<double> isnp[NIF + 1][2];  // normals of the forward program's reflection
  DBuf<IsecDev> isecs_prog;       // isecs with those normals
  bool fwd_program = false;       // last forward ran as the forward program (FMT2D_FWD_PROGRAM)
  // the fields on the device are those of the current model for the last
  // forward's sources and flags (set by a line search that ends on its
  // accepted trial; cleared by every model change and forward)
  bool fields_current = false;
  int fields_flags = 0;
  DBuf<int> hw_slot;
  DBuf<double> hw_vhw;
  DBuf<int> hw_vslot;                 // region-6 velocity slot of each hw_vhw entry (-1: none)
  DBuf<HeadwaveMap> hw_map;
  // fields of the last forward
  int nsrc = 0, nfield = 0;
  std::vector<int> field_of;           // [isrc*12 + itf-1] -> field index
  DBuf<double> fT, isT, fGX, fGZ, fOGX, fOGZ, isGX, isGZ;
  DBuf<uint8_t> fTIE;
  DBuf<FieldDesc> descs;
  DBuf<StageLink> links;
  DBuf<unsigned int> tin;             // tile in active set
  DBuf<unsigned long long> tkey;      // tile arrival key
  DBuf<unsigned long long> tfront;    // per-field front key of active lists 0 and 1
  DBuf<int2> tlist[3];                // active set, its next state, due tiles
  DBuf<int> tcount;                   // due counters, sizes of lists 0 / 1, launches (LevelSched)
  DBuf<int> ttrace;                   // due tiles per launch of a level
  DBuf<int> tmax;                     // most sweeps of a tile visit per launch of a level
  DBuf<unsigned long long> tspan;     // device clock at the start / ends of each launch of a level
  double clock_hz = 0.0;              // rate of the device's constant clock (wall_clock64)
  DBuf<uint8_t> theavy;               // per (region, tile): cut by an interface (slow visits)
  double window_tiles = 3.0;          // window in tile crossing times
  int max_sweeps = 16;                // a visit over the cap stays active and continues later
  int max_sweeps_heavy = 8;           // the cap for tiles cut by an interface
  int colors = 2;                     // Gauss-Seidel colours in a tile (4, 2) or 1: Jacobi
  int tile_x = 8, tile_z = 8;         // tile shape of the solver (8x8, 16x8, 8x16)
  int field_f32 = 0;                  // 1: the tile solver stores grid-node T and gradients in 32 bits
  int fwd_f32 = 0;                    // the last forward did
  DBuf<unsigned> f32T;                // its storage for one level's fields: T (fixed point)
  DBuf<float> f32GX, f32GZ;           // and the gradients (FP32)
  DBuf<double> snapT;                 // fmt2d_field_snapshot: T of every field of a forward
  int snap_nfield = 0;
  std::vector<double> snap_src;
  int heavy_tx = 0, heavy_tz = 0;     // shape the slow-tile flags (theavy) were built for
  int launch_cap = 0;                 // solver launches per stage level before FMT2D_ENOCONV (0: by grid size)
  int solver_launch = 0;              // FMT2D_LAUNCH_GRAPH or FMT2D_LAUNCH_LEVEL
  int level_grid[4] = {0, 0, 0, 0};   // cooperative grid of the level kernel (8x8, 16x8, 8x16, 8x8 FP32)
  DBuf<unsigned> tbar;                // its grid barrier's counters
  size_t adj_cap_bytes = (size_t)4 << 30;  // adjoint sweep: edge records of at most this many bytes at once
  int lsqr_fast = 0;                  // LSQR reductions: 0 in the reference's order, 1 tree / wave reductions
  DBuf<unsigned long long> acc;       // sweep statistics [sum, max, capped visits]
  DBuf<int> stop_flag;                // set when a time is NaN (the reference's stop case)
  DBuf<int> sel_which, sel_up;
  DBuf<double> hw_scratch;
  DBuf<double> hw_save;               // Moho-node times of the PmP up-going fields before the head-wave pass
  std::vector<int> hw_fields;         // their field indices, in hw_save's order
  // heap-exhaustion stop work arrays
  DBuf<ExitField> x_ef;
  DBuf<int> x_cnt, x_diff, x_off, x_total;
  DBuf<unsigned long long> x_max, x_m;
  DBuf<ExitEntry> x_ent;
  // ---- rays and Frechet rows
  std::vector<double> src_x, src_z;    // sources of the last forward (on the seafloor)
  DBuf<double> d_srcx, d_srcz;
  bool rgeom_ok = false;               // ray geometry built for the current model
  RayGeomDev hgeom{};
  DBuf<RayGeomDev> rgeom;
  DBuf<int> rorder[NREG + 1];
  DBuf<int> rjlo[NREG + 1], rjhi[NREG + 1];        // nearest-node search: rows per column
  DBuf<double> rxlo[NREG + 1], rxhi[NREG + 1];     // and x bounds of the columns from / up to c
  int ray_scan = 0;                                // 1: nearest-node search by the full scan
  long long ray_nearest_calls = 0;                 // nearest-node searches of the last trace
  DBuf<double> risx[NIF + 1], risz[NIF + 1], rifx_in[NIF + 1];
  DBuf<double> rsfx, rsfz, rsfz0, rpvel, rvx, rvz;
  DBuf<int> rinv;
  std::vector<int> h_inv;              // parameter number of each velocity-grid node (0: none)
  DBuf<RayField> rfields;
  int nray = 0;
  std::vector<RayJob> hjobs;
  std::vector<long long> hpoff;
  DBuf<RayJob> rjobs;
  CsrWork asm_work;                   // row assembly: chunk offsets and summands (grown, kept)
  CsrWork csr_work;                   // scratch of build_csr / build_csc (grown, kept)
  DBuf<int> rperm;                    // trace order of the rays (longest first; by source and path within)
  std::vector<int> hperm;
  std::vector<int> hwork;             // interpolations per ray in the last trace of hjobs
  DBuf<RayOut> rout;
  DBuf<long long> rpoff;
  DBuf<double2> rpts;
  DBuf<double> rtsf;
  std::vector<double> h_tcalc;         // receiver times of the last trace
  std::vector<int> h_valid;
  // Frechet rows of the last trace on the device (0-based columns)
  DBuf<int> g_rp, g_col;
  DBuf<double> g_val;
  DBuf<int> gt_cp, gt_row;             // column copy of G for G^T x (fmt2d_frechet_apply)
  DBuf<double> gt_val;
  bool gt_ok = false;
  int g_nnz = -1;                      // -1: not built for the current trace
  long long frechet_chunk = 1LL << 28; // summands sorted at once by the row assembly
  // tomography system (selected, weighted, scaled rows) and its column copy
  int sys_ndata = -1, sys_nnz = 0;
  double sys_eps_min = 0.0;
  double sys_rms = 0.0;                // RMS of the selected residuals (before weighting)
  double sys_crit_chi = 0.0;
  int sys_alt_mode = 0;
  std::vector<double> h_mvs_model;     // mvscale of the model update (alternating zeroing undone)
  DBuf<int> s_rp, s_col, s_cp, s_crow;
  DBuf<double> s_val, s_cval, s_res, s_gtg, s_mvs;
  std::vector<double> h_gtg;
  // LSQR work arrays
  DBuf<double> l_u, l_v, l_w, l_dm, l_gv, l_gtu, l_sm, l_sm2, l_dmsum, l_ws, l_wd;
  // Gaussian smoothing operators, built once per parameterisation and
  // correlation lengths (row and column copies, 0-based)
  std::vector<double> h_mvs_smooth;    // mvscale the operators carry (velocity_init based)
  std::vector<double> v_init;          // velocity grid when the inversion was set up
  std::vector<double> topo_init;       // seafloor then (fixes the velocity parameter numbering)
  double gs_key[3] = {-1, -1, -1};     // corr_x, corr_z of R_h/R_v; [2] corr_x_iface of R_if
  int gs_nparm = -1, gsi_nparm = -1, gsi_nnz = 0;
  DBuf<int> gh_rp, gh_col, gh_cp, gh_row, gv_rp, gv_col, gv_cp, gv_row, gi_rp, gi_col;
  DBuf<double> gh_val, gh_cval, gv_val, gv_cval, gi_val;
  DBuf<uint8_t> l_mask;
  DBuf<int4> l_lapnb;
  DBuf<int2> l_ifnb;
  DBuf<LsqrScalars> l_sc;
  // second solve running beside the first (damping searches)
  DBuf<double> l2_u, l2_v, l2_w, l2_dm, l2_gv, l2_gtu, l2_sm, l2_sm2, l2_wd;
  DBuf<LsqrScalars> l2_sc;
  DBuf<double> l_part, l2_part;       // partial sums of the tree-reduced norms
  hipStream_t stream2 = nullptr;
  // inversion parameterisation (frechet.in): velocity grid and interfaces;
  // n_vel_par / n_par count the Frechet columns (grid nodes, interface
  // input nodes), mesh the parameters the system solves for
  int inv_vel = 1;
  std::vector<int> inv_ifaces;
  int n_vel_par = 0, n_par = 0;
  InvMesh mesh;
  bool mesh_ok = false;                // mesh and P built for the parameterisation
  DBuf<int> p_rp, p_col;               // prolongation P by Frechet column (0-based)
  DBuf<double> p_w;
  DBuf<int> gp_rp, gp_col;             // G P of the last trace (0-based columns)
  DBuf<double> gp_val;
  int gp_nnz = -1;                     // -1: not formed for the current trace
  int* h_count = nullptr;              // pinned, 16 ints
  fmt2d_forward_stats stats{};
  std::vector<int> trace;              // active tiles per launch (all levels)
  std::vector<int> trace_sweeps;       // most sweeps of a tile visit per launch
  long long sweep_hist[64] = {};       // tile visits by sweep count (last bin: 64 or more)
  ~fmt2d_ctx() {
    if (h_count) (void)hipHostFree(h_count);
    if (stream) (void)hipStreamDestroy(stream);
    if (stream2) (void)hipStreamDestroy(stream2);
  }
};

template <typename T>
static hipError_t upload(DBuf<T>& b, const std::vector<T>& v, hipStream_t st) {
  hipError_t e = b.alloc(v.size());
  if (e != hipSuccess) return e;
  return hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st);
}

// Frechet rows of the last trace on the device (capi_rays.cpp)
int frechet_build(fmt2d_ctx* ctx);
// G P: the Frechet rows over the mesh parameters (capi_rays.cpp)
int system_rows_build(fmt2d_ctx* ctx);
int model_upload(fmt2d_ctx* ctx);
int lsqr_solve(fmt2d_ctx* ctx, const fmt2d_lsqr_params* p, const double* dm_sum, double* dm_out,
               fmt2d_lsqr_stats* stats);
// one or two solves of the built system (damping weights per solve) side by side
int lsqr_solve_many(fmt2d_ctx* ctx, const fmt2d_lsqr_params* p, const double* dm_sum, int n, const double* wdv,
                    const double* wdv_if, double* const* dm_out, fmt2d_lsqr_stats* stats);
