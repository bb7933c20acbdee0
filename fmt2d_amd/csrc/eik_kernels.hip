// MI355X kernels of the multistage eikonal path.
//
//  * fill_fields      : reset a batch of fields (T=huge, gradients 0)
//  * transfer_reflect : interface times of the previous stage, Snell
//                       re-projection (src/propagate2d.f90:2414-2741)
//  * seed_tiles       : first active tiles of every field
//  * fim_tiles        : block fast-iterative solve, one wavefront per 8x8
//                       tile with a 2-node halo staged in LDS; tiles whose
//                       border changed activate their neighbours for the
//                       next launch (deduplicated by per-launch flag arrays)
//  * refit_gradients  : the reference's post-pass refit of one-sided
//                       gradients (src/propagate2d.f90:119-123)
//  * water_top_fix    : top-row gradient fix of water-layer stages
//                       (src/inversion2d.f90:4264-4292)
//  * pmp_headwave     : head-wave min-plus sweep along the Moho on the PmP
//                       up-going stage (src/inversion2d.f90:4294-4409)
//
// Fields are FP64 and every update follows the reference's operation order
// (this file is compiled with -ffp-contract=off): an FP32 traveltime field
// accumulates ~1.5e-4 s of rounding over 400 nodes, above the 1e-4 s bar, and
// exact FP64 reproduces the reference's tie decisions (gradient refits).
#include <hip/hip_runtime.h>

#include "eik_kernels.h"
#include "eik_update.h"
#include "eik_regular.h"
#include "eik_field.h"

namespace fmt2d {

// Load through a global-address-space pointer (a pointer read from memory is
// generic to the compiler, which then orders it against LDS stores).
template <typename T>
__device__ inline T gload(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}

// Tile shapes.  A wavefront solves one tile of TX x TZ grid nodes with a
// 2-node halo staged in LDS.  8 x 8: one node per lane.  16 x 8 and 8 x 16:
// two nodes per lane, one of each checkerboard colour (a horizontal pair),
// so every lane has a node in each two-colour phase.
constexpr int HALO = 2;

template <int TX, int TZ>
struct Shape {
  static constexpr int NN = TX * TZ;      // tile nodes
  static constexpr int NPL = NN / 64;     // nodes per lane (= 64-bit words of a node mask)
  static constexpr int WX = TX + 2 * HALO, WZ = TZ + 2 * HALO, W2 = WX * WZ;
  static_assert(NN == 64 || NN == 128, "tile of 64 or 128 nodes");
  static_assert(2 * TX <= 32, "interface nodes of a tile fit a 32-bit mask");
  // tile position (a, b) of node k of a lane
  __device__ static int node_a(int lane, int k) {
    if (NPL == 1) return lane % TX;
    const int b = node_b(lane, k);
    return 2 * (lane % (TX / 2)) + ((b + k) & 1);
  }
  __device__ static int node_b(int lane, int k) {
    if (NPL == 1) return lane / TX;
    return lane / (TX / 2);
  }
  // node index (lane + 64 k) of tile position (a, b)
  __device__ static int node_index(int a, int b) {
    if (NPL == 1) return b * TX + a;
    return (b * (TX / 2) + (a >> 1)) + 64 * ((a + b) & 1);
  }
};

template <int NPL>
struct NodeMask {
  unsigned long long w[NPL];
  __device__ void clear() {
#pragma unroll
    for (int k = 0; k < NPL; k++) w[k] = 0ull;
  }
  __device__ void fill() {
#pragma unroll
    for (int k = 0; k < NPL; k++) w[k] = ~0ull;
  }
  __device__ void set(int n) { w[n >> 6] |= 1ull << (n & 63); }
  __device__ void operator|=(const NodeMask& o) {
#pragma unroll
    for (int k = 0; k < NPL; k++) w[k] |= o.w[k];
  }
  __device__ bool meets(const NodeMask& o) const {
    unsigned long long x = 0ull;
#pragma unroll
    for (int k = 0; k < NPL; k++) x |= w[k] & o.w[k];
    return x != 0ull;
  }
};

// LDS view of one tile window plus the interface nodes of its columns.
// Accessors clamp to the window; the operator never reaches beyond two
// nodes from an inner node, which the halo covers.
template <int TX, int TZ>
struct TileLds {
  using S = Shape<TX, TZ>;
  double T[S::W2];
  double GX[S::W2], GZ[S::W2], vel[S::W2];
  uint16_t flags[S::W2];
  uint8_t tie[S::W2];
  double xg[S::WX], zg[S::WZ];
  double iT[2 * S::WX];
  double iGX[2 * S::WX], iGZ[2 * S::WX], ivel[2 * S::WX], ifx[2 * S::WX], ifz[2 * S::WX];
  int ifcc[2 * S::WX];
  uint8_t ifok[2 * S::WX];
  // tiles of two nodes per lane: the nodes' seeds and fast-path model data
  // (one node per lane keeps them in registers)
  static constexpr int NS = S::NPL > 1 ? S::NN : 1;
  RegStatic rs[NS];
  double t0[NS], gx0[NS], gz0[NS];
};

template <int TX, int TZ>
struct LView {
  using S = Shape<TX, TZ>;
  const RegionDev* R;
  const FieldDesc* D;
  const TileLds<TX, TZ>* L;
  int wx0, wz0;
  __device__ int w(int i, int j) const {
    int a = i - wx0, b = j - wz0;
    a = a < 0 ? 0 : (a >= S::WX ? S::WX - 1 : a);
    b = b < 0 ? 0 : (b >= S::WZ ? S::WZ - 1 : b);
    return b * S::WX + a;
  }
  __device__ int wc(int s, int c) const {
    int a = c - wx0;
    a = a < 0 ? 0 : (a >= S::WX ? S::WX - 1 : a);
    return s * S::WX + a;
  }
  __device__ const RegionDev& reg() const { return *R; }
  __device__ bool in(int i, int j) const {
    return i >= 0 && i < R->nx && j >= 0 && j < R->nz && (L->flags[w(i, j)] & RF_IN);
  }
  __device__ uint16_t flags(int i, int j) const { return L->flags[w(i, j)]; }
  __device__ double T(int i, int j) const { return L->T[w(i, j)]; }
  __device__ double GX(int i, int j) const { return L->GX[w(i, j)]; }
  __device__ double GZ(int i, int j) const { return L->GZ[w(i, j)]; }
  __device__ int tie(int i, int j) const { return L->tie[w(i, j)]; }
  __device__ double vel(int i, int j) const { return L->vel[w(i, j)]; }
  __device__ double xg(int i) const {
    int a = i - wx0; a = a < 0 ? 0 : (a >= S::WX ? S::WX - 1 : a); return L->xg[a];
  }
  __device__ double zg(int j) const {
    int a = j - wz0; a = a < 0 ? 0 : (a >= S::WZ ? S::WZ - 1 : a); return L->zg[a];
  }
  __device__ double iT(int s, int c) const { return L->iT[wc(s, c)]; }
  __device__ double iGX(int s, int c) const { return L->iGX[wc(s, c)]; }
  __device__ double iGZ(int s, int c) const { return L->iGZ[wc(s, c)]; }
  __device__ double ivel(int s, int c) const { return L->ivel[wc(s, c)]; }
  __device__ double ifx(int s, int c) const { return L->ifx[wc(s, c)]; }
  __device__ double ifz(int s, int c) const { return L->ifz[wc(s, c)]; }
  __device__ int ifcc(int s, int c) const { return L->ifcc[wc(s, c)]; }
  __device__ int ifok(int s, int c) const { return L->ifok[wc(s, c)]; }
  __device__ double T0(int i, int j) const { double t, a, b; seed_grid(*D, *R, i, j, t, a, b); return t; }
  __device__ double GX0(int i, int j) const { double t, a, b; seed_grid(*D, *R, i, j, t, a, b); return a; }
  __device__ double GZ0(int i, int j) const { double t, a, b; seed_grid(*D, *R, i, j, t, a, b); return b; }
  __device__ double iT0(int s, int c) const { double t, a, b; seed_iface(*D, s, c, t, a, b); return t; }
  __device__ double iGX0(int s, int c) const { double t, a, b; seed_iface(*D, s, c, t, a, b); return a; }
  __device__ double iGZ0(int s, int c) const { double t, a, b; seed_iface(*D, s, c, t, a, b); return b; }
  __device__ bool fixed(const Member& m) const {
    return m.kind == 0 && (m.j * R->nx + m.i) == D->fixed_key;
  }
};

// ---------------------------------------------------------------- kernels

__global__ void fill_fields_kernel(const FieldDesc* __restrict__ descs, int nfield, int nslot) {
  const long long total = (long long)nfield * nslot;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (long long)gridDim.x * blockDim.x) {
    const int f = (int)(k / nslot), q = (int)(k % nslot);
    const FieldDesc& d = descs[f];
    d.T[q] = FMT_HUGE_TIME;
    d.GX[q] = 0.0; d.GZ[q] = 0.0;
    d.OGX[q] = 0.0; d.OGZ[q] = 0.0;
    d.TIE[q] = 0;
  }
}

// Interface times of the previous stage copied onto the start interface
// (transfer_times_to_intersection mapped branch), then re-projected by
// Snell's law into the target region when requested.  One thread per
// (field, column).
__global__ void transfer_reflect_kernel(const FieldDesc* __restrict__ descs_all, int f0,
                                        const StageLink* __restrict__ links_all, int nfield, int nx, int nz,
                                        const IsecDev* __restrict__ isecs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nfield * nx) return;
  const int f = f0 + k / nx, c = k % nx;
  const FieldDesc& d = descs_all[f];
  const StageLink& L = links_all[f];
  if (d.kind != SEED_IFACE) return;
  const FieldDesc& p = descs_all[L.prev];
  const int q = (nz + L.prev_side) * nx + c;
  double t = p.T[q];
  // the previous stage's gradient after its post-pass (the reference copies
  // the field as saved, i.e. refitted)
  double gx = p.OGX[q], gz = p.OGZ[q];
  if (L.reflect) {
    const IsecDev& I = isecs[L.isec];
    const double n0 = I.nx[c], n1 = I.nz[c];
    if (t >= 0.5e20) { gx = 0.0; gz = 0.0; }
    else {
      const double gp = gx * n0 + gz * n1;
      const double p0 = gx - gp * n0, p1 = gz - gp * n1;
      int side = L.target_side;                    // +1 top of target, -1 bottom, 0 decide
      if (side == 0) side = gp < 0.0 ? -1 : 1;
      double vt = side > 0 ? I.vel_bot[c] : I.vel_top[c];
      if (!(vt > 0.01)) vt = L.target_has_iface ? (double)L.target_vel[c] : 0.0;
      if (vt <= 0.01) { t = 1.0e20; gx = 0.0; gz = 0.0; }
      else {
        const double pt = 1.0 / vt, pn2 = pt * pt - (p0 * p0 + p1 * p1);
        if (pn2 < -1.0e-10) { t = 1.0e20; gx = 0.0; gz = 0.0; }
        else {
          const double sq = sqrt(pn2 < 0.0 ? 0.0 : pn2);
          gx = p0 + side * sq * n0; gz = p1 + side * sq * n1;
        }
      }
    }
  }
  double* isT = const_cast<double*>(d.isT);
  double* isGX = const_cast<double*>(d.isGX);
  double* isGZ = const_cast<double*>(d.isGZ);
  isT[c] = t; isGX[c] = gx; isGZ[c] = gz;
}

// ---------------------------------------------------------------- tile schedule
//
// Active tiles form a set (in_set flag per field x tile) with an arrival key
// (smallest time of the change that activated it, as ordered bits of a
// non-negative double).  Each launch processes only the tiles whose key lies
// within `window` seconds of their field's front (smallest key): tiles are
// solved roughly in arrival-time order, as the reference's heap does, so most
// tiles are final after one or two visits instead of being re-solved for
// every transient upstream value.  Tiles outside the window are carried over.

// "no key": a byte-repeated pattern (about 3.4e296 as a double) so the key
// array can be reset with a memset
constexpr unsigned long long KEY_INF = 0x7F7F7F7F7F7F7F7Full;

__device__ inline unsigned long long key_of(double t) {
  return t <= 0.0 ? 0ull : (unsigned long long)__double_as_longlong(t);
}

// Position for this lane's item in an append list: one atomic per wave.
// Every lane of the wave must call it (pred false for lanes with no item).
__device__ inline int wave_append(int* counter, bool pred) {
  const unsigned long long m = __ballot(pred);
  if (m == 0ull) return -1;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, __popcll(m));
  base = __shfl(base, leader, 64);
  return pred ? base + __popcll(m & ((1ull << lane) - 1ull)) : -1;
}

// The field's front over the next active set is kept as the tiles enter it
// (front_next[f]: smallest key of field f's tiles in the set).  It only
// falls while a launch runs, so a plain read that is not yet current reads
// a value at or above the true one and never skips a needed atomic.
__device__ inline void lower_front(unsigned long long* front_next, int f, unsigned long long key) {
  if (key < gload(front_next + f)) atomicMin(front_next + f, key);
}

__device__ inline void add_tile(int f, int t, int ntiles, unsigned long long key, unsigned int* in_set,
                                unsigned long long* tkey, int2* list_next, int* count_next,
                                unsigned long long* front_next) {
  const long long k = (long long)f * ntiles + t;
  atomicMin(tkey + k, key);
  lower_front(front_next, f, key);
  if (atomicExch(in_set + k, 1u) == 0u) {
    const int pos = atomicAdd(count_next, 1);
    list_next[pos] = make_int2(f, t);
  }
}

// First active tiles: the point-source cell, or every tile within three rows
// of a start-interface node that carries a time.
__global__ void seed_tiles_kernel(const FieldDesc* __restrict__ descs, const RegionDev* __restrict__ regs,
                                  int nfield, int ntx, int ntz, int tx, int tz, unsigned int* in_set,
                                  unsigned long long* tkey, int2* list_next, int* count_next,
                                  unsigned long long* front_next) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const RegionDev& R0 = regs[0];
  const int nx = R0.nx, nz = R0.nz;
  if (k >= nfield * nx) return;
  const int f = k / nx, c = k % nx;
  const FieldDesc& d = descs[f];
  const RegionDev& R = regs[d.region];
  const int ntiles = ntx * ntz;
  if (d.kind == SEED_POINT) {
    if (c != 0) return;
    if (d.fixed_key >= 0) {  // on-node source: alive from the start (:2366-2369)
      d.T[d.fixed_key] = d.t0; d.GX[d.fixed_key] = 0.0; d.GZ[d.fixed_key] = 0.0;
    }
    for (int i = d.ix0; i <= d.ix0 + 1; i++)
      for (int j = d.iz0; j <= d.iz0 + 1; j++)
        if (i >= 0 && i < nx && j >= 0 && j < nz)
          add_tile(f, (j / tz) * ntx + i / tx, ntiles, key_of(d.t0), in_set, tkey, list_next, count_next,
                   front_next);
    return;
  }
  if (!(d.isT[c] < 0.5e20)) return;
  const int cc = R.ifcc[d.side][c];
  if (cc < 0) return;
  const int j0 = max(0, cc - 3), j1 = min(nz - 1, cc + 4);
  const unsigned long long key = key_of(d.isT[c]);
  for (int i = max(0, c - 1); i <= min(nx - 1, c + 1); i++)
    for (int j = j0; j <= j1; j++)
      if ((j % tz) == 0 || j == j0)
        add_tile(f, (j / tz) * ntx + i / tx, ntiles, key, in_set, tkey, list_next, count_next, front_next);
}

// Due tiles of a launch: count_due[0] tiles cut by an interface (placed
// from the front of the due list), count_due[3] the others (placed from its
// end), so the slow tiles start first.
__device__ inline int due_count(const int* count_due) { return count_due[0] + count_due[3]; }

// Clears the counters of a launch: due tiles and the next active list, and
// the fronts of the next list (its tiles lower them as they enter).
// The previous launch's due count goes to its trace slot (when given).
// launch_no counts the launches of a level (the kernels of a launch read it,
// so a batch of launches has the same arguments every time it is replayed).
__global__ void launch_reset_kernel(int* count_due, int* count_next, unsigned long long* front_next, int nfield,
                                    int* launch_no, int* trace, int trace_cap) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f == 0) {
    const int prev = *launch_no - 1;
    if (prev >= 0 && prev < trace_cap) trace[prev] = due_count(count_due);
    *launch_no = prev + 2;
    count_due[0] = 0;
    count_due[3] = 0;
    *count_next = 0;
  }
  if (f < nfield) front_next[f] = KEY_INF;
}

// Splits the active set into tiles due this launch and tiles carried over.
// Tiles of one checkerboard colour run per launch, so two edge-adjacent
// tiles never iterate against each other's stale halos (which can lock a
// pair into a two-cycle of the non-monotone second-order switch).  Threads
// tid, tid + stride, ... take the n entries of the list; every thread of
// the grid calls it (wave_append).
__device__ inline void split_pass(const int2* list, int n, int ntx, int ntiles, unsigned long long* tkey,
                                  const unsigned long long* front, unsigned long long* front_next, double window,
                                  int color, unsigned int* in_set, int2* due, int* count_due, int2* carry,
                                  int* count_carry, const FieldDesc* __restrict__ descs,
                                  const uint8_t* __restrict__ heavy, int due_cap, int tid, int stride) {
  // every lane of a wave runs the same number of rounds
  const int rounds = (n + stride - 1) / stride;
  for (int r = 0; r < rounds; r++) {
    const int k = r * stride + tid;
    bool is_due = false, is_carry = false;
    int2 e = make_int2(0, 0);
    unsigned long long kb = KEY_INF;
    if (k < n) {
      e = list[k];
      const long long q = (long long)e.x * ntiles + e.y;
      kb = tkey[q];
      const double key = __longlong_as_double((long long)kb);
      const double fr = __longlong_as_double((long long)front[e.x]);
      const int tc = ((e.y % ntx) + (e.y / ntx)) & 1;
      is_due = key <= fr + window && tc == color;
      is_carry = !is_due;
      if (is_due) { tkey[q] = KEY_INF; in_set[q] = 0u; }
    }
    const bool hv = is_due && heavy[(long long)descs[e.x].region * ntiles + e.y] != 0;
    const int ph = wave_append(count_due, hv);
    const int pl = wave_append(count_due + 3, is_due && !hv);
    const int pc = wave_append(count_carry, is_carry);
    if (hv) due[ph] = e;
    else if (is_due) due[due_cap - 1 - pl] = e;
    if (is_carry) {
      carry[pc] = e;
      lower_front(front_next, e.x, kb);
    }
  }
}

__global__ void split_kernel(const int2* __restrict__ list, const int* __restrict__ count, int ntx, int ntiles,
                             unsigned long long* tkey, const unsigned long long* __restrict__ front,
                             unsigned long long* front_next, double window, int color, unsigned int* in_set,
                             int2* due, int* count_due, int2* carry, int* count_carry,
                             const FieldDesc* __restrict__ descs, const uint8_t* __restrict__ heavy, int due_cap) {
  split_pass(list, *count, ntx, ntiles, tkey, front, front_next, window, color, in_set, due, count_due, carry,
             count_carry, descs, heavy, due_cap, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// Tile nodes of the box [a0,a1] x [b0,b1], clipped to the tile.
template <int TX, int TZ>
__device__ inline NodeMask<Shape<TX, TZ>::NPL> tile_box(int a0, int a1, int b0, int b1) {
  NodeMask<Shape<TX, TZ>::NPL> m;
  m.clear();
  a0 = max(a0, 0); a1 = min(a1, TX - 1); b0 = max(b0, 0); b1 = min(b1, TZ - 1);
  for (int b = b0; b <= b1; b++)
    for (int a = a0; a <= a1; a++) m.set(Shape<TX, TZ>::node_index(a, b));
  return m;
}

// Interface nodes (bit s*TX + column) of columns [a0,a1], both sides.
template <int TX>
__device__ inline unsigned tile_iface_cols(int a0, int a1) {
  a0 = max(a0, 0); a1 = min(a1, TX - 1);
  if (a0 > a1) return 0u;
  const unsigned m = (unsigned)(((1ull << (a1 - a0 + 1)) - 1ull) << a0);
  return m | (m << TX);
}

// 32-bit storage of T: the fixed-point grid.  A reached time beyond its
// range sets bit 2 of *flag (the forward then fails) and is kept as none.
__device__ inline unsigned t32_encode(double t) {
  if (!(t <= T32_MAX)) return T32_NONE;
  return (unsigned)llrint(t * T32_SCALE);
}
__device__ inline double t32_round(double t, int* flag) {
  if (!(t <= T32_MAX)) {
    if (t < FMT_HUGE_HALF && flag) atomicOr(flag, 2);
    return t;
  }
  return (double)llrint(t * T32_SCALE) * T32_UNIT;
}

// One visit of tile ft: stage the window, sweep to the tile's local fixed
// point, write back, activate the neighbours whose halo changed.
template <int TX, int TZ, bool F32>
__device__ __attribute__((always_inline)) inline void fim_tile(const FieldDesc* __restrict__ descs, const RegionDev* __restrict__ regs, int2 ft, int ntx,
                         int ntz, unsigned int* in_set, unsigned long long* tkey, int2* list_next, int* count_next,
                         unsigned long long* front_next, int max_sweeps, int colors, unsigned long long* acc,
                         int* launch_max, int* flag, TileLds<TX, TZ>& L) {
  using S = Shape<TX, TZ>;
  constexpr int NPL = S::NPL, WX = S::WX, WZ = S::WZ, W2 = S::W2;
  const int f = ft.x, t = ft.y;
  const FieldDesc& d = descs[f];
  const RegionDev& R = regs[d.region];
  const int nx = R.nx, nz = R.nz, lane = threadIdx.x;
  const int ntiles = ntx * ntz;
  const int tx0 = (t % ntx) * TX, tz0 = (t / ntx) * TZ;
  const int wx0 = tx0 - HALO, wz0 = tz0 - HALO;
  // ---- stage the window.  Every load of the window and of the interface
  // columns is issued into registers before the first LDS write: the field
  // pointers are loaded from memory, so the compiler cannot tell them from
  // LDS addresses and would otherwise wait for each load before the next
  // LDS store (one memory round trip per value).
  constexpr int NIT = (W2 + 63) / 64;
  double sT[NIT], sGX[NIT], sGZ[NIT], sV[NIT];
  uint16_t sF[NIT];
  uint8_t sTie[NIT];
  bool sOk[NIT];
  // addresses are clamped into the grid so that every lane loads without a
  // branch; values outside the grid are replaced afterwards
#pragma unroll
  for (int it = 0; it < NIT; it++) {
    const int k = lane + 64 * it;
    const int gi = wx0 + k % WX, gj = wz0 + k / WX;
    const bool ok = k < W2 && gi >= 0 && gi < nx && gj >= 0 && gj < nz;
    const int q = min(max(gj, 0), nz - 1) * nx + min(max(gi, 0), nx - 1);
    if constexpr (F32) {
      const unsigned tv = gload(d.T32 + q);
      const double gx = (double)gload(d.GX32 + q), gz = (double)gload(d.GZ32 + q);
      const bool fx = q == d.fixed_key;
      sT[it] = fx ? d.t0 : (tv == T32_NONE ? FMT_HUGE_TIME : (double)tv * T32_UNIT);
      sGX[it] = fx ? 0.0 : gx; sGZ[it] = fx ? 0.0 : gz;
    } else {
      sT[it] = gload(d.T + q); sGX[it] = gload(d.GX + q); sGZ[it] = gload(d.GZ + q);
    }
    sV[it] = gload(R.vel + q); sF[it] = gload(R.flags + q); sTie[it] = gload(d.TIE + q);
    sOk[it] = ok;
  }
  double cx;
  {
    const int gi = min(max(wx0 + lane, 0), nx - 1), gj = min(max(wz0 + lane - WX, 0), nz - 1);
    const double xv = gload(R.xg + gi), zv = gload(R.zg + gj);
    cx = lane < WX ? ((wx0 + lane >= 0 && wx0 + lane < nx) ? xv : 0.0)
                   : ((wz0 + lane - WX >= 0 && wz0 + lane - WX < nz) ? zv : 0.0);
  }
  double iT, iGX, iGZ, iV, ifx, ifz;
  int ifcc;
  uint8_t ifok;
  {
    const int s = lane / WX > 1 ? 1 : lane / WX, gr = wx0 + lane % WX;
    const int gi = min(max(gr, 0), nx - 1);
    const int q = (nz + s) * nx + gi;
    iT = gload(d.T + q); iGX = gload(d.GX + q); iGZ = gload(d.GZ + q);
    iV = gload(R.vel + q); ifx = gload(R.ifx[s] + gi); ifz = gload(R.ifz[s] + gi);
    ifcc = gload(R.ifcc[s] + gi); ifok = gload(R.ifok[s] + gi);
    if (!(lane < 2 * WX && gr >= 0 && gr < nx)) {
      iT = FMT_HUGE_TIME; iGX = 0.0; iGZ = 0.0; iV = 0.0; ifx = 0.0; ifz = 0.0; ifcc = -1; ifok = 0;
    }
  }
#pragma unroll
  for (int it = 0; it < NIT; it++) {
    const int k = lane + 64 * it;
    if (k < W2) {
      const bool ok = sOk[it];
      L.T[k] = ok ? sT[it] : FMT_HUGE_TIME; L.GX[k] = ok ? sGX[it] : 0.0; L.GZ[k] = ok ? sGZ[it] : 0.0;
      L.vel[k] = ok ? sV[it] : 0.0; L.flags[k] = ok ? sF[it] : (uint16_t)0; L.tie[k] = ok ? sTie[it] : (uint8_t)0;
    }
  }
  if (lane < WX) L.xg[lane] = cx;
  else if (lane < WX + WZ) L.zg[lane - WX] = cx;
  if (lane < 2 * WX) {
    L.iT[lane] = iT; L.iGX[lane] = iGX; L.iGZ[lane] = iGZ; L.ivel[lane] = iV;
    L.ifx[lane] = ifx; L.ifz[lane] = ifz; L.ifcc[lane] = ifcc; L.ifok[lane] = ifok;
  }
  __syncthreads();
  LView<TX, TZ> v{&R, &d, &L, wx0, wz0};
  // ---- my grid nodes (one per colour when a lane holds two) and (lanes
  // below 2 TX) my interface node
  int gi[NPL], gj[NPL], wg[NPL], g_phase[NPL];
  bool g_act[NPL], g_fast[NPL], g_dirty[NPL];
  // seed and fast-path model data of the node (one node per lane: registers)
  double g_t0 = FMT_HUGE_TIME, g_gx0 = 0.0, g_gz0 = 0.0;
  RegStatic g_rs;
  NodeMask<NPL> g_dep[NPL], g_pend[NPL];
  unsigned g_idep[NPL], g_ipend[NPL];
#pragma unroll
  for (int k = 0; k < NPL; k++) {
    const int a = S::node_a(lane, k), b = S::node_b(lane, k);
    gi[k] = tx0 + a; gj[k] = tz0 + b;
    wg[k] = (b + HALO) * WX + (a + HALO);
    g_act[k] = gi[k] < nx && gj[k] < nz && (L.flags[wg[k]] & RF_IN) && (gj[k] * nx + gi[k]) != d.fixed_key;
    g_phase[k] = colors == 1 ? 0 : (colors == 2 ? ((gi[k] + gj[k]) & 1) : ((gi[k] & 1) | ((gj[k] & 1) << 1)));
    double t0 = FMT_HUGE_TIME, gx0 = 0.0, gz0 = 0.0;
    if (g_act[k]) seed_grid(d, R, gi[k], gj[k], t0, gx0, gz0);
    // nodes with four regular cells and no coincident interface node take
    // the fast path (eik_regular.h); what it reads of the model is formed once
    g_fast[k] = g_act[k] && reg_fast_node(L.flags[wg[k]]);
    if constexpr (NPL == 1) {
      g_t0 = t0; g_gx0 = gx0; g_gz0 = gz0;
      if (g_fast[k]) g_rs = reg_static(v, gi[k], gj[k]);
    } else {
      const int n = lane + 64 * k;
      L.t0[n] = t0; L.gx0[n] = gx0; L.gz0[n] = gz0;
      if (g_fast[k]) L.rs[n] = reg_static(v, gi[k], gj[k]);
    }
    g_dirty[k] = false;
    // Nodes whose inputs did not change since their last evaluation give
    // the same result again and are skipped.  A grid node reads tile nodes
    // at most two rows/columns away and the interface nodes of its own and
    // adjacent columns; an interface node reads the grid nodes within two
    // columns and two rows of its cell (rows cz-2 .. cz+3 with its
    // coincident partner) and the interface nodes of its column.
    // Changed-node masks of every phase accumulate until a node is
    // evaluated again; the first sweep evaluates everything.
    g_dep[k] = tile_box<TX, TZ>(a - 2, a + 2, b - 2, b + 2);
    g_idep[k] = tile_iface_cols<TX>(a - 2, a + 2);
    g_pend[k].fill();
    g_ipend[k] = 0xFFFFFFFFu;
  }
  const int is = lane / TX, ic = tx0 + lane % TX;
  bool i_act = false;
  double i_t0 = FMT_HUGE_TIME, i_gx0 = 0.0, i_gz0 = 0.0;
  int i_cc = -1;
  int i_home = -1;  // row the interface node is solved in
  if (lane < 2 * TX && ic < nx - 1) {
    i_cc = L.ifcc[is * WX + ic - wx0];
    if (i_cc >= tz0 - 1 && i_cc < tz0 + TZ) i_home = iface_home_row(v, is, ic, i_cc);
    i_act = i_home >= tz0 && i_home < tz0 + TZ;
    if (i_act) seed_iface(d, is, ic, i_t0, i_gx0, i_gz0);
  }
  const int wi = is * WX + (ic - wx0);
  bool i_dirty = false;
  NodeMask<NPL> i_dep, i_pend;
  i_dep.clear();
  const int ia = lane % TX;
  if (i_act) i_dep = tile_box<TX, TZ>(ia - 2, ia + 2, i_cc - tz0 - 2, i_cc - tz0 + 3);
  const unsigned i_idep = tile_iface_cols<TX>(ia - 2, ia + 2);
  i_pend.fill();
  unsigned i_ipend = 0xFFFFFFFFu;
  constexpr unsigned IMASK = (unsigned)((1ull << (2 * TX)) - 1ull);
  bool more = true;
  // Sweeps.  colors == 4: Gauss-Seidel in five phases, the four parity
  // colours of the grid nodes (a node's x/z neighbours and cell diagonals
  // have other colours), then the interface nodes.  colors == 2: the two
  // checkerboard colours (x/z neighbours differ), then the interface nodes.
  // colors == 1: Jacobi, all grid nodes from the same state.  Within a
  // phase the single wave computes every active node before any writes, so
  // same-colour reads see the phase's starting state.
  const int nphase = colors + 1;
  int sweep = 0;
  for (; sweep < max_sweeps; sweep++) {
    bool ch = false;
    for (int phase = 0; phase < nphase; phase++) {
      const bool grid_phase = phase < nphase - 1;
      bool upd_any = false;
      if (grid_phase) {
        bool upd[NPL];
        double rT[NPL], rGX[NPL], rGZ[NPL];
        int rTie[NPL];
#pragma unroll
        for (int k = 0; k < NPL; k++) { upd[k] = false; rT[k] = 0.0; rGX[k] = 0.0; rGZ[k] = 0.0; rTie[k] = 0; }
        // the lane's nodes of this phase, one per round (two only for
        // Jacobi with two nodes per lane): the update code is emitted once
        const int rounds = (NPL > 1 && colors == 1) ? NPL : 1;
#pragma unroll 1
        for (int r = 0; r < rounds; r++) {
          int sel = -1, cnt = 0;
#pragma unroll
          for (int k = 0; k < NPL; k++)
            if (g_act[k] && g_phase[k] == phase) { if (cnt == r) sel = k; cnt++; }
          bool go = false, fast = false;
          int ci = 0, cj = 0, cw = 0;
#pragma unroll
          for (int k = 0; k < NPL; k++)
            if (k == sel) {
              go = g_pend[k].meets(g_dep[k]) || (g_ipend[k] & g_idep[k]);
              if (go) { g_pend[k].clear(); g_ipend[k] = 0u; }
              fast = g_fast[k]; ci = gi[k]; cj = gj[k]; cw = wg[k];
            }
          if (!go) continue;
          double t0, gx0, gz0;
          if constexpr (NPL == 1) { t0 = g_t0; gx0 = g_gx0; gz0 = g_gz0; }
          else { const int n = lane + 64 * sel; t0 = L.t0[n]; gx0 = L.gx0[n]; gz0 = L.gz0[n]; }
          Best<double> b;
          if (fast) {
            RegDyn y;
            for (int q = 0; q < 4; q++) {
              const int st = (q < 2 ? (q == 0 ? 1 : -1) : (q == 2 ? WX : -WX));
              const int wm = cw + st, ws = cw + 2 * st;  // member, second node (inside the window)
              y.mT[q] = L.T[wm]; y.mGX[q] = L.GX[wm]; y.mGZ[q] = L.GZ[wm]; y.mTie[q] = L.tie[wm];
              y.sT[q] = L.T[ws]; y.sGX[q] = L.GX[ws]; y.sGZ[q] = L.GZ[ws];
            }
            if constexpr (NPL == 1) reg_update(g_rs, y, t0, gx0, gz0, b.T, b.gx, b.gz, b.tie);
            else reg_update(L.rs[lane + 64 * sel], y, t0, gx0, gz0, b.T, b.gx, b.gz, b.tie);
          } else {
            b.T = t0; b.gx = gx0; b.gz = gz0; b.susp = 0; b.reg_set = 0; b.tie = 0; b.indep = 1;
            update_grid_node<double>(v, ci, cj, b);
          }
          if constexpr (F32) {
            // the value as stored: T on the fixed-point grid, FP32 gradient
            b.T = t32_round(b.T, flag);
            b.gx = (double)(float)b.gx;
            b.gz = (double)(float)b.gz;
          }
          const bool u = __double_as_longlong(b.T) != __double_as_longlong(L.T[cw]) ||
                         __double_as_longlong(b.gx) != __double_as_longlong(L.GX[cw]) ||
                         __double_as_longlong(b.gz) != __double_as_longlong(L.GZ[cw]) || b.tie != (int)L.tie[cw];
#pragma unroll
          for (int k = 0; k < NPL; k++)
            if (k == sel) { upd[k] = u; rT[k] = b.T; rGX[k] = b.gx; rGZ[k] = b.gz; rTie[k] = b.tie; }
        }
        if (colors == 1) __syncthreads();
        NodeMask<NPL> chg;
#pragma unroll
        for (int k = 0; k < NPL; k++) {
          if (upd[k]) {
            L.T[wg[k]] = rT[k]; L.GX[wg[k]] = rGX[k]; L.GZ[wg[k]] = rGZ[k]; L.tie[wg[k]] = (uint8_t)rTie[k];
            ch = true; g_dirty[k] = true;
          }
          chg.w[k] = __ballot(upd[k]);
        }
#pragma unroll
        for (int k = 0; k < NPL; k++) g_pend[k] |= chg;
        i_pend |= chg;
      } else {
        if (i_act && (i_pend.meets(i_dep) || (i_ipend & i_idep))) {
          i_pend.clear(); i_ipend = 0u;
          Best<double> b;
          b.T = i_t0; b.gx = i_gx0; b.gz = i_gz0; b.susp = 0; b.reg_set = 0; b.tie = 0; b.indep = 1;
          update_iface_node<double>(v, is, ic, b);
          if (__double_as_longlong(b.T) != __double_as_longlong(L.iT[wi]) ||
              __double_as_longlong(b.gx) != __double_as_longlong(L.iGX[wi]) ||
              __double_as_longlong(b.gz) != __double_as_longlong(L.iGZ[wi])) {
            L.iT[wi] = b.T; L.iGX[wi] = b.gx; L.iGZ[wi] = b.gz; ch = true; i_dirty = true; upd_any = true;
          }
        }
        const unsigned ichg = (unsigned)(__ballot(upd_any) & IMASK);
#pragma unroll
        for (int k = 0; k < NPL; k++) g_ipend[k] |= ichg;
        i_ipend |= ichg;
      }
      __syncthreads();
    }
    more = __syncthreads_or(ch);
    if (!more) break;
  }
  if (lane == 0 && acc) {
    // acc: [0] sweeps, [1] most sweeps of a visit, [2] visits that hit the
    // cap, [3 + k] visits of k + 1 sweeps (last bin: 64 or more)
    const int ns = min(sweep + 1, 64);
    atomicAdd(acc, (unsigned long long)(sweep + 1));
    atomicMax(acc + 1, (unsigned long long)(sweep + 1));
    if (more) atomicAdd(acc + 2, 1ull);
    atomicAdd(acc + 3 + ns - 1, 1ull);
    if (launch_max) atomicMax(launch_max, sweep + 1);
  }
  // ---- write back and activate neighbours whose halo changed
  auto edge = [&](int a, int b) {
    unsigned m = 0;
    if (a < 0) return m;
    const bool L0 = a < HALO, R0 = a >= TX - HALO, T0 = b < HALO, B0 = b >= TZ - HALO;
    if (L0) m |= 1; if (R0) m |= 2; if (T0) m |= 4; if (B0) m |= 8;
    if (L0 && T0) m |= 16; if (R0 && T0) m |= 32; if (L0 && B0) m |= 64; if (R0 && B0) m |= 128;
    return m;
  };
  unsigned mask = 0;
  // activation key: earliest time among this tile's changed nodes
  double kmin = 1.0e300;
#pragma unroll
  for (int k = 0; k < NPL; k++) {
    if (!g_dirty[k]) continue;
    const int q = gj[k] * nx + gi[k];
    if constexpr (F32) {
      d.T32[q] = t32_encode(L.T[wg[k]]);
      d.GX32[q] = (float)L.GX[wg[k]]; d.GZ32[q] = (float)L.GZ[wg[k]];
    } else {
      d.T[q] = L.T[wg[k]]; d.GX[q] = L.GX[wg[k]]; d.GZ[q] = L.GZ[wg[k]];
    }
    d.TIE[q] = L.tie[wg[k]];
    mask |= edge(gi[k] - tx0, gj[k] - tz0);
    kmin = fmin(kmin, L.T[wg[k]]);
  }
  if (i_dirty) {
    const int q = (nz + is) * nx + ic;
    d.T[q] = L.iT[wi]; d.GX[q] = L.iGX[wi]; d.GZ[q] = L.iGZ[wi];
    // border position of a dirty interface node: its home row
    mask |= edge(ic - tx0, i_home - tz0);
    kmin = fmin(kmin, L.iT[wi]);
  }
  if (more) mask |= 256;  // sweep cap reached: run this tile again
  for (int off = 32; off > 0; off >>= 1) {
    mask |= (unsigned)__shfl_xor((int)mask, off, 64);
    kmin = fmin(kmin, __shfl_xor(kmin, off, 64));
  }
  if (lane < 9 && (mask & (1u << lane))) {
    const int tix = t % ntx, tiz = t / ntx;
    // mask bit k -> tile offset: left, right, up, down, the four corners, self
    // (2-bit fields of offset + 1)
    const int ddx = (int)((0x18858u >> (2 * lane)) & 3u) - 1, ddz = (int)((0x1a085u >> (2 * lane)) & 3u) - 1;
    const int nxi = tix + ddx, nzi = tiz + ddz;
    if (nxi >= 0 && nxi < ntx && nzi >= 0 && nzi < ntz)
      add_tile(f, nzi * ntx + nxi, ntiles, key_of(kmin), in_set, tkey, list_next, count_next, front_next);
  }
}

// The due tiles of launch lno (nh cut by an interface at the front of the
// list, the rest from its end): blocks take them in turn.
template <int TX, int TZ, bool F32>
__device__ __attribute__((always_inline)) inline void tiles_pass(const FieldDesc* __restrict__ descs, const RegionDev* __restrict__ regs,
                                  const int2* list, int nh, int n, int ntx, int ntz, unsigned int* in_set,
                                  unsigned long long* tkey, int2* list_next, int* count_next,
                                  unsigned long long* front_next, int max_sweeps, int max_sweeps_heavy, int colors,
                                  unsigned long long* acc, int lno, int* sweep_max, int trace_cap,
                                  unsigned long long* span, int due_cap, int* flag, int* grab, TileLds<TX, TZ>& L) {
  // most sweeps of a visit in this launch, recorded under the launch number
  const bool slot = lno >= 0 && lno < trace_cap;
  int* launch_max = slot ? sweep_max + lno : nullptr;
  // the launch's span on the device's constant-rate clock: block 0 (the
  // first dispatched) stamps the start in slot 0; every block that visited
  // a tile, and block 0, stamps its end into one of SPAN_SLOTS slots
  // (atomicMax, spread over slots to keep the atomics uncontended)
  unsigned long long* sp = span + (size_t)lno * (SPAN_SLOTS + 1);
  if (slot && blockIdx.x == 0 && threadIdx.x == 0) sp[0] = wall_clock64();
  // tiles blockIdx.x, then (with grab, one-wave blocks) the next untaken
  // one from a shared counter, asked for once the visit is done (asking
  // before it would hold the next tile behind a slow visit), else every
  // gridDim.x-th.  A block stops when handed a number past the due tiles,
  // which comes only after every due tile was handed out.
  for (int k = blockIdx.x; k < n;) {
    const int2 ft = k < nh ? list[k] : list[due_cap - 1 - (k - nh)];
    // a tile cut by an interface gets its own sweep cap: its visits are the
    // slow ones, and a capped visit continues in a later launch
    fim_tile<TX, TZ, F32>(descs, regs, ft, ntx, ntz, in_set, tkey, list_next, count_next, front_next,
                          k < nh ? max_sweeps_heavy : max_sweeps, colors, acc, launch_max, flag, L);
    __syncthreads();
    if (grab) {
      int nk = 0;
      if (threadIdx.x == 0) nk = (int)gridDim.x + atomicAdd(grab, 1);
      k = __shfl(nk, 0, 64);
    } else
      k += gridDim.x;
  }
  if (slot && threadIdx.x == 0 && ((int)blockIdx.x < n || blockIdx.x == 0))
    atomicMax(sp + 1 + blockIdx.x % SPAN_SLOTS, (unsigned long long)wall_clock64());
}

// One launch's tile visits (the due count is known only on the device).
template <int TX, int TZ, bool F32>
__global__ void __launch_bounds__(64, (TX * TZ > 64) ? 2 : 1) fim_tiles_kernel(const FieldDesc* __restrict__ descs,
                                                       const RegionDev* __restrict__ regs, const int2* __restrict__ list,
                                                       const int* __restrict__ count_due, int ntx, int ntz,
                                                       unsigned int* in_set, unsigned long long* tkey,
                                                       int2* list_next, int* count_next,
                                                       unsigned long long* front_next, int max_sweeps,
                                                       int max_sweeps_heavy, int colors, unsigned long long* acc,
                                                       const int* launch_no, int* sweep_max, int trace_cap,
                                                       unsigned long long* span, int due_cap, int* flag) {
  __shared__ TileLds<TX, TZ> L;
  const int nh = count_due[0];
  tiles_pass<TX, TZ, F32>(descs, regs, list, nh, nh + count_due[3], ntx, ntz, in_set, tkey, list_next, count_next,
                          front_next, max_sweeps, max_sweeps_heavy, colors, acc, *launch_no - 1, sweep_max,
                          trace_cap, span, due_cap, flag, nullptr, L);
}

// Counters the level kernel and the host share (ints): [1] / [2] sizes of
// active lists 0 / 1, [4] launches run; due tiles of launch b in
// [8 (b & 1)] (cut by an interface) and [8 (b & 1) + 3] (the others); tiles
// handed out past the first gridDim.x in [12 + (b & 1)].
constexpr int LVL_LAUNCHES = 4, LVL_GRAB = 12;

// An agent-scope read of a counter other blocks of the grid write
__device__ inline int load_count(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid-wide barrier of the level kernel (one-wave blocks, all resident).
// Blocks arrive on a counter of their group of LVL_BAR_GROUP blocks (each
// counter on its own 256-byte line); the last of a group arrives on the top
// counter, which every block polls.  Counters only grow: barrier number
// `epoch` (from 1) is complete when the top counter reaches epoch x groups.
// A release fence before the arrival and an acquire fence after the wait
// order the blocks' memory around the barrier at agent scope; the last
// block of a group fences (acquire, release) between its two arrivals, so
// that the top arrival carries every block of the group.
constexpr int LVL_BAR_GROUP = 32, LVL_BAR_STRIDE = 64;
__device__ inline void level_barrier(unsigned* bar, unsigned epoch) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  if (threadIdx.x == 0) {
    const unsigned g = blockIdx.x / LVL_BAR_GROUP;
    const unsigned gsize = min((unsigned)LVL_BAR_GROUP, gridDim.x - g * LVL_BAR_GROUP);
    const unsigned ngroups = (gridDim.x + LVL_BAR_GROUP - 1) / LVL_BAR_GROUP;
    const unsigned old = __hip_atomic_fetch_add(bar + LVL_BAR_STRIDE * (1 + g), 1u, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
    if (old == epoch * gsize - 1) {
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
      __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch * ngroups)
      __builtin_amdgcn_s_sleep(1);
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// All launches of one dependency level in one dispatch.  Every block of the
// cooperative grid is resident; the phases of a launch are separated by
// grid-wide barriers.  Launch b: split active list b & 1 into the due tiles
// of colour b & 1 and the tiles carried into list (b & 1) ^ 1 | barrier |
// the tile visits, which append to list (b & 1) ^ 1; block 0 meanwhile
// clears what launch b + 1 fills (list b & 1, its front, the other due and
// hand-out counters) | barrier | stop when the next list is empty or
// max_launch launches have run.  Every block reads the same count after the
// barrier and takes the same number of barriers.
template <int TX, int TZ, bool F32>
__global__ void __launch_bounds__(64, 2) fim_level_kernel(LevelSched a) {
  __shared__ TileLds<TX, TZ> L;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  unsigned epoch = 0;
  for (int b = 0;; b++) {
    const int cur = b & 1, nxt = cur ^ 1;
    int* cdue = a.count + 8 * cur;
    split_pass(a.list[cur], load_count(a.count + 1 + cur), a.ntx, a.ntiles, a.tkey, a.front[cur], a.front[nxt],
               a.window, cur, a.in_set, a.due, cdue, a.list[nxt], a.count + 1 + nxt, a.descs, a.heavy, a.due_cap,
               tid, stride);
    level_barrier(a.bar, ++epoch);
    const int nh = load_count(cdue), n = nh + load_count(cdue + 3);
    if (tid == 0) {
      if (b < a.trace_cap) a.trace[b] = n;
      a.count[1 + cur] = 0;
      a.count[8 * nxt] = 0;
      a.count[8 * nxt + 3] = 0;
      a.count[LVL_GRAB + nxt] = 0;
    }
    for (int f = tid; f < a.nfield; f += stride) a.front[cur][f] = KEY_INF;
    tiles_pass<TX, TZ, F32>(a.descs, a.regs, a.due, nh, n, a.ntx, a.ntz, a.in_set, a.tkey, a.list[nxt],
                            a.count + 1 + nxt, a.front[nxt], a.max_sweeps, a.max_sweeps_heavy, a.colors, a.acc, b,
                            a.sweep_max, a.trace_cap, a.span, a.due_cap, a.flag, a.count + LVL_GRAB + cur, L);
    level_barrier(a.bar, ++epoch);
    if (load_count(a.count + 1 + nxt) == 0 || b + 1 >= a.max_launch) {
      if (tid == 0) a.count[LVL_LAUNCHES] = b + 1;
      break;
    }
  }
}

// Post-pass: refit one-sided gradients of grid nodes; everything else keeps
// the raw gradient.  One thread per (field, slot).
__global__ void refit_gradients_kernel(const FieldDesc* __restrict__ descs, const RegionDev* __restrict__ regs,
                                       int nfield, int nx, int nz, int* stop_flag) {
  const int nslot = (nz + 2) * nx;
  const long long total = (long long)nfield * nslot;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (long long)gridDim.x * blockDim.x) {
    const int f = (int)(k / nslot), q = (int)(k % nslot);
    const FieldDesc& d = descs[f];
    const RegionDev& R = regs[d.region];
    double ox = d.GX[q], oz = d.GZ[q];
    if (q < nz * nx) {
      const int i = q % nx, j = q / nx;
      // a time the update left as NaN: the reference stops there
      // ("negative determinant too large", propagate2d.f90:750-753)
      if (d.T[q] != d.T[q]) atomicOr(stop_flag, 1);
      if ((R.flags[q] & RF_IN) && q != d.fixed_key && d.T[q] < 0.5e20) {
        GView v{&R, &d};
        double t0, a, b;
        seed_grid(d, R, i, j, t0, a, b);
        bool susp;
        if (reg_fast_node(R.flags[q])) {
          // four regular cells, no partner: the fast path's walk gives the flag
          const RegStatic rs = reg_static(v, i, j);
          susp = reg_suspect(rs, reg_dyn(v, rs, i, j), t0);
        } else {
          susp = grid_suspect<double>(v, i, j, t0) != 0;
        }
        if (susp) {
          double gx, gz;
          fit_gradient_grid<double>(v, i, j, gx, gz);
          ox = gx; oz = gz;
        }
      }
    }
    d.OGX[q] = ox; d.OGZ[q] = oz;
  }
}

// Top-row gradient fix of water-layer stages: an up-going stage needs
// dT/dz < 0 at the sea surface, a down-going one dT/dz > 0.
__global__ void water_top_fix_kernel(const FieldDesc* __restrict__ descs, const RegionDev* __restrict__ regs,
                                     const int* __restrict__ which, const int* __restrict__ up, int nsel) {
  const RegionDev& R0 = regs[0];
  const int nx = R0.nx;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nsel * nx) return;
  const int f = which[k / nx], c = k % nx;
  const FieldDesc& d = descs[f];
  const RegionDev& R = regs[d.region];
  if (!(R.flags[c] & RF_IN) || !(R.flags[nx + c] & RF_IN)) return;
  const double g = d.OGZ[c];
  const bool bad = up[k / nx] ? (g > 0.0 || fabs(g) < 0.01) : (g < 0.0 || fabs(g) < 0.01);
  if (bad) { d.OGX[c] = d.OGX[nx + c]; d.OGZ[c] = d.OGZ[nx + c]; }
}

// PmP head-wave post-pass (one thread per field): min-plus sweeps of the
// Moho node times with the fastest velocity just below, then gradient
// refit of the alive Moho nodes.  The node ids come from the Moho's
// rright_node_id / rleft_node_id maps of the reference (region 4's top nodes
// first), evaluated against the PmP region's own node numbering exactly as
// the reference does.
__global__ void pmp_headwave_kernel(const FieldDesc* __restrict__ descs, const RegionDev* __restrict__ regs,
                                    const int* __restrict__ which, int nsel, const HeadwaveMap* __restrict__ hm,
                                    double* __restrict__ scratch, double* __restrict__ save) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nsel) return;
  const FieldDesc& d = descs[which[k]];
  const RegionDev& R = regs[d.region];
  const int nx = R.nx, n = hm->n;
  double* thw = scratch + (size_t)k * 2 * n;
  double* vhw = thw + n;
  for (int c = 0; c < n; c++) {
    const int q = hm->slot[c];
    thw[c] = q >= 0 ? d.T[q] : 1.0e20;
    vhw[c] = hm->vhw[c];
    // the solved times of the Moho nodes, before the pass moves them (the
    // adjoint sweep needs the field its updates were evaluated on)
    if (save) save[(size_t)k * n + c] = thw[c];
  }
  const double dxif = hm->dxif;
  for (int c = 1; c < n; c++) { const double v = vhw[c]; if (v > 0.0) { const double x = thw[c - 1] + dxif / v; if (x < thw[c]) thw[c] = x; } }
  for (int c = n - 2; c >= 0; c--) { const double v = vhw[c]; if (v > 0.0) { const double x = thw[c + 1] + dxif / v; if (x < thw[c]) thw[c] = x; } }
  for (int c = 0; c < n; c++) {
    const int q = hm->slot[c];
    if (q >= 0 && thw[c] < d.T[q] - 1.0e-10) d.T[q] = thw[c];
  }
  GView v{&R, &d};
  for (int c = 0; c < n; c++) {
    const int q = hm->slot[c];
    if (q < 0 || !(d.T[q] < 0.5e20)) continue;   // alive nodes only
    if (q < R.nz * nx) {
      double gx, gz;
      fit_gradient_grid<double>(v, q % nx, q / nx, gx, gz);
      d.OGX[q] = gx; d.OGZ[q] = gz;
    } else {
      double dtdx;
      if (c > 0 && c < n - 1) dtdx = (thw[c + 1] - thw[c - 1]) / (2.0 * dxif);
      else if (c == 0 && n > 1) dtdx = (thw[1] - thw[0]) / dxif;
      else if (c == n - 1 && n > 1) dtdx = (thw[n - 1] - thw[n - 2]) / dxif;
      else dtdx = 0.0;
      double dtdz = 0.0;
      const double vv = vhw[c];
      if (vv > 0.0) { const double s = 1.0 / (vv * vv) - dtdx * dtdx; dtdz = s > 0.0 ? sqrt(s) : 0.0; }
      d.OGX[q] = dtdx; d.OGZ[q] = dtdz;
    }
  }
}

// Per-field summary (diagnostics): reached slots, an order-free hash of the
// bits of T and the output gradient over all slots, and the eikonal residual
// | |grad T| v - 1 | over fully regular reached grid nodes whose four
// neighbours are reached.  One block per field.
__device__ inline unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) field_summary_kernel(const FieldDesc* __restrict__ descs,
                                                            const RegionDev* __restrict__ regs, int nx, int nz,
                                                            FieldSummary* out) {
  const int f = blockIdx.x;
  const FieldDesc& d = descs[f];
  const RegionDev& R = regs[d.region];
  const int nslot = (nz + 2) * nx;
  long long reached = 0, nres = 0, nbig = 0;
  unsigned long long h = 0;
  double rmax = 0.0, rsum = 0.0;
  for (int q = threadIdx.x; q < nslot; q += blockDim.x) {
    const double t = d.T[q];
    const unsigned long long bt = (unsigned long long)__double_as_longlong(t);
    const unsigned long long bx = (unsigned long long)__double_as_longlong(d.OGX[q]);
    const unsigned long long bz = (unsigned long long)__double_as_longlong(d.OGZ[q]);
    h += mix64(bt ^ mix64(bx ^ mix64(bz ^ (unsigned long long)q)));
    if (!(t < 0.5e20)) continue;
    reached++;
    if (q >= nz * nx) continue;
    const int i = q % nx, j = q / nx;
    if (!(R.flags[q] & RF_FULLREG) || i == 0 || j == 0 || i == nx - 1 || j == nz - 1) continue;
    if (!(d.T[q - 1] < 0.5e20 && d.T[q + 1] < 0.5e20 && d.T[q - nx] < 0.5e20 && d.T[q + nx] < 0.5e20)) continue;
    const double gx = d.OGX[q], gz = d.OGZ[q];
    const double r = fabs(sqrt(gx * gx + gz * gz) * R.vel[q] - 1.0);
    nres++;
    rsum += r;
    rmax = fmax(rmax, r);
    if (r > 0.05) nbig++;
  }
  __shared__ long long s_reached[256], s_nres[256], s_nbig[256];
  __shared__ unsigned long long s_h[256];
  __shared__ double s_rmax[256], s_rsum[256];
  const int k = threadIdx.x;
  s_reached[k] = reached; s_nres[k] = nres; s_nbig[k] = nbig; s_h[k] = h; s_rmax[k] = rmax; s_rsum[k] = rsum;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (k < w) {
      s_reached[k] += s_reached[k + w]; s_nres[k] += s_nres[k + w]; s_nbig[k] += s_nbig[k + w];
      s_h[k] += s_h[k + w]; s_rmax[k] = fmax(s_rmax[k], s_rmax[k + w]); s_rsum[k] += s_rsum[k + w];
    }
    __syncthreads();
  }
  if (k == 0) {
    FieldSummary o;
    o.reached = s_reached[0]; o.hash = s_h[0]; o.max_resid = s_rmax[0]; o.sum_resid = s_rsum[0];
    o.n_resid = s_nres[0]; o.n_resid_big = s_nbig[0];
    out[f] = o;
  }
}

// Largest time among each node's connected nodes (written into OGX, which the
// refit pass overwrites afterwards), the input of the heap-exhaustion rule.
__global__ void member_time_kernel(const FieldDesc* __restrict__ descs, const RegionDev* __restrict__ regs,
                                   int nfield, int nx, int nz) {
  const int nslot = (nz + 2) * nx;
  const long long total = (long long)nfield * nslot;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (long long)gridDim.x * blockDim.x) {
    const int f = (int)(k / nslot), q = (int)(k % nslot);
    const FieldDesc& d = descs[f];
    const RegionDev& R = regs[d.region];
    GView v{&R, &d};
    double mx = -1.0;
    if (q < nz * nx) {
      if (R.flags[q] & RF_IN) mx = max_member_time_grid<double>(v, q % nx, q / nx);
    } else {
      mx = max_member_time_iface<double>(v, q / nx - nz, q % nx);
    }
    d.OGX[q] = mx;
  }
}

// ---------------------------------------------------------------- heap-exhaustion stop
// Binned evaluation of the stop rule stated in heap_exit.h.  Every kernel
// below runs on a grid (ceil(nslot/256), nfield) so a block never straddles
// two fields, or one block per field for the bin scan.

// Node q of field d for the stop rule: false when q is no node of the
// region.  `seed` marks the nodes the reference puts in its heap before
// marching: the point-source cell corners and the diagonal neighbours of an
// on-node source (:2337-2383), or the start-interface nodes and the grid
// nodes 1-3 rows into the region that its sweep seeds (:2141-2222).
__device__ inline bool exit_node(const FieldDesc& d, const RegionDev& R, int q, bool& seed, bool& fixed) {
  const int nx = R.nx, nz = R.nz, i = q % nx, j = q / nx;
  seed = false;
  fixed = q == d.fixed_key;
  if (j < nz) {
    if (!(R.flags[q] & RF_IN)) return false;
    if (d.kind == SEED_POINT) {
      const int fi = d.fixed_key >= 0 ? d.fixed_key % nx : -10, fj = d.fixed_key >= 0 ? d.fixed_key / nx : -10;
      seed = (i >= d.ix0 && i <= d.ix0 + 1 && j >= d.iz0 && j <= d.iz0 + 1) || (abs(i - fi) == 1 && abs(j - fj) == 1);
    } else {
      const int* cc = R.ifcc[d.side];
      for (int n = i - 1; n <= i + 1 && !seed; n++) {
        if (n < 0 || n >= nx) continue;
        const int c = cc[n];
        if (c < 0) continue;
        seed = (d.side == 1) ? (j <= c && j >= c - 2) : (j >= c + 1 && j <= c + 3);
      }
    }
  } else {
    seed = d.kind == SEED_IFACE && (j - nz) == d.side;
  }
  return true;
}

__global__ void exit_init_kernel(ExitField* ef, int nfield, int* bcnt, int* bdiff, unsigned long long* bmax,
                                 unsigned long long* bm, int* total) {
  const long long nb = (long long)nfield * EXIT_NBIN;
  const unsigned long long kneg = ord_key(-INFINITY);
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nb; k += (long long)gridDim.x * blockDim.x) {
    bcnt[k] = 0; bdiff[k] = 0; bmax[k] = kneg; bm[k] = kneg;
    if (k < nfield) {
      ExitField e;
      e.lo = ~0ull; e.hi = 0ull; e.smax = kneg; e.texit = ord_key(INFINITY);
      e.count = 0; e.bad = 0; e.ncand = 0; e.pad = 0;
      ef[k] = e;
    }
    if (k == 0) *total = 0;
  }
}

__global__ void __launch_bounds__(256) exit_stats_kernel(const FieldDesc* __restrict__ descs,
                                                         const RegionDev* __restrict__ regs, int nslot, ExitField* ef) {
  __shared__ unsigned long long s_lo[256], s_hi[256], s_sm[256];
  __shared__ int s_cnt[256], s_bad[256];
  const int f = blockIdx.y, q = blockIdx.x * 256 + threadIdx.x;
  const FieldDesc& d = descs[f];
  unsigned long long lo = ~0ull, hi = 0ull, sm = ord_key(-INFINITY);
  int cnt = 0, bad = 0;
  bool seed, fixed;
  if (q < nslot && exit_node(d, regs[d.region], q, seed, fixed) && !fixed) {
    const double t = d.T[q];
    const bool listed = t < FMT_HUGE_HALF;
    if (seed) {
      if (!listed) bad = 1;
      else sm = ord_key(t);
    }
    if (listed) { lo = hi = ord_key(t); cnt = 1; }
  }
  const int x = threadIdx.x;
  s_lo[x] = lo; s_hi[x] = hi; s_sm[x] = sm; s_cnt[x] = cnt; s_bad[x] = bad;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (x < o) {
      s_lo[x] = min(s_lo[x], s_lo[x + o]);
      s_hi[x] = max(s_hi[x], s_hi[x + o]);
      s_sm[x] = max(s_sm[x], s_sm[x + o]);
      s_cnt[x] += s_cnt[x + o];
      s_bad[x] |= s_bad[x + o];
    }
    __syncthreads();
  }
  if (x == 0 && s_cnt[0] + s_bad[0] > 0) {
    atomicMin(&ef[f].lo, s_lo[0]);
    atomicMax(&ef[f].hi, s_hi[0]);
    atomicMax(&ef[f].smax, s_sm[0]);
    atomicAdd(&ef[f].count, s_cnt[0]);
    if (s_bad[0]) atomicOr(&ef[f].bad, 1);
  }
}

// per-bin node count, latest time, latest connection, and the difference
// array of the bins one interval (T_i, maxnb_i) covers completely
__global__ void __launch_bounds__(256) exit_bins_kernel(const FieldDesc* __restrict__ descs,
                                                        const RegionDev* __restrict__ regs, int nslot,
                                                        const ExitField* __restrict__ ef, int* bcnt, int* bdiff,
                                                        unsigned long long* bmax, unsigned long long* bm) {
  const int f = blockIdx.y, q = blockIdx.x * 256 + threadIdx.x;
  const ExitField e = ef[f];
  if (e.bad || e.count < 3 || q >= nslot) return;
  const FieldDesc& d = descs[f];
  bool seed, fixed;
  if (!exit_node(d, regs[d.region], q, seed, fixed) || fixed) return;
  const double t = d.T[q];
  if (!(t < FMT_HUGE_HALF)) return;
  const double m = d.OGX[q];
  const double lo = ord_val(e.lo), hi = ord_val(e.hi), sc = exit_scale(lo, hi);
  const int b = exit_bin(t, lo, sc);
  const size_t base = (size_t)f * EXIT_NBIN;
  atomicAdd(&bcnt[base + b], 1);
  atomicMax(&bmax[base + b], ord_key(t));
  atomicMax(&bm[base + b], ord_key(m));
  const int end = exit_cover_end(m, lo, hi, sc);
  if (end > b + 1) {
    atomicAdd(&bdiff[base + b + 1], 1);
    if (end < EXIT_NBIN) atomicAdd(&bdiff[base + end], -1);
  }
}

// One block per field: running cover count and exclusive prefix maximum of
// the per-bin latest connection; open bins get a slot range in the entry
// list (bdiff becomes the gather fill counter, bm the prefix maximum).
__global__ void __launch_bounds__(256) exit_scan_kernel(ExitField* ef, int* bcnt, int* bdiff,
                                                        const unsigned long long* __restrict__ bmax,
                                                        unsigned long long* bm, int* boff, int* total) {
  constexpr int PER = EXIT_NBIN / 256;
  __shared__ int s_a[256], s_c[256];
  __shared__ double s_m[256];
  __shared__ int s_base;
  const int f = blockIdx.x, x = threadIdx.x;
  const size_t base = (size_t)f * EXIT_NBIN + (size_t)x * PER;
  const ExitField e = ef[f];
  const bool skip = e.bad || e.count < 3;
  const double smax = ord_val(e.smax);
  // pass 1: per-thread sums of the difference array and max of connections
  int dsum = 0;
  double mmax = -INFINITY;
  for (int k = 0; k < PER; k++) {
    dsum += bdiff[base + k];
    mmax = fmax(mmax, ord_val(bm[base + k]));
  }
  s_a[x] = dsum; s_m[x] = mmax;
  __syncthreads();
  // inclusive scans over the 256 thread totals
  for (int o = 1; o < 256; o <<= 1) {
    const int a = x >= o ? s_a[x - o] : 0;
    const double m = x >= o ? s_m[x - o] : -INFINITY;
    __syncthreads();
    s_a[x] += a; s_m[x] = fmax(s_m[x], m);
    __syncthreads();
  }
  int cover = x > 0 ? s_a[x - 1] : 0;
  double run = x > 0 ? s_m[x - 1] : -INFINITY;
  __syncthreads();
  // pass 2: open bins and their node counts
  int ncand = 0;
  for (int k = 0; k < PER; k++) {
    cover += bdiff[base + k];
    const double bmk = ord_val(bm[base + k]);
    const bool open = !skip && exit_bin_open(bcnt[base + k], cover, run, ord_val(bmax[base + k]), smax);
    bm[base + k] = ord_key(run);
    bdiff[base + k] = 0;
    if (open) ncand += bcnt[base + k];
    else bcnt[base + k] = 0;  // closed bin: no slot range
    run = fmax(run, bmk);
  }
  s_c[x] = ncand;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const int a = x >= o ? s_c[x - o] : 0;
    __syncthreads();
    s_c[x] += a;
    __syncthreads();
  }
  if (x == 255) {
    s_base = s_c[255] > 0 ? atomicAdd(total, s_c[255]) : 0;
    ef[f].ncand = s_c[255];
  }
  __syncthreads();
  int off = s_base + (x > 0 ? s_c[x - 1] : 0);
  for (int k = 0; k < PER; k++) {
    const int c = bcnt[base + k];
    if (c > 0) { boff[base + k] = off; off += c; }
    else boff[base + k] = -1;
  }
}

__global__ void __launch_bounds__(256) exit_gather_kernel(const FieldDesc* __restrict__ descs,
                                                          const RegionDev* __restrict__ regs, int nslot,
                                                          const ExitField* __restrict__ ef,
                                                          const int* __restrict__ boff, int* bfill, ExitEntry* ent,
                                                          long long cap) {
  const int f = blockIdx.y, q = blockIdx.x * 256 + threadIdx.x;
  const ExitField e = ef[f];
  if (e.ncand == 0 || q >= nslot) return;
  const FieldDesc& d = descs[f];
  bool seed, fixed;
  if (!exit_node(d, regs[d.region], q, seed, fixed) || fixed) return;
  const double t = d.T[q];
  if (!(t < FMT_HUGE_HALF)) return;
  const double lo = ord_val(e.lo), hi = ord_val(e.hi), sc = exit_scale(lo, hi);
  const int b = exit_bin(t, lo, sc);
  const size_t bb = (size_t)f * EXIT_NBIN + b;
  const int o = boff[bb];
  if (o < 0) return;
  const long long p = (long long)o + atomicAdd(&bfill[bb], 1);
  if (p >= cap) return;
  ExitEntry en;
  en.t = t; en.m = d.OGX[q]; en.f = f; en.b = b;
  ent[p] = en;
}

// exact check of every node of an open bin against the nodes of its bin
__global__ void __launch_bounds__(256) exit_check_kernel(const ExitEntry* __restrict__ ent,
                                                         const int* __restrict__ total, long long cap,
                                                         const int* __restrict__ bcnt, const int* __restrict__ boff,
                                                         const unsigned long long* __restrict__ bm, ExitField* ef) {
  const long long n = min((long long)*total, cap);
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
    const ExitEntry c = ent[k];
    const ExitField e = ef[c.f];
    const double t = c.t;
    if (!(t >= ord_val(e.smax) && t > ord_val(e.lo) && t < ord_val(e.hi))) continue;
    const size_t bb = (size_t)c.f * EXIT_NBIN + c.b;
    const long long j0 = boff[bb], j1 = min(j0 + bcnt[bb], cap);
    double mx = ord_val(bm[bb]);
    bool dup = false;
    for (long long j = j0; j < j1 && !dup; j++) {
      if (j == k) continue;
      const double tj = ent[j].t;
      dup = tj == t;
      if (tj < t) mx = fmax(mx, ent[j].m);
    }
    if (!dup && mx <= t) atomicMin(&ef[c.f].texit, ord_key(t));
  }
}

// Nodes the reference never reaches once its heap ran empty (times above the
// stop time of their field) go back to huge_time with a zero gradient.
__global__ void truncate_kernel(const FieldDesc* __restrict__ descs, const ExitField* __restrict__ ef, int nfield,
                                int nslot) {
  const long long total = (long long)nfield * nslot;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (long long)gridDim.x * blockDim.x) {
    const int f = (int)(k / nslot), q = (int)(k % nslot);
    const double te = ord_val(ef[f].texit);
    if (te == INFINITY) continue;
    const FieldDesc& d = descs[f];
    const double t = d.T[q];
    if (t > te && t < FMT_HUGE_HALF) { d.T[q] = FMT_HUGE_TIME; d.GX[q] = 0.0; d.GZ[q] = 0.0; }
  }
}

// ---------------------------------------------------------------- launchers

hipError_t launch_member_time(const FieldDesc* descs, const RegionDev* regs, int nfield, int nx, int nz,
                              hipStream_t st) {
  const long long total = (long long)nfield * (nz + 2) * nx;
  int blocks = (int)std::min<long long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(member_time_kernel, dim3(blocks), dim3(256), 0, st, descs, regs, nfield, nx, nz);
  return hipGetLastError();
}

hipError_t launch_exit_bins(const FieldDesc* descs, const RegionDev* regs, int nfield, int nslot, ExitWork& w,
                            hipStream_t st) {
  const long long nb = (long long)nfield * EXIT_NBIN;
  hipLaunchKernelGGL(exit_init_kernel, dim3((unsigned)std::min<long long>((nb + 255) / 256, 65536)), dim3(256), 0, st,
                     w.ef, nfield, w.bcnt, w.bdiff, w.bmax, w.bm, w.total);
  const dim3 g((nslot + 255) / 256, nfield);
  hipLaunchKernelGGL(exit_stats_kernel, g, dim3(256), 0, st, descs, regs, nslot, w.ef);
  hipLaunchKernelGGL(exit_bins_kernel, g, dim3(256), 0, st, descs, regs, nslot, w.ef, w.bcnt, w.bdiff, w.bmax, w.bm);
  hipLaunchKernelGGL(exit_scan_kernel, dim3(nfield), dim3(256), 0, st, w.ef, w.bcnt, w.bdiff, w.bmax, w.bm, w.boff,
                     w.total);
  return hipGetLastError();
}

hipError_t launch_exit_check(const FieldDesc* descs, const RegionDev* regs, int nfield, int nslot, ExitWork& w,
                             long long ncand_bound, hipStream_t st) {
  const dim3 g((nslot + 255) / 256, nfield);
  hipLaunchKernelGGL(exit_gather_kernel, g, dim3(256), 0, st, descs, regs, nslot, w.ef, w.boff, w.bdiff, w.ent, w.cap);
  if (ncand_bound > 0) {
    const int blocks = (int)std::min<long long>((ncand_bound + 255) / 256, 16384);
    hipLaunchKernelGGL(exit_check_kernel, dim3(blocks), dim3(256), 0, st, w.ent, w.total, w.cap, w.bcnt, w.boff, w.bm,
                       w.ef);
  }
  const long long total = (long long)nfield * nslot;
  int blocks = (int)std::min<long long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(truncate_kernel, dim3(blocks), dim3(256), 0, st, descs, w.ef, nfield, nslot);
  return hipGetLastError();
}

hipError_t launch_fill(const FieldDesc* descs, int nfield, int nslot, hipStream_t st) {
  const long long total = (long long)nfield * nslot;
  int blocks = (int)std::min<long long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(fill_fields_kernel, dim3(blocks), dim3(256), 0, st, descs, nfield, nslot);
  return hipGetLastError();
}

hipError_t launch_transfer_reflect(const FieldDesc* descs_all, int f0, const StageLink* links_all, int nfield, int nx,
                                   int nz, const IsecDev* isecs, hipStream_t st) {
  const int total = nfield * nx;
  hipLaunchKernelGGL(transfer_reflect_kernel, dim3((total + 255) / 256), dim3(256), 0, st, descs_all, f0, links_all,
                     nfield, nx, nz, isecs);
  return hipGetLastError();
}

hipError_t launch_seed_tiles(const FieldDesc* descs, const RegionDev* regs, int nfield, int nx, int ntx, int ntz,
                             int tx, int tz, unsigned int* in_set, unsigned long long* tkey, int2* list_next,
                             int* count_next, unsigned long long* front_next, hipStream_t st) {
  const int total = nfield * nx;
  hipLaunchKernelGGL(seed_tiles_kernel, dim3((total + 255) / 256), dim3(256), 0, st, descs, regs, nfield, ntx, ntz,
                     tx, tz, in_set, tkey, list_next, count_next, front_next);
  return hipGetLastError();
}

hipError_t launch_schedule(const int2* list, const int* count, int nfield, int ntx, int ntiles,
                           unsigned long long* tkey, const unsigned long long* front, unsigned long long* front_next,
                           double window, int color, unsigned int* in_set, int2* due, int* count_due, int2* carry,
                           int* count_carry, int* launch_no, int* trace, int trace_cap, const FieldDesc* descs,
                           const uint8_t* heavy, int due_cap, int grid, hipStream_t st) {
  hipLaunchKernelGGL(launch_reset_kernel, dim3((nfield + 255) / 256), dim3(256), 0, st, count_due, count_carry,
                     front_next, nfield, launch_no, trace, trace_cap);
  hipLaunchKernelGGL(split_kernel, dim3(grid), dim3(256), 0, st, list, count, ntx, ntiles, tkey, front, front_next,
                     window, color, in_set, due, count_due, carry, count_carry, descs, heavy, due_cap);
  return hipGetLastError();
}

// Per field: the largest |T - S| over slots both reach and the slots only
// one of them reaches (S: an earlier forward's T of the same fields).  One
// block per field.
__global__ void __launch_bounds__(256) field_compare_kernel(const FieldDesc* __restrict__ descs,
                                                            const double* __restrict__ snap, long long nslot,
                                                            double tol, double* max_dt, long long* n_diff,
                                                            long long* n_over) {
  const int f = blockIdx.x;
  const double* T = descs[f].T;
  const double* S = snap + (long long)f * nslot;
  double m = 0.0;
  long long nd = 0, no = 0;
  for (long long q = threadIdx.x; q < nslot; q += blockDim.x) {
    const bool a = T[q] < 0.5e20, b = S[q] < 0.5e20;
    if (a != b) nd++;
    else if (a) {
      const double e = fabs(T[q] - S[q]);
      m = fmax(m, e);
      if (e > tol) no++;
    }
  }
  __shared__ double sm[256];
  __shared__ long long sn[256], so[256];
  sm[threadIdx.x] = m;
  sn[threadIdx.x] = nd;
  so[threadIdx.x] = no;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
      sn[threadIdx.x] += sn[threadIdx.x + o];
      so[threadIdx.x] += so[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { max_dt[f] = sm[0]; n_diff[f] = sn[0]; n_over[f] = so[0]; }
}

hipError_t launch_field_compare(const FieldDesc* descs, int nfield, const double* snap, long long nslot, double tol,
                                double* max_dt, long long* n_diff, long long* n_over, hipStream_t st) {
  if (nfield <= 0) return hipSuccess;
  hipLaunchKernelGGL(field_compare_kernel, dim3(nfield), dim3(256), 0, st, descs, snap, nslot, tol, max_dt, n_diff,
                     n_over);
  return hipGetLastError();
}

// 32-bit storage of a level: every grid node unreached, the gradients 0
__global__ void f32_reset_kernel(const FieldDesc* __restrict__ descs, int nfield, int nslotg) {
  const long long total = (long long)nfield * nslotg;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (long long)gridDim.x * blockDim.x) {
    const int f = (int)(k / nslotg), q = (int)(k % nslotg);
    const FieldDesc& d = descs[f];
    d.T32[q] = T32_NONE; d.GX32[q] = 0.0f; d.GZ32[q] = 0.0f;
  }
}

// The solved level in FP64 for everything after the solve (the on-node
// source keeps t0 and its zero gradient from the seeding).
__global__ void f32_export_kernel(const FieldDesc* __restrict__ descs, int nfield, int nslotg) {
  const long long total = (long long)nfield * nslotg;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (long long)gridDim.x * blockDim.x) {
    const int f = (int)(k / nslotg), q = (int)(k % nslotg);
    const FieldDesc& d = descs[f];
    if (q == d.fixed_key) continue;
    const unsigned tv = d.T32[q];
    d.T[q] = tv == T32_NONE ? FMT_HUGE_TIME : (double)tv * T32_UNIT;
    d.GX[q] = (double)d.GX32[q];
    d.GZ[q] = (double)d.GZ32[q];
  }
}

hipError_t launch_f32_reset(const FieldDesc* descs, int nfield, int nslotg, hipStream_t st) {
  const long long total = (long long)nfield * nslotg;
  const int blocks = (int)((total + 255) / 256 < 65536 ? (total + 255) / 256 : 65536);
  hipLaunchKernelGGL(f32_reset_kernel, dim3(blocks), dim3(256), 0, st, descs, nfield, nslotg);
  return hipGetLastError();
}

hipError_t launch_f32_export(const FieldDesc* descs, int nfield, int nslotg, hipStream_t st) {
  const long long total = (long long)nfield * nslotg;
  const int blocks = (int)((total + 255) / 256 < 65536 ? (total + 255) / 256 : 65536);
  hipLaunchKernelGGL(f32_export_kernel, dim3(blocks), dim3(256), 0, st, descs, nfield, nslotg);
  return hipGetLastError();
}

hipError_t launch_fim(const FieldDesc* descs, const RegionDev* regs, const int2* list, int grid, const int* count_due,
                      int ntx, int ntz, int tx, int tz, unsigned int* in_set, unsigned long long* tkey,
                      int2* list_next, int* count_next, unsigned long long* front_next, int max_sweeps,
                      int max_sweeps_heavy, int colors, unsigned long long* acc, const int* launch_no, int* sweep_max,
                      int trace_cap, unsigned long long* span, int due_cap, int f32, int* flag, hipStream_t st) {
#define FIM_ARGS                                                                                                     \
  descs, regs, list, count_due, ntx, ntz, in_set, tkey, list_next, count_next, front_next, max_sweeps,               \
      max_sweeps_heavy, colors, acc, launch_no, sweep_max, trace_cap, span, due_cap, flag
  if (f32) {
    if (tx != 8 || tz != 8) return hipErrorInvalidValue;
    hipLaunchKernelGGL((fim_tiles_kernel<8, 8, true>), dim3(grid), dim3(64), 0, st, FIM_ARGS);
  } else if (tx == 8 && tz == 8)
    hipLaunchKernelGGL((fim_tiles_kernel<8, 8, false>), dim3(grid), dim3(64), 0, st, FIM_ARGS);
  else if (tx == 16 && tz == 8)
    hipLaunchKernelGGL((fim_tiles_kernel<16, 8, false>), dim3(grid), dim3(64), 0, st, FIM_ARGS);
  else if (tx == 8 && tz == 16)
    hipLaunchKernelGGL((fim_tiles_kernel<8, 16, false>), dim3(grid), dim3(64), 0, st, FIM_ARGS);
  else
    return hipErrorInvalidValue;
#undef FIM_ARGS
  return hipGetLastError();
}

// the level kernel's instantiation for a tile shape and storage width
static const void* level_kernel(int tx, int tz, int f32) {
  if (f32) return tx == 8 && tz == 8 ? (const void*)fim_level_kernel<8, 8, true> : nullptr;
  if (tx == 8 && tz == 8) return (const void*)fim_level_kernel<8, 8, false>;
  if (tx == 16 && tz == 8) return (const void*)fim_level_kernel<16, 8, false>;
  if (tx == 8 && tz == 16) return (const void*)fim_level_kernel<8, 16, false>;
  return nullptr;
}

hipError_t level_grid(int device, int tx, int tz, int f32, int* grid) {
  const void* k = level_kernel(tx, tz, f32);
  if (!k) return hipErrorInvalidValue;
  int coop = 0, ncu = 0, per_cu = 0;
  hipError_t e = hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device);
  if (e != hipSuccess) return e;
  if (!coop) return hipErrorNotSupported;
  e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64, 0);
  if (e != hipSuccess) return e;
  *grid = per_cu * ncu;
  return *grid > 0 ? hipSuccess : hipErrorInvalidConfiguration;
}

size_t level_bar_words(int grid) {
  return (size_t)LVL_BAR_STRIDE * (1 + (grid + LVL_BAR_GROUP - 1) / LVL_BAR_GROUP);
}

hipError_t launch_fim_level(const LevelSched& a, int tx, int tz, int f32, int grid, hipStream_t st) {
  const void* k = level_kernel(tx, tz, f32);
  if (!k) return hipErrorInvalidValue;
  LevelSched arg = a;
  void* args[] = {&arg};
  return hipLaunchCooperativeKernel(k, dim3(grid), dim3(64), args, 0, st);
}

hipError_t launch_field_summary(const FieldDesc* descs, const RegionDev* regs, int nfield, int nx, int nz,
                                FieldSummary* out, hipStream_t st) {
  if (nfield <= 0) return hipSuccess;
  hipLaunchKernelGGL(field_summary_kernel, dim3(nfield), dim3(256), 0, st, descs, regs, nx, nz, out);
  return hipGetLastError();
}

hipError_t launch_refit(const FieldDesc* descs, const RegionDev* regs, int nfield, int nx, int nz, int* stop_flag,
                        hipStream_t st) {
  const long long total = (long long)nfield * (nz + 2) * nx;
  int blocks = (int)std::min<long long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(refit_gradients_kernel, dim3(blocks), dim3(256), 0, st, descs, regs, nfield, nx, nz, stop_flag);
  return hipGetLastError();
}

hipError_t launch_water_fix(const FieldDesc* descs, const RegionDev* regs, const int* which, const int* up, int nsel,
                            int nx, hipStream_t st) {
  const int total = nsel * nx;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(water_top_fix_kernel, dim3((total + 255) / 256), dim3(256), 0, st, descs, regs, which, up, nsel);
  return hipGetLastError();
}

hipError_t launch_headwave(const FieldDesc* descs, const RegionDev* regs, const int* which, int nsel,
                           const HeadwaveMap* hm, double* scratch, double* save, hipStream_t st) {
  if (nsel == 0) return hipSuccess;
  hipLaunchKernelGGL(pmp_headwave_kernel, dim3((nsel + 63) / 64), dim3(64), 0, st, descs, regs, which, nsel, hm,
                     scratch, save);
  return hipGetLastError();
}

// Exchanges the Moho-node times of the listed fields with `save` (twice:
// back where they were).
__global__ void headwave_swap_kernel(const FieldDesc* __restrict__ descs, const int* __restrict__ which, int nsel,
                                     const HeadwaveMap* __restrict__ hm, double* save) {
  const int n = hm->n;
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)nsel * n) return;
  const int f = which[k / n], c = (int)(k % n);
  const int q = hm->slot[c];
  if (q < 0) return;
  double* T = descs[f].T;
  const double t = T[q];
  T[q] = save[k];
  save[k] = t;
}

hipError_t launch_headwave_swap(const FieldDesc* descs, const int* which, int nsel, const HeadwaveMap* hm, int n,
                                double* save, hipStream_t st) {
  if (nsel == 0) return hipSuccess;
  const long long total = (long long)nsel * n;
  hipLaunchKernelGGL(headwave_swap_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, descs, which, nsel,
                     hm, save);
  return hipGetLastError();
}

}  // namespace fmt2d
