// Adjoint-state gradient of a receiver-time misfit: a reverse sweep over the
// frozen dependency order of the solved traveltime fields.
//
// Every reached node's time is one local update of the fixed point the
// forward solved (eik_update.h): the first candidate that reaches the node's
// final minimum (a regular_update2d of one or two members, a
// time_from_1_node / time_from_2_nodes of an irregular cell, the copy of a
// coincident partner's own minimum) or the node's seed.  That candidate is a
// smooth function of a handful of inputs: member times, region velocities,
// interface-node depths and, for seeded nodes, the previous stage's
// interface times.  Its partial derivatives are the edges of a directed
// acyclic graph from every node to the nodes it rests on (times only
// decrease along it).  The adjoint of the receiver times (d t_r / d T is
// the interpolation weight at the ray's entry point) is swept back over that
// graph in reverse topological order, stage level by stage level (last
// stages first: a stage's seeds hand their adjoint to the previous stage's
// interface nodes), and accumulated onto the region velocities and
// interface depths, which the host maps to the propagation-grid velocities
// (velocities_to_region2d / velocities_to_intersection2d are linear in
// them) and the interface node depths.
//
// Per node the winning candidate is found with the forward's own operator
// (same code, FP64, same order), then evaluated once more with sparse
// forward-mode derivatives (the value part reproduces the time bit for bit,
// which is checked); the reverse sweep runs in rounds: a node's adjoint is
// final when every node resting on it has passed its share (a counter per
// node), and the nodes that become final form the next round.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <chrono>
#include <mutex>
#include <string>
#include <vector>

#include "ctx.h"
#include "eik_field.h"

namespace fmt2d {

// ---------------------------------------------------------------- sparse derivatives

constexpr int SD_K = 8;  // non-zero partials a value carries (the winner needs at most 8)
enum { AK_T = 0, AK_TPREV = 1, AK_V = 2, AK_Z = 3 };
__device__ inline int acode(int kind, int idx) { return (kind << 28) | idx; }
__device__ inline int akind(int code) { return (code >> 28) & 3; }
__device__ inline int aidx(int code) { return code & 0x0FFFFFFF; }

__device__ unsigned int g_adj_overflow;  // partials dropped because a value held SD_K already

// A value and its partial derivatives with respect to coded inputs.  The
// value part is computed with exactly the operations of the plain code.
struct SD {
  double v;
  int n;
  int c[SD_K];
  double d[SD_K];
  __device__ SD() : v(0.0), n(0) {}
  __device__ SD(double x) : v(x), n(0) {}
  __device__ explicit operator double() const { return v; }
};

__device__ inline SD sd_input(double x, int code) {
  SD r(x);
  r.n = 1; r.c[0] = code; r.d[0] = 1.0;
  return r;
}

__device__ inline void sd_acc(SD& r, int code, double x) {
  if (x == 0.0) return;
  for (int k = 0; k < r.n; k++)
    if (r.c[k] == code) { r.d[k] += x; return; }
  if (r.n < SD_K) { r.c[r.n] = code; r.d[r.n] = x; r.n++; }
  else atomicAdd(&g_adj_overflow, 1u);
}

__device__ inline SD sd_scaled(const SD& a, double s) {
  SD r(0.0);
  for (int k = 0; k < a.n; k++) if (a.d[k] * s != 0.0) { r.c[r.n] = a.c[k]; r.d[r.n] = a.d[k] * s; r.n++; }
  return r;
}

__device__ inline SD operator+(const SD& a, const SD& b) {
  SD r = sd_scaled(a, 1.0);
  r.v = a.v + b.v;
  for (int k = 0; k < b.n; k++) sd_acc(r, b.c[k], b.d[k]);
  return r;
}
__device__ inline SD operator-(const SD& a, const SD& b) {
  SD r = sd_scaled(a, 1.0);
  r.v = a.v - b.v;
  for (int k = 0; k < b.n; k++) sd_acc(r, b.c[k], -b.d[k]);
  return r;
}
__device__ inline SD operator-(const SD& a) {
  SD r = sd_scaled(a, -1.0);
  r.v = -a.v;
  return r;
}
__device__ inline SD operator*(const SD& a, const SD& b) {
  SD r = sd_scaled(a, b.v);
  r.v = a.v * b.v;
  for (int k = 0; k < b.n; k++) sd_acc(r, b.c[k], b.d[k] * a.v);
  return r;
}
__device__ inline SD operator/(const SD& a, const SD& b) {
  const double inv = 1.0 / b.v;
  SD r = sd_scaled(a, inv);
  r.v = a.v / b.v;
  const double q = r.v * inv;
  for (int k = 0; k < b.n; k++) sd_acc(r, b.c[k], -b.d[k] * q);
  return r;
}
__device__ inline SD sqrt(const SD& a) {
  const double v = ::sqrt(a.v);
  SD r = sd_scaled(a, v > 0.0 ? 0.5 / v : 0.0);
  r.v = v;
  return r;
}
__device__ inline SD fabs(const SD& a) {
  SD r = sd_scaled(a, a.v < 0.0 ? -1.0 : 1.0);
  r.v = ::fabs(a.v);
  return r;
}
__device__ inline SD& operator+=(SD& a, const SD& b) { a = a + b; return a; }
__device__ inline SD& operator-=(SD& a, const SD& b) { a = a - b; return a; }
__device__ inline SD& operator*=(SD& a, const SD& b) { a = a * b; return a; }
__device__ inline SD& operator/=(SD& a, const SD& b) { a = a / b; return a; }
#define SD_CMP(op)                                                                       \
  __device__ inline bool operator op(const SD& a, const SD& b) { return a.v op b.v; }     \
  __device__ inline bool operator op(const SD& a, double b) { return a.v op b; }          \
  __device__ inline bool operator op(double a, const SD& b) { return a op b.v; }
SD_CMP(<)
SD_CMP(>)
SD_CMP(<=)
SD_CMP(>=)
SD_CMP(==)
SD_CMP(!=)
#undef SD_CMP

// ---------------------------------------------------------------- views

// The field as the operator sees it, with the inputs the adjoint tracks
// coded: times of this field's slots (AK_T), region velocities (AK_V), the
// depths of interface nodes (AK_Z, interface id * nx + column) and the
// previous stage's interface times a seed copies (AK_TPREV, column).
struct AdjView {
  const RegionDev* R;
  const FieldDesc* D;
  int zid[2];  // interface ids of the region's top and bottom
  __device__ const RegionDev& reg() const { return *R; }
  __device__ bool in(int i, int j) const {
    return i >= 0 && i < R->nx && j >= 0 && j < R->nz && (R->flags[j * R->nx + i] & RF_IN);
  }
  __device__ uint16_t flags(int i, int j) const { return R->flags[j * R->nx + i]; }
  __device__ SD T(int i, int j) const { const int q = j * R->nx + i; return sd_input(D->T[q], acode(AK_T, q)); }
  __device__ double GX(int i, int j) const { return D->GX[j * R->nx + i]; }
  __device__ double GZ(int i, int j) const { return D->GZ[j * R->nx + i]; }
  __device__ int tie(int i, int j) const { return D->TIE[j * R->nx + i]; }
  __device__ SD vel(int i, int j) const { const int q = j * R->nx + i; return sd_input(R->vel[q], acode(AK_V, q)); }
  __device__ double xg(int i) const { return R->xg[i]; }
  __device__ double zg(int j) const { return R->zg[j]; }
  __device__ SD iT(int s, int c) const {
    const int q = (R->nz + s) * R->nx + c;
    return sd_input(D->T[q], acode(AK_T, q));
  }
  __device__ double iGX(int s, int c) const { return D->GX[(R->nz + s) * R->nx + c]; }
  __device__ double iGZ(int s, int c) const { return D->GZ[(R->nz + s) * R->nx + c]; }
  __device__ SD ivel(int s, int c) const {
    const int q = (R->nz + s) * R->nx + c;
    return sd_input(R->vel[q], acode(AK_V, q));
  }
  __device__ double ifx(int s, int c) const { return R->ifx[s][c]; }
  __device__ SD ifz(int s, int c) const { return sd_input(R->ifz[s][c], acode(AK_Z, zid[s] * R->nx + c)); }
  __device__ int ifcc(int s, int c) const { return R->ifcc[s][c]; }
  __device__ int ifok(int s, int c) const { return R->ifok[s][c]; }
  // seeds (seed_grid / seed_iface with their inputs)
  __device__ SD T0(int i, int j) const {
    const FieldDesc& d = *D;
    if (!(R->flags[j * R->nx + i] & RF_IN)) return SD(FMT_HUGE_TIME);
    SD v = vel(i, j);
    if (v.v < 0.01) v = SD(0.01);
    if (d.kind == SEED_POINT) {
      if (i < d.ix0 || i > d.ix0 + 1 || j < d.iz0 || j > d.iz0 + 1) return SD(FMT_HUGE_TIME);
      const double dx = R->xg[i] - d.sx, dz = R->zg[j] - d.sz;
      const double dist = ::sqrt(dx * dx + dz * dz);
      if (dist < 0.01 * R->tol) return SD(d.t0);
      return SD(d.t0) + SD(dist) / v;
    }
    const int s = d.side;
    for (int n = i - 1; n <= i + 1; n++) {
      if (n < 0 || n >= R->nx) continue;
      const int cc = R->ifcc[s][n];
      if (cc < 0) continue;
      const bool hit = (s == 1) ? (j <= cc && j >= cc - 2) : (j >= cc + 1 && j <= cc + 3);
      if (!hit) continue;
      const double px = R->ifx[s][n];
      const SD pz = ifz(s, n);
      const double dx = R->xg[i] - px;
      const SD dz = SD(R->zg[j]) - pz;
      const SD dist = sqrt(SD(dx * dx) + dz * dz);
      return sd_input(d.isT[n], acode(AK_TPREV, n)) + dist / v;
    }
    return SD(FMT_HUGE_TIME);
  }
  __device__ double GX0(int i, int j) const { double t, a, b; seed_grid(*D, *R, i, j, t, a, b); return a; }
  __device__ double GZ0(int i, int j) const { double t, a, b; seed_grid(*D, *R, i, j, t, a, b); return b; }
  __device__ SD iT0(int s, int c) const {
    if (D->kind == SEED_IFACE && s == D->side) return sd_input(D->isT[c], acode(AK_TPREV, c));
    return SD(FMT_HUGE_TIME);
  }
  __device__ double iGX0(int s, int c) const { double t, a, b; seed_iface(*D, s, c, t, a, b); return a; }
  __device__ double iGZ0(int s, int c) const { double t, a, b; seed_iface(*D, s, c, t, a, b); return b; }
  __device__ bool fixed(const Member& m) const { return m.kind == 0 && (m.j * R->nx + m.i) == D->fixed_key; }
};

// ---------------------------------------------------------------- visitors

// What an offered update is (eik_update.h announces it with note()).
struct NoteRec {
  int kind;
  Member a, b;
};

// The forward's minimum (MinVisitor's rule), remembering which offered
// update first reached it and which of its adds (time_from_2_nodes may add
// two roots) did.
struct IndexVisitor {
  Best<double> b;
  NoteRec pending, cur, win;
  int nadd = 0, win_add = -1;
  bool have = false;
  template <class M>
  __device__ void note(int k, const M& a, const M& c) { pending.kind = k; pending.a = a; pending.b = c; }
  __device__ bool want(double evt) {
    const bool w = evt < b.T;
    if (w) { cur = pending; nadd = 0; }
    return w;
  }
  __device__ void add(double T, double, double, int, int, int) {
    if (T < b.T || T != T) { b.T = T; win = cur; win_add = nadd; have = true; }
    nadd++;
  }
};

// Keeps the target-th add of one update evaluated with partials.
struct AddSink {
  int target;
  int n;
  bool found;
  SD T;
  __device__ void add(const SD& t, const SD&, const SD&, int, int, int) {
    if (n == target) { T = t; found = true; }
    n++;
  }
};

// Position and velocity of the node an update is for: grid node (i, j) or
// interface node (s, i) of the region.
struct NodeRef {
  bool grid;
  int i, j, s;
};

// The update `w` (not a partner copy) for node X with partials.
__device__ bool eval_update(const AdjView& av, const NodeRef& X, const NoteRec& w, int win_add, SD& out) {
  if (w.kind == CN_REG1 || w.kind == CN_REG2) {
    int mx[2] = {w.a.i, w.b.i}, mz[2] = {w.a.j, w.b.j};
    SD T, gx, gz;
    int susp;
    if (!regular_cand<SD>(av, X.i, X.j, mx, mz, w.kind == CN_REG1 ? 1 : 2, T, gx, gz, susp)) return false;
    out = T;
    return true;
  }
  // position and velocity of the node (if/else, not ?: between values of
  // class type, which the device compiler has mis-selected)
  SD xu, zu, vu;
  if (X.grid) { xu = SD(av.xg(X.i)); zu = SD(av.zg(X.j)); vu = av.vel(X.i, X.j); }
  else { xu = SD(av.ifx(X.s, X.i)); zu = av.ifz(X.s, X.i); vu = av.ivel(X.s, X.i); }
  const SD tol = SD(av.reg().tol);
  SD xa, za, ta, gax, gaz, va;
  member_data<SD>(av, w.a, xa, za, ta, gax, gaz, va);
  AddSink sink;
  sink.target = win_add; sink.n = 0; sink.found = false;
  if (w.kind == CN_C1) {
    cand1<SD>(xu, zu, vu, xa, za, ta, gax, gaz, va, tol, sink);
  } else if (w.kind == CN_C2) {
    SD xc, zc, tc, gcx, gcz, vc;
    member_data<SD>(av, w.b, xc, zc, tc, gcx, gcz, vc);
    cand2<SD>(xu, zu, vu, xa, za, ta, gax, gaz, va, xc, zc, tc, gcx, gcz, vc, tol, sink);
  } else {
    return false;
  }
  out = sink.T;
  return sink.found;
}

// The winning update of node X with partials; a partner copy is the
// partner's own winning update (or its seed), found the way
// iface_own_min / grid_own_min find it.
__device__ bool eval_winner(const AdjView& av, const GView& gv, const NodeRef& X, const NoteRec& w, int win_add,
                            SD& out) {
  if (w.kind != CN_PARTNER) return eval_update(av, X, w, win_add, out);
  const Member& P = w.a;
  IndexVisitor pv;
  pv.b.susp = 0; pv.b.reg_set = 0; pv.b.tie = 0; pv.b.indep = 1;
  NodeRef Y;
  if (P.kind >= 1) {
    const int s = P.kind - 1, c = P.i;
    Y.grid = false; Y.i = c; Y.j = 0; Y.s = s;
    pv.b.T = gv.iT0(s, c); pv.b.gx = gv.iGX0(s, c); pv.b.gz = gv.iGZ0(s, c);
    Member ex; ex.kind = 0; ex.i = X.i; ex.j = X.j;
    own_iface_candidates<double>(gv, s, c, ex, pv);
    if (!pv.have) { out = av.iT0(s, c); return true; }
  } else {
    Y.grid = true; Y.i = P.i; Y.j = P.j; Y.s = 0;
    pv.b.T = gv.T0(P.i, P.j); pv.b.gx = gv.GX0(P.i, P.j); pv.b.gz = gv.GZ0(P.i, P.j);
    Member ex; ex.kind = 1 + X.s; ex.i = X.i; ex.j = 0;
    Member gm; gm.kind = 0; gm.i = P.i; gm.j = P.j;
    if (!gv.fixed(gm)) own_grid_candidates<double>(gv, P.i, P.j, ex, pv);
    if (!pv.have) { out = av.T0(P.i, P.j); return true; }
  }
  if (pv.win.kind == CN_PARTNER) return false;
  return eval_update(av, Y, pv.win, pv.win_add, out);
}

// ---------------------------------------------------------------- kernels

struct AdjRec {
  int* n;          // entries per slot
  int* code;       // [slot * SD_K + e]
  double* w;
};

// The edges of every reached node of the listed fields; counts the nodes
// resting on each node (cnt).  bad[0]: nodes whose winner did not
// reproduce the stored time (left as leaves: the head-wave post-pass
// changes tf5's Moho nodes after the solve).
__global__ void __launch_bounds__(256) adj_record_kernel(const FieldDesc* __restrict__ descs,
                                                         const RegionDev* __restrict__ regs,
                                                         const int* __restrict__ flist, int nf, long long nslot,
                                                         const int* __restrict__ zids, AdjRec rec, int* cnt,
                                                         unsigned long long* bad) {
  const long long total = (long long)nf * nslot;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (long long)gridDim.x * blockDim.x) {
    const int lf = (int)(k / nslot), q = (int)(k % nslot);
    rec.n[k] = 0;
    const FieldDesc& d = descs[flist[lf]];
    const RegionDev& R = regs[d.region];
    const int nx = R.nx, nz = R.nz;
    const double T = d.T[q];
    if (!(T < FMT_HUGE_HALF) || q == d.fixed_key) continue;
    const bool grid = q < nz * nx;
    const int i = q % nx, j = q / nx, s = j - nz;
    if (grid && !(R.flags[q] & RF_IN)) continue;
    GView gv{&R, &d};
    IndexVisitor iv;
    if (grid) seed_grid(d, R, i, j, iv.b.T, iv.b.gx, iv.b.gz);
    else seed_iface(d, s, i, iv.b.T, iv.b.gx, iv.b.gz);
    iv.b.susp = 0; iv.b.reg_set = 0; iv.b.tie = 0; iv.b.indep = 1;
    if (grid) grid_candidates<double>(gv, i, j, iv);
    else iface_candidates<double>(gv, s, i, iv);
    if (__double_as_longlong(iv.b.T) != __double_as_longlong(T)) { atomicAdd(bad, 1ull); continue; }
    AdjView av{&R, &d, {zids[2 * d.region], zids[2 * d.region + 1]}};
    NodeRef X;
    X.grid = grid; X.i = i; X.j = grid ? j : 0; X.s = grid ? 0 : s;
    SD t;
    bool ok = true;
    if (!iv.have) {
      if (grid) t = av.T0(i, j);
      else t = av.iT0(s, i);
    } else ok = eval_winner(av, gv, X, iv.win, iv.win_add, t);
    if (!ok) {
      atomicAdd(bad + 1, 1ull);
      continue;
    }
    if (__double_as_longlong(t.v) != __double_as_longlong(T)) { atomicAdd(bad + 2, 1ull); continue; }
    int ne = 0;
    for (int e = 0; e < t.n; e++) {
      if (t.d[e] == 0.0 || t.d[e] != t.d[e]) continue;
      rec.code[k * SD_K + ne] = t.c[e];
      rec.w[k * SD_K + ne] = t.d[e];
      ne++;
      if (akind(t.c[e]) == AK_T) atomicAdd(cnt + (long long)lf * nslot + aidx(t.c[e]), 1);
    }
    rec.n[k] = ne;
  }
}

__device__ inline int adj_wave_append(int* counter, bool pred) {
  const unsigned long long m = __ballot(pred);
  if (m == 0ull) return -1;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, __popcll(m));
  base = __shfl(base, leader, 64);
  return pred ? base + __popcll(m & ((1ull << lane) - 1ull)) : -1;
}

// First round: every reached node nobody rests on.
__global__ void adj_frontier_kernel(const FieldDesc* __restrict__ descs, const int* __restrict__ flist, int nf,
                                    long long nslot, const int* __restrict__ cnt, int* list, int* count) {
  const long long total = (long long)nf * nslot;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long rounds = (total + stride - 1) / stride;
  for (long long r = 0; r < rounds; r++) {
    const long long k = r * stride + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    bool go = false;
    if (k < total) {
      const int lf = (int)(k / nslot), q = (int)(k % nslot);
      const FieldDesc& d = descs[flist[lf]];
      go = d.T[q] < FMT_HUGE_HALF && cnt[k] == 0;
    }
    const int p = adj_wave_append(count, go);
    if (go) list[p] = (int)k;
  }
}

// One round: each listed node passes its adjoint along its edges; a node
// whose last dependant has passed joins the next round.
__global__ void adj_round_kernel(const FieldDesc* __restrict__ descs, const StageLink* __restrict__ links,
                                 const int* __restrict__ flist, long long nslot, int nx, int nz, AdjRec rec,
                                 int* cnt, double* lam, int* active, double* gvr, double* gz, const int* __restrict__ list,
                                 const int* __restrict__ count, int* next, int* count_next) {
  const int n = *count;
  const int stride = gridDim.x * blockDim.x;
  const int rounds = (n + stride - 1) / stride;
  for (int r = 0; r < rounds; r++) {
    const int i = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
    int cand[SD_K];
    int ncand = 0;
    if (i < n) {
      const long long k = list[i];
      const int lf = (int)(k / nslot);
      const int f = flist[lf];
      const FieldDesc& d = descs[f];
      const double L = lam[(long long)f * nslot + (k % nslot)];
      const int ne = rec.n[k];
      for (int e = 0; e < ne; e++) {
        const int c = rec.code[k * SD_K + e];
        const double w = rec.w[k * SD_K + e];
        const int kind = akind(c), idx = aidx(c);
        if (kind == AK_T) {
          if (L != 0.0) atomicAdd(lam + (long long)f * nslot + idx, L * w);
          const long long kk = (long long)lf * nslot + idx;
          if (atomicSub(cnt + kk, 1) == 1) cand[ncand++] = (int)kk;
        } else if (L != 0.0) {
          if (kind == AK_TPREV) {
            const StageLink& lk = links[f];
            atomicAdd(lam + (long long)lk.prev * nslot + (long long)(nz + lk.prev_side) * nx + idx, L * w);
            atomicOr(active + lk.prev, 1);
          } else if (kind == AK_V) {
            atomicAdd(gvr + (long long)d.region * nslot + idx, L * w);
          } else {
            atomicAdd(gz + idx, L * w);
          }
        }
      }
    }
    // every lane runs SD_K appends (wave-level ballots)
    for (int e = 0; e < SD_K; e++) {
      const bool go = e < ncand;
      const int p = adj_wave_append(count_next, go);
      if (go) next[p] = cand[e];
    }
  }
}

__global__ void adj_reset_kernel(int* count) { *count = 0; }

// The head-wave pass (pmp_headwave_kernel) set every Moho-node time of a
// PmP up-going field that it lowered to a min-plus chain along the Moho:
// T(c) = T(c0) + sum over the chain's nodes c' after c0 of dxif / vhw(c'),
// c0 the chain's origin (a node the pass left alone).  The adjoint such a
// node holds (from the stages seeded from it and from receivers) moves to
// the origin, and every vhw(c') on the way, a region velocity of the PmP
// region, collects its share; the node's solved value, which the field's
// own updates read, then starts the field's sweep with no adjoint.  One
// thread per field; `post` holds the times after the pass, the field the
// times before it (exchanged for the sweep).  bad counts nodes whose
// recomputed chain does not give the stored time.
__global__ void adj_headwave_kernel(const FieldDesc* __restrict__ descs, const int* __restrict__ which,
                                    const int* __restrict__ widx, int nsel, const HeadwaveMap* __restrict__ hm,
                                    const int* __restrict__ vslot, const double* __restrict__ post, long long nslot,
                                    double* thw_s, int* pred_s, double* lam, double* gvr,
                                    unsigned long long* bad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nsel) return;
  const int f = which[k];
  const FieldDesc& d = descs[f];
  const int n = hm->n;
  const double dxif = hm->dxif;
  const double* pst = post + (size_t)widx[k] * n;
  double* thw = thw_s + (size_t)k * n;
  int* pred = pred_s + (size_t)k * n;
  for (int c = 0; c < n; c++) {
    const int q = hm->slot[c];
    thw[c] = q >= 0 ? d.T[q] : 1.0e20;
    pred[c] = -1;
  }
  for (int c = 1; c < n; c++) {
    const double v = hm->vhw[c];
    if (v > 0.0) { const double x = thw[c - 1] + dxif / v; if (x < thw[c]) { thw[c] = x; pred[c] = c - 1; } }
  }
  for (int c = n - 2; c >= 0; c--) {
    const double v = hm->vhw[c];
    if (v > 0.0) { const double x = thw[c + 1] + dxif / v; if (x < thw[c]) { thw[c] = x; pred[c] = c + 1; } }
  }
  double* lf = lam + (long long)f * nslot;
  for (int c = 0; c < n; c++) {
    const int q = hm->slot[c];
    if (q < 0) continue;
    const bool moved = thw[c] < d.T[q] - 1.0e-10;
    if (moved != (pst[c] != d.T[q]) || (moved && __double_as_longlong(pst[c]) != __double_as_longlong(thw[c]))) {
      atomicAdd(bad, 1ull);
      continue;
    }
    if (!moved) continue;
    const double L = lf[q];
    lf[q] = 0.0;
    if (L == 0.0) continue;
    int e = c;
    while (pred[e] >= 0) {
      const double v = hm->vhw[e];
      if (vslot[e] >= 0) atomicAdd(gvr + (long long)d.region * nslot + vslot[e], -L * dxif / (v * v));
      e = pred[e];
    }
    if (hm->slot[e] >= 0) atomicAdd(lf + hm->slot[e], L);
  }
}

}  // namespace fmt2d

// ---------------------------------------------------------------- host side

using namespace fmt2d;

namespace {

template <typename T>
hipError_t zero_buf(DBuf<T>& b, size_t n, hipStream_t st) {
  if (b.n < n) {
    hipError_t e = b.alloc(n);
    if (e != hipSuccess) return e;
  }
  return hipMemsetAsync(b.p, 0, sizeof(T) * std::max<size_t>(n, 1), st);
}

// weights of velocities_to_intersection2d (model.cpp, propagate2d.f90:1698-1785):
// the interface node's velocity above (side 0) or below (side 1) as a
// bilinear combination of grid velocities, and its derivative with respect
// to the node's depth
void iface_vel_weights(const HostModel& m, const HostInterface& is, int c, int side, int* g, double* w,
                       double& dvdz) {
  const double vdx = m.dx, dz = m.dz;
  const double px = is.x[c], pz = is.z[c];
  int ivx = (int)floor((px - m.vx[0]) / vdx) + 1;
  if (ivx < 1) ivx = 1;
  if (ivx >= m.nx) ivx = m.nx - 1;
  double wx = (px - m.vx[ivx - 1]) / vdx;
  wx = std::min(1.0, std::max(0.0, wx));
  const double po = side == 0 ? pz - dz : pz + dz;
  int ivz;
  double wz, up = 0.0;
  bool inside = false;
  if (po < m.zg[0]) { ivz = 1; wz = 0.0; }
  else if (po >= m.zg[m.nz - 1]) { ivz = m.nz - 1; wz = 1.0; }
  else {
    for (ivz = 1; ivz <= m.nz - 1; ivz++)
      if (po >= m.zg[ivz - 1] && po < m.zg[ivz]) break;
    if (ivz >= m.nz) ivz = m.nz - 1;
    const double lo = po - m.zg[ivz - 1];
    up = m.zg[ivz] - m.zg[ivz - 1];
    wz = up > 0.0 ? lo / up : 0.0;
    inside = up > 0.0 && wz >= 0.0 && wz <= 1.0;
  }
  wz = std::min(1.0, std::max(0.0, wz));
  auto G = [&](int i, int j) { return (j - 1) * m.nx + (i - 1); };
  g[0] = G(ivx, ivz); w[0] = (1.0 - wx) * (1.0 - wz);
  g[1] = G(ivx + 1, ivz); w[1] = wx * (1.0 - wz);
  g[2] = G(ivx, ivz + 1); w[2] = (1.0 - wx) * wz;
  g[3] = G(ivx + 1, ivz + 1); w[3] = wx * wz;
  dvdz = 0.0;
  if (inside) {
    const auto& V = m.vgrid;
    dvdz = ((1.0 - wx) * (V[g[2]] - V[g[0]]) + wx * (V[g[3]] - V[g[1]])) / up;
  }
}

}  // namespace

extern "C" {

int fmt2d_set_adjoint_memory(fmt2d_ctx* ctx, long long bytes) {
  if (!ctx) return fail(FMT2D_EINVAL, "NULL context");
  if (bytes < 0) return fail(FMT2D_EINVAL, "memory cap must be >= 0");
  ctx->adj_cap_bytes = bytes > 0 ? (size_t)bytes : ((size_t)4 << 30);
  return FMT2D_OK;
}

int fmt2d_adjoint_gradient(fmt2d_ctx* ctx, const double* dJdt, double* grad_vel, double* grad_iface,
                           fmt2d_adjoint_stats* stats) {
  if (!ctx || !dJdt) return fail(FMT2D_EINVAL, "NULL argument");
  if (ctx->nray < 1) return fail(FMT2D_ESTATE, "no rays traced");
  if (ctx->fwd_program) return fail(FMT2D_ESTATE, "the adjoint needs a forward without FMT2D_FWD_PROGRAM");
  if (ctx->fwd_f32) return fail(FMT2D_ESTATE, "the adjoint needs a forward with FP64 field storage");
  HIPCHK(hipSetDevice(ctx->device));
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t st = ctx->stream;
  const HostModel& m = ctx->m;
  const int nx = m.nx, nz = m.nz, nray = ctx->nray, nfield = ctx->nfield;
  const long long nslot = (long long)(nz + 2) * nx;
  fmt2d_adjoint_stats S{};
  // adjoint fields of every resident field, region-velocity and depth sums
  DBuf<double> lam, gvr, gzs, dg;
  DBuf<int> active, fo, zids;
  HIPCHK(zero_buf(lam, (size_t)nfield * nslot, st));
  HIPCHK(zero_buf(gvr, (size_t)(NREG + 1) * nslot, st));
  HIPCHK(zero_buf(gzs, (size_t)(NIF + 1) * nx, st));
  HIPCHK(zero_buf(active, (size_t)nfield, st));
  HIPCHK(upload(fo, ctx->field_of, st));
  std::vector<int> zid(2 * (NREG + 1), 0);
  for (int r = 1; r <= NREG; r++) { zid[2 * r] = m.reg[r].itop; zid[2 * r + 1] = m.reg[r].ibot; }
  HIPCHK(upload(zids, zid, st));
  {
    std::vector<double> g(dJdt, dJdt + nray);
    HIPCHK(upload(dg, g, st));
  }
  HIPCHK(launch_adjoint_sources(ctx->rgeom.p, ctx->rfields.p, ctx->rjobs.p, ctx->rout.p, dg.p, nray, fo.p, nslot,
                                lam.p, active.p, st));
  unsigned int zero = 0;
  HIPCHK(hipMemcpyToSymbolAsync(HIP_SYMBOL(g_adj_overflow), &zero, sizeof(zero), 0, hipMemcpyHostToDevice, st));
  // field index ranges of the stage levels (fields are stored level by level)
  std::vector<int> lev_of(nfield, 0);
  for (int s = 0; s < ctx->nsrc; s++)
    for (int itf = 1; itf <= FMT2D_NTF; itf++) {
      int L = 0;
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < LEVEL_N[a]; b++)
          if (LEVELS[a][b] == itf) L = a;
      lev_of[ctx->field_of[s * FMT2D_NTF + itf - 1]] = L;
    }
  DBuf<int> rn, rcode, cnt, flist, lists[2], counts;
  DBuf<double> rw;
  DBuf<unsigned long long> bad;
  HIPCHK(zero_buf(bad, 4, st));
  // the PmP up-going fields' Moho nodes as solved (before the head-wave
  // pass moved them): the updates the sweep recomputes were evaluated on
  // them; put back on every return
  struct HwSwap {
    fmt2d_ctx* c;
    DBuf<int> which;
    bool on = false;
    hipError_t swap() {
      if (c->hw_fields.empty()) return hipSuccess;
      on = !on;
      return launch_headwave_swap(c->descs.p, which.p, (int)c->hw_fields.size(), c->hw_map.p, c->m.isec[4].n,
                                  c->hw_save.p, c->stream);
    }
    ~HwSwap() {
      if (on) { (void)swap(); (void)hipStreamSynchronize(c->stream); }
    }
  } hws{ctx};
  if (!ctx->hw_fields.empty()) {
    HIPCHK(upload(hws.which, ctx->hw_fields, st));
    HIPCHK(hws.swap());
  }
  HIPCHK(zero_buf(counts, 4, st));
  int* hc = ctx->h_count;
  double ms_rec = 0.0, ms_rev = 0.0;
  size_t peak = lam.n * 8 + gvr.n * 8;
  for (int L = 3; L >= 0; L--) {
    std::vector<int> act(nfield);
    HIPCHK(hipMemcpyAsync(act.data(), active.p, sizeof(int) * nfield, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<int> fl;
    for (int f = 0; f < nfield; f++)
      if (act[f] && lev_of[f] == L) fl.push_back(f);
    if (fl.empty()) continue;
    const int nf = (int)fl.size();
    // the head-wave chains of this level's PmP up-going fields first: their
    // moved Moho nodes hand what the later stages and receivers gave them on
    std::vector<int> hwf, hwi;
    for (size_t w = 0; w < ctx->hw_fields.size(); w++)
      if (lev_of[ctx->hw_fields[w]] == L && act[ctx->hw_fields[w]]) {
        hwf.push_back(ctx->hw_fields[w]);
        hwi.push_back((int)w);
      }
    if (!hwf.empty()) {
      const int nm = m.isec[4].n, nh = (int)hwf.size();
      DBuf<int> hwl, hwx, hpred;
      DBuf<double> hthw;
      HIPCHK(upload(hwl, hwf, st));
      HIPCHK(upload(hwx, hwi, st));
      HIPCHK(hpred.alloc((size_t)nh * nm));
      HIPCHK(hthw.alloc((size_t)nh * nm));
      hipLaunchKernelGGL(adj_headwave_kernel, dim3((nh + 63) / 64), dim3(64), 0, st, ctx->descs.p, hwl.p, hwx.p, nh,
                         ctx->hw_map.p, ctx->hw_vslot.p, ctx->hw_save.p, nslot, hthw.p, hpred.p, lam.p, gvr.p,
                         bad.p + 3);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(st));
    }
    // the level's fields in chunks whose edge records fit the memory cap
    // (fields of one level do not rest on each other)
    const size_t per_field = (size_t)nslot * (4 + SD_K * 12 + 4 + 8);
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)nf, ctx->adj_cap_bytes / per_field));
    for (int c0 = 0; c0 < nf; c0 += chunk) {
      const int nc = std::min(chunk, nf - c0);
      const size_t ns = (size_t)nc * nslot;
      const auto tc0 = std::chrono::steady_clock::now();
      HIPCHK(upload(flist, std::vector<int>(fl.begin() + c0, fl.begin() + c0 + nc), st));
      if (rn.n < ns) {
        HIPCHK(rn.alloc(ns));
        HIPCHK(rcode.alloc(ns * SD_K));
        HIPCHK(rw.alloc(ns * SD_K));
        HIPCHK(cnt.alloc(ns));
        HIPCHK(lists[0].alloc(ns));
        HIPCHK(lists[1].alloc(ns));
      }
      peak = std::max(peak, lam.n * 8 + gvr.n * 8 + ns * (4 + SD_K * 12 + 4 + 8));
      HIPCHK(hipMemsetAsync(cnt.p, 0, sizeof(int) * ns, st));
      AdjRec rec{rn.p, rcode.p, rw.p};
      const int grid = (int)std::min<long long>(((long long)ns + 255) / 256, 32768);
      hipLaunchKernelGGL(adj_record_kernel, dim3(grid), dim3(256), 0, st, ctx->descs.p, ctx->regs.p, flist.p, nc,
                         nslot, zids.p, rec, cnt.p, bad.p);
      HIPCHK(hipGetLastError());
      // list 0 / 1 with their sizes in counts[1] / counts[2]
      hipLaunchKernelGGL(adj_reset_kernel, dim3(1), dim3(1), 0, st, counts.p + 1);
      hipLaunchKernelGGL(adj_frontier_kernel, dim3(grid), dim3(256), 0, st, ctx->descs.p, flist.p, nc, nslot, cnt.p,
                         lists[0].p, counts.p + 1);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(st));
      const auto tc1 = std::chrono::steady_clock::now();
      ms_rec += std::chrono::duration<double, std::milli>(tc1 - tc0).count();
      // rounds, queued in batches; the host reads the list size once per batch
      int cur = 0, rounds = 0;
      const int max_rounds = 8 * (int)(nx + nz) + 20000;
      for (;;) {
        constexpr int BATCH = 16;
        for (int b = 0; b < BATCH; b++) {
          const int nxt = cur ^ 1;
          hipLaunchKernelGGL(adj_reset_kernel, dim3(1), dim3(1), 0, st, counts.p + 1 + nxt);
          hipLaunchKernelGGL(adj_round_kernel, dim3(1024), dim3(256), 0, st, ctx->descs.p, ctx->links.p, flist.p,
                             nslot, nx, nz, rec, cnt.p, lam.p, active.p, gvr.p, gzs.p, lists[cur].p,
                             counts.p + 1 + cur, lists[nxt].p, counts.p + 1 + nxt);
          cur = nxt;
          rounds++;
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(hc, counts.p + 1 + cur, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (hc[0] == 0) break;
        if (rounds > max_rounds) return fail(FMT2D_ENOCONV, "adjoint sweep did not drain within its round cap");
      }
      S.rounds += rounds;
      ms_rev += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tc1).count();
    }
    S.fields += nf;
  }
  if (hws.on) HIPCHK(hws.swap());
  unsigned long long hbad[4] = {0, 0, 0, 0};
  unsigned int nover = 0;
  HIPCHK(hipMemcpy(hbad, bad.p, sizeof(hbad), hipMemcpyDeviceToHost));
  const unsigned long long nbad = hbad[0] + hbad[1] + hbad[2] + hbad[3];
  if (getenv("FMT2D_ADJ_DEBUG"))
    fprintf(stderr, "adjoint: unmatched nodes: operator %llu, winner %llu, value %llu, head-wave %llu\n", hbad[0],
            hbad[1], hbad[2], hbad[3]);
  HIPCHK(hipMemcpyFromSymbol(&nover, HIP_SYMBOL(g_adj_overflow), sizeof(nover), 0, hipMemcpyDeviceToHost));
  // region velocities and interface depths -> grid velocities and depths
  std::vector<double> hv((size_t)(NREG + 1) * nslot), hz((size_t)(NIF + 1) * nx);
  HIPCHK(hipMemcpy(hv.data(), gvr.p, sizeof(double) * hv.size(), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(hz.data(), gzs.p, sizeof(double) * hz.size(), hipMemcpyDeviceToHost));
  std::vector<double> gv((size_t)nx * nz, 0.0);
  auto add_iface = [&](int k, int side, int c, double gval) {
    int g[4];
    double w[4], dvdz;
    iface_vel_weights(m, m.isec[k], c, side, g, w, dvdz);
    for (int a = 0; a < 4; a++) gv[g[a]] += gval * w[a];
    hz[(size_t)k * nx + c] += gval * dvdz;
  };
  for (int r = 1; r <= NREG; r++) {
    const HostRegion& R = m.reg[r];
    const double* gr = hv.data() + (size_t)r * nslot;
    auto in = [&](int i, int j) { return (R.flags[(size_t)j * nx + i] & RF_IN) != 0; };
    for (int j = 0; j < nz; j++)
      for (int i = 0; i < nx; i++) {
        const double a = gr[(size_t)j * nx + i];
        if (a == 0.0 || !in(i, j)) continue;
        // velocities_to_region2d (model.cpp): boundary nodes take the
        // interface's velocity of the region's side
        int k = 0, side = 0, c = i;
        if (j < nz - 1 && !in(i, j + 1)) { k = R.ibot; side = 0; }
        if (j > 0 && !in(i, j - 1)) { k = R.itop; side = 1; }
        if (k) {
          const HostInterface& is = m.isec[k];
          c = (int)lround((m.xg[i] - is.x[0]) / m.dx);
          c = std::min(std::max(c, 0), is.n - 1);
          add_iface(k, side, c, a);
        } else {
          gv[(size_t)j * nx + i] += a;
        }
      }
    for (int c = 0; c < nx; c++) {
      const double a = gr[(size_t)nz * nx + c], b = gr[(size_t)(nz + 1) * nx + c];
      if (a != 0.0) add_iface(R.itop, 1, c, a);     // vel_bot of the top interface
      if (b != 0.0) add_iface(R.ibot, 0, c, b);     // vel_top of the bottom interface
    }
  }
  if (grad_vel) std::copy(gv.begin(), gv.end(), grad_vel);
  if (grad_iface) std::copy(hz.begin() + (size_t)nx, hz.begin() + (size_t)(NIF + 1) * nx, grad_iface);
  S.leaves_unmatched = (long long)nbad;
  S.dropped_partials = (long long)nover;
  S.ms_record = ms_rec;
  S.ms_reverse = ms_rev;
  S.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  S.peak_bytes = (long long)peak;
  if (stats) *stats = S;
  return FMT2D_OK;
}

}  // extern "C"
