// LSQR on the augmented, smoothed and damped tomography system, on the
// device (run_lsqr, src/inversion2d.f90:6578-7167, smooth_type 1, 2, 3).
//
// Krylov iterations amplify rounding differences (reordering a single dot
// product moves the reference's model update by ~1%), so every reduction
// here keeps the reference's FP64 order: G*x one thread per row and G^T*u
// one thread per column, both front to back over ascending indices; the
// Laplacian adds its neighbours in the reference's order; the two norms of
// an iteration are front-to-back sums.  Element-wise work is parallel (it
// does not round differently), and the scalar recurrences run in one
// thread.  The host launches a fixed sequence per iteration and polls the
// stop flag every few iterations; kernels of iterations after the stop do
// nothing, so the iteration count is the reference's.  A context setting
// (fmt2d_set_lsqr_mode) switches the norms and sparse products to tree and
// wave reductions instead: faster, with rounding of their own.
#include "ordered_sum.h"
#include <hip/hip_runtime.h>
#include <limits.h>
#include <math.h>

#include <algorithm>
#include <vector>

#include "lsqr.h"

namespace fmt2d {

__device__ inline bool stopped(const LsqrScalars* s) { return s != nullptr && s->done != 0; }

// One wave per row: y[i] = sum_k M(i,k) x(idx k), ascending k.  The lanes
// form the products of 64 entries at a time (coalesced loads, gathers of x)
// into LDS; lane 0 adds them front to back, so the sum rounds as the
// reference's sequential loop does.  The next chunk's loads are issued
// before lane 0's additions and complete behind them.
__global__ void __launch_bounds__(64) ordered_spmv_kernel(const int* __restrict__ ptr, const int* __restrict__ idx,
                                                         const double* __restrict__ val, int nrow,
                                                         const double* __restrict__ x, double* y,
                                                         const LsqrScalars* s) {
  __shared__ double buf[64];
  const int i = blockIdx.x;
  if (i >= nrow || stopped(s)) return;
  const int lane = threadIdx.x;
  const int a = ptr[i], b = ptr[i + 1];
  double acc = 0.0;
  double p = 0.0;
  if (a + lane < b) p = val[a + lane] * x[idx[a + lane]];
  for (int c = a; c < b; c += 64) {
    buf[lane] = p;
    __syncthreads();
    const int k = c + 64 + lane;
    double q = 0.0;
    if (k < b) q = val[k] * x[idx[k]];
    if (lane == 0) {
      const int m = min(64, b - c);
      for (int t = 0; t < m; t++) acc = acc + buf[t];
    }
    __syncthreads();
    p = q;
  }
  if (lane == 0) y[i] = acc;
}

// G x by rows (sparse_Ax, :2997-3018) and G^T u by columns of the stable
// column copy (sparse_ATx, :3021-3042: each column's rows ascending, the
// order in which the reference's scatter accumulates) use ordered_spmv_kernel.

// row j of a Gaussian smoothing matrix times x, entries in stored order
// (apply_gauss_smooth_h / _v / apply_iface_gauss_smooth_h, :1405-1521)
__device__ inline double row_dot(const int* __restrict__ rp, const int* __restrict__ col,
                                 const double* __restrict__ val, int j, const double* __restrict__ x) {
  double a = 0.0;
  for (int k = rp[j]; k < rp[j + 1]; k++) a = a + val[k] * x[col[k]];
  return a;
}

// apply_laplacian_smooth_2d (:1161-1249), then the mask and weight of the
// smoothing block: out = masked ? 0 : w_j * (L vin)_j.  smooth_type 3: the
// interface rows are w_if * (R_if vin)_j instead (0 without w_if).
__global__ void smooth_kernel(const LsqrSys S, const double* __restrict__ vin, double* out, const LsqrScalars* s) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= S.nparm || stopped(s)) return;
  if (S.stype == 3 && j >= S.n_vel) {
    double r = 0.0;
    if (S.rif_rp) r = row_dot(S.rif_rp, S.rif_col, S.rif_val, j, vin);
    out[j] = S.w_smooth_iface > 0.0 ? S.w_smooth_iface * r : 0.0;
    return;
  }
  double acc = 0.0;
  if (j < S.n_vel) {
    const int4 nb = S.lap_nb[j];
    int deg = 0;
    if (nb.x >= 0) { acc = acc + vin[nb.x]; deg++; }
    if (nb.y >= 0) { acc = acc + vin[nb.y]; deg++; }
    if (nb.z >= 0) { acc = acc + vin[nb.z]; deg++; }
    if (nb.w >= 0) { acc = acc + vin[nb.w]; deg++; }
    acc = acc - (double)deg * vin[j];
  } else {
    const int2 e = S.if_nb[j - S.n_vel];  // (kind, neighbour): kind 1 one-sided, 2 central, 0 none
    if (e.x == 1) acc = acc + vin[e.y] - vin[j];
    else if (e.x == 2) {
      acc = acc + vin[j - 1];
      acc = acc + vin[j + 1];
      acc = acc - 2.0 * vin[j];
    }
  }
  out[j] = S.mask[j] ? 0.0 : S.ws[j] * acc;
}

// smooth_type 2 blocks of A x: H = w R_h x, V = w R_v x (masked columns 0),
// R_h block's interface rows w_if R_if x when R_if is in use, else 0.  With
// rhs, the jumping right-hand side -H, -V written into u (the reference
// leaves -w*0 in the interface rows when R_if is not in use).
__global__ void gauss_fwd_kernel(const LsqrSys S, const double* __restrict__ x, double* H, double* V, int rhs,
                                 const LsqrScalars* s) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= S.nparm || stopped(s)) return;
  const bool masked = S.mask[j] != 0;
  double h = 0.0, v = 0.0;
  if (S.w_smooth > 0.0) {
    double a = row_dot(S.rh_rp, S.rh_col, S.rh_val, j, x);
    double b = row_dot(S.rv_rp, S.rv_col, S.rv_val, j, x);
    if (masked) { a = 0.0; b = 0.0; }
    h = S.w_smooth * a;
    v = S.w_smooth * b;
    if (rhs) { h = -h; v = -v; }
  }
  if (j >= S.n_vel) {
    if (S.gsi) {
      const double r = S.w_smooth_iface * row_dot(S.rif_rp, S.rif_col, S.rif_val, j, x);
      h = rhs ? -r : r;
    } else if (!rhs) {
      h = 0.0;
    }
  }
  H[j] = h;
  V[j] = v;
}

// front-to-back sum of a[i]^2 into *dst, the same bits as the loop
// s = s + a[i] * a[i], in the bulk form of ordered_sum.h.  The first 512
// squares (where the sum changes binade every few terms) are added one by
// one by lane 0 from LDS.  Then one block walks the vector in chunks of
// 4096 squares (8 per thread, the next chunk's loads in flight meanwhile):
// each thread forms its terms' mantissa increments at the running sum's
// exponent, a block scan of the thread totals finds the first term that
// changes the sum's binade, ties or is not finite; the terms before it are
// added as one integer, that term the ordinary way, and the chunk goes on
// after it.  Past the head a sum of squares changes binade a few times per
// doubling of the term count, so nearly every chunk is one pass.
constexpr int SEQ_THREADS = 1024;
constexpr int SEQ_PER = 4;
constexpr int SEQ_CHUNK = SEQ_THREADS * SEQ_PER;
constexpr int SEQ_WAVES = SEQ_THREADS / 64;
constexpr int SEQ_HEAD = 512;

// n >= 1; loads past the end read a[n - 1] and are replaced by 0 (no
// branches around the loads, so that later waits count them exactly)
__device__ __forceinline__ void seq_load(const double* __restrict__ a, long long n, long long b0,
                                         double (&r)[SEQ_PER]) {
#pragma unroll
  for (int i = 0; i < SEQ_PER; i++) r[i] = __builtin_nontemporal_load(a + min(b0 + i, n - 1));
#pragma unroll
  for (int i = 0; i < SEQ_PER; i++) r[i] = b0 + i < n ? r[i] : 0.0;
}

struct SeqShared {
  double wtot[2][SEQ_WAVES];  // by pass parity: one barrier per pass
  int wbad[2][SEQ_WAVES];
  int wbrk[SEQ_WAVES];
  double acc;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = v + __shfl_xor(v, o, 64);
  return v;
}

// One chunk [c0, c0 + SEQ_CHUNK) of squares, this thread's in r (squared
// here; +0 past the end).  False when the sum has become inf or NaN: then
// thread 0 has added the rest one by one and written *dst.  Positions are
// chunk offsets: this thread's terms are j0 .. j0 + SEQ_PER - 1, terms
// before lo are summed already, terms from lim on are past the end.
__device__ __forceinline__ bool seq_chunk(const double* __restrict__ a, long long n, long long c0,
                                          const double (&r)[SEQ_PER], double& acc, double* dst, SeqShared& sh,
                                          int& parity) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int j0 = t * SEQ_PER;
  const int lim = (int)min(n - c0, (long long)SEQ_CHUNK);
  double x[SEQ_PER];
#pragma unroll
  for (int i = 0; i < SEQ_PER; i++) x[i] = r[i] * r[i];
  int lo = 0;
  for (;;) {
    if (!isfinite(acc)) {  // inf or NaN stays so: the rest one by one
      if (t == 0) {
        for (long long k = c0 + lo; k < n; k++) {
          const double v = a[k];
          acc = acc + v * v;
        }
        *dst = acc;
      }
      return false;
    }
    const SumBase B = sum_base(acc);
    const double room = sum_room(B);
    double d[SEQ_PER];
    unsigned ok = 0;
#pragma unroll
    for (int i = 0; i < SEQ_PER; i++) {
      d[i] = 0.0;
      if (j0 + i < lo || sum_incr(x[i], B.e, &d[i])) ok |= 1u << i;
    }
    // the thread's total up to its first term that must go the ordinary way
    double tot = 0.0;
#pragma unroll
    for (int i = 0; i < SEQ_PER; i++) {
      if (!((ok >> i) & 1)) break;
      tot = tot + d[i];
    }
    const bool bad = ok != (1u << SEQ_PER) - 1;
    const double wt = wave_sum(tot);
    const int wb = __any(bad) ? 1 : 0;
    if (lane == 0) {
      sh.wtot[parity][w] = wt;
      sh.wbad[parity][w] = wb;
    }
    __syncthreads();
    double total = 0.0;
    int anybad = 0;
    for (int v = 0; v < SEQ_WAVES; v++) {
      total = total + sh.wtot[parity][v];
      anybad |= sh.wbad[parity][v];
    }
    if (!anybad && total < room) {  // every term in bulk (total exact below 2^53)
      acc = sum_value(B.m + (uint64_t)total, B.e);
      parity ^= 1;
      return true;
    }
    // locate the first term that breaks: exclusive prefix of the totals
    double inc = tot;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const double y = __shfl_up(inc, o, 64);
      if (lane >= o) inc = inc + y;
    }
    double P = __shfl_up(inc, 1, 64);
    if (lane == 0) P = 0.0;
    for (int v = 0; v < w; v++) P = P + sh.wtot[parity][v];
    int mine = INT_MAX;
    double before = 0.0, xb = 0.0;
#pragma unroll
    for (int i = 0; i < SEQ_PER; i++) {
      if (mine != INT_MAX || j0 + i < lo || j0 + i >= lim) continue;
      const double Q = P + d[i];
      if (!((ok >> i) & 1) || Q >= room) {
        mine = j0 + i;
        before = P;
        xb = x[i];
      } else {
        P = Q;
      }
    }
    int wm = mine;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) wm = min(wm, __shfl_xor(wm, o, 64));
    if (lane == 0) sh.wbrk[w] = wm;
    __syncthreads();
    int kb = sh.wbrk[0];
    for (int v = 1; v < SEQ_WAVES; v++) kb = min(kb, sh.wbrk[v]);
    if (mine == kb) sh.acc = sum_value(B.m + (uint64_t)before, B.e) + xb;
    __syncthreads();
    acc = sh.acc;
    parity ^= 1;
    lo = kb + 1;
    if (lo >= lim) return true;
  }
}

__global__ void __launch_bounds__(SEQ_THREADS) seqnorm_kernel(const double* __restrict__ a, long long n, double* dst,
                                                             const LsqrScalars* s) {
  if (stopped(s)) return;
  __shared__ double head[SEQ_HEAD];
  __shared__ SeqShared sh;
  const int t = threadIdx.x;
  const long long nh = min(n, (long long)SEQ_HEAD);
  for (int k = t; k < nh; k += SEQ_THREADS) {
    const double v = a[k];
    head[k] = v * v;
  }
  // the next chunk's loads are in flight while the current one is summed;
  // they land in this thread's own LDS slots (strided by the block size: no
  // bank conflicts) after the chunk (no values
  // carried in registers from one iteration to the next, whose loads the
  // compiler would wait for at the top of the loop)
  __shared__ double stage[SEQ_CHUNK];
  const int j0 = t * SEQ_PER;
  if (nh < n) {
    double r[SEQ_PER];
    seq_load(a, n, nh + j0, r);
#pragma unroll
    for (int i = 0; i < SEQ_PER; i++) stage[i * SEQ_THREADS + t] = r[i];
  }
  __syncthreads();
  if (t == 0) {
    double h = 0.0;
    for (int k = 0; k < (int)nh; k++) h = h + head[k];
    sh.acc = h;
  }
  __syncthreads();
  double acc = sh.acc;
  int parity = 0;
  for (long long c0 = nh; c0 < n; c0 += SEQ_CHUNK) {
    double nxt[SEQ_PER], r[SEQ_PER];
    seq_load(a, n, c0 + SEQ_CHUNK + j0, nxt);
#pragma unroll
    for (int i = 0; i < SEQ_PER; i++) r[i] = stage[i * SEQ_THREADS + t];
    if (!seq_chunk(a, n, c0, r, acc, dst, sh, parity)) return;
#pragma unroll
    for (int i = 0; i < SEQ_PER; i++) stage[i * SEQ_THREADS + t] = nxt[i];
  }
  if (t == 0) *dst = acc;
}

hipError_t launch_ordered_sumsq(const double* a, long long n, double* dst, hipStream_t st) {
  hipLaunchKernelGGL(seqnorm_kernel, dim3(1), dim3(SEQ_THREADS), 0, st, a, n, dst, nullptr);
  return hipGetLastError();
}

// Parallel-reduction mode: the norms as trees (blocks of 256, then the
// block sums) and the sparse products with a wave reduction per row.  Every
// partial carries its rounding error (two-sum and the fma residue of each
// product: double-double accumulation), so the sums are close to exact and
// differ from the ordered ones by about the ordered sums' own rounding.
struct DD {
  double hi, lo;
};
__device__ inline DD dd_add(DD a, double b) {
  const double s = a.hi + b, bb = s - a.hi;
  const double e = (a.hi - (s - bb)) + (b - bb);
  return DD{s, a.lo + e};
}
__device__ inline DD dd_add(DD a, DD b) {
  DD r = dd_add(a, b.hi);
  r.lo += b.lo;
  return r;
}
__device__ inline DD dd_fma(DD acc, double x, double y) {
  const double p = x * y;
  const double e = fma(x, y, -p);
  DD r = dd_add(acc, p);
  r.lo += e;
  return r;
}
__device__ inline DD dd_shfl_xor(DD a, int o) { return DD{__shfl_xor(a.hi, o, 64), __shfl_xor(a.lo, o, 64)}; }

__global__ void __launch_bounds__(256) tree_norm_part_kernel(const double* __restrict__ a, long long n,
                                                            double* part, const LsqrScalars* s) {
  if (stopped(s)) return;
  __shared__ double rh[256], rl[256];
  DD acc{0.0, 0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double x = a[i];
    acc = dd_fma(acc, x, x);
  }
  rh[threadIdx.x] = acc.hi;
  rl[threadIdx.x] = acc.lo;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      const DD r = dd_add(DD{rh[threadIdx.x], rl[threadIdx.x]}, DD{rh[threadIdx.x + o], rl[threadIdx.x + o]});
      rh[threadIdx.x] = r.hi;
      rl[threadIdx.x] = r.lo;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = rh[0]; part[2 * blockIdx.x + 1] = rl[0]; }
}

__global__ void __launch_bounds__(256) tree_norm_final_kernel(const double* __restrict__ part, int np, double* dst,
                                                             const LsqrScalars* s) {
  if (stopped(s)) return;
  __shared__ double rh[256], rl[256];
  DD acc{0.0, 0.0};
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc = dd_add(acc, DD{part[2 * i], part[2 * i + 1]});
  rh[threadIdx.x] = acc.hi;
  rl[threadIdx.x] = acc.lo;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      const DD r = dd_add(DD{rh[threadIdx.x], rl[threadIdx.x]}, DD{rh[threadIdx.x + o], rl[threadIdx.x + o]});
      rh[threadIdx.x] = r.hi;
      rl[threadIdx.x] = r.lo;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *dst = rh[0] + rl[0];
}

// y[i] = sum_k M(i,k) x(idx k): one wave per row, lanes stride the row,
// then a butterfly over the wave
__global__ void __launch_bounds__(64) wave_spmv_kernel(const int* __restrict__ ptr, const int* __restrict__ idx,
                                                      const double* __restrict__ val, int nrow,
                                                      const double* __restrict__ x, double* y, const LsqrScalars* s) {
  const int i = blockIdx.x;
  if (i >= nrow || stopped(s)) return;
  const int a = ptr[i], b = ptr[i + 1];
  DD acc{0.0, 0.0};
  for (int k = a + (int)threadIdx.x; k < b; k += 64) acc = dd_fma(acc, val[k], x[idx[k]]);
  for (int o = 32; o > 0; o >>= 1) acc = dd_add(acc, dd_shfl_xor(acc, o));
  if (threadIdx.x == 0) y[i] = acc.hi + acc.lo;
}

// ---- element-wise steps

// blocks of the augmented system: data, one (Laplacian / hybrid) or two
// (Gaussian) smoothing blocks, damping
__device__ inline int nblk(const LsqrSys& S) { return S.stype == 2 ? 2 : 1; }

// initial right-hand side: u = [r; -smooth(dm_sum); masked ? 0 : -wd dm_sum]
// (smooth_type 2: the smoothing blocks were written by gauss_fwd_kernel)
__global__ void rhs_kernel(const LsqrSys S, const double* __restrict__ sm, double* u) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nd = S.ndata, np = S.nparm, nb = nblk(S);
  if (i >= nd + (nb + 1) * np) return;
  if (i < nd) u[i] = S.res[i];
  else if (i < nd + nb * np) { if (nb == 1) u[i] = S.smooth_on ? -sm[i - nd] : 0.0; }
  else {
    const int j = (int)(i - nd - nb * np);
    u[i] = (S.damp_on && !S.mask[j]) ? -S.wd[j] * S.dm_sum[j] : 0.0;
  }
}

// u = (what == 0) u / beta ; (what == 1) u := Av - alpha u, Av assembled from
// G v (gv), the smoothing block (sm) and the damping block
__global__ void u_update_kernel(const LsqrSys S, const double* __restrict__ gv, const double* __restrict__ sm,
                                const double* __restrict__ sm2, const double* __restrict__ v, double* u,
                                const LsqrScalars* s) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nd = S.ndata, np = S.nparm, nb = nblk(S);
  if (i >= nd + (nb + 1) * np || stopped(s)) return;
  double av;
  if (i < nd) av = gv[i];
  else if (i < nd + np) av = (nb == 2 || S.smooth_on) ? sm[i - nd] : 0.0;
  else if (i < nd + nb * np) av = sm2[i - nd - np];
  else {
    const int j = (int)(i - nd - nb * np);
    av = (S.damp_on && !S.mask[j]) ? S.wd[j] * v[j] : 0.0;
  }
  u[i] = av - s->alpha * u[i];
}

__global__ void scale_kernel(double* a, long long n, const double* __restrict__ by, int only_positive,
                             const LsqrScalars* s) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || stopped(s)) return;
  const double b = *by;
  if (only_positive && !(b > 0.0)) return;
  a[i] = a[i] / b;
}

// column j of a Gaussian matrix (its transpose's row), rows ascending
// (apply_gauss_smooth_h_T / _v_T scatter in row order)
__device__ inline double col_dot(const int* __restrict__ cp, const int* __restrict__ row,
                                 const double* __restrict__ val, int j, const double* __restrict__ x) {
  double a = 0.0;
  for (int k = cp[j]; k < cp[j + 1]; k++) a = a + val[k] * x[row[k]];
  return a;
}

// v = (G^T u + smoothing blocks^T) + damping(u_D) - beta_factor * v, masked
// to 0.  smooth_type 2: + w R_h^T u_H + w R_v^T u_V, then w_if R_if u_H on
// the interface rows (R_if is applied untransposed, as the reference does).
__global__ void v_update_kernel(const LsqrSys S, const double* __restrict__ gtu, const double* __restrict__ sm,
                                const double* __restrict__ u, double* v, int first, const LsqrScalars* s) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= S.nparm || stopped(s)) return;
  const long long nd = S.ndata, np = S.nparm;
  double a = gtu[j];
  if (S.stype == 2) {
    const double* uh = u + nd;
    if (S.w_smooth > 0.0) {
      double h = col_dot(S.rh_cp, S.rh_crow, S.rh_cval, j, uh);
      double w = col_dot(S.rv_cp, S.rv_crow, S.rv_cval, j, u + nd + np);
      if (S.mask[j]) { h = 0.0; w = 0.0; }
      a = a + S.w_smooth * h;
      a = a + S.w_smooth * w;
    }
    if (S.gsi && j >= S.n_vel) a = a + S.w_smooth_iface * row_dot(S.rif_rp, S.rif_col, S.rif_val, j, uh);
  } else if (S.smooth_on) {
    a = a + sm[j];
  }
  if (S.damp_on && !S.mask[j]) a = a + S.wd[j] * u[nd + (long long)nblk(S) * np + j];
  if (!first) a = a - s->beta * v[j];
  v[j] = S.mask[j] ? 0.0 : a;
}

// v is masked before G v and L v at the top of every iteration
__global__ void mask_kernel(const LsqrSys S, double* v, const LsqrScalars* s) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= S.nparm || stopped(s)) return;
  if (S.mask[j]) v[j] = 0.0;
}

__global__ void w_init_kernel(const LsqrSys S, const double* __restrict__ v, double* w, double* dm) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= S.nparm) return;
  w[j] = v[j];
  dm[j] = 0.0;
}

// z = masked w; w = masked (v - theta/rho w); dm = dm + phi/rho z
__global__ void w_update_kernel(const LsqrSys S, const double* __restrict__ v, double* w, double* dm,
                                const LsqrScalars* s) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= S.nparm || stopped(s)) return;
  const double z = S.mask[j] ? 0.0 : w[j];
  const double wn = v[j] - (s->theta / s->rho) * w[j];
  w[j] = S.mask[j] ? 0.0 : wn;
  dm[j] = dm[j] + (s->phi / s->rho) * z;
}

// ---- scalar steps (one thread)

// step 0: beta = bnorm = sqrt(norm)   step 1: alpha = sqrt(norm) (initial)
// step 2: beta = sqrt(norm)           step 3: alpha = sqrt(norm) + recurrences
__global__ void scalar_kernel(LsqrScalars* s, int step, int iter, int max_iter, double atol) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || s->done) return;
  if (step == 0) {
    s->bnorm = sqrt(s->norm);
    s->beta = s->bnorm;
    if (s->bnorm < 1.0e-30) { s->done = 1; s->iters = 0; s->rnorm = 0.0; s->test2 = 0.0; s->trivial = 1; }
  } else if (step == 1) {
    s->alpha = sqrt(s->norm);
    if (s->alpha < 1.0e-30) { s->done = 1; s->iters = 0; s->rnorm = s->bnorm; s->test2 = 0.0; s->trivial = 1; return; }
    s->phibar = s->beta;
    s->rhobar = s->alpha;
    s->anorm2 = s->alpha * s->alpha + s->beta * s->beta;
  } else if (step == 2) {
    s->beta = sqrt(s->norm);
  } else {
    s->alpha = sqrt(s->norm);
    s->anorm2 = s->anorm2 + s->alpha * s->alpha + s->beta * s->beta;
    s->rho = sqrt(s->rhobar * s->rhobar + s->beta * s->beta);
    const double c = s->rhobar / s->rho, sn = s->beta / s->rho;
    s->theta = sn * s->alpha;
    s->rhobar = -c * s->alpha;
    s->phi = c * s->phibar;
    s->phibar = sn * s->phibar;
    s->c = c;
  }
}

// after the w/dm update: convergence test of iteration `iter`
__global__ void test_kernel(LsqrScalars* s, int iter, int max_iter, double atol) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || s->done) return;
  s->rnorm = fabs(s->phibar);
  const double arnorm = s->phibar * s->alpha * fabs(s->c);
  s->test2 = arnorm / (sqrt(s->anorm2) * s->rnorm + 1.0e-30);
  s->iters = iter;
  if (s->test2 < atol || 1.0 + s->test2 <= 1.0 || iter >= max_iter) s->done = 1;
}

// ---------------------------------------------------------------- driver

static inline unsigned nb(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// sum of squares of a[0..n) into *dst, in the mode of the system
static void launch_norm(const LsqrSys& S, LsqrWork& W, const double* a, long long n, double* dst, hipStream_t st) {
  if (S.fast) {
    const int g = (int)std::min<long long>(LSQR_NPART, std::max<long long>(1, (n + 255) / 256));
    hipLaunchKernelGGL(tree_norm_part_kernel, dim3(g), dim3(256), 0, st, a, n, W.part, W.sc);
    hipLaunchKernelGGL(tree_norm_final_kernel, dim3(1), dim3(256), 0, st, W.part, g, dst, W.sc);
  } else {
    hipLaunchKernelGGL(seqnorm_kernel, dim3(1), dim3(SEQ_THREADS), 0, st, a, n, dst, W.sc);
  }
}

static void launch_spmv(const LsqrSys& S, const int* ptr, const int* idx, const double* val, int nrow,
                        const double* x, double* y, const LsqrScalars* sc, hipStream_t st) {
  if (S.fast)
    hipLaunchKernelGGL(wave_spmv_kernel, dim3((unsigned)nrow), dim3(64), 0, st, ptr, idx, val, nrow, x, y, sc);
  else
    hipLaunchKernelGGL(ordered_spmv_kernel, dim3((unsigned)nrow), dim3(64), 0, st, ptr, idx, val, nrow, x, y, sc);
}

// The kernels of the start of a solve: u = [r; -w L dm_sum; -wd dm_sum]
// (Gaussian: [r; -w R_h dm_sum; -w R_v dm_sum; -wd dm_sum]), beta, u /= beta,
// v = A^T u masked, alpha, v /= alpha, w = v.
static hipError_t launch_start(const LsqrSys& S, LsqrWork& W, int max_iter, double atol, hipStream_t st) {
  const bool gauss = S.stype == 2;
  const long long nd = S.ndata, np = S.nparm, na = nd + (gauss ? 3 : 2) * np;
  LsqrScalars* sc = W.sc;
  (void)hipMemsetAsync(sc, 0, sizeof(LsqrScalars), st);
  if (gauss)
    hipLaunchKernelGGL(gauss_fwd_kernel, dim3(nb(np, 128)), dim3(128), 0, st, S, S.dm_sum, W.u + nd, W.u + nd + np, 1,
                       sc);
  else if (S.smooth_on)
    hipLaunchKernelGGL(smooth_kernel, dim3(nb(np, 256)), dim3(256), 0, st, S, S.dm_sum, W.sm, sc);
  hipLaunchKernelGGL(rhs_kernel, dim3(nb(na, 256)), dim3(256), 0, st, S, W.sm, W.u);
  launch_norm(S, W, W.u, na, &sc->norm, st);
  hipLaunchKernelGGL(scalar_kernel, dim3(1), dim3(64), 0, st, sc, 0, 0, max_iter, atol);
  hipLaunchKernelGGL(scale_kernel, dim3(nb(na, 256)), dim3(256), 0, st, W.u, na, &sc->beta, 0, sc);
  launch_spmv(S, S.cp, S.crow, S.cval, S.nparm, W.u, W.gtu, sc, st);
  if (!gauss && S.smooth_on)
    hipLaunchKernelGGL(smooth_kernel, dim3(nb(np, 256)), dim3(256), 0, st, S, W.u + nd, W.sm, sc);
  hipLaunchKernelGGL(v_update_kernel, dim3(nb(np, 256)), dim3(256), 0, st, S, W.gtu, W.sm, W.u, W.v, 1, sc);
  launch_norm(S, W, W.v, np, &sc->norm, st);
  hipLaunchKernelGGL(scalar_kernel, dim3(1), dim3(64), 0, st, sc, 1, 0, max_iter, atol);
  hipLaunchKernelGGL(scale_kernel, dim3(nb(np, 256)), dim3(256), 0, st, W.v, np, &sc->alpha, 0, sc);
  hipLaunchKernelGGL(w_init_kernel, dim3(nb(np, 256)), dim3(256), 0, st, S, W.v, W.w, W.dm);
  return hipGetLastError();
}

// The kernels of LSQR iteration `it` (each returns at once after the stop).
static hipError_t launch_iteration(const LsqrSys& S, LsqrWork& W, int it, int max_iter, double atol,
                                   hipStream_t st) {
  const bool gauss = S.stype == 2;
  const long long nd = S.ndata, np = S.nparm, na = nd + (gauss ? 3 : 2) * np;
  LsqrScalars* sc = W.sc;
  hipLaunchKernelGGL(mask_kernel, dim3(nb(np, 256)), dim3(256), 0, st, S, W.v, sc);
  if (nd > 0) launch_spmv(S, S.rp, S.col, S.val, S.ndata, W.v, W.gv, sc, st);
  if (gauss)
    hipLaunchKernelGGL(gauss_fwd_kernel, dim3(nb(np, 128)), dim3(128), 0, st, S, W.v, W.sm, W.sm2, 0, sc);
  else if (S.smooth_on)
    hipLaunchKernelGGL(smooth_kernel, dim3(nb(np, 256)), dim3(256), 0, st, S, W.v, W.sm, sc);
  hipLaunchKernelGGL(u_update_kernel, dim3(nb(na, 256)), dim3(256), 0, st, S, W.gv, W.sm, W.sm2, W.v, W.u, sc);
  launch_norm(S, W, W.u, na, &sc->norm, st);
  hipLaunchKernelGGL(scalar_kernel, dim3(1), dim3(64), 0, st, sc, 2, it, max_iter, atol);
  hipLaunchKernelGGL(scale_kernel, dim3(nb(na, 256)), dim3(256), 0, st, W.u, na, &sc->beta, 1, sc);
  launch_spmv(S, S.cp, S.crow, S.cval, S.nparm, W.u, W.gtu, sc, st);
  if (!gauss && S.smooth_on)
    hipLaunchKernelGGL(smooth_kernel, dim3(nb(np, 256)), dim3(256), 0, st, S, W.u + nd, W.sm, sc);
  hipLaunchKernelGGL(v_update_kernel, dim3(nb(np, 256)), dim3(256), 0, st, S, W.gtu, W.sm, W.u, W.v, 0, sc);
  launch_norm(S, W, W.v, np, &sc->norm, st);
  hipLaunchKernelGGL(scalar_kernel, dim3(1), dim3(64), 0, st, sc, 3, it, max_iter, atol);
  hipLaunchKernelGGL(scale_kernel, dim3(nb(np, 256)), dim3(256), 0, st, W.v, np, &sc->alpha, 1, sc);
  hipLaunchKernelGGL(w_update_kernel, dim3(nb(np, 256)), dim3(256), 0, st, S, W.v, W.w, W.dm, sc);
  hipLaunchKernelGGL(test_kernel, dim3(1), dim3(64), 0, st, sc, it, max_iter, atol);
  return hipGetLastError();
}

static hipError_t poll_scalars(LsqrWork& W, LsqrScalars* host_out, hipStream_t st) {
  hipError_t e = hipMemcpyAsync(host_out, W.sc, sizeof(LsqrScalars), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e;
}

hipError_t lsqr_run_many(int n, const LsqrSys* S, LsqrWork* W, int max_iter, double atol, int poll,
                         const hipStream_t* st, LsqrScalars* host_out) {
  std::vector<char> active(n, 1);
  for (int q = 0; q < n; q++) {
    const hipError_t e = launch_start(S[q], W[q], max_iter, atol, st[q]);
    if (e != hipSuccess) return e;
  }
  int left = n;
  for (int it = 1; it <= max_iter && left > 0; it++) {
    // the solves' kernels go to their own streams, iteration by iteration,
    // so independent solves run side by side
    for (int q = 0; q < n; q++) {
      if (!active[q]) continue;
      const hipError_t e = launch_iteration(S[q], W[q], it, max_iter, atol, st[q]);
      if (e != hipSuccess) return e;
    }
    if (it % poll == 0 || it == max_iter) {
      for (int q = 0; q < n; q++) {
        if (!active[q]) continue;
        const hipError_t e = poll_scalars(W[q], &host_out[q], st[q]);
        if (e != hipSuccess) return e;
        if (host_out[q].done) { active[q] = 0; left--; }
      }
    }
  }
  for (int q = 0; q < n; q++) {
    const hipError_t e = poll_scalars(W[q], &host_out[q], st[q]);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t lsqr_run(const LsqrSys& S, LsqrWork& W, int max_iter, double atol, int poll, hipStream_t st,
                    LsqrScalars* host_out) {
  return lsqr_run_many(1, &S, &W, max_iter, atol, poll, &st, host_out);
}

// ---- system assembly (read_frechet_data2d + scaling, :280-486)

// rows kept: copy, scaled by mvscale(j) * data_wt(i) (gval*mvscale first)
__global__ void scale_rows_kernel(const int* __restrict__ src_rp, const int* __restrict__ src_col,
                                  const double* __restrict__ src_val, const int* __restrict__ keep_row,
                                  const double* __restrict__ wt, const int* __restrict__ rp, int ndata, int base,
                                  const double* __restrict__ mvs, int nparm, int* col, double* val) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ndata) return;
  const int r = keep_row[i];
  int o = rp[i];
  for (int k = src_rp[r]; k < src_rp[r + 1]; k++, o++) {
    const int j = src_col[k] - base;
    double g = src_val[k];
    if (j >= 0 && j < nparm) g = g * mvs[j] * wt[i];
    col[o] = j;
    val[o] = g;
  }
}

// gtg_diag(j) = sum over rows (ascending) of G(i,j)^2, from the column copy
__global__ void gtg_kernel(const int* __restrict__ cp, const double* __restrict__ cval, int nparm, double* gtg) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nparm) return;
  double a = 0.0;
  for (int k = cp[j]; k < cp[j + 1]; k++) a = a + cval[k] * cval[k];
  gtg[j] = a;
}

hipError_t launch_scale_rows(const int* src_rp, const int* src_col, const double* src_val, const int* keep_row,
                             const double* wt, const int* rp, int ndata, int base, const double* mvs, int nparm,
                             int* col, double* val, hipStream_t st) {
  if (ndata == 0) return hipSuccess;
  hipLaunchKernelGGL(scale_rows_kernel, dim3(nb(ndata, 128)), dim3(128), 0, st, src_rp, src_col, src_val, keep_row,
                     wt, rp, ndata, base, mvs, nparm, col, val);
  return hipGetLastError();
}

hipError_t launch_gtg(const int* cp, const double* cval, int nparm, double* gtg, hipStream_t st) {
  hipLaunchKernelGGL(gtg_kernel, dim3(nb(nparm, 128)), dim3(128), 0, st, cp, cval, nparm, gtg);
  return hipGetLastError();
}

hipError_t launch_ordered_spmv(const int* ptr, const int* idx, const double* val, int nrow, const double* x,
                               double* y, hipStream_t st) {
  if (nrow <= 0) return hipSuccess;
  hipLaunchKernelGGL(ordered_spmv_kernel, dim3((unsigned)nrow), dim3(64), 0, st, ptr, idx, val, nrow, x, y,
                     (const LsqrScalars*)nullptr);
  return hipGetLastError();
}

}  // namespace fmt2d
