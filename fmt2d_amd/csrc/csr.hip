// CSR assembly of the Frechet rows from per-ray contributions.
//
// The reference accumulates every ray into a dense scratch row and compacts
// it by parameter index (compute_frechet_for_ray2d, inversion2d.f90:4573-4577,
// 4761-4768; read_frechet_data2d packs the rows, :3965-4052).  Here each
// contribution is a (ray << 32 | parameter, value) pair written in the
// reference's accumulation order; a stable radix sort groups equal keys
// without reordering them, so summing each run front to back reproduces the
// reference's sums bit for bit.  Entries with |sum| <= 1e-30 are dropped as
// the reference drops them.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "csr.h"
#include "lsqr.h"

namespace fmt2d {

__global__ void run_heads_kernel(const unsigned long long* __restrict__ k, long long n, int* head) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  head[i] = (i == 0 || k[i] != k[i - 1]) ? 1 : 0;
}

__global__ void run_start_kernel(const int* __restrict__ head, const int* __restrict__ rid, long long n,
                                 long long* start) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (head[i]) start[rid[i]] = i;
}

// one thread per run: ordered sum, keep flag, per-ray count
__global__ void run_sum_kernel(const unsigned long long* __restrict__ k, const double* __restrict__ v,
                               const long long* __restrict__ start, int nrun, long long n, double* sum, int* keep,
                               int* rowcnt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrun) return;
  const long long a = start[r], b = (r + 1 < nrun) ? start[r + 1] : n;
  double s = 0.0;
  for (long long i = a; i < b; i++) s = s + v[i];
  const int kp = fabs(s) > 1.0e-30 ? 1 : 0;
  sum[r] = s;
  keep[r] = kp;
  if (kp) atomicAdd(&rowcnt[(int)(k[a] >> 32)], 1);
}

__global__ void run_scatter_kernel(const unsigned long long* __restrict__ k, const long long* __restrict__ start,
                                   const double* __restrict__ sum, const int* __restrict__ keep,
                                   const int* __restrict__ pos, int nrun, int base, int* col, double* val) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrun || !keep[r]) return;
  col[pos[r]] = (int)(k[start[r]] & 0xffffffffull) + base;
  val[pos[r]] = sum[r];
}

template <typename T>
static hipError_t dalloc(T** p, size_t n) { return hipMalloc((void**)p, (n ? n : 1) * sizeof(T)); }

__global__ void prolong_rows_kernel(const int* __restrict__ rp, const int* __restrict__ col,
                                    const double* __restrict__ val, int nrow, const int* __restrict__ p_rp,
                                    const int* __restrict__ p_col, const double* __restrict__ p_w,
                                    const long long* __restrict__ off, long long* cnt, unsigned long long* keys,
                                    double* vals) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrow) return;
  if (!keys) {
    long long n = 0;
    for (int k = rp[r]; k < rp[r + 1]; k++) n += p_rp[col[k] + 1] - p_rp[col[k]];
    cnt[r] = n;
    return;
  }
  long long o = off[r];
  const unsigned long long hi = (unsigned long long)r << 32;
  for (int k = rp[r]; k < rp[r + 1]; k++) {
    const int j = col[k];
    const double g = val[k];
    for (int e = p_rp[j]; e < p_rp[j + 1]; e++, o++) {
      keys[o] = hi | (unsigned long long)(unsigned)p_col[e];
      vals[o] = g * p_w[e];
    }
  }
}

hipError_t launch_prolong_rows(const int* rp, const int* col, const double* val, int nrow, const int* p_rp,
                               const int* p_col, const double* p_w, const long long* off, long long* cnt,
                               unsigned long long* keys, double* vals, hipStream_t st) {
  if (nrow <= 0) return hipSuccess;
  hipLaunchKernelGGL(prolong_rows_kernel, dim3((nrow + 127) / 128), dim3(128), 0, st, rp, col, val, nrow, p_rp,
                     p_col, p_w, off, cnt, keys, vals);
  return hipGetLastError();
}

#define CK(x) do { hipError_t _e = (x); if (_e != hipSuccess) { err = _e; goto done; } } while (0)

hipError_t build_csr(unsigned long long* keys, double* vals, long long n, int nray, int key_bits, int base,
                     CsrOut& out, CsrWork& w, hipStream_t st) {
  hipError_t err = hipSuccess;
  unsigned long long* k2 = nullptr;
  double* v2 = nullptr;
  int *head = nullptr, *rid = nullptr, *keep = nullptr, *pos = nullptr;
  long long* start = nullptr;
  double* sum = nullptr;
  unsigned char* tmp = nullptr;
  size_t tb = 0, tb2 = 0;
  int nrun = 0, nnz = 0;
  int hl[2];
  out.nnz = 0;
  CK(dalloc(&out.rowptr, (size_t)nray + 1));
  CK(hipMemsetAsync(out.rowptr, 0, sizeof(int) * ((size_t)nray + 1), st));
  if (n == 0) {
    CK(dalloc(&out.col, 1));
    CK(dalloc(&out.val, 1));
    goto done;
  }
  CK(w.get(0, n, &k2));
  CK(w.get(1, n, &v2));
  CK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys, k2, vals, v2, n, 0, key_bits, st));
  CK(w.get(2, n, &head));
  CK(w.get(3, n, &rid));
  CK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, head, rid, n, st));
  CK(w.get(4, std::max(tb, tb2), &tmp));
  CK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, keys, k2, vals, v2, n, 0, key_bits, st));
  hipLaunchKernelGGL(run_heads_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, k2, n, head);
  CK(hipcub::DeviceScan::ExclusiveSum(tmp, tb2, head, rid, n, st));
  CK(hipMemcpyAsync(&hl[0], rid + n - 1, sizeof(int), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(&hl[1], head + n - 1, sizeof(int), hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  nrun = hl[0] + hl[1];
  CK(w.get(5, nrun, &start));
  CK(w.get(6, nrun, &sum));
  CK(w.get(7, nrun, &keep));
  CK(w.get(8, nrun, &pos));
  hipLaunchKernelGGL(run_start_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, head, rid, n, start);
  // row counts in rowptr[0..nray), scanned below
  hipLaunchKernelGGL(run_sum_kernel, dim3((nrun + 255) / 256), dim3(256), 0, st, k2, v2, start, nrun, n, sum, keep,
                     out.rowptr);
  {
    size_t tb3 = 0;
    CK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, keep, pos, nrun, st));
    CK(w.get(4, tb3, &tmp));
    CK(hipcub::DeviceScan::ExclusiveSum(tmp, tb3, keep, pos, nrun, st));
    CK(hipMemcpyAsync(&hl[0], pos + nrun - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    CK(hipMemcpyAsync(&hl[1], keep + nrun - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    nnz = hl[0] + hl[1];
    CK(dalloc(&out.col, nnz));
    CK(dalloc(&out.val, nnz));
    hipLaunchKernelGGL(run_scatter_kernel, dim3((nrun + 255) / 256), dim3(256), 0, st, k2, start, sum, keep, pos, nrun,
                       base, out.col, out.val);
    // rowptr: exclusive scan of the counts (nray+1 entries, last count 0)
    int* rp = nullptr;
    size_t tb4 = 0;
    CK(dalloc(&rp, (size_t)nray + 1));
    CK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb4, out.rowptr, rp, nray + 1, st));
    CK(w.get(4, tb4, &tmp));
    CK(hipcub::DeviceScan::ExclusiveSum(tmp, tb4, out.rowptr, rp, nray + 1, st));
    CK(hipStreamSynchronize(st));
    (void)hipFree(out.rowptr);
    out.rowptr = rp;
  }
  out.nnz = nnz;
done:
  return err;
}

}  // namespace fmt2d

namespace fmt2d {

__global__ void entry_rows_kernel(const int* __restrict__ rp, int nrow, int* erow, int* eidx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrow) return;
  for (int k = rp[i]; k < rp[i + 1]; k++) { erow[k] = i; eidx[k] = k; }
}

__global__ void col_count_kernel(const int* __restrict__ col, int nnz, int* cnt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  atomicAdd(&cnt[col[k]], 1);
}

__global__ void csc_gather_kernel(const int* __restrict__ perm, const int* __restrict__ erow,
                                  const double* __restrict__ val, int nnz, int* crow, double* cval) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const int e = perm[k];
  crow[k] = erow[e];
  cval[k] = val[e];
}

hipError_t build_csc(const int* rp, const int* col, const double* val, int nrow, int ncol, int nnz, int* cp, int* crow,
                     double* cval, CsrWork& w, hipStream_t st) {
  hipError_t err = hipSuccess;
  int *erow = nullptr, *eidx = nullptr, *kout = nullptr, *perm = nullptr, *cnt = nullptr;
  unsigned char* tmp = nullptr;
  size_t tb = 0, tb2 = 0;
  int bits = 1;
  while ((1 << bits) < ncol + 1) bits++;
  CK(w.get(0, nnz, &erow)); CK(w.get(1, nnz, &eidx)); CK(w.get(2, nnz, &kout)); CK(w.get(3, nnz, &perm));
  CK(w.get(5, (size_t)ncol + 1, &cnt));
  CK(hipMemsetAsync(cnt, 0, sizeof(int) * ((size_t)ncol + 1), st));
  if (nnz > 0) {
    hipLaunchKernelGGL(entry_rows_kernel, dim3((nrow + 127) / 128), dim3(128), 0, st, rp, nrow, erow, eidx);
    hipLaunchKernelGGL(col_count_kernel, dim3((nnz + 255) / 256), dim3(256), 0, st, col, nnz, cnt);
    CK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, col, kout, eidx, perm, nnz, 0, bits, st));
  }
  CK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, cnt, cp, ncol + 1, st));
  CK(w.get(4, std::max<size_t>(std::max(tb, tb2), 1), &tmp));
  if (nnz > 0) CK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, col, kout, eidx, perm, nnz, 0, bits, st));
  CK(hipcub::DeviceScan::ExclusiveSum(tmp, tb2, cnt, cp, ncol + 1, st));
  if (nnz > 0)
    hipLaunchKernelGGL(csc_gather_kernel, dim3((nnz + 255) / 256), dim3(256), 0, st, perm, erow, val, nnz, crow, cval);
  CK(hipStreamSynchronize(st));
done:
  return err;
}

}  // namespace fmt2d
