// G x / G^T x with no device allocation: the staging of the host vectors
// lives in a fixed device array of this library instead of buffers
// allocated per call.  For callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops).
//
// The array is one per device (each GPU that runs the library holds its own
// copy until the process ends).  Host code never passes the array's name to
// a kernel or copy: the name is a host-side stand-in, so its device address
// is taken once per device with hipGetSymbolAddress and kept.  Every call on
// one device shares that device's array, which keeps the last call's
// contents; calls that overlap on one device race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "ctx.h"

namespace {

// 2^21 doubles (16 MiB): x and y of one product together
constexpr size_t SCRATCH_DOUBLES = size_t(1) << 21;
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) double scratch[SCRATCH_DOUBLES];

std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on `device` (which must be current).
int scratch_on(int device, double** out) {
  if (device < 0 || device >= MAX_DEVICES) return fail(FMT2D_EINVAL, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FMT2D_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[device];
  return FMT2D_OK;
}

}  // namespace

extern "C" {

int fmt2d_frechet_apply_noalloc(fmt2d_ctx* ctx, int transpose, const double* x, double* y) {
  if (!ctx || !x || !y) return fail(FMT2D_EINVAL, "NULL argument");
  if (transpose != 0 && transpose != 1) return fail(FMT2D_EINVAL, "transpose must be 0 or 1");
  if (ctx->nray < 1) return fail(FMT2D_ESTATE, "no rays traced");
  // building the rows or their column copy allocates: both must exist
  if (ctx->g_nnz < 0)
    return fail(FMT2D_ESTATE, "Frechet rows of the last trace not built (call fmt2d_frechet_apply once first)");
  if (transpose && !ctx->gt_ok)
    return fail(FMT2D_ESTATE, "column copy of the Frechet rows not built (call fmt2d_frechet_apply with "
                              "transpose 1 once first)");
  const int nray = ctx->nray, np = ctx->n_par;
  const size_t nin = transpose ? nray : np, nout = transpose ? np : nray;
  const size_t off = (nin + 31) & ~size_t(31);  // y starts 256-byte aligned
  if (off + nout > SCRATCH_DOUBLES)
    return fail(FMT2D_EINVAL, "x and y need " + std::to_string(off + nout) + " doubles of staging; the limit is " +
                                  std::to_string(SCRATCH_DOUBLES));
  HIPCHK(hipSetDevice(ctx->device));
  double* s = nullptr;
  const int rc = scratch_on(ctx->device, &s);
  if (rc) return rc;
  double* dx = s;
  double* dy = s + off;
  hipStream_t st = ctx->stream;
  HIPCHK(hipMemcpyAsync(dx, x, sizeof(double) * nin, hipMemcpyHostToDevice, st));
  if (transpose)
    HIPCHK(launch_ordered_spmv(ctx->gt_cp.p, ctx->gt_row.p, ctx->gt_val.p, np, dx, dy, st));
  else
    HIPCHK(launch_ordered_spmv(ctx->g_rp.p, ctx->g_col.p, ctx->g_val.p, nray, dx, dy, st));
  HIPCHK(hipMemcpyAsync(y, dy, sizeof(double) * nout, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return FMT2D_OK;
}

}  // extern "C"
