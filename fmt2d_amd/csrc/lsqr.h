// Device LSQR of the tomography system (see lsqr.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "csr.h"

namespace fmt2d {

struct LsqrScalars {
  double norm;  // last front-to-back sum of squares
  double alpha, beta, bnorm, rnorm, test2, anorm2, rho, rhobar, phi, phibar, theta, c;
  int done, iters, trivial, pad;
};

struct LsqrSys {
  int ndata, nparm, n_vel;
  const int* rp;          // G rows (0-based), ndata+1
  const int* col;
  const double* val;
  const int* cp;          // G columns, nparm+1
  const int* crow;
  const double* cval;
  const double* res;      // weighted residuals
  const double* dm_sum;   // accumulated model update (jumping strategy)
  const uint8_t* mask;    // velocity columns with gtg_diag < 0.01 eps_min
  const double* ws;       // smoothing weight per parameter
  const double* wd;       // damping weight per parameter
  const int4* lap_nb;     // velocity parameters: neighbours x-1, x+1, z-1, z+1 (-1 none)
  const int2* if_nb;      // interface parameters: (1 one-sided | 2 central | 0 none, neighbour)
  int smooth_on, damp_on;
  // smoothing operator: 1 Laplacian, 2 Gaussian R_h / R_v blocks (R_if in
  // the R_h block's interface rows), 3 Laplacian with R_if on interface rows
  int stype;
  double w_smooth, w_smooth_iface;
  const int *rh_rp, *rh_col, *rh_cp, *rh_crow;   // R_h rows and columns (0-based)
  const double *rh_val, *rh_cval;
  const int *rv_rp, *rv_col, *rv_cp, *rv_crow;   // R_v
  const double *rv_val, *rv_cval;
  const int *rif_rp, *rif_col;                   // R_if rows (nullptr: not built)
  const double* rif_val;
  int gsi;                                       // smooth_type 2: R_if in use
  // 0: every reduction in the reference's order (bit-exact model update);
  // 1: tree-reduced norms and wave-reduced sparse products (other rounding)
  int fast;
};

struct LsqrWork {
  double *u, *v, *w, *dm, *gv, *gtu, *sm, *sm2;
  LsqrScalars* sc;
  double* part;   // partial sums of the tree-reduced norms (2 LSQR_NPART: high and low parts)
};
constexpr int LSQR_NPART = 1024;

hipError_t lsqr_run(const LsqrSys& S, LsqrWork& W, int max_iter, double atol, int poll, hipStream_t st,
                    LsqrScalars* host_out);
// n independent solves (own systems, work buffers and streams) launched
// iteration by iteration side by side; host_out[q] receives solve q's scalars
hipError_t lsqr_run_many(int n, const LsqrSys* S, LsqrWork* W, int max_iter, double atol, int poll,
                         const hipStream_t* st, LsqrScalars* host_out);
hipError_t launch_scale_rows(const int* src_rp, const int* src_col, const double* src_val, const int* keep_row,
                             const double* wt, const int* rp, int ndata, int base, const double* mvs, int nparm,
                             int* col, double* val, hipStream_t st);
// y[i] = sum over the stored entries of row i of M times x, in stored order
// (one wave per row)
hipError_t launch_ordered_spmv(const int* ptr, const int* idx, const double* val, int nrow, const double* x,
                               double* y, hipStream_t st);
// front-to-back sum of a[i]^2 into *dst (device pointers), the ordered
// norms' kernel on its own
hipError_t launch_ordered_sumsq(const double* a, long long n, double* dst, hipStream_t st);
hipError_t launch_gtg(const int* cp, const double* cval, int nparm, double* gtg, hipStream_t st);
// column-major copy of a CSR matrix (stable: rows ascending within a
// column); its scratch comes from w
hipError_t build_csc(const int* rp, const int* col, const double* val, int nrow, int ncol, int nnz, int* cp, int* crow,
                     double* cval, CsrWork& w, hipStream_t st);

}  // namespace fmt2d
