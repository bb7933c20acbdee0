// C ABI of the tomography system and its LSQR solve: data selection and
// weighting (read_frechet_data2d and the scaling of lsqr_inversion2d,
// src/inversion2d.f90:280-486, 3965-4052), the Gaussian smoothing operators
// (:1252-1636) and run_lsqr (:6578-7167).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <utility>
#include <string>
#include <vector>

#include "ctx.h"

namespace {

// CSR rows (1-based columns as in the reference, converted at upload)
struct RowsCsr {
  std::vector<int> rp, col;
  std::vector<double> val;
};

RowsCsr pack_rows(std::vector<std::vector<std::pair<int, double>>>& rows) {
  RowsCsr r;
  r.rp.assign(rows.size() + 1, 0);
  for (size_t i = 0; i < rows.size(); i++) r.rp[i + 1] = r.rp[i] + (int)rows[i].size();
  r.col.reserve(r.rp.back());
  r.val.reserve(r.rp.back());
  for (auto& row : rows)
    for (auto& e : row) { r.col.push_back(e.first - 1); r.val.push_back(e.second); }
  return r;
}

// build_gauss_smooth_matrix (inversion2d.f90:1252-1402): for every inverted
// velocity node, R_h holds its row neighbours within int(3 corr_x / dx) + 1
// columns and R_v its column neighbours within 3 corr_z, ascending, each
// exp(-d^2 / corr^2) * mvscale(neighbour), then the diagonal
// -sum(beta) * mvscale(node).
void gauss_rows(const std::vector<double>& vx, const std::vector<double>& vz, const std::vector<int>& inv, int vnx,
                int vnz, double corr_x, double corr_z, const std::vector<double>& mvs, RowsCsr& rh, RowsCsr& rv) {
  const int nparm = (int)mvs.size();
  std::vector<std::vector<std::pair<int, double>>> h(nparm), v(nparm);
  const double dx = vx[1] - vx[0];
  const int kmin = -(int)(3.0 * corr_x / dx) - 1, kmax = (int)(3.0 * corr_x / dx) + 1;
  auto at = [&](int ix, int iz) { return inv[(size_t)(iz - 1) * vnx + (ix - 1)]; };  // 1-based
  for (int iz = 1; iz <= vnz; iz++)
    for (int ix = 1; ix <= vnx; ix++) {
      const int idx = at(ix, iz);
      if (idx <= 0) continue;
      double bsum = 0.0;
      auto& rowh = h[idx - 1];
      for (int jx = std::max(1, ix + kmin); jx <= std::min(vnx, ix + kmax); jx++) {
        if (jx == ix) continue;
        const int jdx = at(jx, iz);
        if (jdx <= 0) continue;
        const double d = vx[jx - 1] - vx[ix - 1];
        const double beta = std::exp(-(d * d) / (corr_x * corr_x));
        bsum = bsum + beta;
        rowh.push_back({jdx, beta * mvs[jdx - 1]});
      }
      rowh.push_back({idx, -bsum * mvs[idx - 1]});
      bsum = 0.0;
      auto& rowv = v[idx - 1];
      for (int jz = 1; jz <= vnz; jz++) {
        if (jz == iz) continue;
        const double d = vz[jz - 1] - vz[iz - 1];
        if (std::fabs(d) > 3.0 * corr_z) continue;
        const int jdx = at(ix, jz);
        if (jdx <= 0) continue;
        const double beta = std::exp(-(d * d) / (corr_z * corr_z));
        bsum = bsum + beta;
        rowv.push_back({jdx, beta * mvs[jdx - 1]});
      }
      rowv.push_back({idx, -bsum * mvs[idx - 1]});
    }
  rh = pack_rows(h);
  rv = pack_rows(v);
}

// build_iface_gauss_smooth_matrix (inversion2d.f90:1524-1616): the same
// 1-D rows along each inverted interface's input nodes.
RowsCsr iface_gauss_rows(const fmt2d_ctx* ctx, double corr, const std::vector<double>& mvs) {
  const int nparm = (int)mvs.size();
  std::vector<std::vector<std::pair<int, double>>> rows(nparm);
  for (int k : ctx->inv_ifaces) {
    const int start = ctx->mesh.if_start[k], nxi = ctx->mesh.if_nnode[k];
    const std::vector<double>& x = ctx->mesh.ifx[k];
    const double dx = x[1] - x[0];
    const int kmin = -(int)(3.0 * corr / dx) - 1, kmax = (int)(3.0 * corr / dx) + 1;
    for (int ix = 1; ix <= nxi; ix++) {
      const int idx = start + ix;
      if (idx < 1 || idx > nparm) continue;
      double bsum = 0.0;
      auto& row = rows[idx - 1];
      for (int jx = std::max(1, ix + kmin); jx <= std::min(nxi, ix + kmax); jx++) {
        if (jx == ix) continue;
        const int jdx = start + jx;
        if (jdx < 1 || jdx > nparm) continue;
        const double d = x[jx - 1] - x[ix - 1];
        const double beta = std::exp(-(d * d) / (corr * corr));
        bsum = bsum + beta;
        row.push_back({jdx, beta * mvs[jdx - 1]});
      }
      row.push_back({idx, -bsum * mvs[idx - 1]});
    }
  }
  return pack_rows(rows);
}

// column copy of a square CSR, rows ascending within a column (the order in
// which the reference's transposed products scatter)
void csc_of(const RowsCsr& a, int n, std::vector<int>& cp, std::vector<int>& row, std::vector<double>& val) {
  cp.assign(n + 1, 0);
  for (int c : a.col) cp[c + 1]++;
  for (int j = 0; j < n; j++) cp[j + 1] += cp[j];
  row.resize(a.col.size());
  val.resize(a.col.size());
  std::vector<int> fill(cp.begin(), cp.end() - 1);
  for (int i = 0; i < n; i++)
    for (int k = a.rp[i]; k < a.rp[i + 1]; k++) {
      const int o = fill[a.col[k]]++;
      row[o] = i;
      val[o] = a.val[k];
    }
}

}  // namespace

extern "C" {

int fmt2d_build_system(fmt2d_ctx* ctx, const double* t_obs, const fmt2d_system_params* p, int* ndata, int* nnz,
                       double* eps_min) {
  if (!ctx || !t_obs || !p || !ndata || !nnz || !eps_min) return fail(FMT2D_EINVAL, "NULL argument");
  if (ctx->nray < 1) return fail(FMT2D_ESTATE, "no rays traced");
  if (p->alt_mode < 0 || p->alt_mode > 2) return fail(FMT2D_EINVAL, "alt_mode must be 0, 1 or 2");
  HIPCHK(hipSetDevice(ctx->device));
  const int rc0 = system_rows_build(ctx);
  if (rc0) return rc0;
  hipStream_t st = ctx->stream;
  const HostModel& m = ctx->m;
  const InvMesh& M = ctx->mesh;
  const int nray = ctx->nray, nparm = M.n_par, n_vel = M.n_vel;
  // rows kept: valid ray, observed pick, |obs - calc| <= crit_chi
  std::vector<int> rp_all((size_t)nray + 1);
  HIPCHK(hipMemcpy(rp_all.data(), ctx->gp_rp.p, sizeof(int) * (nray + 1), hipMemcpyDeviceToHost));
  std::vector<int> keep, rp(1, 0);
  std::vector<double> wt, res;
  for (int r = 0; r < nray; r++) {
    if (!ctx->h_valid[r] || t_obs[r] < 0.0) continue;
    const double rc = t_obs[r] - ctx->h_tcalc[r];
    if (fabs(rc) > p->crit_chi) continue;
    const int path = ctx->hjobs[r].path;
    double w = 1.0;
    if (path == 2) w = w * p->phase_wt_psed;
    else if (path == 3) w = w * p->phase_wt_pmp;
    else if (path == 4 || path == 5) w = w * p->phase_wt_pgss;
    keep.push_back(r);
    wt.push_back(w);
    res.push_back(rc * w);
    rp.push_back(rp.back() + (rp_all[r + 1] - rp_all[r]));
  }
  const int nd = (int)keep.size(), nz_ = rp.back();
  // RMS of the selected, unweighted residuals (compute_rms2d, :332, :3819-3839)
  {
    double ss = 0.0;
    for (int i = 0; i < nd; i++) {
      const double rc = t_obs[keep[i]] - ctx->h_tcalc[keep[i]];
      ss = ss + rc * rc;
    }
    ctx->sys_rms = nd > 0 ? sqrt(ss / (double)nd) : 0.0;
    ctx->sys_crit_chi = p->crit_chi;
    ctx->sys_alt_mode = p->alt_mode;
  }
  // parameter scaling: 1/max(v_init,1) at the velocity mesh nodes,
  // refl_weight on interface nodes, zero for the parameters the alternating
  // mode fixes
  std::vector<double> mvs(nparm, 1.0);
  {
    if (ctx->v_init.size() != m.vgrid.size()) ctx->v_init = m.vgrid;
    const std::vector<double>& vi = ctx->v_init;
    const size_t fnx = (size_t)ctx->hgeom.vnx, sv = (size_t)M.vstride;
    for (int b = 0; b < M.vnz; b++)
      for (int a = 0; a < M.vnx; a++) {
        const int k = M.inv[(size_t)b * M.vnx + a];
        if (k > 0) mvs[k - 1] = 1.0 / std::max(vi[b * sv * fnx + a * sv], 1.0);
      }
    for (int j = n_vel; j < nparm; j++) mvs[j] = p->refl_weight;
    if (p->alt_mode == 1) for (int j = n_vel; j < nparm; j++) mvs[j] = 0.0;
    // the smoothing operators carry the velocity scaling even when the
    // velocity columns are fixed (lsqr_inversion2d recomputes it, :451-463)
    ctx->h_mvs_smooth = mvs;
    ctx->h_mvs_model = mvs;
    for (int j = n_vel; j < nparm; j++) ctx->h_mvs_model[j] = p->refl_weight;
    if (p->alt_mode == 2) for (int j = 0; j < n_vel; j++) mvs[j] = 0.0;
  }
  DBuf<int> dkeep;
  DBuf<double> dwt;
  HIPCHK(upload(dkeep, keep, st));
  HIPCHK(upload(dwt, wt, st));
  HIPCHK(upload(ctx->s_rp, rp, st));
  HIPCHK(upload(ctx->s_res, res, st));
  HIPCHK(upload(ctx->s_mvs, mvs, st));
  HIPCHK(ctx->s_col.alloc((size_t)std::max(nz_, 1)));
  HIPCHK(ctx->s_val.alloc((size_t)std::max(nz_, 1)));
  HIPCHK(launch_scale_rows(ctx->gp_rp.p, ctx->gp_col.p, ctx->gp_val.p, dkeep.p, dwt.p, ctx->s_rp.p, nd, 0, ctx->s_mvs.p,
                           nparm, ctx->s_col.p, ctx->s_val.p, st));
  HIPCHK(ctx->s_cp.alloc((size_t)nparm + 1));
  HIPCHK(ctx->s_crow.alloc((size_t)std::max(nz_, 1)));
  HIPCHK(ctx->s_cval.alloc((size_t)std::max(nz_, 1)));
  HIPCHK(build_csc(ctx->s_rp.p, ctx->s_col.p, ctx->s_val.p, nd, nparm, nz_, ctx->s_cp.p, ctx->s_crow.p,
                   ctx->s_cval.p, ctx->csr_work, st));
  HIPCHK(ctx->s_gtg.alloc((size_t)std::max(nparm, 1)));
  HIPCHK(launch_gtg(ctx->s_cp.p, ctx->s_cval.p, nparm, ctx->s_gtg.p, st));
  ctx->h_gtg.resize(nparm);
  HIPCHK(hipMemcpyAsync(ctx->h_gtg.data(), ctx->s_gtg.p, sizeof(double) * nparm, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  double gmax = 0.0;
  for (int j = 0; j < nparm; j++) gmax = std::max(gmax, ctx->h_gtg[j]);
  ctx->sys_eps_min = std::max(1.0e-6, 1.0e-4 * gmax);
  ctx->sys_ndata = nd;
  ctx->sys_nnz = nz_;
  *ndata = nd;
  *nnz = nz_;
  *eps_min = ctx->sys_eps_min;
  return FMT2D_OK;
}

int fmt2d_system_get(fmt2d_ctx* ctx, double* gtg_diag, double* residual) {
  if (!ctx) return fail(FMT2D_EINVAL, "NULL argument");
  if (ctx->sys_ndata < 0) return fail(FMT2D_ESTATE, "no system built");
  HIPCHK(hipSetDevice(ctx->device));
  if (gtg_diag) memcpy(gtg_diag, ctx->h_gtg.data(), sizeof(double) * ctx->mesh.n_par);
  if (residual && ctx->sys_ndata > 0)
    HIPCHK(hipMemcpy(residual, ctx->s_res.p, sizeof(double) * ctx->sys_ndata, hipMemcpyDeviceToHost));
  return FMT2D_OK;
}

int fmt2d_set_lsqr_mode(fmt2d_ctx* ctx, int mode) {
  if (!ctx || (mode != 0 && mode != 1)) return fail(FMT2D_EINVAL, "LSQR mode must be 0 (ordered) or 1 (parallel)");
  ctx->lsqr_fast = mode;
  return FMT2D_OK;
}

int fmt2d_ordered_sumsq(fmt2d_ctx* ctx, const double* x, long long n, double* out) {
  if (!ctx || !out || n < 0 || (n > 0 && !x)) return fail(FMT2D_EINVAL, "NULL argument or negative length");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  DBuf<double> dx, dsum;
  HIPCHK(dx.alloc((size_t)n));
  HIPCHK(dsum.alloc(1));
  if (n > 0) HIPCHK(hipMemcpyAsync(dx.p, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(launch_ordered_sumsq(dx.p, n, dsum.p, st));
  HIPCHK(hipMemcpyAsync(out, dsum.p, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return FMT2D_OK;
}

int fmt2d_lsqr(fmt2d_ctx* ctx, const fmt2d_lsqr_params* p, const double* dm_sum, double* dm_out,
               fmt2d_lsqr_stats* stats) {
  return lsqr_solve(ctx, p, dm_sum, dm_out, stats);
}

}  // extern "C"

// Everything the solves of one system share: checks, column mask, smoothing
// weights and tables, Gaussian operators, dm_sum; fills S (without the
// damping weights).  p's damping weights are not used here.
static int lsqr_prepare(fmt2d_ctx* ctx, const fmt2d_lsqr_params* p, const double* dm_sum, LsqrSys& S,
                        size_t& na_out) {
  if (!ctx || !p || !dm_sum) return fail(FMT2D_EINVAL, "NULL argument");
  if (ctx->sys_ndata < 0) return fail(FMT2D_ESTATE, "no system built");
  if (p->smooth_type < 1 || p->smooth_type > 3)
    return fail(FMT2D_EINVAL, "smooth_type must be 1 (Laplacian), 2 (Gaussian) or 3 (hybrid)");
  if (p->max_iter < 1) return fail(FMT2D_EINVAL, "max_iter must be >= 1");
  if (p->smooth_type == 2 && p->w_smooth > 0.0 && !(p->corr_x > 0.0 && p->corr_z > 0.0))
    return fail(FMT2D_EINVAL, "Gaussian smoothing needs corr_x > 0 and corr_z > 0");
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const InvMesh& M = ctx->mesh;
  const int nparm = M.n_par, n_vel = M.n_vel, nd = ctx->sys_ndata;
  const double cut = ctx->sys_eps_min * 0.01;
  std::vector<uint8_t> mask(nparm, 0);
  std::vector<double> ws(nparm);
  for (int j = 0; j < nparm; j++) {
    const bool vel = j < n_vel;
    mask[j] = (vel && ctx->h_gtg[j] < cut) ? 1 : 0;
    ws[j] = vel ? p->w_smooth : p->w_smooth_iface;
  }
  // Laplacian tables (apply_laplacian_smooth_2d, inversion2d.f90:1161-1249)
  std::vector<int4> lap(std::max(n_vel, 1));
  std::vector<int2> ifn(std::max(nparm - n_vel, 1), make_int2(0, 0));
  {
    const int vnx = M.vnx, vnz = M.vnz;
    const std::vector<int>& inv = M.inv;
    auto at = [&](int ix, int iz) { return inv[(size_t)iz * vnx + ix]; };
    for (int iz = 0; iz < vnz; iz++)
      for (int ix = 0; ix < vnx; ix++) {
        const int idx = at(ix, iz);
        if (idx <= 0) continue;
        int4 nb = make_int4(-1, -1, -1, -1);
        if (ix > 0 && at(ix - 1, iz) > 0) nb.x = at(ix - 1, iz) - 1;
        if (ix < vnx - 1 && at(ix + 1, iz) > 0) nb.y = at(ix + 1, iz) - 1;
        if (iz > 0 && at(ix, iz - 1) > 0) nb.z = at(ix, iz - 1) - 1;
        if (iz < vnz - 1 && at(ix, iz + 1) > 0) nb.w = at(ix, iz + 1) - 1;
        lap[idx - 1] = nb;
      }
    for (int k : ctx->inv_ifaces) {
      const int start = M.if_start[k], nxi = M.if_nnode[k];
      for (int ix = 1; ix <= nxi; ix++) {
        const int idx = start + ix;  // 1-based
        if (idx < 1 || idx > nparm || idx <= n_vel) continue;
        int2 e = make_int2(0, 0);
        if (ix == 1) { if (idx + 1 <= nparm) e = make_int2(1, idx); }
        else if (ix == nxi) { if (idx - 1 >= 1) e = make_int2(1, idx - 2); }
        else e = make_int2(2, 0);
        ifn[idx - 1 - n_vel] = e;
      }
    }
  }
  // Gaussian operators: built on first use for this parameterisation and
  // these correlation lengths, with the scaling of that system, then kept
  // (the reference builds them once, inversion2d.f90:516-522)
  const bool want_g = p->smooth_type == 2 && p->w_smooth > 0.0;
  const bool want_i = (p->smooth_type == 2 || p->smooth_type == 3) && p->w_smooth_iface > 0.0 && p->corr_x_iface > 0.0;
  if (want_g && !(ctx->gs_nparm == nparm && ctx->gs_key[0] == p->corr_x && ctx->gs_key[1] == p->corr_z)) {
    RowsCsr rh, rv;
    gauss_rows(M.vx, M.vz, M.inv, M.vnx, M.vnz, p->corr_x, p->corr_z, ctx->h_mvs_smooth, rh, rv);
    std::vector<int> cp, row;
    std::vector<double> cv;
    HIPCHK(upload(ctx->gh_rp, rh.rp, st)); HIPCHK(upload(ctx->gh_col, rh.col, st)); HIPCHK(upload(ctx->gh_val, rh.val, st));
    csc_of(rh, nparm, cp, row, cv);
    HIPCHK(upload(ctx->gh_cp, cp, st)); HIPCHK(upload(ctx->gh_row, row, st)); HIPCHK(upload(ctx->gh_cval, cv, st));
    HIPCHK(upload(ctx->gv_rp, rv.rp, st)); HIPCHK(upload(ctx->gv_col, rv.col, st)); HIPCHK(upload(ctx->gv_val, rv.val, st));
    csc_of(rv, nparm, cp, row, cv);
    HIPCHK(upload(ctx->gv_cp, cp, st)); HIPCHK(upload(ctx->gv_row, row, st)); HIPCHK(upload(ctx->gv_cval, cv, st));
    HIPCHK(hipStreamSynchronize(st));
    ctx->gs_nparm = nparm;
    ctx->gs_key[0] = p->corr_x;
    ctx->gs_key[1] = p->corr_z;
  }
  if (want_i && !(ctx->gsi_nparm == nparm && ctx->gs_key[2] == p->corr_x_iface)) {
    RowsCsr ri = iface_gauss_rows(ctx, p->corr_x_iface, ctx->h_mvs_smooth);
    HIPCHK(upload(ctx->gi_rp, ri.rp, st)); HIPCHK(upload(ctx->gi_col, ri.col, st)); HIPCHK(upload(ctx->gi_val, ri.val, st));
    HIPCHK(hipStreamSynchronize(st));
    ctx->gsi_nparm = nparm;
    ctx->gsi_nnz = (int)ri.col.size();
    ctx->gs_key[2] = p->corr_x_iface;
  }
  std::vector<double> dms(dm_sum, dm_sum + nparm);
  HIPCHK(upload(ctx->l_mask, mask, st));
  HIPCHK(upload(ctx->l_ws, ws, st));
  HIPCHK(upload(ctx->l_lapnb, lap, st));
  HIPCHK(upload(ctx->l_ifnb, ifn, st));
  HIPCHK(upload(ctx->l_dmsum, dms, st));
  na_out = (size_t)nd + (p->smooth_type == 2 ? 3 : 2) * (size_t)nparm;
  S = LsqrSys{};
  S.ndata = nd; S.nparm = nparm; S.n_vel = n_vel;
  S.rp = ctx->s_rp.p; S.col = ctx->s_col.p; S.val = ctx->s_val.p;
  S.cp = ctx->s_cp.p; S.crow = ctx->s_crow.p; S.cval = ctx->s_cval.p;
  S.res = ctx->s_res.p; S.dm_sum = ctx->l_dmsum.p; S.mask = ctx->l_mask.p; S.ws = ctx->l_ws.p; S.wd = nullptr;
  S.lap_nb = ctx->l_lapnb.p; S.if_nb = ctx->l_ifnb.p;
  S.smooth_on = (p->w_smooth > 0.0 || p->w_smooth_iface > 0.0) ? 1 : 0;
  S.damp_on = 0;
  S.stype = p->smooth_type;
  S.w_smooth = p->w_smooth;
  S.w_smooth_iface = p->w_smooth_iface;
  S.rh_rp = S.rh_col = S.rh_cp = S.rh_crow = nullptr; S.rh_val = S.rh_cval = nullptr;
  S.rv_rp = S.rv_col = S.rv_cp = S.rv_crow = nullptr; S.rv_val = S.rv_cval = nullptr;
  S.rif_rp = S.rif_col = nullptr; S.rif_val = nullptr;
  if (want_g) {
    S.rh_rp = ctx->gh_rp.p; S.rh_col = ctx->gh_col.p; S.rh_val = ctx->gh_val.p;
    S.rh_cp = ctx->gh_cp.p; S.rh_crow = ctx->gh_row.p; S.rh_cval = ctx->gh_cval.p;
    S.rv_rp = ctx->gv_rp.p; S.rv_col = ctx->gv_col.p; S.rv_val = ctx->gv_val.p;
    S.rv_cp = ctx->gv_cp.p; S.rv_crow = ctx->gv_row.p; S.rv_cval = ctx->gv_cval.p;
  }
  if (want_i) { S.rif_rp = ctx->gi_rp.p; S.rif_col = ctx->gi_col.p; S.rif_val = ctx->gi_val.p; }
  S.gsi = (p->smooth_type == 2 && want_i) ? 1 : 0;
  S.fast = ctx->lsqr_fast;
  return FMT2D_OK;
}

// Work buffers of solve slot q (0 or 1) for na augmented rows.
static int lsqr_work(fmt2d_ctx* ctx, int q, size_t na, LsqrWork& W) {
  const int nparm = ctx->mesh.n_par, nd = ctx->sys_ndata;
  DBuf<double>* b[8] = {&ctx->l_u, &ctx->l_v, &ctx->l_w, &ctx->l_dm, &ctx->l_gv, &ctx->l_gtu, &ctx->l_sm, &ctx->l_sm2};
  DBuf<double>* b2[8] = {&ctx->l2_u, &ctx->l2_v, &ctx->l2_w, &ctx->l2_dm, &ctx->l2_gv, &ctx->l2_gtu, &ctx->l2_sm,
                         &ctx->l2_sm2};
  DBuf<double>** B = q == 0 ? b : b2;
  DBuf<LsqrScalars>& sc = q == 0 ? ctx->l_sc : ctx->l2_sc;
  const size_t need[8] = {na, (size_t)nparm, (size_t)nparm, (size_t)nparm, (size_t)std::max(nd, 1),
                          (size_t)nparm, (size_t)nparm, (size_t)nparm};
  for (int k = 0; k < 8; k++)
    if (B[k]->n < need[k]) HIPCHK(B[k]->alloc(need[k]));
  if (!sc.p) HIPCHK(sc.alloc(1));
  DBuf<double>& part = q == 0 ? ctx->l_part : ctx->l2_part;
  if (!part.p) HIPCHK(part.alloc(2 * LSQR_NPART));
  W = LsqrWork{B[0]->p, B[1]->p, B[2]->p, B[3]->p, B[4]->p, B[5]->p, B[6]->p, B[7]->p, sc.p, part.p};
  return FMT2D_OK;
}

// n (1 or 2) solves of the built system with damping weights (wdv[q],
// wdv_if[q]), side by side on two streams.
int lsqr_solve_many(fmt2d_ctx* ctx, const fmt2d_lsqr_params* p, const double* dm_sum, int n, const double* wdv,
                    const double* wdv_if, double* const* dm_out, fmt2d_lsqr_stats* stats) {
  if (n < 1 || n > 2) return fail(FMT2D_EINVAL, "one or two solves");
  for (int q = 0; q < n; q++)
    if (!dm_out[q]) return fail(FMT2D_EINVAL, "NULL argument");
  const auto t0 = std::chrono::steady_clock::now();
  LsqrSys S0;
  size_t na = 0;
  int rc = lsqr_prepare(ctx, p, dm_sum, S0, na);
  if (rc) return rc;
  const int nparm = ctx->mesh.n_par, n_vel = ctx->mesh.n_vel;
  if (n == 2 && !ctx->stream2) HIPCHK(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
  LsqrSys S[2];
  LsqrWork W[2];
  hipStream_t st[2] = {ctx->stream, ctx->stream2};
  for (int q = 0; q < n; q++) {
    std::vector<double> wd(nparm);
    for (int j = 0; j < nparm; j++) wd[j] = j < n_vel ? wdv[q] : wdv_if[q];
    HIPCHK(upload(q == 0 ? ctx->l_wd : ctx->l2_wd, wd, ctx->stream));
    rc = lsqr_work(ctx, q, na, W[q]);
    if (rc) return rc;
    S[q] = S0;
    S[q].wd = q == 0 ? ctx->l_wd.p : ctx->l2_wd.p;
    S[q].damp_on = (wdv[q] > 0.0 || wdv_if[q] > 0.0) ? 1 : 0;
  }
  // the second stream reads what the first one uploaded
  if (n == 2) HIPCHK(hipStreamSynchronize(ctx->stream));
  LsqrScalars out[2] = {};
  HIPCHK(lsqr_run_many(n, S, W, p->max_iter, 1.0e-4, 8, st, out));
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  for (int q = 0; q < n; q++) {
    HIPCHK(hipMemcpy(dm_out[q], W[q].dm, sizeof(double) * nparm, hipMemcpyDeviceToHost));
    if (stats) {
      stats[q].iterations = out[q].iters;
      stats[q].rnorm = out[q].rnorm;
      stats[q].bnorm = out[q].bnorm;
      stats[q].test2 = out[q].test2;
      stats[q].ms = ms;
    }
  }
  return FMT2D_OK;
}

// One run_lsqr solve of the built system (fmt2d_lsqr; the damping search
// calls it repeatedly).
int lsqr_solve(fmt2d_ctx* ctx, const fmt2d_lsqr_params* p, const double* dm_sum, double* dm_out,
               fmt2d_lsqr_stats* stats) {
  if (!p || !dm_out) return fail(FMT2D_EINVAL, "NULL argument");
  double* outs[1] = {dm_out};
  return lsqr_solve_many(ctx, p, dm_sum, 1, &p->wdv, &p->wdv_iface, outs, stats);
}


