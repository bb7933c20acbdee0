// C ABI of the tomography system and its LSQR solve: data selection and
// weighting (read_frechet_data2d and the scaling of lsqr_inversion2d,
// src/inversion2d.f90:280-486, 3965-4052) and run_lsqr (:6578-7167).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "ctx.h"

extern "C" {

int fmt2d_build_system(fmt2d_ctx* ctx, const double* t_obs, const fmt2d_system_params* p, int* ndata, int* nnz,
                       double* eps_min) {
  if (!ctx || !t_obs || !p || !ndata || !nnz || !eps_min) return fail(FMT2D_EINVAL, "NULL argument");
  if (ctx->nray < 1) return fail(FMT2D_ESTATE, "no rays traced");
  if (p->alt_mode < 0 || p->alt_mode > 2) return fail(FMT2D_EINVAL, "alt_mode must be 0, 1 or 2");
  HIPCHK(hipSetDevice(ctx->device));
  const int rc0 = frechet_build(ctx);
  if (rc0) return rc0;
  hipStream_t st = ctx->stream;
  const HostModel& m = ctx->m;
  const int nray = ctx->nray, nparm = ctx->n_par, n_vel = ctx->n_vel_par;
  // rows kept: valid ray, observed pick, |obs - calc| <= crit_chi
  std::vector<int> rp_all((size_t)nray + 1);
  HIPCHK(hipMemcpy(rp_all.data(), ctx->g_rp.p, sizeof(int) * (nray + 1), hipMemcpyDeviceToHost));
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
  // parameter scaling: 1/max(v_init,1) on the velocity grid, refl_weight on
  // interface nodes, zero for the parameters the alternating mode fixes
  std::vector<double> mvs(nparm, 1.0);
  {
    std::vector<int> inv((size_t)ctx->hgeom.vnx * ctx->hgeom.vnz);
    HIPCHK(hipMemcpy(inv.data(), ctx->rinv.p, sizeof(int) * inv.size(), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < inv.size(); q++)
      if (inv[q] > 0) mvs[inv[q] - 1] = 1.0 / std::max(m.vgrid[q], 1.0);
    for (int j = n_vel; j < nparm; j++) mvs[j] = p->refl_weight;
    if (p->alt_mode == 1) for (int j = n_vel; j < nparm; j++) mvs[j] = 0.0;
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
  HIPCHK(launch_scale_rows(ctx->g_rp.p, ctx->g_col.p, ctx->g_val.p, dkeep.p, dwt.p, ctx->s_rp.p, nd, 0, ctx->s_mvs.p,
                           nparm, ctx->s_col.p, ctx->s_val.p, st));
  HIPCHK(ctx->s_cp.alloc((size_t)nparm + 1));
  HIPCHK(ctx->s_crow.alloc((size_t)std::max(nz_, 1)));
  HIPCHK(ctx->s_cval.alloc((size_t)std::max(nz_, 1)));
  HIPCHK(build_csc(ctx->s_rp.p, ctx->s_col.p, ctx->s_val.p, nd, nparm, nz_, ctx->s_cp.p, ctx->s_crow.p,
                   ctx->s_cval.p, st));
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
  if (gtg_diag) memcpy(gtg_diag, ctx->h_gtg.data(), sizeof(double) * ctx->n_par);
  if (residual && ctx->sys_ndata > 0)
    HIPCHK(hipMemcpy(residual, ctx->s_res.p, sizeof(double) * ctx->sys_ndata, hipMemcpyDeviceToHost));
  return FMT2D_OK;
}

int fmt2d_lsqr(fmt2d_ctx* ctx, const fmt2d_lsqr_params* p, const double* dm_sum, double* dm_out,
               fmt2d_lsqr_stats* stats) {
  if (!ctx || !p || !dm_sum || !dm_out) return fail(FMT2D_EINVAL, "NULL argument");
  if (ctx->sys_ndata < 0) return fail(FMT2D_ESTATE, "no system built");
  if (p->smooth_type != 1)
    return fail(FMT2D_EINVAL, "smooth_type 1 (Laplacian) is supported; Gaussian smoothing (2, 3) is not");
  if (p->max_iter < 1) return fail(FMT2D_EINVAL, "max_iter must be >= 1");
  HIPCHK(hipSetDevice(ctx->device));
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t st = ctx->stream;
  const int nparm = ctx->n_par, n_vel = ctx->n_vel_par, nd = ctx->sys_ndata;
  const double cut = ctx->sys_eps_min * 0.01;
  std::vector<uint8_t> mask(nparm, 0);
  std::vector<double> ws(nparm), wd(nparm);
  for (int j = 0; j < nparm; j++) {
    const bool vel = j < n_vel;
    mask[j] = (vel && ctx->h_gtg[j] < cut) ? 1 : 0;
    ws[j] = vel ? p->w_smooth : p->w_smooth_iface;
    wd[j] = vel ? p->wdv : p->wdv_iface;
  }
  // Laplacian tables (apply_laplacian_smooth_2d, inversion2d.f90:1161-1249)
  std::vector<int4> lap(std::max(n_vel, 1));
  std::vector<int2> ifn(std::max(nparm - n_vel, 1), make_int2(0, 0));
  {
    const int vnx = ctx->hgeom.vnx, vnz = ctx->hgeom.vnz;
    std::vector<int> inv((size_t)vnx * vnz);
    HIPCHK(hipMemcpy(inv.data(), ctx->rinv.p, sizeof(int) * inv.size(), hipMemcpyDeviceToHost));
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
      const int start = ctx->hgeom.if_start[k], nxi = ctx->hgeom.if_nnode[k];
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
  std::vector<double> dms(dm_sum, dm_sum + nparm);
  HIPCHK(upload(ctx->l_mask, mask, st));
  HIPCHK(upload(ctx->l_ws, ws, st));
  HIPCHK(upload(ctx->l_wd, wd, st));
  HIPCHK(upload(ctx->l_lapnb, lap, st));
  HIPCHK(upload(ctx->l_ifnb, ifn, st));
  HIPCHK(upload(ctx->l_dmsum, dms, st));
  const size_t na = (size_t)nd + 2 * (size_t)nparm;
  if (ctx->l_u.n < na) HIPCHK(ctx->l_u.alloc(na));
  if (ctx->l_v.n < (size_t)nparm) {
    HIPCHK(ctx->l_v.alloc(nparm)); HIPCHK(ctx->l_w.alloc(nparm)); HIPCHK(ctx->l_dm.alloc(nparm));
    HIPCHK(ctx->l_gtu.alloc(nparm)); HIPCHK(ctx->l_sm.alloc(nparm));
  }
  if (ctx->l_gv.n < (size_t)std::max(nd, 1)) HIPCHK(ctx->l_gv.alloc((size_t)std::max(nd, 1)));
  if (!ctx->l_sc.p) HIPCHK(ctx->l_sc.alloc(1));
  LsqrSys S;
  S.ndata = nd; S.nparm = nparm; S.n_vel = n_vel;
  S.rp = ctx->s_rp.p; S.col = ctx->s_col.p; S.val = ctx->s_val.p;
  S.cp = ctx->s_cp.p; S.crow = ctx->s_crow.p; S.cval = ctx->s_cval.p;
  S.res = ctx->s_res.p; S.dm_sum = ctx->l_dmsum.p; S.mask = ctx->l_mask.p; S.ws = ctx->l_ws.p; S.wd = ctx->l_wd.p;
  S.lap_nb = ctx->l_lapnb.p; S.if_nb = ctx->l_ifnb.p;
  S.smooth_on = (p->w_smooth > 0.0 || p->w_smooth_iface > 0.0) ? 1 : 0;
  S.damp_on = (p->wdv > 0.0 || p->wdv_iface > 0.0) ? 1 : 0;
  LsqrWork W{ctx->l_u.p, ctx->l_v.p, ctx->l_w.p, ctx->l_dm.p, ctx->l_gv.p, ctx->l_gtu.p, ctx->l_sm.p, ctx->l_sc.p};
  LsqrScalars out{};
  HIPCHK(lsqr_run(S, W, p->max_iter, 1.0e-4, 8, st, &out));
  HIPCHK(hipMemcpy(dm_out, ctx->l_dm.p, sizeof(double) * nparm, hipMemcpyDeviceToHost));
  if (stats) {
    stats->iterations = out.iters;
    stats->rnorm = out.rnorm;
    stats->bnorm = out.bnorm;
    stats->test2 = out.test2;
    stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return FMT2D_OK;
}

}  // extern "C"
