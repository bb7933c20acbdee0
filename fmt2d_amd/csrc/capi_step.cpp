// C ABI of the outer-loop pieces of one inversion iteration
// (lsqr_inversion2d, src/inversion2d.f90:112-930): the damping choice with
// its secant search over LSQR solves (:533-739), the clean-up and cap of the
// update (:741-773), the model update (apply_model_step, :7173-7240) and the
// line search on the RMS after a forward (:775-889, compute_rms_after_forward
// :3298-3339).  The LSQR solves, forwards and rays run on the device; the
// few scalars of the searches are host arithmetic in the reference's order.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "ctx.h"

namespace {

// calc_avg_dmv (:6523-6543): RMS of the active velocity parameters' update
double avg_dmv(const std::vector<double>& dm, const std::vector<double>& gtg, double eps_min, int n_vel) {
  double a = 0.0;
  int n = 0;
  for (int j = 0; j < n_vel; j++)
    if (gtg[j] >= eps_min * 0.01) { n++; a = a + dm[j] * dm[j]; }
  return n > 0 ? sqrt(a / (double)n) : a;
}

// calc_avg_dmd (:6549-6567): RMS of the interface depth updates mvscale * dm
double avg_dmd(const std::vector<double>& dm, const std::vector<double>& mvs, int n_vel, int nparm) {
  double a = 0.0;
  int n = 0;
  for (int j = n_vel; j < nparm; j++) {
    n++;
    const double dz = mvs[j] * dm[j];
    a = a + dz * dz;
  }
  return n > 0 ? sqrt(a / (double)n) : a;
}

std::vector<double> plus(const std::vector<double>& a, const std::vector<double>& b) {
  std::vector<double> c(a.size());
  for (size_t i = 0; i < a.size(); i++) c[i] = a[i] + b[i];
  return c;
}

// The velocity grid and inverted interface depths a line search returns to.
struct ModelState {
  std::vector<double> vel;
  std::vector<std::vector<double>> ifz;
};

ModelState save_state(const fmt2d_ctx* ctx) {
  ModelState s;
  s.vel = ctx->m.vgrid;
  for (int k : ctx->inv_ifaces) s.ifz.push_back(ctx->m.if_z[k]);
  return s;
}

inline double step_scale(double s, double mvs, double dm) { return s * mvs * dm; }

int rebuild(fmt2d_ctx* ctx) {
  const std::string e = rebuild_geometry(ctx->m);
  if (!e.empty()) return fail(FMT2D_EINVAL, e);
  return model_upload(ctx);
}

// restore_model_state (:7242-7276); the region velocities follow the grid,
// so the geometry is rebuilt, unless a model step follows at once: the step
// reads only the grid velocities and interface depths, and rebuilds itself.
int restore_state(fmt2d_ctx* ctx, const ModelState& s, bool step_follows = false) {
  ctx->m.vgrid = s.vel;
  for (size_t q = 0; q < ctx->inv_ifaces.size(); q++) ctx->m.if_z[ctx->inv_ifaces[q]] = s.ifz[q];
  return step_follows ? FMT2D_OK : rebuild(ctx);
}

// apply_model_step (:7173-7240) on the host model.  The scaled update of
// each mesh parameter, step * mvscale * dm, reaches a grid node through its
// row of P (weights in row order); at strides 1 that row is one weight 1.0,
// the reference's update itself.
int model_step(fmt2d_ctx* ctx, double step, double step_iface, const double* dm) {
  HostModel& m = ctx->m;
  const RayGeomDev& G = ctx->hgeom;
  const InvMesh& M = ctx->mesh;
  const std::vector<double>& mvs = ctx->h_mvs_model;
  const int nparm = M.n_par;
  if ((int)mvs.size() != nparm) return fail(FMT2D_ESTATE, "no system built for this parameterisation");
  if ((int)M.p_rp.size() != ctx->n_par + 1) return fail(FMT2D_ESTATE, "no inversion mesh built");
  auto prolong = [&](int col, double s) {   // Frechet column col (0-based)
    double acc = 0.0;
    bool first = true;
    for (int e = M.p_rp[col]; e < M.p_rp[col + 1]; e++) {
      const int k = M.p_col[e];
      const double t = step_scale(s, mvs[k], dm[k]) * M.p_w[e];
      acc = first ? t : acc + t;
      first = false;
    }
    return acc;
  };
  if (step > 0.0) {
    const std::vector<int>& inv = ctx->h_inv;
    if (inv.size() != (size_t)G.vnx * G.vnz) return fail(FMT2D_ESTATE, "no parameterisation set up");
    const std::vector<double>& vi = ctx->v_init.size() == m.vgrid.size() ? ctx->v_init : m.vgrid;
    for (size_t q = 0; q < inv.size(); q++) {
      const int idx = inv[q];
      if (idx <= 0) continue;
      const double cur_s = 1.0 / m.vgrid[q];
      const double trial_s = cur_s + prolong(idx - 1, step);
      double tv = 2.0;
      if (trial_s > 0.0) tv = 1.0 / trial_s;
      if (tv < 2.0) tv = 2.0;
      if (tv > 10.0) tv = 10.0;
      const double v0 = vi[q];
      if (v0 >= 2.0) {
        if (tv > v0 * 1.30) tv = v0 * 1.30;
        if (tv < v0 * 0.70) tv = v0 * 0.70;
      }
      m.vgrid[q] = tv;
    }
  }
  if (!ctx->inv_ifaces.empty() && step_iface > 0.0) {
    for (int k : ctx->inv_ifaces) {
      const int start = G.if_start[k], nxi = G.if_nnode[k];
      for (int ix = 1; ix <= nxi; ix++) {
        const int idx = start + ix;
        if (idx < 1 || idx > ctx->n_par) continue;
        double dz = prolong(idx - 1, step_iface);
        if (dz > 0.5) dz = 0.5;
        if (dz < -0.5) dz = -0.5;
        m.if_z[k][ix - 1] = m.if_z[k][ix - 1] + dz;
      }
    }
  }
  return rebuild(ctx);
}

}  // namespace

extern "C" {

int fmt2d_damped_update(fmt2d_ctx* ctx, const fmt2d_lsqr_params* lp, const fmt2d_damping_params* d,
                        const double* dm_sum, double* dm_out, fmt2d_damping_result* res) {
  if (!ctx || !lp || !d || !dm_sum || !dm_out || !res) return fail(FMT2D_EINVAL, "NULL argument");
  if (ctx->sys_ndata < 0) return fail(FMT2D_ESTATE, "no system built");
  const int nparm = ctx->mesh.n_par, n_vel = ctx->mesh.n_vel, n_if = nparm - n_vel;
  const std::vector<double>& gtg = ctx->h_gtg;
  const double eps = ctx->sys_eps_min;
  const std::vector<double>& mvs = ctx->h_mvs_smooth;
  const std::vector<double> dms(dm_sum, dm_sum + nparm);
  fmt2d_lsqr_params p = *lp;
  int solves = 0, iters = 0;
  std::vector<double> dm(nparm), trial(nparm);
  // The reference's closing solve repeats the search's last trial (same
  // system, same two weights, :726-731): the solve is deterministic, so the
  // last result is reused instead of solving again.
  std::vector<double> last_out;
  double last_wv = 0.0, last_wi = 0.0;
  int last_iters = 0;
  auto solve = [&](double wdv, double wdv_if, std::vector<double>& out) {
    if (!last_out.empty() && wdv == last_wv && wdv_if == last_wi) {
      out = last_out;
      iters = last_iters;
      return 0;
    }
    p.wdv = wdv;
    p.wdv_iface = wdv_if;
    fmt2d_lsqr_stats st{};
    const int rc = lsqr_solve(ctx, &p, dms.data(), out.data(), &st);
    if (rc) return rc;
    solves++;
    iters = st.iterations;
    last_out = out;
    last_wv = wdv;
    last_wi = wdv_if;
    last_iters = iters;
    return 0;
  };
  // two independent trials (the first two of a secant search) side by side
  auto solve2 = [&](const double* wv, const double* wi, std::vector<double>& o0, std::vector<double>& o1) {
    double* outs[2] = {o0.data(), o1.data()};
    fmt2d_lsqr_stats st[2] = {};
    const int rc = lsqr_solve_many(ctx, &p, dms.data(), 2, wv, wi, outs, st);
    if (rc) return rc;
    solves += 2;
    iters = st[1].iterations;
    last_out = o1;
    last_wv = wv[1];
    last_wi = wi[1];
    last_iters = iters;
    return 0;
  };
  double target_dv = d->target_dv > 0.0 ? d->target_dv : 0.03;
  double target_dd = d->target_dd > 0.0 ? d->target_dd : 0.5;
  double wd_best = 0.0, wd_best_if = 0.0, a_dmv = 0.0, a_dmd = 0.0;
  if (d->lam_damp > 0.0) {
    // fixed damping (:533-556)
    const double wif = (!ctx->inv_ifaces.empty() && d->lam_damp_iface > 0.0) ? d->lam_damp_iface : d->lam_damp;
    wd_best_if = wif;
    int rc = solve(d->lam_damp, wif, dm);
    if (rc) return rc;
    trial = plus(dm, dms);
    a_dmv = avg_dmv(trial, gtg, eps, n_vel);
    a_dmd = n_if > 0 ? avg_dmd(trial, mvs, n_vel, nparm) : 0.0;
    wd_best = d->lam_damp;
  } else {
    // automatic damping (:557-739): no damping first, then a secant search on
    // the interface and the velocity damping weights for the target averages
    int rc = solve(0.0, 0.0, dm);
    if (rc) return rc;
    trial = plus(dm, dms);
    a_dmv = avg_dmv(trial, gtg, eps, n_vel);
    a_dmd = n_if > 0 ? avg_dmd(trial, mvs, n_vel, nparm) : 0.0;
    // secant search of one damping weight; f(w) = average(w) - target
    auto secant = [&](bool iface, double target, double& w_out) -> int {
      const double ddm = target * 0.3;
      auto eval = [&](double w, double& f) -> int {
        const int r = iface ? solve(0.0, w, trial) : solve(w, wd_best_if, trial);
        if (r) return r;
        trial = plus(trial, dms);
        f = (iface ? avg_dmd(trial, mvs, n_vel, nparm) : avg_dmv(trial, gtg, eps, n_vel)) - target;
        return 0;
      };
      const double wd1 = 1.0, wd2 = 100.0;
      double fl, fv;
      // f at wd1 and at wd2: independent solves, run together
      std::vector<double> t2(nparm);
      const double wv[2] = {iface ? 0.0 : wd1, iface ? 0.0 : wd2};
      const double wi[2] = {iface ? wd1 : wd_best_if, iface ? wd2 : wd_best_if};
      int r = solve2(wv, wi, trial, t2);
      if (r) return r;
      auto favg = [&](std::vector<double>& out) {
        out = plus(out, dms);
        return (iface ? avg_dmd(out, mvs, n_vel, nparm) : avg_dmv(out, gtg, eps, n_vel)) - target;
      };
      fl = favg(trial);
      fv = favg(t2);
      double rts, xl;
      if (fabs(fl) < fabs(fv)) {
        rts = wd1; xl = wd2;
        const double t = fl; fl = fv; fv = t;
      } else {
        xl = wd1; rts = wd2;
      }
      for (int k = 1; k <= 10; k++) {
        const double dx = (xl - rts) * fv / (fv - fl);
        if (rts + dx <= 0.0) { xl = rts; fl = fv; rts = rts * 0.5; }
        else if (rts + dx > 1.0e7) { xl = rts; fl = fv; rts = rts + (1.0e7 - rts) * 0.5; }
        else { xl = rts; fl = fv; rts = rts + dx; }
        r = eval(rts, fv);
        if (r) return r;
        if (rts > 1.0e7) break;
        if (fabs(fv) < ddm) break;
      }
      w_out = rts;
      return 0;
    };
    if (n_if > 0 && a_dmd > target_dd) {
      rc = secant(true, target_dd, wd_best_if);
      if (rc) return rc;
    }
    if (a_dmv <= target_dv) {
      wd_best = 0.0;
    } else {
      rc = secant(false, target_dv, wd_best);
      if (rc) return rc;
    }
    rc = solve(wd_best, wd_best_if, dm);
    if (rc) return rc;
    trial = plus(dm, dms);
    a_dmv = avg_dmv(trial, gtg, eps, n_vel);
    a_dmd = n_if > 0 ? avg_dmd(trial, mvs, n_vel, nparm) : 0.0;
  }
  // inactive velocity nodes, parameters the alternating mode fixes, cap
  // (:741-773)
  for (int j = 0; j < n_vel; j++)
    if (gtg[j] < eps * 0.01) dm[j] = 0.0;
  if (ctx->sys_alt_mode == 1) for (int j = n_vel; j < nparm; j++) dm[j] = 0.0;
  if (ctx->sys_alt_mode == 2) for (int j = 0; j < n_vel; j++) dm[j] = 0.0;
  for (int j = 0; j < n_vel; j++)
    if (fabs(dm[j]) > d->max_rel_pert) dm[j] = dm[j] >= 0.0 ? fabs(d->max_rel_pert) : -fabs(d->max_rel_pert);
  memcpy(dm_out, dm.data(), sizeof(double) * nparm);
  res->wd = wd_best;
  res->wd_iface = wd_best_if;
  res->avg_dmv = a_dmv;
  res->avg_dmd = a_dmd;
  res->solves = solves;
  res->iterations = iters;
  return FMT2D_OK;
}

int fmt2d_model_get(fmt2d_ctx* ctx, int what, double* out, int cap, int* n) {
  if (!ctx || !n) return fail(FMT2D_EINVAL, "NULL argument");
  if (!ctx->have_model) return fail(FMT2D_ESTATE, "no model set");
  const std::vector<double>* v = nullptr;
  if (what == 0) v = &ctx->m.vgrid;
  else if (what >= 1 && what <= NIF) v = &ctx->m.if_z[what];
  else return fail(FMT2D_EINVAL, "what must be 0 (velocity grid) or an interface 1..5");
  *n = (int)v->size();
  if (out && cap > 0) {
    if (cap < *n) return fail(FMT2D_EINVAL, "output buffer too small");
    memcpy(out, v->data(), sizeof(double) * v->size());
  }
  return FMT2D_OK;
}

int fmt2d_model_put(fmt2d_ctx* ctx, int what, const double* in, int n) {
  if (!ctx || !in) return fail(FMT2D_EINVAL, "NULL argument");
  if (!ctx->have_model) return fail(FMT2D_ESTATE, "no model set");
  std::vector<double>* v = nullptr;
  if (what == 0) v = &ctx->m.vgrid;
  else if (what >= 2 && what <= 4) v = &ctx->m.if_z[what];
  else return fail(FMT2D_EINVAL, "what must be 0 (velocity grid) or an interface 2..4");
  if (n != (int)v->size()) return fail(FMT2D_EINVAL, "size does not match the model");
  HIPCHK(hipSetDevice(ctx->device));
  memcpy(v->data(), in, sizeof(double) * v->size());
  return rebuild(ctx);
}

int fmt2d_apply_model_step(fmt2d_ctx* ctx, double step, double step_iface, const double* dm) {
  if (!ctx || !dm) return fail(FMT2D_EINVAL, "NULL argument");
  if (!ctx->have_model) return fail(FMT2D_ESTATE, "no model set");
  HIPCHK(hipSetDevice(ctx->device));
  return model_step(ctx, step, step_iface, dm);
}

int fmt2d_line_search(fmt2d_ctx* ctx, const fmt2d_line_search_params* p, const double* t_obs, double* dm,
                      int* valid_io, fmt2d_line_search_result* res) {
  if (!ctx || !p || !t_obs || !dm || !res) return fail(FMT2D_EINVAL, "NULL argument");
  if (ctx->sys_ndata < 0) return fail(FMT2D_ESTATE, "no system built");
  if (p->first_step < 1 || p->first_step > 6) return fail(FMT2D_EINVAL, "first_step must be 1..6");
  HIPCHK(hipSetDevice(ctx->device));
  const int nparm = ctx->mesh.n_par, n_vel = ctx->mesh.n_vel, nray = ctx->nray;
  const double rms_orig = ctx->sys_rms;
  // what a trial forward needs: the sources and the rays of this iteration
  const std::vector<double> src_x = ctx->src_x;
  const std::vector<RayJob> jobs = ctx->hjobs;
  std::vector<int> js(nray), jp(nray), valid(nray);
  // a ray lost by one trial forward stays lost for the rest of the iteration
  // (ray%valid is only set again when the next iteration starts, :283-288)
  std::vector<int> alive = valid_io ? std::vector<int>(valid_io, valid_io + nray) : ctx->h_valid;
  if ((int)alive.size() != nray) return fail(FMT2D_ESTATE, "ray validity does not match the last trace");
  std::vector<double> jx(nray), jz(nray), tc(nray);
  for (int r = 0; r < nray; r++) { js[r] = jobs[r].src; jp[r] = jobs[r].path; jx[r] = jobs[r].rx; jz[r] = jobs[r].rz; }
  const ModelState saved = save_state(ctx);
  static const double LADDER[7] = {0.0, 1.0, 0.75, 0.5, 0.25, 0.125, 0.0625};
  double step_len = 0.0, rms_best = rms_orig;
  int idx_ok = 1, forwards = 0;
  ModelState trial_model;
  for (int li = p->first_step; li <= 6; li++) {
    const double s = LADDER[li];
    int rc = restore_state(ctx, saved, true);
    if (!rc) rc = model_step(ctx, s, s * p->iface_step_ratio, dm);
    if (!rc) trial_model = save_state(ctx);
    if (!rc) rc = fmt2d_forward(ctx, (int)src_x.size(), src_x.data(), 0);
    if (!rc) rc = fmt2d_trace(ctx, nray, js.data(), jp.data(), jx.data(), jz.data(), tc.data(), valid.data());
    if (rc) {
      // the caller's model is put back (restore_model_state); the trial's
      // error is the one reported
      const std::string msg = fmt2d_last_error();
      (void)restore_state(ctx, saved);
      return fail(rc, msg);
    }
    forwards++;
    // compute_rms_after_forward (:3298-3339)
    int n = 0;
    double ss = 0.0;
    for (int r = 0; r < nray; r++) alive[r] = alive[r] && valid[r];
    for (int r = 0; r < nray; r++)
      if (alive[r] && t_obs[r] >= 0.0) {
        const double rc2 = t_obs[r] - tc[r];
        if (fabs(rc2) <= p->crit_chi) { n++; ss = ss + rc2 * rc2; }
      }
    const double rms_try = n > 0 ? sqrt(ss / (double)n) : FMT_HUGE_TIME;
    if (rms_try < rms_best) { rms_best = rms_try; step_len = s; idx_ok = li; }
    if (rms_try < rms_orig) break;
  }
  int next = p->first_step;
  if (step_len > 0.0) next = idx_ok;
  int rc = restore_state(ctx, saved, step_len > 0.0);
  if (!rc && step_len > 0.0) rc = model_step(ctx, step_len, step_len * p->iface_step_ratio, dm);
  if (rc) return rc;
  // an accepted step is the last trial: when the committed model equals that
  // trial's byte for byte, its fields on the device are the model's fields
  if (step_len > 0.0) {
    const ModelState now = save_state(ctx);
    bool same = now.vel.size() == trial_model.vel.size() &&
                memcmp(now.vel.data(), trial_model.vel.data(), sizeof(double) * now.vel.size()) == 0 &&
                now.ifz.size() == trial_model.ifz.size();
    for (size_t k = 0; same && k < now.ifz.size(); k++)
      same = now.ifz[k].size() == trial_model.ifz[k].size() &&
             memcmp(now.ifz[k].data(), trial_model.ifz[k].data(), sizeof(double) * now.ifz[k].size()) == 0;
    ctx->fields_current = same;
    ctx->fields_flags = 0;
  }
  const double step_if = ctx->inv_ifaces.empty() ? 0.0 : step_len * p->iface_step_ratio;
  for (int j = 0; j < nparm; j++)
    dm[j] = (j < n_vel || ctx->inv_ifaces.empty()) ? step_len * dm[j] : step_if * dm[j];
  if (valid_io) std::copy(alive.begin(), alive.end(), valid_io);
  res->step = step_len;
  res->rms_orig = rms_orig;
  res->rms_new = step_len > 0.0 ? rms_best : rms_orig;
  res->next_first_step = next;
  res->forwards = forwards;
  return FMT2D_OK;
}

}  // extern "C"
