// C ABI of the ray tracer and the Frechet rows: the parameterisation of the
// inversion (frechet.in), ray backtracing on the resident fields of the last
// forward, CSR assembly.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "csr.h"
#include "ctx.h"

namespace {

const int PATH_NSEC[6] = {0, 1, 2, 2, 4, 4};

// Inversion parameters of initialize_inversion2d (frechet2d.f90:14-118):
// the resampled velocity grid's active nodes (below the seafloor,
// invert2d_main.f90:346-363 and :461-475), numbered iz-outer / ix-inner, then
// each inverted interface's input nodes as one block.
int build_mesh(fmt2d_ctx* ctx);

int build_ray_geom(fmt2d_ctx* ctx) {
  HostModel& m = ctx->m;
  hipStream_t st = ctx->stream;
  RayGeomDev& G = ctx->hgeom;
  memset(&G, 0, sizeof(G));
  G.nx = m.nx; G.nz = m.nz; G.dx = m.dx; G.dz = m.dz; G.x0 = m.x0; G.z0 = m.z0;
  G.xg = ctx->xg.p; G.zg = ctx->zg.p;
  for (int k = 1; k <= NIF; k++) {
    HIPCHK(upload(ctx->risx[k], m.isec[k].x, st));
    HIPCHK(upload(ctx->risz[k], m.isec[k].z, st));
    G.isx[k] = ctx->risx[k].p; G.isz[k] = ctx->risz[k].p; G.isn[k] = (int)m.isec[k].x.size();
  }
  for (int r = 1; r <= NREG; r++) {
    const HostRegion& R = m.reg[r];
    HIPCHK(upload(ctx->rorder[r], R.ref_order, st));
    RayRegionDev& d = G.reg[r];
    d.id = r; d.nnode = R.nnode; d.itop = R.itop; d.ibot = R.ibot;
    d.order = ctx->rorder[r].p; d.flags = ctx->rflags[r].p; d.vel = ctx->rvel[r].p;
  }
  G.reg[0] = G.reg[1];
  HIPCHK(upload(ctx->rsfx, m.if_x[2], st));
  HIPCHK(upload(ctx->rsfz, m.if_z[2], st));
  HIPCHK(upload(ctx->rsfz0, m.sf_setup, st));
  G.sfx = ctx->rsfx.p; G.sfz = ctx->rsfz.p; G.sfz_setup = ctx->rsfz0.p; G.nsf = (int)m.if_x[2].size();
  G.sfdx = m.if_x[2][1] - m.if_x[2][0];
  G.vwater = m.vwater;
  HIPCHK(upload(ctx->rpvel, m.vgrid, st));
  G.pvel = ctx->rpvel.p;
  // inversion velocity grid
  const int vnx = (int)m.vx.size(), vnz = m.nz;
  G.vnx = vnx; G.vnz = vnz; G.vdx = m.vx[1] - m.vx[0]; G.vdz = m.dz;
  std::vector<double> vz(vnz);
  for (int j = 0; j < vnz; j++) vz[j] = (double)j * m.dz;
  HIPCHK(upload(ctx->rvx, m.vx, st));
  HIPCHK(upload(ctx->rvz, vz, st));
  G.vx = ctx->rvx.p; G.vz = ctx->rvz.p;
  std::vector<int> inv((size_t)vnx * vnz, 0);
  int np = 0;
  // the numbering is fixed when the inversion is set up (a later seafloor
  // update does not renumber, frechet2d.f90:44-60)
  if (ctx->topo_init.size() != m.if_z[2].size()) ctx->topo_init = m.if_z[2];
  const std::vector<double>& topo = ctx->topo_init;
  if (ctx->inv_vel) {
    for (int iz = 1; iz <= vnz; iz++)
      for (int ix = 1; ix <= vnx; ix++) {
        int ivx = ix;
        if (ivx >= vnx) ivx = vnx - 1;
        double wx = (double)(ix - ivx);
        wx = std::min(1.0, std::max(0.0, wx));
        const double zsf = (1.0 - wx) * topo[ivx - 1] + wx * topo[ivx];
        if (vz[iz - 1] - zsf >= 0.0) inv[(size_t)(iz - 1) * vnx + (ix - 1)] = ++np;
      }
  }
  HIPCHK(upload(ctx->rinv, inv, st));
  ctx->h_inv = inv;
  G.inv_idx = ctx->rinv.p;
  G.n_vel_par = np;
  int group = ctx->inv_vel ? 1 : 0, last = np;
  for (int k = 0; k <= NIF; k++) G.if_start[k] = -1;
  for (int k : ctx->inv_ifaces) {
    group++;
    const int start = group == 1 ? 0 : last;
    const int nn = (int)m.if_x[k].size();
    HIPCHK(upload(ctx->rifx_in[k], m.if_x[k], st));
    G.if_start[k] = start; G.if_nnode[k] = nn; G.ifx_in[k] = ctx->rifx_in[k].p;
    G.ifdx_in[k] = m.if_x[k][1] - m.if_x[k][0];
    last = start + nn;
    np += nn;
  }
  G.n_par = np;
  ctx->n_vel_par = G.n_vel_par;
  ctx->n_par = np;
  std::vector<RayGeomDev> one(1, G);
  HIPCHK(upload(ctx->rgeom, one, st));
  HIPCHK(hipStreamSynchronize(st));
  if (!ctx->mesh_ok) {
    const int rc = build_mesh(ctx);
    if (rc) return rc;
  }
  ctx->rgeom_ok = true;
  return FMT2D_OK;
}

// The inversion mesh and its prolongation P onto the Frechet columns.  The
// numbering rules are those of the grid parameters: velocity nodes that are
// parameters at set-up, iz-outer / ix-inner, then each inverted interface's
// nodes as one block.  At strides 1 the mesh is the grid and P = I.
int build_mesh(fmt2d_ctx* ctx) {
  InvMesh& M = ctx->mesh;
  const RayGeomDev& G = ctx->hgeom;
  const HostModel& m = ctx->m;
  const int sv = M.vstride, si = M.istride, fnx = G.vnx, fnz = G.vnz;
  if ((fnx - 1) % sv != 0 || (fnz - 1) % sv != 0)
    return fail(FMT2D_EINVAL, "velocity mesh stride " + std::to_string(sv) + " does not divide the grid's " +
                                  std::to_string(fnx - 1) + " x " + std::to_string(fnz - 1) + " cells");
  M.vnx = (fnx - 1) / sv + 1;
  M.vnz = (fnz - 1) / sv + 1;
  M.vx.resize(M.vnx);
  M.vz.resize(M.vnz);
  for (int a = 0; a < M.vnx; a++) M.vx[a] = m.vx[(size_t)a * sv];
  for (int b = 0; b < M.vnz; b++) M.vz[b] = (double)(b * sv) * m.dz;
  const std::vector<int>& finv = ctx->h_inv;
  M.inv.assign((size_t)M.vnx * M.vnz, 0);
  int np = 0;
  for (int b = 0; b < M.vnz; b++)
    for (int a = 0; a < M.vnx; a++)
      if (finv[(size_t)b * sv * fnx + (size_t)a * sv] > 0) M.inv[(size_t)b * M.vnx + a] = ++np;
  M.n_vel = np;
  int group = ctx->inv_vel ? 1 : 0, last = np;
  for (int k = 0; k <= NIF; k++) { M.if_start[k] = -1; M.if_nnode[k] = 0; M.ifx[k].clear(); }
  for (int k : ctx->inv_ifaces) {
    const int nn = G.if_nnode[k];
    if ((nn - 1) % si != 0)
      return fail(FMT2D_EINVAL, "interface mesh stride " + std::to_string(si) + " does not divide interface " +
                                    std::to_string(k) + "'s " + std::to_string(nn - 1) + " node spacings");
    group++;
    const int start = group == 1 ? 0 : last;
    const int nc = (nn - 1) / si + 1;
    M.ifx[k].resize(nc);
    for (int c = 0; c < nc; c++) M.ifx[k][c] = m.if_x[k][(size_t)c * si];
    M.if_start[k] = start;
    M.if_nnode[k] = nc;
    last = start + nc;
    np += nc;
  }
  M.n_par = np;
  // P, one row per Frechet column in column order
  M.p_rp.assign(1, 0);
  M.p_col.clear();
  M.p_w.clear();
  auto put = [&](int k, double w) {
    if (w != 0.0) { M.p_col.push_back(k); M.p_w.push_back(w); }
  };
  for (int iz = 0; iz < fnz; iz++)
    for (int ix = 0; ix < fnx; ix++) {
      if (finv[(size_t)iz * fnx + ix] <= 0) continue;
      const int a = ix / sv, b = iz / sv;
      const double wx = (double)(ix % sv) / (double)sv, wz = (double)(iz % sv) / (double)sv;
      const int a1 = std::min(a + 1, M.vnx - 1), b1 = std::min(b + 1, M.vnz - 1);
      auto par = [&](int aa, int bb) { return M.inv[(size_t)bb * M.vnx + aa] - 1; };
      if (par(a, b) >= 0) put(par(a, b), (1.0 - wx) * (1.0 - wz));
      if (wx != 0.0 && par(a1, b) >= 0) put(par(a1, b), wx * (1.0 - wz));
      if (wz != 0.0 && par(a, b1) >= 0) put(par(a, b1), (1.0 - wx) * wz);
      if (wx != 0.0 && wz != 0.0 && par(a1, b1) >= 0) put(par(a1, b1), wx * wz);
      M.p_rp.push_back((int)M.p_col.size());
    }
  for (int k : ctx->inv_ifaces) {
    for (int ix = 0; ix < G.if_nnode[k]; ix++) {
      const int c = ix / si;
      const double w = (double)(ix % si) / (double)si;
      put(M.if_start[k] + c, 1.0 - w);
      if (w != 0.0) put(M.if_start[k] + c + 1, w);
      M.p_rp.push_back((int)M.p_col.size());
    }
  }
  if ((int)M.p_rp.size() != ctx->n_par + 1) return fail(FMT2D_EINVAL, "prolongation does not cover the Frechet columns");
  hipStream_t st = ctx->stream;
  HIPCHK(upload(ctx->p_rp, M.p_rp, st));
  HIPCHK(upload(ctx->p_col, M.p_col, st));
  HIPCHK(upload(ctx->p_w, M.p_w, st));
  HIPCHK(hipStreamSynchronize(st));
  ctx->gp_nnz = -1;
  ctx->mesh_ok = true;
  return FMT2D_OK;
}

}  // namespace

// Frechet rows of the last trace, assembled on the device once per trace.
int frechet_build(fmt2d_ctx* ctx) {
  if (ctx->g_nnz >= 0) return FMT2D_OK;
  if (ctx->fwd_program) return fail(FMT2D_ESTATE, "Frechet rows need a forward without FMT2D_FWD_PROGRAM");
  hipStream_t st = ctx->stream;
  const int nray = ctx->nray;
  // pass 1: contributions per ray
  DBuf<long long> cnt, coff;
  HIPCHK(cnt.alloc(nray));
  HIPCHK(launch_frechet_contrib(ctx->rgeom.p, ctx->rfields.p, ctx->rjobs.p, nray, 0, ctx->rpoff.p, ctx->rpts.p,
                                ctx->rout.p, nullptr, cnt.p, nullptr, nullptr, st));
  std::vector<long long> hc(nray), ho(nray);
  HIPCHK(hipMemcpyAsync(hc.data(), cnt.p, sizeof(long long) * nray, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  long long total = 0;
  for (int r = 0; r < nray; r++) { ho[r] = total; total += hc[r]; }
  HIPCHK(upload(coff, ho, st));
  DBuf<unsigned long long> keys;
  DBuf<double> vals;
  HIPCHK(keys.alloc((size_t)std::max<long long>(total, 1)));
  HIPCHK(vals.alloc((size_t)std::max<long long>(total, 1)));
  HIPCHK(launch_frechet_contrib(ctx->rgeom.p, ctx->rfields.p, ctx->rjobs.p, nray, 0, ctx->rpoff.p, ctx->rpts.p,
                                ctx->rout.p, coff.p, nullptr, keys.p, vals.p, st));
  int rbits = 1;
  while ((1LL << rbits) < nray) rbits++;
  CsrOut co;
  const hipError_t e = build_csr(keys.p, vals.p, total, nray, 32 + rbits, 0, co, st);
  if (e != hipSuccess) {
    if (co.rowptr) (void)hipFree(co.rowptr);
    if (co.col) (void)hipFree(co.col);
    if (co.val) (void)hipFree(co.val);
    return fail(FMT2D_EHIP, std::string("CSR assembly: ") + hipGetErrorString(e));
  }
  ctx->g_rp.release(); ctx->g_col.release(); ctx->g_val.release();
  ctx->g_rp.p = co.rowptr; ctx->g_rp.n = (size_t)nray + 1;
  ctx->g_col.p = co.col; ctx->g_col.n = (size_t)std::max(co.nnz, 1);
  ctx->g_val.p = co.val; ctx->g_val.n = (size_t)std::max(co.nnz, 1);
  ctx->g_nnz = co.nnz;
  ctx->gt_ok = false;
  ctx->gp_nnz = -1;
  return FMT2D_OK;
}

// G P on the device: every Frechet entry times its column's P row, written
// in row order (entries in stored order, P rows in order) and summed per
// (row, mesh parameter) front to back by the same stable CSR assembly.  At
// strides 1 each entry is multiplied by 1.0 alone, so G P = G bit for bit.
int system_rows_build(fmt2d_ctx* ctx) {
  const int rc0 = frechet_build(ctx);
  if (rc0) return rc0;
  if (ctx->gp_nnz >= 0) return FMT2D_OK;
  if (!ctx->mesh_ok)
    return fail(FMT2D_ESTATE, "the parameterisation changed after the last trace: run a forward and trace again");
  hipStream_t st = ctx->stream;
  const int nray = ctx->nray;
  DBuf<long long> cnt, coff;
  HIPCHK(cnt.alloc(nray));
  HIPCHK(launch_prolong_rows(ctx->g_rp.p, ctx->g_col.p, ctx->g_val.p, nray, ctx->p_rp.p, ctx->p_col.p, ctx->p_w.p,
                             nullptr, cnt.p, nullptr, nullptr, st));
  std::vector<long long> hc(nray), ho(nray);
  HIPCHK(hipMemcpyAsync(hc.data(), cnt.p, sizeof(long long) * nray, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  long long total = 0;
  for (int r = 0; r < nray; r++) { ho[r] = total; total += hc[r]; }
  HIPCHK(upload(coff, ho, st));
  DBuf<unsigned long long> keys;
  DBuf<double> vals;
  HIPCHK(keys.alloc((size_t)std::max<long long>(total, 1)));
  HIPCHK(vals.alloc((size_t)std::max<long long>(total, 1)));
  HIPCHK(launch_prolong_rows(ctx->g_rp.p, ctx->g_col.p, ctx->g_val.p, nray, ctx->p_rp.p, ctx->p_col.p, ctx->p_w.p,
                             coff.p, nullptr, keys.p, vals.p, st));
  int rbits = 1;
  while ((1LL << rbits) < nray) rbits++;
  CsrOut co;
  const hipError_t e = build_csr(keys.p, vals.p, total, nray, 32 + rbits, 0, co, st);
  if (e != hipSuccess) {
    if (co.rowptr) (void)hipFree(co.rowptr);
    if (co.col) (void)hipFree(co.col);
    if (co.val) (void)hipFree(co.val);
    return fail(FMT2D_EHIP, std::string("G P assembly: ") + hipGetErrorString(e));
  }
  ctx->gp_rp.release(); ctx->gp_col.release(); ctx->gp_val.release();
  ctx->gp_rp.p = co.rowptr; ctx->gp_rp.n = (size_t)nray + 1;
  ctx->gp_col.p = co.col; ctx->gp_col.n = (size_t)std::max(co.nnz, 1);
  ctx->gp_val.p = co.val; ctx->gp_val.n = (size_t)std::max(co.nnz, 1);
  ctx->gp_nnz = co.nnz;
  return FMT2D_OK;
}

extern "C" {

int fmt2d_param_index(fmt2d_ctx* ctx, int* node_param, int cap, int* n) {
  if (!ctx || !n) return fail(FMT2D_EINVAL, "NULL argument");
  if (!ctx->have_model) return fail(FMT2D_ESTATE, "no model set");
  HIPCHK(hipSetDevice(ctx->device));
  if (!ctx->rgeom_ok) { const int rc = build_ray_geom(ctx); if (rc) return rc; }
  *n = (int)ctx->h_inv.size();
  if (node_param && cap > 0) {
    if (cap < *n) return fail(FMT2D_EINVAL, "output buffer too small");
    memcpy(node_param, ctx->h_inv.data(), sizeof(int) * ctx->h_inv.size());
  }
  return FMT2D_OK;
}

int fmt2d_frechet_apply(fmt2d_ctx* ctx, int transpose, const double* x, double* y) {
  if (!ctx || !x || !y) return fail(FMT2D_EINVAL, "NULL argument");
  if (transpose != 0 && transpose != 1) return fail(FMT2D_EINVAL, "transpose must be 0 or 1");
  if (ctx->nray < 1) return fail(FMT2D_ESTATE, "no rays traced");
  HIPCHK(hipSetDevice(ctx->device));
  const int rc0 = frechet_build(ctx);
  if (rc0) return rc0;
  hipStream_t st = ctx->stream;
  const int nray = ctx->nray, np = ctx->n_par, nnz = ctx->g_nnz;
  if (transpose && !ctx->gt_ok) {
    HIPCHK(ctx->gt_cp.alloc((size_t)np + 1));
    HIPCHK(ctx->gt_row.alloc((size_t)std::max(nnz, 1)));
    HIPCHK(ctx->gt_val.alloc((size_t)std::max(nnz, 1)));
    HIPCHK(build_csc(ctx->g_rp.p, ctx->g_col.p, ctx->g_val.p, nray, np, nnz, ctx->gt_cp.p, ctx->gt_row.p,
                     ctx->gt_val.p, st));
    ctx->gt_ok = true;
  }
  const int nin = transpose ? nray : np, nout = transpose ? np : nray;
  DBuf<double> dx, dy;
  HIPCHK(dx.alloc((size_t)std::max(nin, 1)));
  HIPCHK(dy.alloc((size_t)std::max(nout, 1)));
  HIPCHK(hipMemcpyAsync(dx.p, x, sizeof(double) * nin, hipMemcpyHostToDevice, st));
  if (transpose)
    HIPCHK(launch_ordered_spmv(ctx->gt_cp.p, ctx->gt_row.p, ctx->gt_val.p, np, dx.p, dy.p, st));
  else
    HIPCHK(launch_ordered_spmv(ctx->g_rp.p, ctx->g_col.p, ctx->g_val.p, nray, dx.p, dy.p, st));
  HIPCHK(hipMemcpyAsync(y, dy.p, sizeof(double) * nout, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return FMT2D_OK;
}

int fmt2d_set_inversion(fmt2d_ctx* ctx, int invert_velocity, int n_iface, const int* iface_ids) {
  if (!ctx || (n_iface > 0 && !iface_ids) || n_iface < 0 || n_iface > 3)
    return fail(FMT2D_EINVAL, "bad inversion parameterisation");
  std::vector<int> ids;
  for (int k = 0; k < n_iface; k++) {
    if (iface_ids[k] < 2 || iface_ids[k] > 4) return fail(FMT2D_EINVAL, "invertible interfaces are 2, 3, 4");
    ids.push_back(iface_ids[k]);
  }
  ctx->inv_vel = invert_velocity ? 1 : 0;
  ctx->inv_ifaces = ids;
  ctx->rgeom_ok = false;
  ctx->mesh.vstride = 1;
  ctx->mesh.istride = 1;
  ctx->mesh_ok = false;
  // velocity_init of the inversion: parameter scaling uses the model of the
  // set-up, not the updated one (inversion2d.f90:186-191, 451-463)
  ctx->v_init = ctx->m.vgrid;
  ctx->topo_init = ctx->m.if_z[2];
  ctx->gs_nparm = -1;
  ctx->gsi_nparm = -1;
  return FMT2D_OK;
}

int fmt2d_set_inversion_mesh(fmt2d_ctx* ctx, int vel_stride, int iface_stride) {
  if (!ctx) return fail(FMT2D_EINVAL, "NULL argument");
  if (vel_stride < 1 || iface_stride < 1) return fail(FMT2D_EINVAL, "mesh strides must be >= 1");
  if (ctx->have_model) {
    const HostModel& m = ctx->m;
    if ((m.vx.size() - 1) % vel_stride != 0 || (m.nz - 1) % vel_stride != 0)
      return fail(FMT2D_EINVAL, "velocity mesh stride " + std::to_string(vel_stride) + " does not divide the grid's " +
                                    std::to_string(m.vx.size() - 1) + " x " + std::to_string(m.nz - 1) + " cells");
    for (int k : ctx->inv_ifaces)
      if ((m.if_x[k].size() - 1) % iface_stride != 0)
        return fail(FMT2D_EINVAL, "interface mesh stride " + std::to_string(iface_stride) +
                                      " does not divide interface " + std::to_string(k) + "'s node spacings");
  }
  ctx->mesh.vstride = vel_stride;
  ctx->mesh.istride = iface_stride;
  ctx->mesh_ok = false;
  ctx->rgeom_ok = false;
  ctx->gs_nparm = -1;
  ctx->gsi_nparm = -1;
  ctx->sys_ndata = -1;
  return FMT2D_OK;
}

int fmt2d_inversion_info(fmt2d_ctx* ctx, int* n_vel_par, int* n_par) {
  if (!ctx || !n_vel_par || !n_par) return fail(FMT2D_EINVAL, "NULL argument");
  if (!ctx->have_model) return fail(FMT2D_ESTATE, "no model set");
  HIPCHK(hipSetDevice(ctx->device));
  if (!ctx->rgeom_ok) { const int rc = build_ray_geom(ctx); if (rc) return rc; }
  *n_vel_par = ctx->mesh.n_vel;
  *n_par = ctx->mesh.n_par;
  return FMT2D_OK;
}

int fmt2d_mesh_info(fmt2d_ctx* ctx, int* vnx, int* vnz, int* n_frechet) {
  if (!ctx || !vnx || !vnz || !n_frechet) return fail(FMT2D_EINVAL, "NULL argument");
  if (!ctx->have_model) return fail(FMT2D_ESTATE, "no model set");
  HIPCHK(hipSetDevice(ctx->device));
  if (!ctx->rgeom_ok) { const int rc = build_ray_geom(ctx); if (rc) return rc; }
  *vnx = ctx->mesh.vnx;
  *vnz = ctx->mesh.vnz;
  *n_frechet = ctx->n_par;
  return FMT2D_OK;
}

int fmt2d_trace(fmt2d_ctx* ctx, int nray, const int* src, const int* path, const double* rx, const double* rz,
                double* t_rcv, int* valid) {
  if (!ctx || nray < 0 || (nray > 0 && (!src || !path || !rx || !rz || !t_rcv || !valid)))
    return fail(FMT2D_EINVAL, "NULL argument");
  if (ctx->nsrc < 1) return fail(FMT2D_ESTATE, "no forward has run");
  HIPCHK(hipSetDevice(ctx->device));
  if (!ctx->rgeom_ok) { const int rc = build_ray_geom(ctx); if (rc) return rc; }
  HostModel& m = ctx->m;
  hipStream_t st = ctx->stream;
  const size_t nslot = (size_t)(m.nz + 2) * m.nx;
  // stage fields of every source (final gradients after the post-passes)
  std::vector<RayField> fl((size_t)ctx->nsrc * FMT2D_NTF);
  for (int s = 0; s < ctx->nsrc; s++)
    for (int itf = 1; itf <= FMT2D_NTF; itf++) {
      const int f = ctx->field_of[s * FMT2D_NTF + itf - 1];
      RayField& r = fl[(size_t)s * FMT2D_NTF + itf - 1];
      r.T = ctx->fT.p + (size_t)f * nslot;
      r.GX = ctx->fOGX.p + (size_t)f * nslot;
      r.GZ = ctx->fOGZ.p + (size_t)f * nslot;
      r.region = TF_REGION[itf];
    }
  HIPCHK(upload(ctx->rfields, fl, st));
  HIPCHK(upload(ctx->d_srcx, ctx->src_x, st));
  HIPCHK(upload(ctx->d_srcz, ctx->src_z, st));
  ctx->hjobs.resize(nray);
  ctx->hpoff.resize((size_t)nray + 1);
  long long off = 0;
  for (int r = 0; r < nray; r++) {
    if (src[r] < 0 || src[r] >= ctx->nsrc) return fail(FMT2D_EINVAL, "ray source index out of range");
    if (path[r] < 1 || path[r] > 5) return fail(FMT2D_EINVAL, "ray path must be 1..5");
    ctx->hjobs[r] = RayJob{src[r], path[r], rx[r], rz[r]};
    ctx->hpoff[r] = off;
    off += (long long)PATH_NSEC[path[r]] * RAY_MAXP;
  }
  ctx->hpoff[nray] = off;
  ctx->nray = nray;
  ctx->g_nnz = -1;
  ctx->gp_nnz = -1;
  ctx->sys_ndata = -1;
  if (nray == 0) return FMT2D_OK;
  HIPCHK(upload(ctx->rjobs, ctx->hjobs, st));
  HIPCHK(upload(ctx->rpoff, ctx->hpoff, st));
  if (ctx->rpts.n < (size_t)off) HIPCHK(ctx->rpts.alloc((size_t)off));
  if (ctx->rout.n < (size_t)nray) HIPCHK(ctx->rout.alloc((size_t)nray));
  const int nsf = ctx->hgeom.nsf;
  if (ctx->rtsf.n < (size_t)ctx->nsrc * 3 * nsf) HIPCHK(ctx->rtsf.alloc((size_t)ctx->nsrc * 3 * nsf));
  HIPCHK(launch_seafloor_times(ctx->rgeom.p, ctx->rfields.p, ctx->nsrc, nsf, ctx->rtsf.p, st));
  HIPCHK(launch_trace(ctx->rgeom.p, ctx->rfields.p, ctx->d_srcx.p, ctx->d_srcz.p, ctx->rtsf.p, ctx->rjobs.p, nray,
                      ctx->rpoff.p, ctx->rpts.p, ctx->rout.p, ctx->fwd_program ? 1 : 0, st));
  std::vector<RayOut> out(nray);
  HIPCHK(hipMemcpyAsync(out.data(), ctx->rout.p, sizeof(RayOut) * nray, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  ctx->h_tcalc.resize(nray);
  ctx->h_valid.resize(nray);
  for (int r = 0; r < nray; r++) {
    t_rcv[r] = out[r].t; valid[r] = out[r].valid;
    ctx->h_tcalc[r] = out[r].t; ctx->h_valid[r] = out[r].valid;
  }
  ctx->g_nnz = -1;
  ctx->gp_nnz = -1;
  ctx->sys_ndata = -1;
  return FMT2D_OK;
}

int fmt2d_ray_info(fmt2d_ctx* ctx, int iray, double* entry_xz, double* water_time, int* npts, int* flags) {
  if (!ctx || !entry_xz || !water_time || !npts || !flags) return fail(FMT2D_EINVAL, "NULL argument");
  if (iray < 0 || iray >= ctx->nray) return fail(FMT2D_EINVAL, "ray index out of range");
  HIPCHK(hipSetDevice(ctx->device));
  RayOut o;
  HIPCHK(hipMemcpy(&o, ctx->rout.p + iray, sizeof(o), hipMemcpyDeviceToHost));
  entry_xz[0] = o.ex; entry_xz[1] = o.ez;
  *water_time = o.water_time;
  for (int k = 0; k < RAY_MAXSEC; k++) npts[k] = o.npts[k];
  flags[0] = o.headwave; flags[1] = o.diffracted; flags[2] = o.direct;
  return FMT2D_OK;
}

int fmt2d_ray_points(fmt2d_ctx* ctx, int iray, int section, double* x, double* z, int cap, int* npts) {
  if (!ctx || !npts) return fail(FMT2D_EINVAL, "NULL argument");
  if (iray < 0 || iray >= ctx->nray) return fail(FMT2D_EINVAL, "ray index out of range");
  const int nsec = PATH_NSEC[ctx->hjobs[iray].path];
  if (section < 1 || section > nsec) return fail(FMT2D_EINVAL, "section out of range");
  HIPCHK(hipSetDevice(ctx->device));
  RayOut o;
  HIPCHK(hipMemcpy(&o, ctx->rout.p + iray, sizeof(o), hipMemcpyDeviceToHost));
  const int n = o.npts[section - 1];
  *npts = n;
  if (!x || !z || cap <= 0) return FMT2D_OK;
  if (cap < n) return fail(FMT2D_EINVAL, "point buffer too small");
  std::vector<double2> p(n);
  HIPCHK(hipMemcpy(p.data(), ctx->rpts.p + ctx->hpoff[iray] + (long long)(section - 1) * RAY_MAXP,
                   sizeof(double2) * n, hipMemcpyDeviceToHost));
  for (int k = 0; k < n; k++) { x[k] = p[k].x; z[k] = p[k].y; }
  return FMT2D_OK;
}

int fmt2d_frechet_csr(fmt2d_ctx* ctx, int index_base, int* row_ptr, int* col, double* val, int cap, int* nnz) {
  if (!ctx || !nnz) return fail(FMT2D_EINVAL, "NULL argument");
  if (index_base != 0 && index_base != 1) return fail(FMT2D_EINVAL, "index_base must be 0 or 1");
  if (ctx->nray < 1) return fail(FMT2D_ESTATE, "no rays traced");
  HIPCHK(hipSetDevice(ctx->device));
  const int rc0 = frechet_build(ctx);
  if (rc0) return rc0;
  const int nray = ctx->nray;
  *nnz = ctx->g_nnz;
  if (row_ptr) {
    std::vector<int> rp((size_t)nray + 1);
    HIPCHK(hipMemcpy(rp.data(), ctx->g_rp.p, sizeof(int) * (nray + 1), hipMemcpyDeviceToHost));
    for (int r = 0; r <= nray; r++) row_ptr[r] = rp[r] + index_base;
  }
  if (col && val) {
    if (cap < ctx->g_nnz) return fail(FMT2D_EINVAL, "CSR buffers too small");
    HIPCHK(hipMemcpy(col, ctx->g_col.p, sizeof(int) * ctx->g_nnz, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(val, ctx->g_val.p, sizeof(double) * ctx->g_nnz, hipMemcpyDeviceToHost));
    if (index_base) for (int k = 0; k < ctx->g_nnz; k++) col[k] += index_base;
  }
  return FMT2D_OK;
}

}  // extern "C"
