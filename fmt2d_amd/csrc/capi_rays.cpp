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
    // intersections with the same node x (bit for bit) point at one copy:
    // the tracer then computes their cell index and weight once
    for (int j = 1; j < k; j++)
      if (m.isec[j].x.size() == m.isec[k].x.size() &&
          (m.isec[k].x.empty() ||
           memcmp(m.isec[j].x.data(), m.isec[k].x.data(), sizeof(double) * m.isec[k].x.size()) == 0)) {
        G.isx[k] = G.isx[j];
        break;
      }
  }
  for (int r = 1; r <= NREG; r++) {
    const HostRegion& R = m.reg[r];
    HIPCHK(upload(ctx->rorder[r], R.ref_order, st));
    RayRegionDev& d = G.reg[r];
    d.id = r; d.nnode = R.nnode; d.itop = R.itop; d.ibot = R.ibot;
    d.order = ctx->rorder[r].p; d.flags = ctx->rflags[r].p; d.vel = ctx->rvel[r].p;
    // column index of the nearest-node search: the run of region rows per
    // column and bounds on the x of the nodes of the columns from / up to c
    const int nx = m.nx, nz = m.nz;
    std::vector<int> jlo(nx, 1), jhi(nx, 0);
    std::vector<double> xlo(nx), xhi(nx);
    bool runs = true;
    const HostInterface& T = m.isec[R.itop];
    const HostInterface& B = m.isec[R.ibot];
    // rows in storage order; per column: first and last region row, count
    std::vector<int> clo(nx, -1), chi(nx, -1), ccnt(nx, 0);
    for (int j = 0; j < nz; j++)
      for (int c = 0; c < nx; c++)
        if (R.flags[(size_t)j * nx + c] & RF_IN) { if (clo[c] < 0) clo[c] = j; chi[c] = j; ccnt[c]++; }
    for (int c = 0; c < nx; c++) {
      const int lo = clo[c], hi = chi[c], cnt = ccnt[c];
      if (lo >= 0) {
        jlo[c] = lo; jhi[c] = hi;
        if (cnt != hi - lo + 1) runs = false;
      }
      double a = std::min(T.x[c], B.x[c]), b = std::max(T.x[c], B.x[c]);
      if (lo >= 0) { a = std::min(a, m.xg[c]); b = std::max(b, m.xg[c]); }
      xlo[c] = a; xhi[c] = b;
    }
    for (int c = nx - 2; c >= 0; c--) xlo[c] = std::min(xlo[c], xlo[c + 1]);
    for (int c = 1; c < nx; c++) xhi[c] = std::max(xhi[c], xhi[c - 1]);
    HIPCHK(upload(ctx->rjlo[r], jlo, st));
    HIPCHK(upload(ctx->rjhi[r], jhi, st));
    HIPCHK(upload(ctx->rxlo[r], xlo, st));
    HIPCHK(upload(ctx->rxhi[r], xhi, st));
    d.indexed = (runs && (int)T.x.size() == nx && (int)B.x.size() == nx && ctx->ray_scan == 0) ? 1 : 0;
    d.jlo = ctx->rjlo[r].p; d.jhi = ctx->rjhi[r].p; d.xlo_from = ctx->rxlo[r].p; d.xhi_upto = ctx->rxhi[r].p;
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

// Sparse rows assembled in chunks of at most ctx->frechet_chunk summands:
// emit(r0, nr, offsets, keys, vals) writes the (local row << 32 | column,
// value) summands of rows r0 .. r0+nr-1, build_csr sorts and sums each
// chunk, and the chunks' rows are placed one after another.  A row lies in
// one chunk, so its sums do not depend on the chunking; the chunk bounds the
// device memory of the sort (about 40 bytes per summand).
template <class Emit>
static int assemble_rows(fmt2d_ctx* ctx, int nrow, const std::vector<long long>& cnt, Emit emit, const char* what,
                         DBuf<int>& rp_out, DBuf<int>& col_out, DBuf<double>& val_out, int& nnz_out) {
  hipStream_t st = ctx->stream;
  const long long cap = std::max<long long>(ctx->frechet_chunk, 1);
  std::vector<int> bounds(1, 0);
  long long acc = 0;
  for (int r = 0; r < nrow; r++) {
    if (acc > 0 && acc + cnt[r] > cap) { bounds.push_back(r); acc = 0; }
    acc += cnt[r];
  }
  bounds.push_back(nrow);
  std::vector<CsrOut> parts;
  auto free_parts = [&]() {
    for (CsrOut& c : parts) {
      if (c.rowptr) (void)hipFree(c.rowptr);
      if (c.col) (void)hipFree(c.col);
      if (c.val) (void)hipFree(c.val);
    }
    parts.clear();
  };
  for (size_t c = 0; c + 1 < bounds.size(); c++) {
    const int r0 = bounds[c], nr = bounds[c + 1] - r0;
    std::vector<long long> ho(std::max(nr, 1), 0);
    long long total = 0;
    for (int r = 0; r < nr; r++) { ho[r] = total; total += cnt[r0 + r]; }
    // the chunk's offsets and summands in scratch kept between assemblies
    long long* coff = nullptr;
    unsigned long long* keys = nullptr;
    double* vals = nullptr;
    hipError_t e = ctx->asm_work.get(0, ho.size(), &coff);
    if (e == hipSuccess) e = hipMemcpyAsync(coff, ho.data(), sizeof(long long) * ho.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);  // ho is a local vector
    if (e == hipSuccess) e = ctx->asm_work.get(1, (size_t)std::max<long long>(total, 1), &keys);
    if (e == hipSuccess) e = ctx->asm_work.get(2, (size_t)std::max<long long>(total, 1), &vals);
    if (e == hipSuccess) e = emit(r0, nr, coff, keys, vals);
    int rbits = 1;
    while ((1LL << rbits) < nr) rbits++;
    CsrOut co;
    if (e == hipSuccess) e = build_csr(keys, vals, total, nr, 32 + rbits, 0, co, ctx->csr_work, st);
    parts.push_back(co);
    if (e != hipSuccess) {
      free_parts();
      return fail(FMT2D_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    }
  }
  rp_out.release(); col_out.release(); val_out.release();
  if (parts.size() == 1) {
    const CsrOut& co = parts[0];
    rp_out.p = co.rowptr; rp_out.n = rp_out.cap = (size_t)nrow + 1;
    col_out.p = co.col; col_out.n = col_out.cap = (size_t)std::max(co.nnz, 1);
    val_out.p = co.val; val_out.n = val_out.cap = (size_t)std::max(co.nnz, 1);
    nnz_out = co.nnz;
    parts.clear();
    return FMT2D_OK;
  }
  long long nnz = 0;
  for (const CsrOut& co : parts) nnz += co.nnz;
  if (nnz > 2147483647LL) { free_parts(); return fail(FMT2D_EINVAL, std::string(what) + ": more than 2^31 entries"); }
  std::vector<int> rp((size_t)nrow + 1, 0), part_rp;
  hipError_t e = col_out.alloc((size_t)std::max<long long>(nnz, 1));
  if (e == hipSuccess) e = val_out.alloc((size_t)std::max<long long>(nnz, 1));
  long long off = 0;
  for (size_t c = 0; c < parts.size() && e == hipSuccess; c++) {
    const int r0 = bounds[c], nr = bounds[c + 1] - r0;
    part_rp.resize((size_t)nr + 1);
    e = hipMemcpyAsync(part_rp.data(), parts[c].rowptr, sizeof(int) * (nr + 1), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    for (int r = 0; r <= nr && e == hipSuccess; r++) rp[(size_t)r0 + r] = (int)(off + part_rp[r]);
    if (e == hipSuccess && parts[c].nnz > 0)
      e = hipMemcpyAsync(col_out.p + off, parts[c].col, sizeof(int) * parts[c].nnz, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess && parts[c].nnz > 0)
      e = hipMemcpyAsync(val_out.p + off, parts[c].val, sizeof(double) * parts[c].nnz, hipMemcpyDeviceToDevice, st);
    off += parts[c].nnz;
  }
  if (e == hipSuccess) e = upload(rp_out, rp, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  free_parts();
  if (e != hipSuccess) return fail(FMT2D_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  nnz_out = (int)nnz;
  return FMT2D_OK;
}

// Frechet rows of the last trace, assembled on the device once per trace.
int frechet_build(fmt2d_ctx* ctx) {
  if (ctx->g_nnz >= 0) return FMT2D_OK;
  if (ctx->fwd_program) return fail(FMT2D_ESTATE, "Frechet rows need a forward without FMT2D_FWD_PROGRAM");
  hipStream_t st = ctx->stream;
  const int nray = ctx->nray;
  // pass 1: contributions per ray
  DBuf<long long> cnt;
  HIPCHK(cnt.alloc(nray));
  HIPCHK(launch_frechet_contrib(ctx->rgeom.p, ctx->rfields.p, ctx->rjobs.p, nray, 0, ctx->rpoff.p, ctx->rpts.p,
                                ctx->rout.p, nullptr, cnt.p, nullptr, nullptr, st));
  std::vector<long long> hc(nray);
  HIPCHK(hipMemcpyAsync(hc.data(), cnt.p, sizeof(long long) * nray, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  auto emit = [&](int r0, int nr, const long long* coff, unsigned long long* keys, double* vals) {
    return launch_frechet_contrib(ctx->rgeom.p, ctx->rfields.p, ctx->rjobs.p, nr, r0, ctx->rpoff.p, ctx->rpts.p,
                                  ctx->rout.p, coff, nullptr, keys, vals, st);
  };
  int nnz = 0;
  const int rc = assemble_rows(ctx, nray, hc, emit, "CSR assembly", ctx->g_rp, ctx->g_col, ctx->g_val, nnz);
  if (rc) return rc;
  ctx->g_nnz = nnz;
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
  DBuf<long long> cnt;
  HIPCHK(cnt.alloc(nray));
  HIPCHK(launch_prolong_rows(ctx->g_rp.p, ctx->g_col.p, ctx->g_val.p, nray, ctx->p_rp.p, ctx->p_col.p, ctx->p_w.p,
                             nullptr, cnt.p, nullptr, nullptr, st));
  std::vector<long long> hc(nray);
  HIPCHK(hipMemcpyAsync(hc.data(), cnt.p, sizeof(long long) * nray, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  auto emit = [&](int r0, int nr, const long long* coff, unsigned long long* keys, double* vals) {
    return launch_prolong_rows(ctx->g_rp.p + r0, ctx->g_col.p, ctx->g_val.p, nr, ctx->p_rp.p, ctx->p_col.p,
                               ctx->p_w.p, coff, nullptr, keys, vals, st);
  };
  int nnz = 0;
  const int rc = assemble_rows(ctx, nray, hc, emit, "G P assembly", ctx->gp_rp, ctx->gp_col, ctx->gp_val, nnz);
  if (rc) return rc;
  ctx->gp_nnz = nnz;
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
                     ctx->gt_val.p, ctx->csr_work, st));
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

int fmt2d_set_frechet_chunk(fmt2d_ctx* ctx, long long max_summands) {
  if (!ctx || max_summands < 1) return fail(FMT2D_EINVAL, "the chunk needs at least one summand");
  ctx->frechet_chunk = max_summands;
  ctx->g_nnz = -1;
  ctx->gp_nnz = -1;
  ctx->gt_ok = false;
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
  // the work per ray of the last trace predicts this one's if the rays are the same
  bool same_rays = (int)ctx->hjobs.size() == nray && (int)ctx->hwork.size() == nray;
  for (int r = 0; same_rays && r < nray; r++) {
    const RayJob& j = ctx->hjobs[r];
    same_rays = j.src == src[r] && j.path == path[r] && j.rx == rx[r] && j.rz == rz[r];
  }
  if (!same_rays) ctx->hwork.clear();
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
  {
    // Trace order: by source, then path, input order within, so that a
    // wave's lanes trace rays of one stage chain through neighbouring
    // receivers.  When the last trace was of the same rays, its
    // interpolations per ray order them first (descending, in buckets; the
    // chain order within a bucket): waves start in order as earlier ones
    // finish, and the longest rays then start first instead of running on
    // alone at the end, and a wave's lanes do about the same work.  Stable
    // counting sorts.
    const int nkey = ctx->nsrc * 6;
    std::vector<int> start((size_t)nkey + 1, 0), by_chain(nray);
    for (int r = 0; r < nray; r++) start[(size_t)src[r] * 6 + path[r] + 1]++;
    for (int k = 0; k < nkey; k++) start[k + 1] += start[k];
    for (int r = 0; r < nray; r++) by_chain[start[(size_t)src[r] * 6 + path[r]]++] = r;
    const int NB = 1024;
    auto bucket = [&](int r) {
      if (ctx->hwork.empty()) return 0;
      const int c = ctx->hwork[r] / 32;
      return NB - 1 - (c >= NB - 1 ? NB - 1 : c);  // descending
    };
    std::vector<int> bstart(NB + 1, 0);
    for (int r = 0; r < nray; r++) bstart[bucket(r) + 1]++;
    for (int k = 0; k < NB; k++) bstart[k + 1] += bstart[k];
    ctx->hperm.resize(nray);
    for (int i = 0; i < nray; i++) {
      const int r = by_chain[i];
      ctx->hperm[bstart[bucket(r)]++] = r;
    }
    HIPCHK(upload(ctx->rperm, ctx->hperm, st));
  }
  if (ctx->rpts.n < (size_t)off) HIPCHK(ctx->rpts.alloc((size_t)off));
  if (ctx->rout.n < (size_t)nray) HIPCHK(ctx->rout.alloc((size_t)nray));
  const int nsf = ctx->hgeom.nsf;
  if (ctx->rtsf.n < (size_t)ctx->nsrc * 3 * nsf) HIPCHK(ctx->rtsf.alloc((size_t)ctx->nsrc * 3 * nsf));
  HIPCHK(nearest_calls_reset(st));
  HIPCHK(launch_seafloor_times(ctx->rgeom.p, ctx->rfields.p, ctx->nsrc, nsf, ctx->rtsf.p, st));
  HIPCHK(launch_trace(ctx->rgeom.p, ctx->rfields.p, ctx->d_srcx.p, ctx->d_srcz.p, ctx->rtsf.p, ctx->rjobs.p,
                      ctx->rperm.p, nray, ctx->rpoff.p, ctx->rpts.p, ctx->rout.p, ctx->fwd_program ? 1 : 0, st));
  std::vector<RayOut> out(nray);
  HIPCHK(hipMemcpyAsync(out.data(), ctx->rout.p, sizeof(RayOut) * nray, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  {
    unsigned long long nc = 0;
    HIPCHK(nearest_calls_read(&nc, st));
    ctx->ray_nearest_calls = (long long)nc;
  }
  ctx->h_tcalc.resize(nray);
  ctx->h_valid.resize(nray);
  ctx->hwork.resize(nray);
  for (int r = 0; r < nray; r++) {
    t_rcv[r] = out[r].t; valid[r] = out[r].valid;
    ctx->h_tcalc[r] = out[r].t; ctx->h_valid[r] = out[r].valid;
    ctx->hwork[r] = out[r].work;
  }
  ctx->g_nnz = -1;
  ctx->gp_nnz = -1;
  ctx->sys_ndata = -1;
  return FMT2D_OK;
}

int fmt2d_set_ray_search(fmt2d_ctx* ctx, int mode) {
  if (!ctx || (mode != 0 && mode != 1)) return fail(FMT2D_EINVAL, "search mode must be 0 (indexed) or 1 (scan)");
  if (ctx->ray_scan != mode) {
    ctx->ray_scan = mode;
    ctx->rgeom_ok = false;
  }
  return FMT2D_OK;
}

int fmt2d_ray_nearest_check(fmt2d_ctx* ctx, int region, int n, const double* px, const double* pz, int* idx,
                            int* scan) {
  if (!ctx || n < 0 || (n > 0 && (!px || !pz || !idx || !scan))) return fail(FMT2D_EINVAL, "NULL argument");
  if (region < 1 || region > NREG) return fail(FMT2D_EINVAL, "region must be 1..6");
  if (!ctx->have_model) return fail(FMT2D_ESTATE, "no model set");
  HIPCHK(hipSetDevice(ctx->device));
  if (!ctx->rgeom_ok) { const int rc = build_ray_geom(ctx); if (rc) return rc; }
  if (n == 0) return FMT2D_OK;
  hipStream_t st = ctx->stream;
  DBuf<double> dx, dz;
  DBuf<int> di, ds;
  HIPCHK(upload(dx, std::vector<double>(px, px + n), st));
  HIPCHK(upload(dz, std::vector<double>(pz, pz + n), st));
  HIPCHK(di.alloc(n));
  HIPCHK(ds.alloc(n));
  HIPCHK(launch_nearest_check(ctx->rgeom.p, region, n, dx.p, dz.p, di.p, ds.p, st));
  HIPCHK(hipMemcpyAsync(idx, di.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(scan, ds.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return FMT2D_OK;
}

int fmt2d_ray_nearest_calls(fmt2d_ctx* ctx, long long* n) {
  if (!ctx || !n) return fail(FMT2D_EINVAL, "NULL argument");
  *n = ctx->ray_nearest_calls;
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
