// C ABI of the MI355X forward path: context, model upload, batched stage
// schedule of the eikonal solver, field read-back.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "ctx.h"

static thread_local std::string g_err;

int fmt2d_fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Tiles (of the current shape) with cut cells, coincident nodes or interface
// nodes of the region: their visits are the slow ones and start first in a
// launch.
static int build_heavy_tiles(fmt2d_ctx* ctx) {
  const HostModel& m = ctx->m;
  const int nx = m.nx, nz = m.nz, tx = ctx->tile_x, tz = ctx->tile_z;
  const int ntx = (nx + tx - 1) / tx, ntz = (nz + tz - 1) / tz, ntiles = ntx * ntz;
  std::vector<uint8_t> heavy((size_t)(NREG + 1) * ntiles, 0);
  for (int r = 1; r <= NREG; r++) {
    const HostRegion& R = m.reg[r];
    uint8_t* h = heavy.data() + (size_t)r * ntiles;
    for (int j = 0; j < nz; j++)
      for (int i = 0; i < nx; i++) {
        const uint16_t fl = R.flags[(size_t)j * nx + i];
        if ((fl & RF_IN) && (fl & ((RF_IRR0 * 15) | (RF_COI0 * 3)))) h[(j / tz) * ntx + i / tx] = 1;
      }
    const int ids[2] = {R.itop, R.ibot};
    for (int s2 = 0; s2 < 2; s2++) {
      const HostInterface& is = m.isec[ids[s2]];
      for (int c = 0; c < nx; c++) {
        const int cz = is.cc[c];
        if (cz < 0) continue;
        for (int j = cz; j <= std::min(cz + 1, nz - 1); j++) h[(j / tz) * ntx + c / tx] = 1;
      }
    }
  }
  HIPCHK(upload(ctx->theavy, heavy, ctx->stream));
  ctx->heavy_tx = tx;
  ctx->heavy_tz = tz;
  return FMT2D_OK;
}

extern "C" {

const char* fmt2d_last_error(void) { return g_err.c_str(); }

int fmt2d_tf_region(int itf) { return (itf >= 1 && itf <= 12) ? TF_REGION[itf] : 0; }

int fmt2d_ctx_create(int device, fmt2d_ctx** out) {
  if (!out) return fail(FMT2D_EINVAL, "out is NULL");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(FMT2D_EINVAL, "no such HIP device");
  HIPCHK(hipSetDevice(device));
  fmt2d_ctx* c = new fmt2d_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_count, sizeof(int) * 16, hipHostMallocDefault);
  if (e != hipSuccess) {
    delete c;
    return fail(FMT2D_EHIP, std::string("stream/pinned allocation: ") + hipGetErrorString(e));
  }
  *out = c;
  return FMT2D_OK;
}

int fmt2d_ctx_destroy(fmt2d_ctx* ctx) {
  if (!ctx) return FMT2D_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
  return FMT2D_OK;
}

int fmt2d_set_model(fmt2d_ctx* ctx, int nx_s, int nz_s, const double* xpos, const double* topo, const double* zrel,
                    const double* vel, double vwater, int nsed, const double* sed_x, const double* sed_z, int nmoho,
                    const double* moho_x, const double* moho_z) {
  if (!ctx || !xpos || !topo || !zrel || !vel || !sed_x || !sed_z || !moho_x || !moho_z)
    return fail(FMT2D_EINVAL, "NULL argument");
  HIPCHK(hipSetDevice(ctx->device));
  ctx->have_model = false;
  ctx->rgeom_ok = false;
  ctx->nsrc = 0;
  ctx->nray = 0;
  ctx->m = HostModel();
  // a new model starts a new inversion: its initial velocities and seafloor
  // are taken again at the next set-up
  ctx->v_init.clear();
  ctx->topo_init.clear();
  ctx->mesh_ok = false;
  std::string e = build_model(ctx->m, nx_s, nz_s, xpos, topo, zrel, vel, vwater, nsed, sed_x, sed_z, nmoho, moho_x,
                              moho_z);
  if (!e.empty()) return fail(FMT2D_EINVAL, e);
  return model_upload(ctx);
}

}  // extern "C"

// Device copies of the host model (grid, regions, intersections, PmP
// head-wave map); after fmt2d_set_model and after a model step.
int model_upload(fmt2d_ctx* ctx) {
  ctx->fields_current = false;
  HostModel& m = ctx->m;
  hipStream_t st = ctx->stream;
  const int nx = m.nx, nz = m.nz;
  HIPCHK(upload(ctx->xg, m.xg, st));
  HIPCHK(upload(ctx->zg, m.zg, st));
  std::vector<RegionDev> rd(NREG + 1);
  for (int r = 1; r <= NREG; r++) {
    HostRegion& R = m.reg[r];
    HIPCHK(upload(ctx->rflags[r], R.flags, st));
    HIPCHK(upload(ctx->rvel[r], R.vel, st));
    RegionDev& d = rd[r];
    d.nx = nx; d.nz = nz; d.dx = m.dx; d.dz = m.dz; d.tol = m.tol;
    d.xg = ctx->xg.p; d.zg = ctx->zg.p; d.flags = ctx->rflags[r].p; d.vel = ctx->rvel[r].p;
    const int ids[2] = {R.itop, R.ibot};
    for (int s = 0; s < 2; s++) {
      const HostInterface& is = m.isec[ids[s]];
      HIPCHK(upload(ctx->ifx[r][s], is.x, st));
      HIPCHK(upload(ctx->ifz[r][s], is.z, st));
      HIPCHK(upload(ctx->ifcc[r][s], is.cc, st));
      HIPCHK(upload(ctx->ifok[r][s], is.ok, st));
      d.ifx[s] = ctx->ifx[r][s].p; d.ifz[s] = ctx->ifz[r][s].p;
      d.ifcc[s] = ctx->ifcc[r][s].p; d.ifok[s] = ctx->ifok[r][s].p;
    }
  }
  rd[0] = rd[1];
  HIPCHK(upload(ctx->regs, rd, st));
  {
    const int rc = build_heavy_tiles(ctx);
    if (rc) return rc;
  }
  std::vector<IsecDev> idv(NIF + 1);
  for (int k = 1; k <= NIF; k++) {
    const HostInterface& is = m.isec[k];
    HIPCHK(upload(ctx->isn[k][0], is.nrm_x, st));
    HIPCHK(upload(ctx->isn[k][1], is.nrm_z, st));
    HIPCHK(upload(ctx->isn[k][2], is.vel_top, st));
    HIPCHK(upload(ctx->isn[k][3], is.vel_bot, st));
    idv[k].nx = ctx->isn[k][0].p; idv[k].nz = ctx->isn[k][1].p;
    idv[k].vel_top = ctx->isn[k][2].p; idv[k].vel_bot = ctx->isn[k][3].p;
  }
  idv[0] = idv[1];
  HIPCHK(upload(ctx->isecs, idv, st));
  for (int k = 1; k <= NIF; k++) {
    HIPCHK(upload(ctx->isnp[k][0], m.isec[k].pnrm_x, st));
    HIPCHK(upload(ctx->isnp[k][1], m.isec[k].pnrm_z, st));
    idv[k].nx = ctx->isnp[k][0].p; idv[k].nz = ctx->isnp[k][1].p;
  }
  idv[0] = idv[1];
  HIPCHK(upload(ctx->isecs_prog, idv, st));
  // PmP head-wave map.  The reference addresses the PmP region (6) with the
  // Moho's rright_node_id, i.e. region 4's top-node numbering
  // (inversion2d.f90:4298-4301, propagate2d.f90:1593-1597).
  {
    const HostRegion& R4 = m.reg[4];
    const HostRegion& R6 = m.reg[6];
    const HostInterface& moho = m.isec[4];
    const int n = moho.n;
    std::vector<int> slot(n, -1);
    std::vector<double> vhw(n, 0.0);
    std::vector<int> vslot(n, -1);  // region-6 slot vhw was read from
    for (int c = 0; c < n; c++) {
      const int mref = R4.ngnode + c + 1;  // 1-based node id
      if (mref < 1 || mref > R6.nnode) continue;
      const int q = R6.ref_order[mref - 1];
      slot[c] = q;
      if (q >= nz * nx) {
        // interface node of region 6: fastest grid velocity of the Moho cut cell
        // of its column, below the Moho node of column c
        const int i3 = q % nx;
        const int ccz = moho.cc[i3];
        if (ccz >= 0) {
          for (int ii = 0; ii <= 1; ii++)
            for (int jj = 0; jj <= 1; jj++) {
              const int ixx = i3 + ii, izz = ccz + jj;
              if (ixx < nx && izz < nz && m.node_region[(size_t)izz * nx + ixx] == 6) {
                if (m.zg[izz] > moho.z[c] + 0.5 * m.dz && R6.vel[(size_t)izz * nx + ixx] > vhw[c]) {
                  vhw[c] = R6.vel[(size_t)izz * nx + ixx];
                  vslot[c] = izz * nx + ixx;
                }
              }
            }
        }
      } else {
        vhw[c] = R6.vel[q];
        vslot[c] = q;
      }
    }
    HIPCHK(upload(ctx->hw_slot, slot, st));
    HIPCHK(upload(ctx->hw_vslot, vslot, st));
    HIPCHK(upload(ctx->hw_vhw, vhw, st));
    HeadwaveMap hm;
    hm.n = n; hm.dxif = moho.x[1] - moho.x[0]; hm.slot = ctx->hw_slot.p; hm.vhw = ctx->hw_vhw.p;
    std::vector<HeadwaveMap> hv(1, hm);
    HIPCHK(upload(ctx->hw_map, hv, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  ctx->have_model = true;
  ctx->rgeom_ok = false;
  return FMT2D_OK;
}

extern "C" {

int fmt2d_grid_info_get(fmt2d_ctx* ctx, fmt2d_grid_info* info) {
  if (!ctx || !info) return fail(FMT2D_EINVAL, "NULL argument");
  if (!ctx->have_model) return fail(FMT2D_ESTATE, "no model set");
  const HostModel& m = ctx->m;
  info->nx = m.nx; info->nz = m.nz; info->dx = m.dx; info->dz = m.dz; info->x0 = m.x0; info->z0 = m.z0;
  info->region_nnode[0] = 0;
  for (int r = 1; r <= NREG; r++) info->region_nnode[r] = m.reg[r].nnode;
  return FMT2D_OK;
}

int fmt2d_forward(fmt2d_ctx* ctx, int nsrc, const double* src_x, int flags) {
  if (!ctx || (nsrc > 0 && !src_x)) return fail(FMT2D_EINVAL, "NULL argument");
  if (!ctx->have_model) return fail(FMT2D_ESTATE, "no model set");
  if (nsrc < 1) return fail(FMT2D_EINVAL, "nsrc must be >= 1");
  if (flags & ~3) return fail(FMT2D_EINVAL, "unknown forward flag");
  HIPCHK(hipSetDevice(ctx->device));
  // A line search that accepts a step ends on the forward of exactly the
  // model it commits; the same forward again would give the same fields bit
  // for bit, so it returns at once (rays of the old trace are dropped as
  // after any forward).
  if (ctx->fields_current && flags == ctx->fields_flags && nsrc == ctx->nsrc &&
      std::equal(src_x, src_x + nsrc, ctx->src_x.begin())) {
    ctx->nray = 0;
    ctx->g_nnz = -1;
    ctx->gp_nnz = -1;
    ctx->sys_ndata = -1;
    const int nfield = ctx->stats.nfield;
    const long long region_nodes = ctx->stats.region_nodes;
    ctx->stats = fmt2d_forward_stats{};
    ctx->stats.nsrc = nsrc;
    ctx->stats.nfield = nfield;
    ctx->stats.region_nodes = region_nodes;
    return FMT2D_OK;
  }
  ctx->fields_current = false;
  // the forward program's variant has no head-wave post-pass
  const bool program = (flags & FMT2D_FWD_PROGRAM) != 0;
  if (program) flags |= FMT2D_FWD_NO_HEADWAVE;
  ctx->fwd_program = program;
  ctx->fwd_f32 = ctx->field_f32;
  HostModel& m = ctx->m;
  hipStream_t st = ctx->stream;
  const int nx = m.nx, nz = m.nz;
  const size_t nslot = (size_t)(nz + 2) * nx;
  const int nfield = nsrc * FMT2D_NTF;
  const auto t_start = std::chrono::steady_clock::now();
  // ---- field storage, ordered by level so every level is one contiguous batch
  std::vector<int> order;  // field index -> (isrc*12 + itf-1)
  std::vector<int> level_off(5, 0);
  ctx->field_of.assign(nfield, -1);
  for (int L = 0; L < 4; L++) {
    level_off[L] = (int)order.size();
    for (int s = 0; s < nsrc; s++)
      for (int k = 0; k < LEVEL_N[L]; k++) {
        const int itf = LEVELS[L][k];
        ctx->field_of[s * FMT2D_NTF + itf - 1] = (int)order.size();
        order.push_back(s * FMT2D_NTF + itf - 1);
      }
  }
  level_off[4] = (int)order.size();
  if (ctx->nfield != nfield || ctx->fT.n != nslot * nfield) {
    HIPCHK(ctx->fT.alloc(nslot * nfield));
    HIPCHK(ctx->fGX.alloc(nslot * nfield));
    HIPCHK(ctx->fGZ.alloc(nslot * nfield));
    HIPCHK(ctx->fOGX.alloc(nslot * nfield));
    HIPCHK(ctx->fOGZ.alloc(nslot * nfield));
    HIPCHK(ctx->fTIE.alloc(nslot * nfield));
    HIPCHK(ctx->isT.alloc((size_t)nx * nfield));
    HIPCHK(ctx->isGX.alloc((size_t)nx * nfield));
    HIPCHK(ctx->isGZ.alloc((size_t)nx * nfield));
  }
  ctx->nsrc = nsrc;
  ctx->nfield = nfield;
  ctx->src_x.assign(src_x, src_x + nsrc);
  ctx->src_z.resize(nsrc);
  for (int s = 0; s < nsrc; s++) ctx->src_z[s] = m.setup_seafloor_at(src_x[s]);
  ctx->nray = 0;
  std::vector<FieldDesc> fd(nfield);
  std::vector<StageLink> lk(nfield);
  long long region_nodes = 0;
  for (int f = 0; f < nfield; f++) {
    const int s = order[f] / FMT2D_NTF, itf = order[f] % FMT2D_NTF + 1;
    FieldDesc& d = fd[f];
    memset(&d, 0, sizeof(d));
    d.region = TF_REGION[itf];
    const HostRegion& R = m.reg[d.region];
    region_nodes += R.nnode;
    d.T = ctx->fT.p + (size_t)f * nslot;
    d.GX = ctx->fGX.p + (size_t)f * nslot;
    d.GZ = ctx->fGZ.p + (size_t)f * nslot;
    d.OGX = ctx->fOGX.p + (size_t)f * nslot;
    d.OGZ = ctx->fOGZ.p + (size_t)f * nslot;
    d.TIE = ctx->fTIE.p + (size_t)f * nslot;
    d.isT = ctx->isT.p + (size_t)f * nx;
    d.isGX = ctx->isGX.p + (size_t)f * nx;
    d.isGZ = ctx->isGZ.p + (size_t)f * nx;
    d.fixed_key = -1;
    if (TF_PREV[itf] == 0) {
      // point source on the seafloor of the set-up (fmt2d_main.f90:545-546,
      // invert2d_main.f90:549-550: a later seafloor update leaves it), cell
      // corners
      d.kind = SEED_POINT;
      d.sx = src_x[s];
      d.sz = ctx->src_z[s];
      d.t0 = 0.0;
      int ix0 = (int)floor((d.sx - m.x0) / m.dx), iz0 = (int)floor((d.sz - m.z0) / m.dz);
      ix0 = std::max(0, std::min(ix0, nx - 2));
      iz0 = std::max(0, std::min(iz0, nz - 2));
      d.ix0 = ix0; d.iz0 = iz0;
      for (int i = ix0; i <= ix0 + 1; i++)
        for (int j = iz0; j <= iz0 + 1; j++) {
          if (!(R.flags[(size_t)j * nx + i] & RF_IN)) continue;
          const double dx = m.xg[i] - d.sx, dz = m.zg[j] - d.sz;
          if (sqrt(dx * dx + dz * dz) < 0.01 * m.tol) d.fixed_key = j * nx + i;
        }
    } else {
      d.kind = SEED_IFACE;
      const int I = TF_ISTART[itf];
      const int pr = TF_PREV[itf];
      const HostRegion& Rp = m.reg[TF_REGION[pr]];
      StageLink& L = lk[f];
      L.prev = ctx->field_of[s * FMT2D_NTF + pr - 1];
      if (I == Rp.ibot) L.prev_side = 1;
      else if (I == Rp.itop) L.prev_side = 0;
      else return fail(FMT2D_EINVAL, "stage transfer from a non-boundary interface is not supported");
      if (I == R.itop) d.side = 0;
      else if (I == R.ibot) d.side = 1;
      else return fail(FMT2D_EINVAL, "stage sweep from a non-boundary interface is not supported");
      L.isec = I;
      L.reflect = 1;
      L.target_side = (I == R.itop) ? 1 : ((I == R.ibot) ? -1 : 0);
      L.target_has_iface = 1;
      L.target_vel = ctx->rvel[d.region].p + (size_t)(nz + d.side) * nx;
    }
  }
  ctx->stats = fmt2d_forward_stats{};
  ctx->hw_fields.clear();
  ctx->trace.clear();
  ctx->trace_sweeps.clear();
  ctx->stats.nsrc = nsrc;
  ctx->stats.nfield = nfield;
  ctx->stats.region_nodes = region_nodes;
  // ---- tile bookkeeping sized for the largest level (32-bit storage: 8x8 tiles)
  const bool f32 = ctx->field_f32 != 0;
  if (f32 && (ctx->tile_x != 8 || ctx->tile_z != 8)) { ctx->tile_x = 8; ctx->tile_z = 8; }
  const int tx = ctx->tile_x, tz = ctx->tile_z;
  if (ctx->heavy_tx != tx || ctx->heavy_tz != tz) {
    const int rc = build_heavy_tiles(ctx);
    if (rc) return rc;
  }
  const int ntx = (nx + tx - 1) / tx, ntz = (nz + tz - 1) / tz, ntiles = ntx * ntz;
  int maxlev = 0;
  for (int L = 0; L < 4; L++) maxlev = std::max(maxlev, level_off[L + 1] - level_off[L]);
  if (f32) {
    // 32-bit working storage for one level's fields, reused level by level
    const size_t ng = (size_t)nx * nz;
    if (ctx->f32T.n < (size_t)maxlev * ng) {
      HIPCHK(ctx->f32T.alloc((size_t)maxlev * ng));
      HIPCHK(ctx->f32GX.alloc((size_t)maxlev * ng));
      HIPCHK(ctx->f32GZ.alloc((size_t)maxlev * ng));
    }
    for (int L = 0; L < 4; L++)
      for (int f = level_off[L]; f < level_off[L + 1]; f++) {
        const size_t pos = (size_t)(f - level_off[L]);
        fd[f].T32 = ctx->f32T.p + pos * ng;
        fd[f].GX32 = ctx->f32GX.p + pos * ng;
        fd[f].GZ32 = ctx->f32GZ.p + pos * ng;
      }
  }
  HIPCHK(upload(ctx->descs, fd, st));
  HIPCHK(upload(ctx->links, lk, st));
  const size_t ntt = (size_t)maxlev * ntiles;
  if (ctx->tin.n < ntt) {
    HIPCHK(ctx->tin.alloc(ntt));
    HIPCHK(ctx->tkey.alloc(ntt));
    for (int b = 0; b < 3; b++) HIPCHK(ctx->tlist[b].alloc(ntt));
  }
  // every tile starts outside the active set with no key: a forward that
  // returned early (launch cap, HIP error) leaves tiles marked active
  HIPCHK(hipMemsetAsync(ctx->tin.p, 0, ntt * sizeof(unsigned int), st));
  HIPCHK(hipMemsetAsync(ctx->tkey.p, 0x7F, ntt * sizeof(unsigned long long), st));
  if (ctx->tfront.n < 2 * (size_t)maxlev) HIPCHK(ctx->tfront.alloc(2 * (size_t)maxlev));
  if (!ctx->tcount.p) HIPCHK(ctx->tcount.alloc(16));
  HIPCHK(launch_fill(ctx->descs.p, nfield, (int)nslot, st));
  if (!ctx->stop_flag.p) HIPCHK(ctx->stop_flag.alloc(1));
  HIPCHK(hipMemsetAsync(ctx->stop_flag.p, 0, sizeof(int), st));
  if (!ctx->acc.p) HIPCHK(ctx->acc.alloc(3 + 64));
  HIPCHK(hipMemsetAsync(ctx->acc.p, 0, (3 + 64) * sizeof(unsigned long long), st));
  const int max_launch = ctx->launch_cap > 0 ? ctx->launch_cap : 64 * (ntx + ntz) + 4000;
  const bool headwave = !(flags & FMT2D_FWD_NO_HEADWAVE);
  // arrival-time window: `window_tiles` tile crossings at 4 km/s
  const double window = ctx->window_tiles * std::max(tx * m.dx, tz * m.dz) / 4.0;
  size_t nev = 0;       // tile-solver dispatches
  double fim_ms = 0.0;  // their time (first wave start to last wave end)
  for (int L = 0; L < 4; L++) {
    const int f0 = level_off[L], nf = level_off[L + 1] - level_off[L];
    const FieldDesc* dl = ctx->descs.p + f0;
    if (L > 0)
      HIPCHK(launch_transfer_reflect(ctx->descs.p, f0, ctx->links.p, nf, nx, nz,
                                     program ? ctx->isecs_prog.p : ctx->isecs.p, st));
    // counters: [0] / [3] due tiles of a launch cut by an interface / not,
    // [1] / [2] sizes of active lists 0 / 1, [4] launches of the level,
    // [8] / [11] the due counters of odd launches in the level kernel,
    // [12] / [13] its tile hand-out counters
    HIPCHK(hipMemsetAsync(ctx->tcount.p, 0, sizeof(int) * 16, st));
    // fronts of lists 0 and 1: no key (seeding lowers list 0's)
    unsigned long long* front[2] = {ctx->tfront.p, ctx->tfront.p + nf};
    HIPCHK(hipMemsetAsync(ctx->tfront.p, 0x7F, sizeof(unsigned long long) * 2 * nf, st));
    if (f32) HIPCHK(launch_f32_reset(dl, nf, nx * nz, st));
    HIPCHK(launch_seed_tiles(dl, ctx->regs.p, nf, nx, ntx, ntz, tx, tz, ctx->tin.p, ctx->tkey.p, ctx->tlist[0].p,
                             ctx->tcount.p + 1, front[0], st));
    constexpr int BATCH = 8;
    const int fim_grid = 8192, sched_grid = 512;
    const int trace_cap = max_launch + BATCH;
    if (ctx->ttrace.n < (size_t)trace_cap) HIPCHK(ctx->ttrace.alloc((size_t)trace_cap));
    if (ctx->tmax.n < (size_t)trace_cap) HIPCHK(ctx->tmax.alloc((size_t)trace_cap));
    const size_t nspan = (size_t)trace_cap * (SPAN_SLOTS + 1);
    if (ctx->tspan.n < nspan) HIPCHK(ctx->tspan.alloc(nspan));
    HIPCHK(hipMemsetAsync(ctx->tmax.p, 0, sizeof(int) * (size_t)trace_cap, st));
    HIPCHK(hipMemsetAsync(ctx->tspan.p, 0, sizeof(unsigned long long) * nspan, st));
    if (ctx->clock_hz <= 0.0) {
      int khz = 0;
      HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
      if (khz <= 0) return fail(FMT2D_EHIP, "the device reports no constant clock rate");
      ctx->clock_hz = 1000.0 * khz;
    }
    // active tiles left after the cap: the error names a few of them
    auto no_convergence = [&](int cur, int n) {
      std::vector<int2> tail(std::min(n, 8));
      (void)hipMemcpy(tail.data(), ctx->tlist[cur].p, tail.size() * sizeof(int2), hipMemcpyDeviceToHost);
      std::string msg = "eikonal tile solver did not converge within the launch cap (level " + std::to_string(L) +
                        ", " + std::to_string(n) + " tiles active; field/tile:";
      for (auto& p : tail) {
        const int itf = order[f0 + p.x] % FMT2D_NTF + 1;
        msg += " tf" + std::to_string(itf) + "/src" + std::to_string(order[f0 + p.x] / FMT2D_NTF) + "/t" +
               std::to_string(p.y) + "(" + std::to_string((p.y % ntx) * tx) + "," + std::to_string((p.y / ntx) * tz) +
               ")";
      }
      return fail(FMT2D_ENOCONV, msg + ")");
    };
    int launches = 0;
    if (ctx->solver_launch == FMT2D_LAUNCH_LEVEL) {
      // one cooperative dispatch runs every launch of the level
      const int gi = f32 ? 3 : (tx == 16 ? 1 : (tz == 16 ? 2 : 0));
      if (ctx->level_grid[gi] <= 0) {
        const hipError_t e = level_grid(ctx->device, tx, tz, f32 ? 1 : 0, &ctx->level_grid[gi]);
        if (e == hipErrorNotSupported) return fail(FMT2D_EHIP, "the device does not run cooperative launches");
        HIPCHK(e);
      }
      const size_t nbar = level_bar_words(ctx->level_grid[gi]);
      if (ctx->tbar.n < nbar) HIPCHK(ctx->tbar.alloc(nbar));
      HIPCHK(hipMemsetAsync(ctx->tbar.p, 0, sizeof(unsigned) * nbar, st));
      LevelSched a{};
      a.descs = dl; a.regs = ctx->regs.p; a.heavy = ctx->theavy.p;
      a.list[0] = ctx->tlist[0].p; a.list[1] = ctx->tlist[1].p; a.due = ctx->tlist[2].p;
      a.count = ctx->tcount.p; a.in_set = ctx->tin.p; a.tkey = ctx->tkey.p;
      a.front[0] = front[0]; a.front[1] = front[1];
      a.acc = ctx->acc.p; a.trace = ctx->ttrace.p; a.sweep_max = ctx->tmax.p; a.span = ctx->tspan.p;
      a.flag = ctx->stop_flag.p; a.bar = ctx->tbar.p; a.window = window;
      a.nfield = nf; a.ntx = ntx; a.ntz = ntz; a.ntiles = ntiles; a.due_cap = (int)ntt;
      a.max_sweeps = ctx->max_sweeps; a.max_sweeps_heavy = ctx->max_sweeps_heavy; a.colors = ctx->colors;
      a.max_launch = max_launch; a.trace_cap = trace_cap;
      HIPCHK(launch_fim_level(a, tx, tz, f32 ? 1 : 0, ctx->level_grid[gi], st));
      nev += 1;
      HIPCHK(hipMemcpyAsync(ctx->h_count, ctx->tcount.p, sizeof(int) * 5, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      launches = ctx->h_count[4];
      const int cur = launches & 1;
      if (ctx->h_count[1 + cur] != 0) return no_convergence(cur, ctx->h_count[1 + cur]);
      if (launches > 0) HIPCHK(hipMemcpy(ctx->h_count, ctx->ttrace.p + launches - 1, sizeof(int), hipMemcpyDeviceToHost));
    } else {
      // Launches run in batches of BATCH, captured once per level in a HIP
      // graph and replayed until the active set is empty: every argument of
      // a batch is the same at each replay (the active lists alternate, and
      // with them the checkerboard colour and the front arrays, with the
      // launch's parity; the launch number the kernels record under is
      // counted on the device).  The host reads the active count once per
      // replay.
      auto record_batch = [&]() -> hipError_t {
        for (int b = 0; b < BATCH; b++) {
          const int cur = b & 1, nxt = cur ^ 1;
          hipError_t e = launch_schedule(ctx->tlist[cur].p, ctx->tcount.p + 1 + cur, nf, ntx, ntiles, ctx->tkey.p,
                                         front[cur], front[nxt], window, b & 1, ctx->tin.p, ctx->tlist[2].p,
                                         ctx->tcount.p, ctx->tlist[nxt].p, ctx->tcount.p + 1 + nxt,
                                         ctx->tcount.p + 4, ctx->ttrace.p, trace_cap, dl, ctx->theavy.p, (int)ntt,
                                         sched_grid, st);
          if (e == hipSuccess)
            e = launch_fim(dl, ctx->regs.p, ctx->tlist[2].p, fim_grid, ctx->tcount.p, ntx, ntz, tx, tz, ctx->tin.p,
                           ctx->tkey.p, ctx->tlist[nxt].p, ctx->tcount.p + 1 + nxt, front[nxt], ctx->max_sweeps,
                           ctx->max_sweeps_heavy, ctx->colors, ctx->acc.p, ctx->tcount.p + 4, ctx->tmax.p,
                           trace_cap, ctx->tspan.p, (int)ntt, f32 ? 1 : 0, ctx->stop_flag.p, st);
          if (e != hipSuccess) return e;
        }
        return hipSuccess;
      };
      hipGraph_t graph = nullptr;
      hipGraphExec_t gexec = nullptr;
      HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      {
        const hipError_t e = record_batch();
        const hipError_t e2 = hipStreamEndCapture(st, &graph);
        if (e != hipSuccess || e2 != hipSuccess) {
          if (graph) (void)hipGraphDestroy(graph);
          HIPCHK(e != hipSuccess ? e : e2);
        }
      }
      {
        const hipError_t e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHK(e);
      }
      struct ExecGuard {
        hipGraphExec_t g;
        ~ExecGuard() { if (g) (void)hipGraphExecDestroy(g); }
      } guard{gexec};
      const int cur = 0;  // a batch has an even number of launches: the active list is list 0 again
      for (;;) {
        HIPCHK(hipGraphLaunch(gexec, st));
        launches += BATCH;
        nev += BATCH;
        HIPCHK(hipMemcpyAsync(ctx->h_count, ctx->tcount.p, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (ctx->h_count[1 + cur] == 0) break;
        if (launches >= max_launch) return no_convergence(cur, ctx->h_count[1 + cur]);
      }
      // the last launch's due count is still in the due counters
      ctx->h_count[0] += ctx->h_count[3];
    }
    {
      // tile-solver time: the launches' spans on the device clock
      std::vector<unsigned long long> sp((size_t)launches * (SPAN_SLOTS + 1));
      HIPCHK(hipMemcpy(sp.data(), ctx->tspan.p, sizeof(unsigned long long) * sp.size(), hipMemcpyDeviceToHost));
      for (int k = 0; k < launches; k++) {
        const unsigned long long* q = sp.data() + (size_t)k * (SPAN_SLOTS + 1);
        const unsigned long long end = *std::max_element(q + 1, q + 1 + SPAN_SLOTS);
        if (end > q[0]) fim_ms += 1e3 * (double)(end - q[0]) / ctx->clock_hz;
      }
      // due tiles per launch (the last one is still in the due counter)
      std::vector<int> tr(launches), mx(launches);
      if (launches > 1)
        HIPCHK(hipMemcpy(tr.data(), ctx->ttrace.p, sizeof(int) * (launches - 1), hipMemcpyDeviceToHost));
      if (launches > 0) {
        tr[launches - 1] = ctx->h_count[0];
        HIPCHK(hipMemcpy(mx.data(), ctx->tmax.p, sizeof(int) * launches, hipMemcpyDeviceToHost));
      }
      // launches queued after the set emptied (rest of the last batch) had nothing due
      int used = 0;
      for (int k = 0; k < launches; k++)
        if (tr[k] > 0) {
          ctx->stats.tiles += tr[k];
          ctx->trace.push_back(tr[k]);
          ctx->trace_sweeps.push_back(mx[k]);
          used++;
        }
      ctx->stats.launches += used;
    }
    // 32-bit storage: the solved level in FP64 for the post-passes and the rays
    if (f32) HIPCHK(launch_f32_export(dl, nf, nx * nz, st));
    // heap-exhaustion stop of the reference (src/propagate2d.f90:74-80)
    {
      HIPCHK(launch_member_time(dl, ctx->regs.p, nf, nx, nz, st));
      const size_t nb = (size_t)nf * EXIT_NBIN;
      if (ctx->x_ef.n < (size_t)nf) HIPCHK(ctx->x_ef.alloc(nf));
      if (ctx->x_cnt.n < nb) {
        HIPCHK(ctx->x_cnt.alloc(nb));
        HIPCHK(ctx->x_diff.alloc(nb));
        HIPCHK(ctx->x_off.alloc(nb));
        HIPCHK(ctx->x_max.alloc(nb));
        HIPCHK(ctx->x_m.alloc(nb));
      }
      if (!ctx->x_total.p) HIPCHK(ctx->x_total.alloc(1));
      ExitWork w{ctx->x_ef.p, ctx->x_cnt.p, ctx->x_diff.p, ctx->x_off.p, ctx->x_max.p, ctx->x_m.p,
                 ctx->x_total.p, ctx->x_ent.p, (long long)ctx->x_ent.n};
      HIPCHK(launch_exit_bins(dl, ctx->regs.p, nf, (int)nslot, w, st));
      HIPCHK(hipMemcpyAsync(ctx->h_count + 2, ctx->x_total.p, sizeof(int), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      const long long ncand = ctx->h_count[2];
      if ((long long)ctx->x_ent.n < ncand) {
        HIPCHK(ctx->x_ent.alloc((size_t)ncand + ncand / 4));
        w.ent = ctx->x_ent.p;
        w.cap = (long long)ctx->x_ent.n;
      }
      HIPCHK(launch_exit_check(dl, ctx->regs.p, nf, (int)nslot, w, ncand, st));
      ctx->stats.exit_nodes += ncand;
    }
    HIPCHK(launch_refit(dl, ctx->regs.p, nf, nx, nz, ctx->stop_flag.p, st));
    // water-layer top-row fixes and the PmP head-wave post-pass
    std::vector<int> which, up, hw;
    for (int f = f0; f < f0 + nf; f++) {
      const int itf = order[f] % FMT2D_NTF + 1;
      if (TF_WATERFIX[itf]) { which.push_back(f - f0); up.push_back(TF_WATERFIX[itf] > 0); }
      if (itf == 5 && headwave) hw.push_back(f - f0);
    }
    if (!which.empty()) {
      HIPCHK(upload(ctx->sel_which, which, st));
      HIPCHK(upload(ctx->sel_up, up, st));
      HIPCHK(launch_water_fix(dl, ctx->regs.p, ctx->sel_which.p, ctx->sel_up.p, (int)which.size(), nx, st));
      HIPCHK(hipStreamSynchronize(st));
    }
    if (!hw.empty()) {
      HIPCHK(upload(ctx->sel_which, hw, st));
      if (ctx->hw_scratch.n < (size_t)hw.size() * 2 * nx) HIPCHK(ctx->hw_scratch.alloc((size_t)hw.size() * 2 * nx));
      if (ctx->hw_save.n < (size_t)hw.size() * nx) HIPCHK(ctx->hw_save.alloc((size_t)hw.size() * nx));
      HIPCHK(launch_headwave(dl, ctx->regs.p, ctx->sel_which.p, (int)hw.size(), ctx->hw_map.p, ctx->hw_scratch.p,
                             ctx->hw_save.p, st));
      for (int w : hw) ctx->hw_fields.push_back(f0 + w);
      HIPCHK(hipStreamSynchronize(st));
    }
  }
  HIPCHK(hipStreamSynchronize(st));
  ctx->stats.ms_total =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  ctx->stats.ms_fim = fim_ms;
  int stop = 0;
  HIPCHK(hipMemcpy(&stop, ctx->stop_flag.p, sizeof(int), hipMemcpyDeviceToHost));
  if (stop & 1) return fail(FMT2D_EINVAL, "negative determinant too large in regular_update2d (the reference stops here)");
  if (stop & 2) return fail(FMT2D_EINVAL, "a traveltime beyond 255 s does not fit the 32-bit field storage");
  ctx->stats.dispatches = (int)nev;
  unsigned long long acc[3 + 64];
  HIPCHK(hipMemcpy(acc, ctx->acc.p, sizeof(acc), hipMemcpyDeviceToHost));
  for (int k = 0; k < 64; k++) ctx->sweep_hist[k] = (long long)acc[3 + k];
  ctx->stats.sweeps = (long long)acc[0];
  ctx->stats.max_sweeps = (int)acc[1];
  ctx->stats.capped = (long long)acc[2];
  return FMT2D_OK;
}

int fmt2d_set_tuning(fmt2d_ctx* ctx, double window_tiles, int max_sweeps, int colors) {
  if (!ctx || !(window_tiles > 0.0) || max_sweeps < 1 || (colors != 1 && colors != 2 && colors != 4))
    return fail(FMT2D_EINVAL, "bad tuning parameters");
  ctx->colors = colors;
  ctx->window_tiles = window_tiles;
  ctx->max_sweeps = max_sweeps;
  return FMT2D_OK;
}

int fmt2d_field_snapshot(fmt2d_ctx* ctx) {
  if (!ctx) return fail(FMT2D_EINVAL, "NULL context");
  if (ctx->nfield < 1) return fail(FMT2D_ESTATE, "no forward has run");
  HIPCHK(hipSetDevice(ctx->device));
  const size_t n = (size_t)(ctx->m.nz + 2) * ctx->m.nx * ctx->nfield;
  if (ctx->snapT.n < n) HIPCHK(ctx->snapT.alloc(n));
  HIPCHK(hipMemcpyAsync(ctx->snapT.p, ctx->fT.p, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->snap_nfield = ctx->nfield;
  ctx->snap_src = ctx->src_x;
  return FMT2D_OK;
}

int fmt2d_field_compare(fmt2d_ctx* ctx, double tol, double* max_dt, long long* n_reach_diff, long long* n_over,
                        int cap, int* n) {
  if (!ctx || !n) return fail(FMT2D_EINVAL, "NULL argument");
  if (ctx->snap_nfield != ctx->nfield || ctx->snap_src != ctx->src_x)
    return fail(FMT2D_ESTATE, "the snapshot is of other sources or no snapshot was taken");
  *n = ctx->nfield;
  if (!max_dt || !n_reach_diff || !n_over || cap <= 0) return FMT2D_OK;
  if (cap < ctx->nfield) return fail(FMT2D_EINVAL, "output arrays too small");
  HIPCHK(hipSetDevice(ctx->device));
  const int nf = ctx->nfield;
  DBuf<double> dm;
  DBuf<long long> dn, dover;
  HIPCHK(dm.alloc(nf));
  HIPCHK(dn.alloc(nf));
  HIPCHK(dover.alloc(nf));
  const long long nslot = (long long)(ctx->m.nz + 2) * ctx->m.nx;
  HIPCHK(launch_field_compare(ctx->descs.p, nf, ctx->snapT.p, nslot, tol, dm.p, dn.p, dover.p, ctx->stream));
  std::vector<double> hm(nf);
  std::vector<long long> hn(nf), ho(nf);
  HIPCHK(hipMemcpyAsync(hm.data(), dm.p, sizeof(double) * nf, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(hn.data(), dn.p, sizeof(long long) * nf, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(ho.data(), dover.p, sizeof(long long) * nf, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  // source-major order: isrc * 12 + itf - 1
  for (int k = 0; k < nf; k++) {
    const int f = ctx->field_of[k];
    max_dt[k] = hm[f];
    n_reach_diff[k] = hn[f];
    n_over[k] = ho[f];
  }
  return FMT2D_OK;
}

int fmt2d_set_field_precision(fmt2d_ctx* ctx, int bits) {
  if (!ctx || (bits != 32 && bits != 64)) return fail(FMT2D_EINVAL, "field precision must be 32 or 64");
  ctx->field_f32 = bits == 32 ? 1 : 0;
  ctx->fields_current = false;
  return FMT2D_OK;
}

int fmt2d_set_tile(fmt2d_ctx* ctx, int tile_x, int tile_z) {
  if (!ctx) return fail(FMT2D_EINVAL, "NULL context");
  const bool ok = (tile_x == 8 && tile_z == 8) || (tile_x == 16 && tile_z == 8) || (tile_x == 8 && tile_z == 16);
  if (!ok) return fail(FMT2D_EINVAL, "tile shape must be 8x8, 16x8 or 8x16");
  ctx->tile_x = tile_x;
  ctx->tile_z = tile_z;
  return FMT2D_OK;
}

int fmt2d_autotune(fmt2d_ctx* ctx, int nsrc, const double* src_x, int flags, int level, fmt2d_tune_result* out) {
  if (!ctx || !out) return fail(FMT2D_EINVAL, "NULL argument");
  if (level < 0 || level > 1) return fail(FMT2D_EINVAL, "level must be 0 or 1");
  struct Cfg {
    int tx, tz, cap, hcap, colors;
    double window;
  };
  auto apply = [&](const Cfg& c) {
    ctx->tile_x = c.tx; ctx->tile_z = c.tz;
    ctx->max_sweeps = c.cap; ctx->max_sweeps_heavy = c.hcap;
    ctx->colors = c.colors; ctx->window_tiles = c.window;
  };
  // one forward per candidate, timed by the forward's own wall clock; every
  // candidate gives the same fields bit for bit, so the context ends solved
  int trials = 0;
  auto time_of = [&](const Cfg& c, double& ms) -> int {
    apply(c);
    ctx->fields_current = false;
    const int rc = fmt2d_forward(ctx, nsrc, src_x, flags);
    if (rc) return rc;
    ms = ctx->stats.ms_total;
    trials++;
    return FMT2D_OK;
  };
  const Cfg start{ctx->tile_x, ctx->tile_z, ctx->max_sweeps, ctx->max_sweeps_heavy, ctx->colors, ctx->window_tiles};
  Cfg best = start;
  double best_ms = 0.0;
  int rc = time_of(best, best_ms);
  if (rc) { apply(start); return rc; }
  const double start_ms = best_ms;
  // coordinate descent: tile shape, light cap, heavy cap, then (level 1)
  // window and in-tile order
  auto try_all = [&](std::vector<Cfg> cands) -> int {
    for (const Cfg& c : cands) {
      double ms;
      const int r = time_of(c, ms);
      if (r) return r;
      if (ms < best_ms) { best_ms = ms; best = c; }
    }
    return FMT2D_OK;
  };
  std::vector<Cfg> cands;
  const int shapes[3][2] = {{8, 8}, {16, 8}, {8, 16}};
  for (auto& sh : shapes) {
    if (sh[0] == best.tx && sh[1] == best.tz) continue;
    if (ctx->field_f32 && (sh[0] != 8 || sh[1] != 8)) continue;  // 32-bit storage solves 8x8 tiles
    Cfg c = best; c.tx = sh[0]; c.tz = sh[1];
    if (c.tx * c.tz > 64) { c.cap = std::max(c.cap, 16); c.hcap = std::max(c.hcap, 16); }
    cands.push_back(c);
  }
  if ((rc = try_all(cands))) { apply(start); return rc; }
  cands.clear();
  for (int cap : {8, 12, 16, 24}) if (cap != best.cap) { Cfg c = best; c.cap = cap; cands.push_back(c); }
  if ((rc = try_all(cands))) { apply(start); return rc; }
  cands.clear();
  for (int cap : {6, 8, 12, 16}) if (cap != best.hcap) { Cfg c = best; c.hcap = cap; cands.push_back(c); }
  if ((rc = try_all(cands))) { apply(start); return rc; }
  if (level >= 1) {
    cands.clear();
    for (double w : {2.0, 3.0, 4.0, 6.0}) if (w != best.window) { Cfg c = best; c.window = w; cands.push_back(c); }
    for (int col : {2, 4}) if (col != best.colors) { Cfg c = best; c.colors = col; cands.push_back(c); }
    if ((rc = try_all(cands))) { apply(start); return rc; }
  }
  apply(best);
  out->tile_x = best.tx; out->tile_z = best.tz;
  out->max_sweeps = best.cap; out->max_sweeps_heavy = best.hcap;
  out->colors = best.colors; out->window_tiles = best.window;
  out->ms_best = best_ms; out->ms_start = start_ms;
  out->trials = trials;
  return FMT2D_OK;
}

int fmt2d_set_solver_launch(fmt2d_ctx* ctx, int mode) {
  if (!ctx || (mode != FMT2D_LAUNCH_GRAPH && mode != FMT2D_LAUNCH_LEVEL))
    return fail(FMT2D_EINVAL, "the solver launch mode must be FMT2D_LAUNCH_GRAPH or FMT2D_LAUNCH_LEVEL");
  ctx->solver_launch = mode;
  return FMT2D_OK;
}

int fmt2d_set_launch_cap(fmt2d_ctx* ctx, int max_launches) {
  if (!ctx || max_launches < 0) return fail(FMT2D_EINVAL, "the launch cap must be >= 0");
  ctx->launch_cap = max_launches;
  return FMT2D_OK;
}

int fmt2d_set_heavy_sweeps(fmt2d_ctx* ctx, int max_sweeps_heavy) {
  if (!ctx || max_sweeps_heavy < 1) return fail(FMT2D_EINVAL, "the sweep cap must be >= 1");
  ctx->max_sweeps_heavy = max_sweeps_heavy;
  return FMT2D_OK;
}

int fmt2d_forward_trace(fmt2d_ctx* ctx, int* counts, int* max_sweeps, int max, int* n) {
  if (!ctx || !n) return fail(FMT2D_EINVAL, "NULL argument");
  *n = (int)ctx->trace.size();
  for (int i = 0; i < std::min(max, *n); i++) {
    if (counts) counts[i] = ctx->trace[i];
    if (max_sweeps) max_sweeps[i] = ctx->trace_sweeps[i];
  }
  return FMT2D_OK;
}

int fmt2d_forward_sweep_hist(fmt2d_ctx* ctx, long long* hist) {
  if (!ctx || !hist) return fail(FMT2D_EINVAL, "NULL argument");
  for (int k = 0; k < 64; k++) hist[k] = ctx->sweep_hist[k];
  return FMT2D_OK;
}

int fmt2d_forward_summary(fmt2d_ctx* ctx, fmt2d_field_summary* out, int cap, int* n) {
  if (!ctx || !n) return fail(FMT2D_EINVAL, "NULL argument");
  if (ctx->nsrc < 1) return fail(FMT2D_ESTATE, "no forward run");
  *n = ctx->nfield;
  if (!out || cap == 0) return FMT2D_OK;
  if (cap < ctx->nfield) return fail(FMT2D_EINVAL, "output buffer too small");
  HIPCHK(hipSetDevice(ctx->device));
  const int nf = ctx->nfield;
  DBuf<FieldSummary> d;
  HIPCHK(d.alloc(nf));
  HIPCHK(launch_field_summary(ctx->descs.p, ctx->regs.p, nf, ctx->m.nx, ctx->m.nz, d.p, ctx->stream));
  std::vector<FieldSummary> h(nf);
  HIPCHK(hipMemcpyAsync(h.data(), d.p, sizeof(FieldSummary) * nf, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < nf; k++) {
    const FieldSummary& s = h[ctx->field_of[k]];
    fmt2d_field_summary& o = out[k];
    o.reached = s.reached;
    o.hash = s.hash;
    o.max_resid = s.max_resid;
    o.mean_resid = s.n_resid > 0 ? s.sum_resid / (double)s.n_resid : 0.0;
    o.n_resid = s.n_resid;
    o.n_resid_big = s.n_resid_big;
  }
  return FMT2D_OK;
}

int fmt2d_forward_stats_get(fmt2d_ctx* ctx, fmt2d_forward_stats* st) {
  if (!ctx || !st) return fail(FMT2D_EINVAL, "NULL argument");
  *st = ctx->stats;
  return FMT2D_OK;
}

int fmt2d_get_field(fmt2d_ctx* ctx, int isrc, int itf, double* T, double* dTdx, double* dTdz) {
  if (!ctx || !T || !dTdx || !dTdz) return fail(FMT2D_EINVAL, "NULL argument");
  if (isrc < 0 || isrc >= ctx->nsrc || itf < 1 || itf > 12) return fail(FMT2D_EINVAL, "field index out of range");
  HIPCHK(hipSetDevice(ctx->device));
  const HostModel& m = ctx->m;
  const size_t nslot = (size_t)(m.nz + 2) * m.nx;
  const int f = ctx->field_of[isrc * FMT2D_NTF + itf - 1];
  std::vector<double> t(nslot), gx(nslot), gz(nslot);
  // FMT2D_RAW_GRADIENTS=1 (diagnostics): the marching gradients before the
  // one-sided refit and the water-layer fixes
  const bool raw = getenv("FMT2D_RAW_GRADIENTS") != nullptr;
  HIPCHK(hipMemcpyAsync(t.data(), ctx->fT.p + (size_t)f * nslot, nslot * sizeof(double), hipMemcpyDeviceToHost,
                        ctx->stream));
  HIPCHK(hipMemcpyAsync(gx.data(), (raw ? ctx->fGX.p : ctx->fOGX.p) + (size_t)f * nslot, nslot * sizeof(double),
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(gz.data(), (raw ? ctx->fGZ.p : ctx->fOGZ.p) + (size_t)f * nslot, nslot * sizeof(double),
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const HostRegion& R = m.reg[TF_REGION[itf]];
  for (int k = 0; k < R.nnode; k++) {
    const int q = R.ref_order[k];
    T[k] = t[q]; dTdx[k] = gx[q]; dTdz[k] = gz[q];
  }
  return FMT2D_OK;
}

}  // extern "C"
