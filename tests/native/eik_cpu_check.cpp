// CPU harness for the shared eikonal update operator (fmt2d_amd/csrc/eik_update.h)
// and the host model build (fmt2d_amd/csrc/model.cpp).  Test infrastructure:
// it iterates the same operator the GPU kernels use, serially, to its fixed
// point so the operator itself can be checked against the oracle on a machine
// without a GPU.  It is not part of the product path.
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <stdlib.h>

#include <string>
#include <vector>

#include "../../fmt2d_amd/csrc/eik_update.h"
#include "../../fmt2d_amd/csrc/eik_regular.h"
#include "../../fmt2d_amd/csrc/model.h"
#include "heap_exit_sorted.h"

using namespace fmt2d;
#ifndef GT
#define GT double
#endif
#ifndef RT
#define RT double
#endif

struct CpuField {
  int nx, nz;
  std::vector<RT> T, T0;
  std::vector<GT> GX, GZ, GX0, GZ0;
  std::vector<uint8_t> TIE;
  int fixed_key = -1;
};

struct CpuView {
  const RegionDev* R;
  const CpuField* F;
  const RegionDev& reg() const { return *R; }
  bool in(int i, int j) const { return i >= 0 && i < R->nx && j >= 0 && j < R->nz && (R->flags[j * R->nx + i] & RF_IN); }
  uint16_t flags(int i, int j) const { return R->flags[j * R->nx + i]; }
  RT T(int i, int j) const { return F->T[j * R->nx + i]; }
  GT GX(int i, int j) const { return F->GX[j * R->nx + i]; }
  GT GZ(int i, int j) const { return F->GZ[j * R->nx + i]; }
  int tie(int i, int j) const { return F->TIE.empty() ? 0 : F->TIE[j * R->nx + i]; }
  double vel(int i, int j) const { return R->vel[j * R->nx + i]; }
  double xg(int i) const { return R->xg[i]; }
  double zg(int j) const { return R->zg[j]; }
  RT iT(int s, int c) const { return F->T[(R->nz + s) * R->nx + c]; }
  GT iGX(int s, int c) const { return F->GX[(R->nz + s) * R->nx + c]; }
  GT iGZ(int s, int c) const { return F->GZ[(R->nz + s) * R->nx + c]; }
  double ivel(int s, int c) const { return R->vel[(R->nz + s) * R->nx + c]; }
  double ifx(int s, int c) const { return R->ifx[s][c]; }
  double ifz(int s, int c) const { return R->ifz[s][c]; }
  int ifcc(int s, int c) const { return R->ifcc[s][c]; }
  int ifok(int s, int c) const { return R->ifok[s][c]; }
  RT T0(int i, int j) const { return F->T0[j * R->nx + i]; }
  GT GX0(int i, int j) const { return F->GX0[j * R->nx + i]; }
  GT GZ0(int i, int j) const { return F->GZ0[j * R->nx + i]; }
  RT iT0(int s, int c) const { return F->T0[(R->nz + s) * R->nx + c]; }
  GT iGX0(int s, int c) const { return F->GX0[(R->nz + s) * R->nx + c]; }
  GT iGZ0(int s, int c) const { return F->GZ0[(R->nz + s) * R->nx + c]; }
  bool fixed(const Member& m) const {
    const int key = m.kind == 0 ? m.j * R->nx + m.i : (R->nz + m.kind - 1) * R->nx + m.i;
    return key == F->fixed_key;
  }
};

// Fast-path check (eik_regular.h): when on, every evaluation of a fast-path
// grid node also runs reg_update on the same state and counts the results
// that differ from update_grid_node in any bit (T, gradient, equal-time mask).
static int g_fast_check = 0;
static long long g_fast_checked = 0, g_fast_bad = 0;

template <class V>
static void fast_check(const V& v, int i, int j, const Best<RT>& seed, const Best<RT>& b) {
  if (!g_fast_check || !reg_fast_node(v.flags(i, j))) return;
  const RegStatic s = reg_static(v, i, j);
  const RegDyn y = reg_dyn(v, s, i, j);
  double T, gx, gz;
  int tie;
  reg_update(s, y, seed.T, seed.gx, seed.gz, T, gx, gz, tie);
  g_fast_checked++;
  auto bits = [](double a) { long long q; memcpy(&q, &a, 8); return q; };
  if (bits(T) != bits(b.T) || bits(gx) != bits(b.gx) || bits(gz) != bits(b.gz) || tie != b.tie) {
    if (g_fast_bad < 5)
      fprintf(stderr, "fast path differs at (%d,%d): T %.17g/%.17g gx %.17g/%.17g gz %.17g/%.17g tie %d/%d\n", i, j, T,
              b.T, gx, b.gx, gz, b.gz, tie, b.tie);
    g_fast_bad++;
  }
}

struct Holder {
  HostModel m;
  std::vector<double> xg, zg, vel[NREG + 1], ifx[NREG + 1][2], ifz[NREG + 1][2];
  std::vector<int> ifcc[NREG + 1][2];
  std::vector<uint8_t> ifok[NREG + 1][2];
  RegionDev rd[NREG + 1];
};

// Emulation of the GPU block schedule: 8x8 tiles, each launch processes the
// active tiles against the state at launch start, a tile sweeps (Jacobi) up
// to 64 times, border changes activate neighbours for the next launch.
static int tile_schedule_run(const RegionDev& R, CpuField& F, int nx, int nz, long long* tiles_done) {
  const int TILE = 8, ntx = (nx + 7) / 8, ntz = (nz + 7) / 8;
  std::vector<char> act(ntx * ntz, 0);
  // seeds: tiles containing finite seeds
  for (int j = 0; j < nz + 2; j++)
    for (int i = 0; i < nx; i++) {
      const size_t g = (size_t)j * nx + i;
      if (F.T0[g] < FMT_HUGE_HALF) {
        int jj = j;
        if (j >= nz) jj = R.ifcc[j - nz][i] >= 0 ? R.ifcc[j - nz][i] : 0;
        act[(jj / TILE) * ntx + i / TILE] = 1;
      }
    }
  if (F.fixed_key >= 0) F.T[F.fixed_key] = 0;
  int launch = 0;
  for (; launch < 100000; launch++) {
    std::vector<int> list;
    for (int t = 0; t < ntx * ntz; t++) if (act[t]) list.push_back(t);
    if (list.empty()) break;
    std::fill(act.begin(), act.end(), 0);
    *tiles_done += list.size();
    CpuField start = F;  // every tile of the launch reads the launch-start state
    for (int t : list) {
      const int tx0 = (t % ntx) * TILE, tz0 = (t / ntx) * TILE;
      CpuField loc = getenv("EMU_GS") ? F : start;
      CpuView vv{&R, &loc};
      std::vector<int> keys;
      for (int j = tz0; j < std::min(nz, tz0 + TILE); j++)
        for (int i = tx0; i < std::min(nx, tx0 + TILE); i++) {
          const int g = j * nx + i;
          if ((R.flags[g] & RF_IN) && g != F.fixed_key) keys.push_back(g);
        }
      for (int s = 0; s < 2; s++)
        for (int c = tx0; c < std::min(nx - 1, tx0 + TILE); c++) {
          const int cc = R.ifcc[s][c];
          if (cc < 0) continue;
          const int home = iface_home_row(vv, s, c, cc);
          if (home >= tz0 && home < tz0 + TILE) keys.push_back((nz + s) * nx + c);
        }
      std::vector<char> dirty(keys.size(), 0);
      int sweep;
      bool more = true;
      for (sweep = 0; sweep < 64; sweep++) {
        CpuField snap = loc;
        CpuView sv{&R, &snap};
        bool ch = false;
        for (size_t k = 0; k < keys.size(); k++) {
          const int g = keys[k], j = g / nx, i = g % nx;
          Best<RT> b{F.T0[g], F.GX0[g], F.GZ0[g], 0, 0, 0, 1};
          const Best<RT> seed = b;
          if (j < nz) { update_grid_node<RT>(sv, i, j, b); fast_check(sv, i, j, seed, b); }
          else update_iface_node<RT>(sv, j - nz, i, b);
          const int tie = j < nz ? b.tie : 0;
          if (b.T != loc.T[g] || (GT)b.gx != loc.GX[g] || (GT)b.gz != loc.GZ[g] || tie != loc.TIE[g]) {
            loc.T[g] = b.T; loc.GX[g] = b.gx; loc.GZ[g] = b.gz; loc.TIE[g] = tie; ch = true; dirty[k] = 1;
          }
        }
        if (!ch) { more = false; break; }
      }
      unsigned mask = more ? 256u : 0u;
      for (size_t k = 0; k < keys.size(); k++) {
        if (!dirty[k]) continue;
        const int g = keys[k];
        F.T[g] = loc.T[g]; F.GX[g] = loc.GX[g]; F.GZ[g] = loc.GZ[g]; F.TIE[g] = loc.TIE[g];
        int j = g / nx, i = g % nx;
        if (j >= nz) j = iface_home_row(vv, j - nz, i, R.ifcc[j - nz][i]);
        const int a = i - tx0, b = j - tz0;
        const bool L0 = a < 2, R0 = a >= TILE - 2, T0 = b < 2, B0 = b >= TILE - 2;
        if (L0) mask |= 1; if (R0) mask |= 2; if (T0) mask |= 4; if (B0) mask |= 8;
        if (L0 && T0) mask |= 16; if (R0 && T0) mask |= 32; if (L0 && B0) mask |= 64; if (R0 && B0) mask |= 128;
      }
      const int dxs[9] = {-1, 1, 0, 0, -1, 1, -1, 1, 0}, dzs[9] = {0, 0, -1, 1, -1, -1, 1, 1, 0};
      for (int q = 0; q < 9; q++)
        if (mask & (1u << q)) {
          const int a = t % ntx + dxs[q], b = t / ntx + dzs[q];
          if (a >= 0 && a < ntx && b >= 0 && b < ntz) act[b * ntx + a] = 1;
        }
    }
  }
  return launch;
}

extern "C" {

// reg_root of the fast path (eik_regular.h), for the discriminant checks
int cpu_reg_root(double a, double b, double c, double tmax, double* t1) { return reg_root(a, b, c, tmax, *t1) ? 1 : 0; }

void cpu_fast_check_set(int on) {
  g_fast_check = on;
  g_fast_checked = 0;
  g_fast_bad = 0;
}

void cpu_fast_check_get(long long* checked, long long* bad) {
  *checked = g_fast_checked;
  *bad = g_fast_bad;
}

void* cpu_model_create(int nx_s, int nz_s, const double* xpos, const double* topo, const double* zrel,
                       const double* vel, double vwater, int nsed, const double* sx, const double* sz,
                       int nm, const double* mx, const double* mz, char* err, int errlen) {
  Holder* h = new Holder();
  std::string e = build_model(h->m, nx_s, nz_s, xpos, topo, zrel, vel, vwater, nsed, sx, sz, nm, mx, mz);
  if (!e.empty()) { snprintf(err, errlen, "%s", e.c_str()); delete h; return nullptr; }
  HostModel& m = h->m;
  h->xg.assign(m.xg.begin(), m.xg.end());
  h->zg.assign(m.zg.begin(), m.zg.end());
  for (int r = 1; r <= NREG; r++) {
    HostRegion& R = m.reg[r];
    h->vel[r].assign(R.vel.begin(), R.vel.end());
    RegionDev& d = h->rd[r];
    d.nx = m.nx; d.nz = m.nz; d.dx = m.dx; d.dz = m.dz; d.tol = m.tol;
    d.xg = h->xg.data(); d.zg = h->zg.data(); d.flags = R.flags.data(); d.vel = h->vel[r].data();
    const int ids[2] = {R.itop, R.ibot};
    for (int s = 0; s < 2; s++) {
      const HostInterface& is = m.isec[ids[s]];
      h->ifx[r][s].assign(is.x.begin(), is.x.end());
      h->ifz[r][s].assign(is.z.begin(), is.z.end());
      h->ifcc[r][s] = is.cc;
      h->ifok[r][s] = is.ok;
      d.ifx[s] = h->ifx[r][s].data(); d.ifz[s] = h->ifz[r][s].data();
      d.ifcc[s] = h->ifcc[r][s].data(); d.ifok[s] = h->ifok[r][s].data();
    }
  }
  return h;
}

int cpu_model_dims(void* hp, int* nx, int* nz) {
  Holder* h = (Holder*)hp;
  *nx = h->m.nx; *nz = h->m.nz;
  return 0;
}

int cpu_region_nnode(void* hp, int r) { return ((Holder*)hp)->m.reg[r].nnode; }

void cpu_region_export(void* hp, int r, int* order, double* vel, int* flags) {
  Holder* h = (Holder*)hp;
  const HostRegion& R = h->m.reg[r];
  for (int k = 0; k < R.nnode; k++) {
    order[k] = R.ref_order[k];
    vel[k] = R.vel[R.ref_order[k]];
  }
  for (size_t g = 0; g < R.flags.size(); g++) flags[g] = R.flags[g];
}

// Point source at (sx, seafloor) in region r, iterated to the fixed point
// with Gauss-Seidel sweeps.  Outputs in the reference node order.
int cpu_point_source(void* hp, int r, double sx, double* outT, double* outGX, double* outGZ, int maxit) {
  Holder* h = (Holder*)hp;
  HostModel& m = h->m;
  const RegionDev& R = h->rd[r];
  const int nx = m.nx, nz = m.nz;
  const size_t N = (size_t)(nz + 2) * nx;
  CpuField F;
  F.nx = nx; F.nz = nz;
  F.T0.assign(N, FMT_HUGE_TIME); F.GX0.assign(N, 0.f); F.GZ0.assign(N, 0.f);
  // initialize_region_from_source2d (src/propagate2d.f90:2307-2411)
  const double sz = m.iface_z_at(2, sx);
  int ix0 = (int)floor((sx - m.x0) / m.dx), iz0 = (int)floor((sz - m.z0) / m.dz);
  ix0 = std::max(0, std::min(ix0, nx - 2));
  iz0 = std::max(0, std::min(iz0, nz - 2));
  for (int i = ix0; i <= ix0 + 1; i++)
    for (int j = iz0; j <= iz0 + 1; j++) {
      const size_t g = (size_t)j * nx + i;
      if (!(R.flags[g] & RF_IN)) continue;
      const double dx = m.xg[i] - sx, dz = m.zg[j] - sz, dist = sqrt(dx * dx + dz * dz);
      double va = m.reg[r].vel[g];
      if (va < 0.01) va = 0.01;
      if (dist < 0.01 * m.tol) { F.T0[g] = 0; F.fixed_key = (int)g; }
      else { F.T0[g] = (RT)(dist / va); F.GX0[g] = (RT)(dx / (dist * va)); F.GZ0[g] = (RT)(dz / (dist * va)); }
    }
  F.T = F.T0; F.GX = F.GX0; F.GZ = F.GZ0;
  F.TIE.assign(N, 0);
  CpuView v{&R, &F};
  int it;
  long long tiles_done = 0;
  int launches = -1;
  if (maxit == 0) { launches = tile_schedule_run(R, F, nx, nz, &tiles_done); maxit = 0; }
  // jacobi != 0: every node of a sweep reads the previous sweep's values
  // (the GPU tile sweeps); otherwise in-place Gauss-Seidel.
  const int jacobi = maxit < 0;
  it = launches;
  if (maxit < 0) maxit = -maxit;
  std::vector<RT> nT(N);
  std::vector<GT> nGX(N), nGZ(N);
  for (it = (launches >= 0 ? launches : 0); launches < 0 && it < maxit; it++) {
    int changed = 0;
    CpuField snap;
    if (jacobi) { snap = F; v.F = &snap; }
    for (int j = 0; j < nz + 2; j++)
      for (int i = 0; i < nx; i++) {
        const size_t g = (size_t)j * nx + i;
        Best<RT> b{F.T0[g], F.GX0[g], F.GZ0[g], 0, 0, 0, 1};
        if (j < nz) {
          if (!(R.flags[g] & RF_IN) || (int)g == F.fixed_key) continue;
          const Best<RT> seed = b;
          update_grid_node<RT>(v, i, j, b);
          fast_check(v, i, j, seed, b);
        } else {
          update_iface_node<RT>(v, j - nz, i, b);
        }
        const int tie = j < nz ? b.tie : 0;
        if (b.T != F.T[g] || (GT)b.gx != F.GX[g] || (GT)b.gz != F.GZ[g] || tie != F.TIE[g]) {
          F.T[g] = b.T; F.GX[g] = b.gx; F.GZ[g] = b.gz; F.TIE[g] = tie; changed++;
        }
      }
    v.F = &F;
    if (!changed) break;
  }
  // heap exhaustion (src/propagate2d.f90:80)
  {
    std::vector<double> TT, MX;
    std::vector<char> SD, FX;
    std::vector<int> slot;
    for (int j = 0; j < nz + 2; j++)
      for (int i = 0; i < nx; i++) {
        const size_t g = (size_t)j * nx + i;
        double mx;
        if (j < nz) {
          if (!(R.flags[g] & RF_IN)) continue;
          mx = max_member_time_grid<RT>(v, i, j);
        } else {
          mx = max_member_time_iface<RT>(v, j - nz, i);
        }
        slot.push_back((int)g);
        TT.push_back(F.T[g]);
        MX.push_back(mx);
        bool sd = F.T0[g] < FMT_HUGE_HALF;
        if (F.fixed_key >= 0 && j < nz) {  // diagonal trial nodes of an on-node source
          const int fi = F.fixed_key % nx, fj = F.fixed_key / nx;
          if (abs(i - fi) == 1 && abs(j - fj) == 1) sd = true;
        }
        SD.push_back(sd);
        FX.push_back((int)g == F.fixed_key);
      }
    const double te = fmt2d_test::heap_exit_time(TT, MX, SD, FX);
    if (getenv("EMU_EXIT_DEBUG")) fprintf(stderr, "heap exit time %.9f\n", te);
    for (size_t k = 0; k < slot.size(); k++)
      if (TT[k] > te && TT[k] < FMT_HUGE_HALF) { F.T[slot[k]] = FMT_HUGE_TIME; F.GX[slot[k]] = 0; F.GZ[slot[k]] = 0; }
  }
  // gradient refit of one-sided nodes (src/propagate2d.f90:119-123)
  std::vector<GT> OX = F.GX, OZ = F.GZ;
  for (int j = 0; j < nz; j++)
    for (int i = 0; i < nx; i++) {
      const size_t g = (size_t)j * nx + i;
      if (!(R.flags[g] & RF_IN) || (int)g == F.fixed_key) continue;
      const bool susp = grid_suspect<RT>(v, i, j, F.T0[g]) != 0;
      if (g_fast_check && reg_fast_node(v.flags(i, j))) {
        // the refit pass takes the flag of fast-path nodes from the fast path's walk
        const RegStatic rs = reg_static(v, i, j);
        g_fast_checked++;
        if (reg_suspect(rs, reg_dyn(v, rs, i, j), F.T0[g]) != susp) {
          if (g_fast_bad < 5) fprintf(stderr, "fast-path suspect flag differs at (%d,%d)\n", i, j);
          g_fast_bad++;
        }
      }
      if (susp) { RT dx, dz; fit_gradient_grid<RT>(v, i, j, dx, dz); OX[g] = dx; OZ[g] = dz; }
    }
  const HostRegion& HR = m.reg[r];
  for (int k = 0; k < HR.nnode; k++) {
    const int g = HR.ref_order[k];
    outT[k] = F.T[g]; outGX[k] = OX[g]; outGZ[k] = OZ[g];
  }
  return it;
}

// One Jacobi pass of the operator over a given field (reference node order):
// where the result differs from the input, the field is no fixed point of
// the operator.  Diagnostics of operator-vs-FMM differences.
int cpu_point_eval(void* hp, int r, double sx, const double* inT, const double* inGX, const double* inGZ,
                   double* outT, double* outGX, double* outGZ) {
  Holder* h = (Holder*)hp;
  HostModel& m = h->m;
  const RegionDev& R = h->rd[r];
  const int nx = m.nx, nz = m.nz;
  const size_t N = (size_t)(nz + 2) * nx;
  CpuField F;
  F.nx = nx; F.nz = nz;
  F.T0.assign(N, FMT_HUGE_TIME); F.GX0.assign(N, 0.f); F.GZ0.assign(N, 0.f);
  const double sz = m.iface_z_at(2, sx);
  int ix0 = (int)floor((sx - m.x0) / m.dx), iz0 = (int)floor((sz - m.z0) / m.dz);
  ix0 = std::max(0, std::min(ix0, nx - 2));
  iz0 = std::max(0, std::min(iz0, nz - 2));
  for (int i = ix0; i <= ix0 + 1; i++)
    for (int j = iz0; j <= iz0 + 1; j++) {
      const size_t g = (size_t)j * nx + i;
      if (!(R.flags[g] & RF_IN)) continue;
      const double dx = m.xg[i] - sx, dz = m.zg[j] - sz, dist = sqrt(dx * dx + dz * dz);
      double va = m.reg[r].vel[g];
      if (va < 0.01) va = 0.01;
      if (dist < 0.01 * m.tol) { F.T0[g] = 0; F.fixed_key = (int)g; }
      else { F.T0[g] = (RT)(dist / va); F.GX0[g] = (RT)(dx / (dist * va)); F.GZ0[g] = (RT)(dz / (dist * va)); }
    }
  F.T.assign(N, FMT_HUGE_TIME); F.GX.assign(N, 0.0); F.GZ.assign(N, 0.0);
  const HostRegion& HR = m.reg[r];
  for (int k = 0; k < HR.nnode; k++) {
    const int g = HR.ref_order[k];
    F.T[g] = inT[k]; F.GX[g] = inGX[k]; F.GZ[g] = inGZ[k];
  }
  CpuView v{&R, &F};
  // pass 1: equal-time support masks of the given field; pass 2: evaluate
  F.TIE.assign(N, 0);
  std::vector<uint8_t> tie(N, 0);
  for (int pass = 0; pass < 2; pass++) {
    for (int k = 0; k < HR.nnode; k++) {
      const int g = HR.ref_order[k], j = g / nx, i = g % nx;
      Best<RT> b{F.T0[g], F.GX0[g], F.GZ0[g], 0, 0, 0, 1};
      if (g == F.fixed_key) { outT[k] = F.T[g]; outGX[k] = F.GX[g]; outGZ[k] = F.GZ[g]; continue; }
      if (j < nz) update_grid_node<RT>(v, i, j, b); else update_iface_node<RT>(v, j - nz, i, b);
      if (pass == 0) { if (j < nz) tie[g] = (uint8_t)b.tie; continue; }
      outT[k] = b.T; outGX[k] = b.gx; outGZ[k] = b.gz;
    }
    F.TIE = tie;
  }
  return 0;
}

struct PrintVisitor {
  Best<double> b;
  double last_evt = -1;
  void note(int, const Member&, const Member&) const {}
  bool want(double evt) { last_evt = evt; return true; }
  void add(double T, double gx, double gz, int kind, int susp, int tie) {
    fprintf(stderr, "    cand kind=%d evt=%.17g T=%.17g g=(%.17g,%.17g) susp=%d tie=%d%s\n", kind, last_evt, T, gx, gz, susp, tie,
            T < b.T ? "  <-- new min" : "");
    if (T < b.T) { b.T = T; b.gx = gx; b.gz = gz; }
  }
};

// Prints every candidate of node k (reference order) of the given field.
int cpu_node_debug(void* hp, int r, double sx, const double* inT, const double* inGX, const double* inGZ, int k) {
  Holder* h = (Holder*)hp;
  HostModel& m = h->m;
  const RegionDev& R = h->rd[r];
  const int nx = m.nx, nz = m.nz;
  const size_t N = (size_t)(nz + 2) * nx;
  CpuField F;
  F.nx = nx; F.nz = nz;
  F.T0.assign(N, FMT_HUGE_TIME); F.GX0.assign(N, 0.f); F.GZ0.assign(N, 0.f);
  F.T.assign(N, FMT_HUGE_TIME); F.GX.assign(N, 0.0); F.GZ.assign(N, 0.0);
  const HostRegion& HR = m.reg[r];
  for (int q = 0; q < HR.nnode; q++) {
    const int g = HR.ref_order[q];
    F.T[g] = inT[q]; F.GX[g] = inGX[q]; F.GZ[g] = inGZ[q];
  }
  CpuView v{&R, &F};
  const int g = HR.ref_order[k], j = g / nx, i = g % nx;
  fprintf(stderr, "node %d slot (%d,%d) T=%.17g flags=%x\n", k, i, j, F.T[g], j < nz ? R.flags[g] : 0);
  PrintVisitor pv;
  pv.b.T = FMT_HUGE_TIME; pv.b.gx = 0; pv.b.gz = 0; pv.b.susp = 0; pv.b.reg_set = 0; pv.b.tie = 0; pv.b.indep = 1;
  if (j < nz) grid_candidates<double>(v, i, j, pv); else iface_candidates<double>(v, j - nz, i, pv);
  return 0;
}

}  // extern "C"
