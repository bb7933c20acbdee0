// Seeds and the global-memory view of one resident traveltime field (device),
// shared by the solver's post-passes and the adjoint-state backward.
#pragma once
#include <hip/hip_runtime.h>

#include "eik_kernels.h"
#include "eik_update.h"

namespace fmt2d {

// Seed of grid node (i,j) of field d: point-source cell corners
// (initialize_region_from_source2d, :2307-2411) or the grid nodes 1-3 rows
// into the region from the start interface, first interface node wins
// (sweep_region_from_interface2d mapped branch, :2141-2222).
__device__ inline void seed_grid(const FieldDesc& d, const RegionDev& R, int i, int j, double& t0, double& gx,
                                 double& gz) {
  t0 = FMT_HUGE_TIME; gx = 0.0; gz = 0.0;
  if (!(R.flags[j * R.nx + i] & RF_IN)) return;
  double v = R.vel[j * R.nx + i];
  if (v < 0.01) v = 0.01;
  if (d.kind == SEED_POINT) {
    if (i < d.ix0 || i > d.ix0 + 1 || j < d.iz0 || j > d.iz0 + 1) return;
    const double dx = R.xg[i] - d.sx, dz = R.zg[j] - d.sz;
    const double dist = sqrt(dx * dx + dz * dz);
    if (dist < 0.01 * R.tol) { t0 = d.t0; return; }
    t0 = d.t0 + dist / v;
    gx = dx / (dist * v); gz = dz / (dist * v);
    return;
  }
  const int s = d.side;
  for (int n = i - 1; n <= i + 1; n++) {
    if (n < 0 || n >= R.nx) continue;
    const int cc = R.ifcc[s][n];
    if (cc < 0) continue;
    const bool hit = (s == 1) ? (j <= cc && j >= cc - 2) : (j >= cc + 1 && j <= cc + 3);
    if (!hit) continue;
    const double px = R.ifx[s][n], pz = R.ifz[s][n];
    const double dx = R.xg[i] - px, dz = R.zg[j] - pz;
    const double dist = sqrt(dx * dx + dz * dz);
    t0 = d.isT[n] + dist / v;
    gx = dx / (dist * v); gz = dz / (dist * v);
    return;
  }
}

__device__ inline void seed_iface(const FieldDesc& d, int s, int c, double& t0, double& gx, double& gz) {
  t0 = FMT_HUGE_TIME; gx = 0.0; gz = 0.0;
  if (d.kind == SEED_IFACE && s == d.side) { t0 = d.isT[c]; gx = d.isGX[c]; gz = d.isGZ[c]; }
}

// Global-memory view of one field (post-passes).
struct GView {
  const RegionDev* R;
  const FieldDesc* D;
  __device__ const RegionDev& reg() const { return *R; }
  __device__ bool in(int i, int j) const {
    return i >= 0 && i < R->nx && j >= 0 && j < R->nz && (R->flags[j * R->nx + i] & RF_IN);
  }
  __device__ uint16_t flags(int i, int j) const { return R->flags[j * R->nx + i]; }
  __device__ double T(int i, int j) const { return D->T[j * R->nx + i]; }
  __device__ double GX(int i, int j) const { return D->GX[j * R->nx + i]; }
  __device__ double GZ(int i, int j) const { return D->GZ[j * R->nx + i]; }
  __device__ int tie(int i, int j) const { return D->TIE[j * R->nx + i]; }
  __device__ double vel(int i, int j) const { return R->vel[j * R->nx + i]; }
  __device__ double xg(int i) const { return R->xg[i]; }
  __device__ double zg(int j) const { return R->zg[j]; }
  __device__ double iT(int s, int c) const { return D->T[(R->nz + s) * R->nx + c]; }
  __device__ double iGX(int s, int c) const { return D->GX[(R->nz + s) * R->nx + c]; }
  __device__ double iGZ(int s, int c) const { return D->GZ[(R->nz + s) * R->nx + c]; }
  __device__ double ivel(int s, int c) const { return R->vel[(R->nz + s) * R->nx + c]; }
  __device__ double ifx(int s, int c) const { return R->ifx[s][c]; }
  __device__ double ifz(int s, int c) const { return R->ifz[s][c]; }
  __device__ int ifcc(int s, int c) const { return R->ifcc[s][c]; }
  __device__ int ifok(int s, int c) const { return R->ifok[s][c]; }
  __device__ double T0(int i, int j) const { double t, a, b; seed_grid(*D, *R, i, j, t, a, b); return t; }
  __device__ double GX0(int i, int j) const { double t, a, b; seed_grid(*D, *R, i, j, t, a, b); return a; }
  __device__ double GZ0(int i, int j) const { double t, a, b; seed_grid(*D, *R, i, j, t, a, b); return b; }
  __device__ double iT0(int s, int c) const { double t, a, b; seed_iface(*D, s, c, t, a, b); return t; }
  __device__ double iGX0(int s, int c) const { double t, a, b; seed_iface(*D, s, c, t, a, b); return a; }
  __device__ double iGZ0(int s, int c) const { double t, a, b; seed_iface(*D, s, c, t, a, b); return b; }
  __device__ bool fixed(const Member& m) const {
    return m.kind == 0 && (m.j * R->nx + m.i) == D->fixed_key;
  }
};


}  // namespace fmt2d
