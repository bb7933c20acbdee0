// Ray backtracing and Frechet contributions, one thread per ray.
//
// Every routine follows its reference counterpart operation for operation in
// FP64 (this file is compiled with -ffp-contract=off), so that the discrete
// decisions of the tracer (interface hits, step halvings, source snapping)
// are taken exactly as the reference takes them.
#include <hip/hip_runtime.h>
#include <math.h>

#include "eik_common.h"
#include "rays.h"

namespace fmt2d {

#define RAY_HUGE 1.0e20

// path tables of invert2d_main.f90:564-681: stage fields per section and the
// intersection sequence (istart of section n = seq[2n-1], iend = seq[2n-2])
__constant__ int c_path_ntf[6] = {0, 1, 2, 2, 4, 4};
__constant__ int c_path_tf[6][4] = {{0, 0, 0, 0}, {1, 0, 0, 0}, {2, 3, 0, 0}, {4, 5, 0, 0}, {1, 6, 7, 8}, {9, 10, 11, 12}};
__constant__ int c_path_seq[6][8] = {{0}, {0, 2}, {0, 3, 3, 0}, {0, 4, 4, 0}, {0, 2, 2, 1, 1, 2, 2, 2},
                                     {0, 1, 1, 2, 2, 2, 2, 2}};
__constant__ int c_tf_region[13] = {0, 5, 2, 2, 6, 6, 1, 1, 1, 1, 1, 5, 1};

__device__ inline int ifloor(double v) {
  const double f = floor(v);
  return f < -1.0e9 ? -1000000000 : (f > 1.0e9 ? 1000000000 : (int)f);
}
__device__ inline double clamp01(double w) { return w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w); }

// Two neighbouring entries of a node array in one load (the cell's corners
// along x): every lane of a wave gathers from its own ray's cell, so fewer,
// wider loads mean fewer cache requests per interpolation.
struct __attribute__((aligned(8))) Pair64 { double lo, hi; };
struct __attribute__((aligned(2))) Pair16 { uint16_t lo, hi; };
__device__ inline Pair64 load_pair(const double* p) { return *(const Pair64*)p; }
__device__ inline Pair16 load_pair(const uint16_t* p) { return *(const Pair16*)p; }

// ---------------------------------------------------------------- region lookups

// region_rnode_id(ix,iz) (1-based): field slot of the grid node, -1 outside
__device__ inline int rnode(const RayGeomDev& G, const RayRegionDev& R, int ix, int iz) {
  if (ix < 1 || ix > G.nx || iz < 1 || iz > G.nz) return -1;
  const int q = (iz - 1) * G.nx + (ix - 1);
  return (R.flags[q] & RF_IN) ? q : -1;
}

// find_nearby_region_node (rays2d.f90:1010-1034): the first region node of
// rows iz+1, iz-1, ..., iz+5, iz-5 in column ix.  The ten flags are read at
// once (one memory round trip instead of up to ten in a row), then taken in
// that order.
__device__ inline int rnode_nearby(const RayGeomDev& G, const RayRegionDev& R, int ix, int iz) {
  if (ix < 1 || ix > G.nx) return -1;
  uint16_t f[10];
#pragma unroll
  for (int k = 0; k < 10; k++) {
    const int j = iz + (k & 1 ? -(k / 2 + 1) : k / 2 + 1);
    f[k] = (j >= 1 && j <= G.nz) ? R.flags[(j - 1) * G.nx + (ix - 1)] : (uint16_t)0;
  }
#pragma unroll
  for (int k = 0; k < 10; k++)
    if (f[k] & RF_IN) return (iz + (k & 1 ? -(k / 2 + 1) : k / 2 + 1) - 1) * G.nx + (ix - 1);
  return -1;
}

__device__ inline void slot_xz(const RayGeomDev& G, const RayRegionDev& R, int q, double& x, double& z) {
  if (q < G.nz * G.nx) { x = G.xg[q % G.nx]; z = G.zg[q / G.nx]; return; }
  const int s = q / G.nx - G.nz, c = q % G.nx, k = s == 0 ? R.itop : R.ibot;
  x = G.isx[k][c]; z = G.isz[k][c];
}

// nearest region node over the whole node list, first minimum in node order
// (the reference's scan, interpolate_arrivaltime_reg / find_position_in_region2d)
__device__ int nearest_scan(const RayGeomDev& G, const RayRegionDev& R, double px, double pz, double dmin0) {
  double dmin = dmin0;
  int best = -1;
  for (int k = 0; k < R.nnode; k++) {
    const int q = R.order[k];
    double x, z;
    slot_xz(G, R, q, x, z);
    const double d = (px - x) * (px - x) + (pz - z) * (pz - z);
    if (d < dmin) { dmin = d; best = q; }
  }
  return best;
}

// The same node from the column index: every node of the region lies on a
// grid column (grid nodes at xg, the column's top and bottom interface nodes
// at their own x), so columns are visited outward from the point's and
// skipped once the x bound of the remaining columns alone exceeds the best
// distance.  In a column the candidates are the region's grid rows next to
// pz (clamped to the column's run of rows) and the two interface nodes.
// Distances are formed as in the scan and ties go to the node that comes
// first in node order (grid nodes ix-outer / iz-inner, then the top
// interface's nodes, then the bottom's), so the result is the scan's.
struct NearBest {
  double d;
  long long rank;
  int q;
};

__device__ inline void near_offer(NearBest& b, double px, double pz, double x, double z, long long rank, int q) {
  const double d = (px - x) * (px - x) + (pz - z) * (pz - z);
  if (d < b.d || (b.q >= 0 && d == b.d && rank < b.rank)) { b.d = d; b.rank = rank; b.q = q; }
}

__device__ inline void near_column(const RayGeomDev& G, const RayRegionDev& R, int c, double px, double pz,
                                   NearBest& b) {
  const long long ng = (long long)G.nx * G.nz;
  const int lo = R.jlo[c], hi = R.jhi[c];
  if (lo <= hi) {
    // rows around pz, clamped to the run (a row offered twice changes nothing)
    const int jr = ifloor((pz - G.z0) / G.dz);
    for (int j = jr - 1; j <= jr + 2; j++) {
      const int jj = j < lo ? lo : (j > hi ? hi : j);
      near_offer(b, px, pz, G.xg[c], G.zg[jj], (long long)c * G.nz + jj, jj * G.nx + c);
    }
  }
  near_offer(b, px, pz, G.isx[R.itop][c], G.isz[R.itop][c], ng + c, G.nz * G.nx + c);
  near_offer(b, px, pz, G.isx[R.ibot][c], G.isz[R.ibot][c], ng + G.nx + c, (G.nz + 1) * G.nx + c);
}

__device__ unsigned long long g_nearest_calls;  // nearest-node searches (diagnostics)

__device__ int nearest_any(const RayGeomDev& G, const RayRegionDev& R, double px, double pz, double dmin0) {
  atomicAdd(&g_nearest_calls, 1ull);
  if (!R.indexed) return nearest_scan(G, R, px, pz, dmin0);
  NearBest b{dmin0, 0, -1};
  int c0 = ifloor((px - G.x0) / G.dx + 0.5);
  c0 = c0 < 0 ? 0 : (c0 >= G.nx ? G.nx - 1 : c0);
  for (int c = c0; c < G.nx; c++) {
    const double xl = R.xlo_from[c];
    if (xl > px && (xl - px) * (xl - px) > b.d) break;
    near_column(G, R, c, px, pz, b);
  }
  for (int c = c0 - 1; c >= 0; c--) {
    const double xh = R.xhi_upto[c];
    if (xh < px && (px - xh) * (px - xh) > b.d) break;
    near_column(G, R, c, px, pz, b);
  }
  return b.q;
}

// find_position_in_region2d (rays2d.f90:1076-1131)
__device__ int position_in_region(const RayGeomDev& G, const RayRegionDev& R, double px, double pz) {
  int ix = ifloor((px - G.x0) / G.dx) + 1, iz = ifloor((pz - G.z0) / G.dz) + 1;
  if (ix < 1) ix = 1;
  if (ix >= G.nx) ix = G.nx - 1;
  if (iz < 1) iz = 1;
  if (iz >= G.nz) iz = G.nz - 1;
  double distmin = 1.0e100;
  int inode = -1;
  for (int ii = ix; ii <= ix + 1; ii++)
    for (int jj = iz; jj <= iz + 1; jj++) {
      const int m = rnode(G, R, ii, jj);
      if (m < 0) continue;
      double x, z;
      slot_xz(G, R, m, x, z);
      const double dist = (px - x) * (px - x) + (pz - z) * (pz - z);
      if (dist < distmin) { inode = m; distmin = dist; }
    }
  if (inode < 0) inode = nearest_any(G, R, px, pz, 1.0e100);
  return inode;
}

// interpolate_time_gradient2d (rays2d.f90:825-894)
__device__ __attribute__((always_inline)) inline void interp_grad(const RayGeomDev& G, const RayRegionDev& R, const RayField& F, double px, double pz,
                            double& dtdx, double& dtdz) {
  int ix = ifloor((px - G.x0) / G.dx) + 1, iz = ifloor((pz - G.z0) / G.dz) + 1;
  if (ix < 1) ix = 1;
  if (ix >= G.nx) ix = G.nx - 1;
  if (iz < 1) iz = 1;
  if (iz >= G.nz) iz = G.nz - 1;
  // grid coordinates of the cell's corners (xg[ix-1], xg[ix], zg[iz-1], zg[iz])
  const Pair64 cx = load_pair(G.xg + ix - 1), cz = load_pair(G.zg + iz - 1);
  const double w1 = clamp01((px - cx.lo) / G.dx), w2 = clamp01((pz - cz.lo) / G.dz);
  // the cell's corners lie in the grid (ix < nx, iz < nz): their gradients
  // are loaded together with their region flags, and a corner outside the
  // region takes the nearby node's instead
  const int q1 = (iz - 1) * G.nx + (ix - 1), q3 = q1 + G.nx;
  const Pair64 x12 = load_pair(F.GX + q1), z12 = load_pair(F.GZ + q1);
  const Pair64 x34 = load_pair(F.GX + q3), z34 = load_pair(F.GZ + q3);
  const Pair16 f12 = load_pair(R.flags + q1), f34 = load_pair(R.flags + q3);
  double g1x = x12.lo, g1z = z12.lo, g2x = x12.hi, g2z = z12.hi;
  double g3x = x34.lo, g3z = z34.lo, g4x = x34.hi, g4z = z34.hi;
  const bool in1 = f12.lo & RF_IN, in2 = f12.hi & RF_IN, in3 = f34.lo & RF_IN, in4 = f34.hi & RF_IN;
  // the corners' own gradients (the nearest-node fallback below takes one)
  const double r1x = g1x, r1z = g1z, r2x = g2x, r2z = g2z, r3x = g3x, r3z = g3z, r4x = g4x, r4z = g4z;
  if (!in1) { const int m = rnode_nearby(G, R, ix, iz); g1x = m >= 0 ? F.GX[m] : 0.0; g1z = m >= 0 ? F.GZ[m] : 0.0; }
  if (!in2) { const int m = rnode_nearby(G, R, ix + 1, iz); g2x = m >= 0 ? F.GX[m] : 0.0; g2z = m >= 0 ? F.GZ[m] : 0.0; }
  if (!in3) { const int m = rnode_nearby(G, R, ix, iz + 1); g3x = m >= 0 ? F.GX[m] : 0.0; g3z = m >= 0 ? F.GZ[m] : 0.0; }
  if (!in4) {
    const int m = rnode_nearby(G, R, ix + 1, iz + 1);
    g4x = m >= 0 ? F.GX[m] : 0.0; g4z = m >= 0 ? F.GZ[m] : 0.0;
  }
  double gx = (1.0 - w1) * (1.0 - w2) * g1x + w1 * (1.0 - w2) * g2x + (1.0 - w1) * w2 * g3x + w1 * w2 * g4x;
  double gz = (1.0 - w1) * (1.0 - w2) * g1z + w1 * (1.0 - w2) * g2z + (1.0 - w1) * w2 * g3z + w1 * w2 * g4z;
  if (sqrt(gx * gx + gz * gz) < 0.01) {
    // interpolate_gradient_reg (rays2d.f90:1036-1073): the nearest node's
    // gradient.  find_position_in_region2d (position_in_region) takes the
    // nearest region node among this cell's corners, in the order (ix, iz),
    // (ix, iz+1), (ix+1, iz), (ix+1, iz+1), first minimum; the corners'
    // flags, coordinates and gradients are at hand.  With no corner in the
    // region it searches the whole region.
    double best = 1.0e100;
    bool found = false;
    auto offer = [&](bool in, double x, double z, double ax, double az) {
      if (!in) return;
      const double dist = (px - x) * (px - x) + (pz - z) * (pz - z);
      if (dist < best) { best = dist; gx = ax; gz = az; found = true; }
    };
    offer(in1, cx.lo, cz.lo, r1x, r1z);
    offer(in3, cx.lo, cz.hi, r3x, r3z);
    offer(in2, cx.hi, cz.lo, r2x, r2z);
    offer(in4, cx.hi, cz.hi, r4x, r4z);
    if (!found) {
      const int m = nearest_any(G, R, px, pz, 1.0e100);
      gx = m >= 0 ? F.GX[m] : 0.0;
      gz = m >= 0 ? F.GZ[m] : 0.0;
    }
  }
  dtdx = gx;
  dtdz = gz;
}

// interpolate_arrivaltime2d (rays2d.f90:898-958)
__device__ __attribute__((always_inline)) inline double interp_time(const RayGeomDev& G, const RayRegionDev& R, const RayField& F, double px, double pz) {
  int ix = ifloor((px - G.x0) / G.dx) + 1, iz = ifloor((pz - G.z0) / G.dz) + 1;
  if (ix < 1) ix = 1;
  if (ix >= G.nx) ix = G.nx - 1;
  if (iz < 1) iz = 1;
  if (iz >= G.nz) iz = G.nz - 1;
  const double w1 = clamp01((px - G.xg[ix - 1]) / G.dx), w2 = clamp01((pz - G.zg[iz - 1]) / G.dz);
  // corners in the grid: times loaded together with the region flags
  const int q1 = (iz - 1) * G.nx + (ix - 1), q3 = q1 + G.nx;
  const Pair64 t12 = load_pair(F.T + q1), t34 = load_pair(F.T + q3);
  const Pair16 f12 = load_pair(R.flags + q1), f34 = load_pair(R.flags + q3);
  double t1 = t12.lo, t2 = t12.hi, t3 = t34.lo, t4 = t34.hi;
  const bool in1 = f12.lo & RF_IN, in2 = f12.hi & RF_IN, in3 = f34.lo & RF_IN, in4 = f34.hi & RF_IN;
  if (!in1) { const int m = rnode_nearby(G, R, ix, iz); t1 = m >= 0 ? F.T[m] : RAY_HUGE; }
  if (!in2) { const int m = rnode_nearby(G, R, ix + 1, iz); t2 = m >= 0 ? F.T[m] : RAY_HUGE; }
  if (!in3) { const int m = rnode_nearby(G, R, ix, iz + 1); t3 = m >= 0 ? F.T[m] : RAY_HUGE; }
  if (!in4) { const int m = rnode_nearby(G, R, ix + 1, iz + 1); t4 = m >= 0 ? F.T[m] : RAY_HUGE; }
  const double half = RAY_HUGE * 0.5;
  if (t1 < half && t2 < half && t3 < half && t4 < half)
    return (1.0 - w1) * (1.0 - w2) * t1 + w1 * (1.0 - w2) * t2 + (1.0 - w1) * w2 * t3 + w1 * w2 * t4;
  // interpolate_arrivaltime_reg (rays2d.f90:960-976): nearest node's time
  const int m = nearest_any(G, R, px, pz, RAY_HUGE);
  return m >= 0 ? F.T[m] : RAY_HUGE;
}

// interpolate_interface2d (rays2d.f90:1228-1262) on intersection k
__device__ __attribute__((always_inline)) inline double iface_z(const RayGeomDev& G, int k, double px) {
  const int n = G.isn[k];
  const double* x = G.isx[k];
  const double* z = G.isz[k];
  if (n < 1) return 0.0;
  if (n < 2) return z[0];
  const int ix = ifloor((px - x[0]) / G.dx) + 1;
  if (ix < 1) return z[0];
  if (ix >= n) return z[n - 1];
  const double w = clamp01((px - x[ix - 1]) / G.dx);
  return (1.0 - w) * z[ix - 1] + w * z[ix];
}

// One intersection's node arrays, read once per ray section: iface_z on it
// is the same function
struct IfaceRef {
  const double* x;
  const double* z;
  int n;
  double x0;
};
__device__ __attribute__((always_inline)) inline IfaceRef iface_ref(const RayGeomDev& G, int k) {
  IfaceRef f;
  f.x = G.isx[k]; f.z = G.isz[k]; f.n = G.isn[k];
  f.x0 = f.n >= 1 ? f.x[0] : 0.0;
  return f;
}
__device__ __attribute__((always_inline)) inline double iface_z(const IfaceRef& f, double dx, double px) {
  if (f.n < 1) return 0.0;
  if (f.n < 2) return f.z[0];
  const int ix = ifloor((px - f.x0) / dx) + 1;
  if (ix < 1) return f.z[0];
  if (ix >= f.n) return f.z[f.n - 1];
  const double w = clamp01((px - f.x[ix - 1]) / dx);
  const Pair64 zz = load_pair(f.z + ix - 1);
  return (1.0 - w) * zz.lo + w * zz.hi;
}

// iface_z of two intersections at one x.  Intersections whose node x are the
// same array (the host shares one copy of equal arrays) share the cell
// index and weight, which are then the same numbers either call computes.
__device__ __attribute__((always_inline)) inline void iface_z2(const IfaceRef& f, const IfaceRef& g, double dx,
                                                               double px, double& zf, double& zg) {
  if (f.x != g.x || f.n != g.n || f.n < 2) {
    zf = iface_z(f, dx, px);
    zg = iface_z(g, dx, px);
    return;
  }
  const int ix = ifloor((px - f.x0) / dx) + 1;
  if (ix < 1) { zf = f.z[0]; zg = g.z[0]; return; }
  if (ix >= f.n) { zf = f.z[f.n - 1]; zg = g.z[g.n - 1]; return; }
  const double w = clamp01((px - f.x[ix - 1]) / dx);
  const Pair64 a = load_pair(f.z + ix - 1), b = load_pair(g.z + ix - 1);
  zf = (1.0 - w) * a.lo + w * a.hi;
  zg = (1.0 - w) * b.lo + w * b.hi;
}

// interface_normal2d (rays2d.f90:1266-1308)
__device__ __attribute__((always_inline)) inline void iface_normal(const RayGeomDev& G, int k, double px, double& nx_, double& nz_, double& h) {
  const int n = G.isn[k];
  const double* x = G.isx[k];
  const double* z = G.isz[k];
  if (n < 2) { nx_ = 0.0; nz_ = 1.0; h = z[0]; return; }
  int ix = ifloor((px - x[0]) / G.dx) + 1;
  if (ix < 1) ix = 1;
  if (ix >= n) ix = n - 1;
  const double w = clamp01((px - x[ix - 1]) / G.dx);
  h = (1.0 - w) * z[ix - 1] + w * z[ix];
  double dzdx;
  if (ix > 1 && ix < n - 1) dzdx = (z[ix] - z[ix - 2]) / (2.0 * G.dx);
  else if (ix == 1 && n >= 3) dzdx = (z[ix] - z[ix - 1]) / G.dx;
  else if (ix == n - 1 && n >= 3) dzdx = (z[ix] - z[ix - 2]) / (2.0 * G.dx);
  else dzdx = 0.0;
  const double nl = sqrt(1.0 + dzdx * dzdx);
  nx_ = -dzdx / nl;
  nz_ = 1.0 / nl;
}

// interpolate_iface_z_inv (inversion2d.f90:6315-6341) on seafloor depths z
// (the current ones, or those at set-up)
__device__ double seafloor_z(const RayGeomDev& G, const double* z, double px) {
  if (G.nsf < 2) return z[0];
  const int i = ifloor((px - G.sfx[0]) / G.sfdx) + 1;
  if (i < 1) return z[0];
  if (i >= G.nsf) return z[G.nsf - 1];
  const double w = clamp01((px - G.sfx[i - 1]) / G.sfdx);
  return (1.0 - w) * z[i - 1] + w * z[i];
}
__device__ double seafloor_z(const RayGeomDev& G, double px) { return seafloor_z(G, G.sfz, px); }

// get_velocity_at2d (rays2d.f90:1312-1348) on the propagation grid
__device__ double velocity_at(const RayGeomDev& G, double px, double pz) {
  int ix = ifloor((px - G.x0) / G.dx) + 1, iz = ifloor((pz - G.z0) / G.dz) + 1;
  if (ix < 1) ix = 1;
  if (ix >= G.nx) ix = G.nx - 1;
  if (iz < 1) iz = 1;
  if (iz >= G.nz) iz = G.nz - 1;
  const double w1 = clamp01((px - G.xg[ix - 1]) / G.dx), w2 = clamp01((pz - G.zg[iz - 1]) / G.dz);
  const double* V = G.pvel;
  const int q = (iz - 1) * G.nx + (ix - 1);
  return (1.0 - w1) * (1.0 - w2) * V[q] + w1 * (1.0 - w2) * V[q + 1] + (1.0 - w1) * w2 * V[q + G.nx] +
         w1 * w2 * V[q + G.nx + 1];
}

// best_diving_return_node2d (rays2d.f90:725-760)
__device__ int diving_return_node(const RayGeomDev& G, const RayRegionDev& R, const RayField& F,
                                  const double* rvel, double x_if, double z_if, double min_depth) {
  int best = -1;
  double best_t = RAY_HUGE;
  for (int k = 0; k < R.nnode; k++) {
    const int q = R.order[k];
    if (F.T[q] >= RAY_HUGE * 0.5) continue;
    double x, z;
    slot_xz(G, R, q, x, z);
    if (z <= z_if + min_depth) continue;
    if (fabs(x - x_if) > 3.0 * min_depth) continue;
    const double dx = x - x_if, dz = z - z_if, dist = sqrt(dx * dx + dz * dz);
    if (dist <= 1.0e-10) continue;
    double vloc = rvel[q];
    if (vloc < 0.01) vloc = 0.01;
    const double cand = F.T[q] + dist / vloc;
    if (cand < best_t) { best_t = cand; best = q; }
  }
  return best;
}

// ---------------------------------------------------------------- sections

struct RayState {
  bool valid, headwave, diffracted;
};

struct SecInfo {
  int seqnum, nsec, path;
  const RayField* F;        // this section's field
  const RayRegionDev* R;
  const RayField* Fn;       // field of section seqnum-1 (toward the source), or null
  const RayRegionDev* Rn;
  int istart, iend;         // intersection ids, 0 none
  int next_reg;             // region id of section seqnum+1 (0 none)
  int prev_reg;             // region id of section seqnum-1 (0 none)
  const double* rvel;       // region node velocities of this section's region
};

// find_interface_crossing2d (rays2d.f90:764-821); false = no crossing
__device__ __attribute__((always_inline)) inline bool interface_crossing(const RayGeomDev& G, double2* pt, int n, const double pp[2], int ifend,
                                   const double grad0[2], double dl) {
  const double tol = G.dx < G.dz ? G.dx * 0.01 : G.dz * 0.01;  // grid%tolerance
  const double zi = iface_z(G, ifend, pp[0]);
  const double x1 = 0.0, x2 = dl;
  const double f = pt[n - 2].y - iface_z(G, ifend, pt[n - 2].x);
  double fmid = pp[1] - zi;
  if (f * fmid > 0.01 * tol) return false;
  if (fabs(f * fmid) <= 0.01 * tol) {
    pt[n - 1] = pt[n - 2];
    pt[n - 1].y = iface_z(G, ifend, pt[n - 1].x);
    return true;
  }
  if (f * fmid < -0.01 * tol) {
    double rtbis, dxb;
    if (f < 0.0) { rtbis = x1; dxb = x2 - x1; } else { rtbis = x2; dxb = x1 - x2; }
    for (int it = 0; it < 10; it++) {
      dxb = dxb * 0.5;
      const double xmid = rtbis + dxb;
      pt[n - 1].x = pt[n - 2].x - xmid * grad0[0];
      pt[n - 1].y = pt[n - 2].y - xmid * grad0[1];
      fmid = pt[n - 1].y - iface_z(G, ifend, pt[n - 1].x);
      if (fmid <= 0.0) rtbis = xmid;
    }
    pt[n - 1].x = pt[n - 2].x - rtbis * grad0[0];
    pt[n - 1].y = pt[n - 2].y - rtbis * grad0[1];
    pt[n - 1].y = iface_z(G, ifend, pt[n - 1].x);
  }
  return true;
}

// Headwave / refraction decision at the section's ending interface
// (rays2d.f90:451-455 and repeats): true = the ray refracts (section ends).
__device__ __attribute__((always_inline)) inline bool refracts(const RayGeomDev& G, const SecInfo& S, int ifend, double px, double zi, double grad0[2],
                         RayState& st) {
  double nx_, nz_, h, g1[2];
  iface_normal(G, ifend, px, nx_, nz_, h);
  interp_grad(G, *S.R, *S.F, px, zi, grad0[0], grad0[1]);
  interp_grad(G, *S.Rn, *S.Fn, px, zi, g1[0], g1[1]);
  const double dn = nx_ * g1[0] + nz_ * g1[1];
  const double a = g1[0] - dn * nx_, b = g1[1] - dn * nz_;
  return sqrt(a * a + b * b) <= sqrt(grad0[0] * grad0[0] + grad0[1] * grad0[1]);
}

// trace_ray_from_receiver2d / trace_ray_section2d (rays2d.f90:7-720) as a
// walk.  The reference traces a ray with nested loops: sections, the step
// loop (ploop) and the corrector's step halvings.  Here a ray is a small
// state machine that needs exactly one gradient interpolation per round
// (the predictor's at the current point, or one corrector try), and the
// kernel's single loop runs rounds until the ray is done: the lanes of a
// wave then share one interpolation per round whatever step, halving or
// section each is in, instead of waiting for each other at every nested
// loop's end.  Every point is computed with the nested form's arithmetic,
// in its order.  The step's points live in registers (a = pt[n-2],
// b = pt[n-3], w = pt[n-1]); every change of pt[n-1] is also stored.
struct RayWalk {
  enum : int { PRED = 0, CORR = 1, SEC_END = 2, DONE = 3 };
  const RayGeomDev& G;
  const RayField* fields;
  int path, nsec, src;
  double sx, sz;
  double2* P;
  RayState st;
  // section
  int sec;
  SecInfo S;
  int iftop, ifbot, ifend, start_direction;
  IfaceRef Itop, Ibot, Iend;
  bool section_is_refraction, same_internal_ifend, left_start_interface, start_on_bot, start_on_top;
  double2* pt;
  int n;
  double dist_headwave;
  double x_in, z_in, x_out, z_out;
  // step
  double dlmax, p_tol, s_tol, dl;
  int iter;
  double grad0[2], pp[2];
  double2 a, b, w;
  // round
  int phase, np_end;
  double qx, qz;

  __device__ __attribute__((always_inline)) void set_w(double x, double z) {
    w = make_double2(x, z);
    pt[n - 1] = w;
  }
  // snap_source on the step's point (the section toward the source)
  __device__ __attribute__((always_inline)) bool snap_w() {
    const double dd = sqrt((w.x - sx) * (w.x - sx) + (w.y - sz) * (w.y - sz));
    if (dd < s_tol) { set_w(sx, sz); return true; }
    return false;
  }
  __device__ __attribute__((always_inline)) void end_section(int np) {
    np_end = np;
    phase = SEC_END;
  }
  __device__ __attribute__((always_inline)) void fail() {
    st.valid = false;
    end_section(0);
  }

  // section set-up (rays2d.f90:157-264): interfaces, start conditions,
  // the first point, the diving phases' seed path
  __device__ __attribute__((always_inline)) void begin_section() {
    const int tf = c_path_tf[path][sec - 1];
    S.seqnum = sec; S.nsec = nsec; S.path = path;
    S.F = &fields[src * 12 + tf - 1];
    S.R = &G.reg[S.F->region];
    S.rvel = G.reg[S.F->region].vel;
    S.istart = sec < nsec ? c_path_seq[path][2 * sec - 1] : 0;
    S.iend = sec > 1 ? c_path_seq[path][2 * sec - 2] : 0;
    S.next_reg = sec < nsec ? c_tf_region[c_path_tf[path][sec]] : 0;
    if (sec > 1) {
      const int tfn = c_path_tf[path][sec - 2];
      S.Fn = &fields[src * 12 + tfn - 1];
      S.Rn = &G.reg[S.Fn->region];
      S.prev_reg = c_tf_region[tfn];
    } else {
      S.Fn = nullptr; S.Rn = nullptr; S.prev_reg = 0;
    }
    const RayRegionDev& R = *S.R;
    iftop = R.itop; ifbot = R.ibot;
    const int seqnum = sec;
    ifend = 0; start_direction = 0;
    section_is_refraction = false; same_internal_ifend = false; left_start_interface = true;
    if (seqnum < nsec && seqnum > 1) {
      ifend = S.iend;
      start_direction = R.id - S.next_reg;
      section_is_refraction = S.prev_reg != R.id;
    }
    if (seqnum == nsec && seqnum > 1) {
      ifend = S.iend;
      start_direction = 0;
      section_is_refraction = S.prev_reg != R.id;
    }
    if (seqnum == 1) {
      ifend = 0;
      start_direction = (seqnum + 1 <= nsec) ? R.id - S.next_reg : 0;
    }
    if (ifend && S.istart) {
      same_internal_ifend = ifend == S.istart && ifend != iftop && ifend != ifbot;
      if (same_internal_ifend) left_start_interface = false;
    }
    Itop = iface_ref(G, iftop);
    Ibot = iface_ref(G, ifbot);
    Iend = iface_ref(G, ifend);
    pt = P + (long long)(sec - 1) * RAY_MAXP;
    n = 1;
    set_w(x_in, z_in);
    dist_headwave = 0.0;
    double zi = iface_z(Ibot, G.dx, x_in);
    start_on_bot = fabs(z_in - zi) < 0.1 * G.dz;
    zi = iface_z(Itop, G.dx, x_in);
    start_on_top = fabs(z_in - zi) < 0.1 * G.dz;
    // converted diving phases start the deep section from the interior
    // node that seeded the return (rays2d.f90:234-264)
    if (seqnum == 2 && start_on_top && S.istart && S.iend && S.istart == iftop && S.iend == iftop &&
        (path == 5 || path == 6 || path == 8)) {
      const int seed = diving_return_node(G, R, *S.F, S.rvel, x_in, z_in, 2.0 * G.dz);
      if (seed >= 0) {
        double qx_, qz_;
        slot_xz(G, R, seed, qx_, qz_);
        const double seed_dist = sqrt((qx_ - x_in) * (qx_ - x_in) + (qz_ - z_in) * (qz_ - z_in));
        if (seed_dist > 1.0e-10) {
          int nseed = (int)ceil(seed_dist / fmax(dlmax, 1.0e-10));
          if (nseed < 1) nseed = 1;
          if (nseed > RAY_MAXP - 2) nseed = RAY_MAXP - 2;
          for (int k = 1; k <= nseed; k++) {
            const double frac = (double)k / (double)nseed;
            a = w;
            n = k + 1;
            set_w(x_in + frac * (qx_ - x_in), z_in + frac * (qz_ - z_in));
          }
        }
      }
    }
  }

  // top of the step loop (rays2d.f90:267-311): next point, point cap,
  // arrival at the source, head-wave length, step length
  __device__ __attribute__((always_inline)) void enter_step() {
    b = a;
    a = w;
    n = n + 1;
    if (n > RAY_MAXP) { fail(); return; }
    if (sec == 1 && n > 2) {
      const double dd = sqrt((a.x - sx) * (a.x - sx) + (a.y - sz) * (a.y - sz));
      if (dd < s_tol) { set_w(sx, sz); end_section(n); return; }
    }
    if (st.headwave || st.diffracted) {
      if (n > 2) {
        const double ax = a.x - b.x, az = a.y - b.y;
        dist_headwave = dist_headwave + sqrt(ax * ax + az * az);
      }
      if (dist_headwave > 200.0) { fail(); return; }
    }
    dl = dlmax;
    if (n == 2 && (start_on_bot || start_on_top)) dl = 10.0 * dlmax;
    qx = a.x; qz = a.y;
    phase = PRED;
  }

  // the step loop left through the ending interface (rays2d.f90:680-718)
  __device__ __attribute__((always_inline)) void exited() {
    if (sec == 1) { fail(); return; }
    if (ifend) {
      if (!interface_crossing(G, pt, n, pp, ifend, grad0, dl)) { fail(); return; }
    }
    x_out = pt[n - 1].x;
    z_out = pt[n - 1].y;
    end_section(n);
  }

  // the step's point put on an interface the predictor or corrector
  // crossed: head wave or diffraction, then the next step
  __device__ __attribute__((always_inline)) void along_interface(double x, double z, bool headwave, bool snap) {
    set_w(x, z);
    if (headwave) st.headwave = true; else st.diffracted = true;
    if (snap && sec == 1 && snap_w()) { end_section(n); return; }
    enter_step();
  }

  // predictor, first part (rays2d.f90:313-434), with grad0 at pt[n-2]:
  // false when the section ended
  __device__ __attribute__((always_inline)) bool predict_start(double gx, double gz) {
    grad0[0] = gx; grad0[1] = gz;
    const int seqnum = sec;
    pp[0] = a.x - dl * grad0[0];
    pp[1] = a.y - dl * grad0[1];
    if (start_on_bot && pp[1] > a.y) pp[1] = 2.0 * a.y - pp[1];
    if (start_on_top && pp[1] < a.y) pp[1] = 2.0 * a.y - pp[1];
    if (n == 2 && start_direction != 0 && !section_is_refraction && seqnum != 1) {
      double nx_, nz_, h;
      iface_normal(G, S.istart, x_in, nx_, nz_, h);
      if (grad0[0] * (start_direction * nx_) + grad0[1] * (start_direction * nz_) > 0.0) { fail(); return false; }
    }
    // ending interface inside the region (rays2d.f90:400-434)
    if (ifend && ifend != iftop && ifend != ifbot) {
      const double zp = iface_z(Iend, G.dx, a.x), zn = iface_z(Iend, G.dx, pp[0]);
      const double fp = a.y - zp, fn = pp[1] - zn;
      if (same_internal_ifend && !left_start_interface) {
        if (fabs(fn) > 0.5 * G.dz) left_start_interface = true;
      } else if (fabs(fp) > 1.0e-3) {
        if (fp * fn <= 0.0) {
          double frac = fabs(fp - fn) > 1.0e-12 ? fp / (fp - fn) : 1.0;
          frac = clamp01(frac);
          const double xe = a.x + frac * (pp[0] - a.x);
          set_w(xe, iface_z(Iend, G.dx, xe));
          if (seqnum == 1) snap_w();
          end_section(n);
          return false;
        }
      }
    }
    return true;
  }

  // predictor, second part (rays2d.f90:436-562): the top and bottom
  // interfaces at pp (depths zt, zb), then the corrector's first try
  __device__ __attribute__((always_inline)) void predict_end(double zt, double zb) {
    const int seqnum = sec;
    if (pp[1] <= zt) {
      if (iftop == ifend) {
        if (section_is_refraction) {
          if (refracts(G, S, ifend, pp[0], zt, grad0, st)) { exited(); return; }
          along_interface(pp[0], zt, true, true);
          return;
        }
        exited();
        return;
      } else if (n == 2 && start_on_top && S.istart && S.istart == iftop) {
        // leave the start interface
      } else {
        along_interface(pp[0], zt, false, true);
        return;
      }
    }
    if (pp[1] >= zb) {
      if (ifbot == ifend) {
        if (section_is_refraction) {
          if (S.path == 4 && seqnum == 2) { exited(); return; }
          if (refracts(G, S, ifend, pp[0], zb, grad0, st)) { exited(); return; }
          along_interface(pp[0], zb, true, true);
          return;
        }
        exited();
        return;
      } else if (n == 2 && start_on_bot && S.istart && S.istart == ifbot) {
      } else {
        along_interface(pp[0], zb, false, true);
        return;
      }
    }
    // corrector (rays2d.f90:564-668): its first try at the full step
    iter = 1;
    pp[0] = a.x - dl * grad0[0];
    pp[1] = a.y - dl * grad0[1];
    qx = pp[0]; qz = pp[1];
    phase = CORR;
  }

  // a corrector try, second part: the top and bottom interfaces at the new
  // point w (depths zt, zb), convergence, halving
  __device__ __attribute__((always_inline)) void correct_end(double zt, double zb) {
    const int seqnum = sec;
    if (w.y <= zt) {
      if (iftop == ifend) {
        if (section_is_refraction) {
          if (refracts(G, S, ifend, w.x, zt, grad0, st)) {
            pp[0] = w.x; pp[1] = w.y;
            exited();
            return;
          }
          along_interface(w.x, zt, true, false);
          return;
        }
        pp[0] = w.x; pp[1] = w.y;
        exited();
        return;
      } else if (n == 2 && start_on_top && S.istart && S.istart == iftop) {
      } else {
        along_interface(w.x, zt, false, false);
        return;
      }
    }
    if (w.y >= zb) {
      if (ifbot == ifend) {
        if (section_is_refraction) {
          if (S.path == 4 && seqnum == 2) {
            pp[0] = w.x; pp[1] = w.y;
            exited();
            return;
          }
          if (refracts(G, S, ifend, w.x, zb, grad0, st)) {
            pp[0] = w.x; pp[1] = w.y;
            exited();
            return;
          }
          along_interface(w.x, zb, true, false);
          return;
        }
        pp[0] = w.x; pp[1] = w.y;
        exited();
        return;
      } else if (n == 2 && start_on_bot && S.istart && S.istart == ifbot) {
      } else {
        along_interface(w.x, zb, false, false);
        return;
      }
    }
    const double ex = w.x - pp[0], ez = w.y - pp[1];
    if (sqrt(ex * ex + ez * ez) < p_tol) {
      if (seqnum == 1 && snap_w()) { end_section(n); return; }
      enter_step();
      return;
    }
    iter = iter + 1;
    if (iter > 10) { fail(); return; }
    dl = 0.5 * dl;
    pp[0] = a.x - dl * grad0[0];
    pp[1] = a.y - dl * grad0[1];
    qx = pp[0]; qz = pp[1];
  }

  // the rest of a round, with the gradient (gx, gz) interpolated at the
  // round's query point: the predictor, or a corrector try (its new point,
  // with grad1 = (gx, gz)).  Both then test one point against the region's
  // top and bottom interfaces, which every lane does at once.
  __device__ __attribute__((always_inline)) void advance(double gx, double gz) {
    const bool pred = phase == PRED;
    double cx, cy;
    if (pred) {
      if (!predict_start(gx, gz)) return;
      cx = pp[0]; cy = pp[1];
    } else {
      set_w(a.x - dl * (grad0[0] + gx) * 0.5, a.y - dl * (grad0[1] + gz) * 0.5);
      cx = w.x; cy = w.y;
    }
    (void)cy;
    double zt, zb;
    iface_z2(Itop, Ibot, G.dx, cx, zt, zb);
    if (pred) predict_end(zt, zb);
    else correct_end(zt, zb);
  }
};

// ---------------------------------------------------------------- kernels

// Arrival time of each path's last stage field at every seafloor node, per
// source (the first sweep of find_optimal_entry_inv): [src][3][nsf] for
// tf1, tf3, tf5.
__global__ void seafloor_times_kernel(const RayGeomDev* __restrict__ Gp, const RayField* __restrict__ fields,
                                      int nsrc, double* tsf) {
  const RayGeomDev& G = *Gp;
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)nsrc * 3 * G.nsf;
  if (k >= total) return;
  const int i = (int)(k % G.nsf), w = (int)((k / G.nsf) % 3), s = (int)(k / G.nsf / 3);
  const int tf = w == 0 ? 1 : (w == 1 ? 3 : 5);
  const RayField& F = fields[s * 12 + tf - 1];
  tsf[k] = interp_time(G, G.reg[F.region], F, G.sfx[i], G.sfz[i]);
}

__global__ void __launch_bounds__(64) trace_kernel(const RayGeomDev* __restrict__ Gp, const RayField* __restrict__ fields,
                             const double* __restrict__ src_x, const double* __restrict__ src_z,
                             const double* __restrict__ tsf, const RayJob* __restrict__ jobs,
                             const int* __restrict__ order, int nray, const long long* __restrict__ pt_off,
                             double2* pts, RayOut* out, int program) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nray) return;
  const int r = order[k];
  const RayGeomDev& G = *Gp;
  const RayJob J = jobs[r];
  RayOut o;
  o.t = 0.0; o.water_time = 0.0; o.valid = 1; o.nsec = c_path_ntf[J.path];
  for (int k = 0; k < RAY_MAXSEC; k++) o.npts[k] = 0;
  o.headwave = 0; o.diffracted = 0; o.direct = 0; o.work = 0;
  const int nsec = o.nsec, s = J.src;
  const double sx = src_x[s], sz = src_z[s];
  // receiver placement (invert2d_main.f90:517-531, inversion2d.f90:4502-4519):
  // in water when above the seafloor of the set-up
  double rx = J.rx, rz = J.rz, wt = 0.0;
  const double zsf = seafloor_z(G, G.sfz_setup, J.rx);
  if (J.rz < zsf) {
    if (J.path == 4 || J.path == 5) {
      rx = J.rx; rz = J.rz; wt = 0.0;
    } else {
      // find_optimal_entry_inv (inversion2d.f90:6344-6438); the forward
      // program's find_optimal_entry (fmt2d_main.f90:1803-1905) also passes
      // over seafloor points within 3 dx of the source
      const double min_src = program ? 3.0 * G.dx : -1.0;
      const int itf = c_path_tf[J.path][nsec - 1];
      const RayField& F = fields[s * 12 + itf - 1];
      const RayRegionDev& R = G.reg[F.region];
      const double* ts = tsf + ((long long)s * 3 + (itf == 1 ? 0 : (itf == 3 ? 1 : 2))) * G.nsf;
      double t_best = RAY_HUGE;
      const double z0 = seafloor_z(G, J.rx);
      double xb = J.rx, zb = z0, wb = (z0 - J.rz) / G.vwater;
      for (int i = 0; i < G.nsf; i++) {
        if (program) {
          const double qx = G.sfx[i] - sx, qz = G.sfz[i] - sz;
          if (sqrt(qx * qx + qz * qz) < min_src) continue;
        }
        const double tf_ = ts[i];
        if (tf_ > RAY_HUGE * 0.5) continue;
        const double ddx = G.sfx[i] - J.rx, ddz = G.sfz[i] - J.rz;
        const double tw = sqrt(ddx * ddx + ddz * ddz) / G.vwater;
        const double tt = tf_ + tw;
        if (tt < t_best) { t_best = tt; xb = G.sfx[i]; zb = G.sfz[i]; wb = tw; }
      }
      if (!(xb <= G.sfx[0] || xb >= G.sfx[G.nsf - 1])) {
        double xl = 0.0, xr = 0.0;
        for (int i = 0; i < G.nsf - 1; i++)
          if (xb >= G.sfx[i] && xb <= G.sfx[i + 1]) { xl = G.sfx[i]; xr = G.sfx[i + 1]; break; }
        const double dxr = (xr - xl) / 10.0;
        for (int k = 1; k <= 9; k++) {
          const double xt = xl + (double)k * dxr;
          const double zt = seafloor_z(G, xt);
          if (program) {
            const double qx = xt - sx, qz = zt - sz;
            if (sqrt(qx * qx + qz * qz) < min_src) continue;
          }
          const double tf_ = interp_time(G, R, F, xt, zt);
          if (tf_ > RAY_HUGE * 0.5) continue;
          const double ddx = xt - J.rx, ddz = zt - J.rz;
          const double tw = sqrt(ddx * ddx + ddz * ddz) / G.vwater;
          const double tt = tf_ + tw;
          if (tt < t_best) { t_best = tt; xb = xt; zb = zt; wb = tw; }
        }
      }
      rx = xb; rz = zb; wt = wb;
      if (program && J.path == 1) {
        // Pg: the direct water wave when it is no later than Pg through the
        // entry point (fmt2d_main.f90:1249-1280)
        const double ddx = J.rx - sx, ddz = J.rz - sz;
        const double t_direct = sqrt(ddx * ddx + ddz * ddz) / G.vwater;
        if (t_direct <= interp_time(G, R, F, rx, rz) + wt) {
          o.ex = rx; o.ez = rz;
          o.t = t_direct; o.water_time = 0.0; o.direct = 1;
          out[r] = o;
          return;
        }
      }
    }
  }
  o.ex = rx; o.ez = rz;
  // the forward program keeps the water leg of PsedP and PmP rays only
  // (fmt2d_main.f90:1290-1294, 1334-1337); it is in the receiver time either way
  o.water_time = (program && J.path == 1) ? 0.0 : wt;
  // trace_ray_from_receiver2d (rays2d.f90:7-153): sections nsec..1
  RayWalk W{G, fields, J.path, nsec, s, sx, sz, pts + pt_off[r]};
  W.st = RayState{true, false, false};
  W.sec = nsec;
  W.x_in = rx; W.z_in = rz; W.x_out = 0.0; W.z_out = 0.0;
  W.dlmax = 0.1 * G.dx * 5.0;  // 0.1 dx v_average
  W.p_tol = 5.0e-4 * G.dx;
  W.s_tol = 2.5 * G.dx;
  W.phase = RayWalk::DONE;
  {
    // receiver time and gradient in the last section's field
    const RayField& F = fields[s * 12 + c_path_tf[J.path][nsec - 1] - 1];
    const RayRegionDev& R = G.reg[F.region];
    const double tr = interp_time(G, R, F, rx, rz);
    double gx, gz;
    interp_grad(G, R, F, rx, rz, gx, gz);
    o.t = tr;
    if (tr > 1.0e10 || (gx == 0.0 && gz == 0.0)) W.st.valid = false;
    else {
      W.begin_section();
      W.enter_step();
    }
  }
  for (;;) {
    // a section that ended: record it, start the next one
    while (W.phase == RayWalk::SEC_END) {
      o.npts[W.sec - 1] = W.np_end;
      if (!W.st.valid || W.sec == 1) { W.phase = RayWalk::DONE; break; }
      W.x_in = W.x_out; W.z_in = W.z_out;
      W.sec = W.sec - 1;
      W.begin_section();
      W.enter_step();
    }
    if (W.phase == RayWalk::DONE) break;
    o.work++;
    double gx, gz;
    interp_grad(G, *W.S.R, *W.S.F, W.qx, W.qz, gx, gz);
    W.advance(gx, gz);
  }
  const RayState st = W.st;
  o.valid = st.valid ? 1 : 0;
  if (o.valid) o.t = o.t + wt;
  else if (program) o.water_time = 0.0;  // set on valid rays only (:1336)
  o.headwave = st.headwave; o.diffracted = st.diffracted;
  out[r] = o;
}

// Frechet contributions of one ray in the order compute_frechet_for_ray2d
// accumulates them (inversion2d.f90:4580-4757): velocity weights of every
// segment midpoint outside the water, sections 1..n, then the interface
// partials at the section starts, sections n..1.
template <bool WRITE>
__device__ long long ray_contrib(const RayGeomDev& G, const RayField* fields, const RayJob& J, int rloc,
                                 const double2* P, const RayOut& o, unsigned long long* keys, double* vals) {
  long long cnt = 0;
  auto emit = [&](int idx1, double v) {  // idx1: 1-based parameter
    if (WRITE) { keys[cnt] = ((unsigned long long)rloc << 32) | (unsigned)(idx1 - 1); vals[cnt] = v; }
    cnt++;
  };
  const int nsec = o.nsec;
  for (int si = 1; si <= nsec; si++) {
    const int np = o.npts[si - 1];
    if (np <= 1) continue;
    const int tf = c_path_tf[J.path][si - 1];
    if (c_tf_region[tf] == 1) continue;  // water sections carry no velocity sensitivity
    const double2* p = P + (long long)(si - 1) * RAY_MAXP;
    for (int k = 0; k < np - 1; k++) {
      const double px = 0.5 * (p[k].x + p[k + 1].x), pz = 0.5 * (p[k].y + p[k + 1].y);
      const double ax = p[k + 1].x - p[k].x, az = p[k + 1].y - p[k].y;
      const double seg = sqrt(ax * ax + az * az);
      if (seg < 1.0e-10) continue;
      int ivx = ifloor((px - G.vx[0]) / G.vdx) + 1;
      if (ivx < 1) ivx = 1;
      if (ivx >= G.vnx) ivx = G.vnx - 1;
      int ivz;
      double wz;
      if (pz < G.vz[0]) continue;
      if (pz >= G.vz[G.vnz - 1]) { ivz = G.vnz - 1; wz = 1.0; }
      else {
        ivz = ifloor((pz - G.vz[0]) / G.vdz) + 1;
        if (ivz < 1) ivz = 1;
        if (ivz >= G.vnz) ivz = G.vnz - 1;
        wz = (pz - G.vz[ivz - 1]) / G.vdz;
      }
      const double wx = clamp01((px - G.vx[ivx - 1]) / G.vdx);
      wz = clamp01(wz);
      const int* I = G.inv_idx + (long long)(ivz - 1) * G.vnx + (ivx - 1);
      if (I[0] > 0) emit(I[0], (1.0 - wx) * (1.0 - wz) * seg);
      if (I[1] > 0) emit(I[1], wx * (1.0 - wz) * seg);
      if (I[G.vnx] > 0) emit(I[G.vnx], (1.0 - wx) * wz * seg);
      if (I[G.vnx + 1] > 0) emit(I[G.vnx + 1], wx * wz * seg);
    }
  }
  // interface partials
  int prev = 0;  // section index of raysec_prev
  bool rec_done = false, do_if = false;
  for (int si = nsec; si >= 1; si--) {
    const int np = o.npts[si - 1];
    if (np <= 2) continue;
    if (!rec_done) { prev = si; rec_done = true; } else do_if = true;
    if (do_if) {
      const int ist = si < nsec ? c_path_seq[J.path][2 * si - 1] : 0;
      if (ist == 0) { prev = si; continue; }
      if (G.if_start[ist] >= 0) {
        const double2* p = P + (long long)(si - 1) * RAY_MAXP;
        const double px = p[0].x, pz = p[0].y;
        const int tf = c_path_tf[J.path][si - 1];
        const int reg = c_tf_region[tf];
        // the region holds the field of the first section (toward the
        // source) that uses it when the Frechet row is built
        int tl = tf;
        for (int m = 1; m <= nsec; m++)
          if (c_tf_region[c_path_tf[J.path][m - 1]] == reg) { tl = c_path_tf[J.path][m - 1]; break; }
        const RayField& F = fields[J.src * 12 + tl - 1];
        double gi[2];
        interp_grad(G, G.reg[reg], F, px, pz, gi[0], gi[1]);
        const double vin = velocity_at(G, px, pz);
        const double nl = sqrt(gi[0] * gi[0] + gi[1] * gi[1]);
        if (nl > 0.0) { gi[0] = gi[0] / (vin * nl); gi[1] = gi[1] / (vin * nl); }
        double nx_, nz_, h;
        iface_normal(G, ist, px, nx_, nz_, h);
        const double perp = gi[0] * nx_ + gi[1] * nz_;
        double out_perp;
        const int prev_reg = c_tf_region[c_path_tf[J.path][prev - 1]];
        if (reg == prev_reg) out_perp = -perp;
        else {
          const double a = gi[0] - perp * nx_, b = gi[1] - perp * nz_;
          const double par = sqrt(a * a + b * b);
          const double vout = velocity_at(G, px, pz);
          double det = 1.0 / (vout * vout) - par * par;
          if (det < 0.0) det = 0.0;
          out_perp = copysign(sqrt(det), perp);
        }
        const double geo = (perp - out_perp) * nz_;
        const int nif = G.if_nnode[ist];
        int ivx = ifloor((px - G.ifx_in[ist][0]) / G.ifdx_in[ist]) + 1;
        if (ivx < 1) ivx = 1;
        if (ivx >= nif) ivx = nif - 1;
        const double wx = clamp01((px - G.ifx_in[ist][ivx - 1]) / G.ifdx_in[ist]);
        int idx = G.if_start[ist] + ivx - 1;
        if (idx >= 1 && idx <= G.n_par) emit(idx + 1, (1.0 - wx) * geo);
        idx = G.if_start[ist] + ivx;
        if (idx >= 1 && idx <= G.n_par) emit(idx + 1, wx * geo);
      }
    }
    prev = si;
  }
  return cnt;
}

__global__ void frechet_contrib_kernel(const RayGeomDev* __restrict__ Gp, const RayField* __restrict__ fields,
                                       const RayJob* __restrict__ jobs, int nray, int ray0,
                                       const long long* __restrict__ pt_off, const double2* __restrict__ pts,
                                       const RayOut* __restrict__ out, const long long* __restrict__ coff,
                                       long long* count, unsigned long long* keys, double* vals) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nray) return;
  const int g = ray0 + r;
  const RayOut o = out[g];
  if (!o.valid) { if (!keys) count[r] = 0; return; }
  const RayGeomDev& G = *Gp;
  if (!keys) {
    count[r] = ray_contrib<false>(G, fields, jobs[g], r, pts + pt_off[g], o, nullptr, nullptr);
  } else {
    ray_contrib<true>(G, fields, jobs[g], r, pts + pt_off[g], o, keys + coff[r], vals + coff[r]);
  }
}

// ---------------------------------------------------------------- launchers

hipError_t launch_seafloor_times(const RayGeomDev* G, const RayField* fields, int nsrc, int nsf, double* tsf,
                                 hipStream_t st) {
  const long long total = (long long)nsrc * 3 * nsf;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(seafloor_times_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, G, fields, nsrc,
                     tsf);
  return hipGetLastError();
}

// Adjoint sources of the receiver times.  A valid ray's time is
// interp_time(last stage field, entry point) + the water leg, and the entry
// point is the minimiser of a discrete search (its own move does not change
// the time to first order), so d t / d T(node) is the node's weight in
// interp_time at the entry point: the bilinear weights of the cell corners
// (or their nearby-node substitutes), or 1 on the nearest node.  lam[f][q]
// += g[r] * weight for field f = field_of[src * 12 + itf - 1]; active[f] is
// set for every field that receives a non-zero source.
__global__ void adjoint_sources_kernel(const RayGeomDev* __restrict__ Gp, const RayField* __restrict__ fields,
                                       const RayJob* __restrict__ jobs, const RayOut* __restrict__ out,
                                       const double* __restrict__ g, int nray, const int* __restrict__ field_of,
                                       long long nslot, double* lam, int* active) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nray) return;
  const double gr = g[r];
  const RayOut& o = out[r];
  if (!o.valid || o.direct || !(gr != 0.0) || gr != gr) return;
  const RayGeomDev& G = *Gp;
  const RayJob J = jobs[r];
  const int itf = c_path_tf[J.path][c_path_ntf[J.path] - 1];
  const RayField& F = fields[J.src * 12 + itf - 1];
  const RayRegionDev& R = G.reg[F.region];
  const int f = field_of[J.src * 12 + itf - 1];
  double* L = lam + (long long)f * nslot;
  const double px = o.ex, pz = o.ez;
  int ix = ifloor((px - G.x0) / G.dx) + 1, iz = ifloor((pz - G.z0) / G.dz) + 1;
  if (ix < 1) ix = 1;
  if (ix >= G.nx) ix = G.nx - 1;
  if (iz < 1) iz = 1;
  if (iz >= G.nz) iz = G.nz - 1;
  const double w1 = clamp01((px - G.xg[ix - 1]) / G.dx), w2 = clamp01((pz - G.zg[iz - 1]) / G.dz);
  int n[4] = {rnode(G, R, ix, iz), rnode(G, R, ix + 1, iz), rnode(G, R, ix, iz + 1), rnode(G, R, ix + 1, iz + 1)};
  const int cx[4] = {ix, ix + 1, ix, ix + 1}, cz[4] = {iz, iz, iz + 1, iz + 1};
  bool all = true;
  for (int k = 0; k < 4; k++) {
    if (n[k] < 0) n[k] = rnode_nearby(G, R, cx[k], cz[k]);
    if (n[k] < 0 || !(F.T[n[k]] < RAY_HUGE * 0.5)) all = false;
  }
  atomicOr(active + f, 1);
  if (all) {
    const double w[4] = {(1.0 - w1) * (1.0 - w2), w1 * (1.0 - w2), (1.0 - w1) * w2, w1 * w2};
    for (int k = 0; k < 4; k++)
      if (w[k] != 0.0) atomicAdd(L + n[k], gr * w[k]);
    return;
  }
  const int m = nearest_any(G, R, px, pz, RAY_HUGE);
  if (m >= 0) atomicAdd(L + m, gr);
}

// Nearest region node of each point by the column index and by the scan
// (diagnostics: the two must agree).
__global__ void nearest_check_kernel(const RayGeomDev* __restrict__ Gp, int region, int n, const double* __restrict__ px,
                                     const double* __restrict__ pz, int* idx, int* scan) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const RayGeomDev& G = *Gp;
  RayRegionDev R = G.reg[region];
  scan[k] = nearest_scan(G, R, px[k], pz[k], RAY_HUGE);
  idx[k] = R.indexed ? nearest_any(G, R, px[k], pz[k], RAY_HUGE) : -2;
}

hipError_t launch_nearest_check(const RayGeomDev* G, int region, int n, const double* px, const double* pz, int* idx,
                                int* scan, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(nearest_check_kernel, dim3((n + 63) / 64), dim3(64), 0, st, G, region, n, px, pz, idx, scan);
  return hipGetLastError();
}

hipError_t nearest_calls_reset(hipStream_t st) {
  const unsigned long long z = 0;
  return hipMemcpyToSymbolAsync(HIP_SYMBOL(g_nearest_calls), &z, sizeof(z), 0, hipMemcpyHostToDevice, st);
}

hipError_t nearest_calls_read(unsigned long long* n, hipStream_t st) {
  hipError_t e = hipMemcpyFromSymbolAsync(n, HIP_SYMBOL(g_nearest_calls), sizeof(*n), 0, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e;
}

hipError_t launch_adjoint_sources(const RayGeomDev* G, const RayField* fields, const RayJob* jobs, const RayOut* out,
                                  const double* g, int nray, const int* field_of, long long nslot, double* lam,
                                  int* active, hipStream_t st) {
  if (nray == 0) return hipSuccess;
  hipLaunchKernelGGL(adjoint_sources_kernel, dim3((nray + 127) / 128), dim3(128), 0, st, G, fields, jobs, out, g, nray,
                     field_of, nslot, lam, active);
  return hipGetLastError();
}

hipError_t launch_trace(const RayGeomDev* G, const RayField* fields, const double* src_x, const double* src_z,
                        const double* tsf, const RayJob* jobs, const int* order, int nray, const long long* pt_off,
                        double2* pts, RayOut* out, int program, hipStream_t st) {
  if (nray == 0) return hipSuccess;
  hipLaunchKernelGGL(trace_kernel, dim3((nray + 63) / 64), dim3(64), 0, st, G, fields, src_x, src_z, tsf, jobs, order,
                     nray, pt_off, pts, out, program);
  return hipGetLastError();
}

hipError_t launch_frechet_contrib(const RayGeomDev* G, const RayField* fields, const RayJob* jobs, int nray,
                                  int ray0, const long long* pt_off, const double2* pts, const RayOut* out,
                                  const long long* coff, long long* count, unsigned long long* keys, double* vals,
                                  hipStream_t st) {
  if (nray == 0) return hipSuccess;
  hipLaunchKernelGGL(frechet_contrib_kernel, dim3((nray + 63) / 64), dim3(64), 0, st, G, fields, jobs, nray, ray0,
                     pt_off, pts, out, coff, count, keys, vals);
  return hipGetLastError();
}

}  // namespace fmt2d
