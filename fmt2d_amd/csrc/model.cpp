// Host model build (propagation grid, intersections, regions, velocities).
// Semantics follow src/fmt2d_main.f90:91-418 and src/propagate2d.f90:1335-1979;
// the data is produced directly in the dense per-region layout of the
// kernels instead of the reference's compacted node lists.
#include "model.h"

#include <math.h>

#include <algorithm>
#include <functional>
#include <sstream>
#include <thread>

#include "eik_common.h"

namespace fmt2d {

static double iface_interp(const std::vector<double>& x, const std::vector<double>& z, double px) {
  const int n = (int)x.size();
  if (n < 2) return z[0];
  const double d = x[1] - x[0];
  const int i = (int)floor((px - x[0]) / d) + 1;
  if (i < 1) return z[0];
  if (i >= n) return z[n - 1];
  double w = (px - x[i - 1]) / d;
  w = std::min(1.0, std::max(0.0, w));
  return (1.0 - w) * z[i - 1] + w * z[i];
}

// linear interpolation on the interface nodes (src/fmt2d_main.f90:1389-1415)
double HostModel::iface_z_at(int k, double px) const { return iface_interp(if_x[k], if_z[k], px); }

double HostModel::setup_seafloor_at(double px) const { return iface_interp(if_x[2], sf_setup, px); }

// interpolate_iface_z2d (src/propagate2d.f90:1615-1645): node spacing is
// taken as the grid dx because interface nodes sit on grid columns.
static double iface_z_grid(const HostModel& m, const HostInterface& is, double px) {
  if (is.n < 2) return is.z[0];
  const int ixx = (int)floor((px - is.x[0]) / m.dx) + 1;
  if (ixx < 1) return is.z[0];
  if (ixx >= is.n) return is.z[is.n - 1];
  double w = (px - is.x[ixx - 1]) / m.dx;
  w = std::min(1.0, std::max(0.0, w));
  return (1.0 - w) * is.z[ixx - 1] + w * is.z[ixx];
}

static double bilinear_vgrid(const HostModel& m, int ivx, int ivz, double wx, double wz) {
  // ivx/ivz 1-based lower corner
  auto V = [&](int i, int j) { return m.vgrid[(size_t)(j - 1) * m.nx + (i - 1)]; };
  return (1.0 - wx) * (1.0 - wz) * V(ivx, ivz) + wx * (1.0 - wz) * V(ivx + 1, ivz) +
         (1.0 - wx) * wz * V(ivx, ivz + 1) + wx * wz * V(ivx + 1, ivz + 1);
}

// velocities_to_intersection2d (src/propagate2d.f90:1698-1785): the velocity
// grid is sampled one dz above (vel_top) and below (vel_bot) each node.
static void intersection_velocities(const HostModel& m, HostInterface& is) {
  is.vel_top.assign(is.n, 0.0);
  is.vel_bot.assign(is.n, 0.0);
  const double vdx = m.dx, dz = m.dz;
  for (int c = 0; c < is.n; c++) {
    const double px = is.x[c], pz = is.z[c];
    int ivx = (int)floor((px - m.vx[0]) / vdx) + 1;
    if (ivx < 1) ivx = 1;
    if (ivx >= m.nx) ivx = m.nx - 1;
    double wx = (px - m.vx[ivx - 1]) / vdx;
    wx = std::min(1.0, std::max(0.0, wx));
    for (int side = 0; side < 2; side++) {
      const double po = side == 0 ? pz - dz : pz + dz;
      int ivz;
      double wz;
      if (po < m.zg[0]) { ivz = 1; wz = 0.0; }
      else if (po >= m.zg[m.nz - 1]) { ivz = m.nz - 1; wz = 1.0; }
      else {
        for (ivz = 1; ivz <= m.nz - 1; ivz++)
          if (po >= m.zg[ivz - 1] && po < m.zg[ivz]) break;
        if (ivz >= m.nz) ivz = m.nz - 1;
        const double lo = po - m.zg[ivz - 1], up = m.zg[ivz] - m.zg[ivz - 1];
        wz = up > 0.0 ? lo / up : 0.0;
      }
      wz = std::min(1.0, std::max(0.0, wz));
      const double v = bilinear_vgrid(m, ivx, ivz, wx, wz);
      (side == 0 ? is.vel_top : is.vel_bot)[c] = v;
    }
  }
}

std::string build_model(HostModel& m, int nx_s, int nz_s, const double* xpos, const double* topo,
                        const double* zrel, const double* vel, double vwater, int nsed,
                        const double* sed_x, const double* sed_z, int nmoho, const double* moho_x,
                        const double* moho_z) {
  std::ostringstream err;
  if (nx_s < 2 || nz_s < 2) return "smesh grid too small";
  m.vwater = vwater;
  const double maxtopo = *std::max_element(topo, topo + nx_s);
  // ---- interfaces 1..5
  std::vector<double> xs(xpos, xpos + nx_s);
  m.if_x[1] = xs; m.if_z[1].assign(nx_s, 0.0);
  m.if_x[2] = xs; m.if_z[2].assign(topo, topo + nx_s);
  m.sf_setup = m.if_z[2];
  m.if_x[3].assign(sed_x, sed_x + nsed); m.if_z[3].assign(sed_z, sed_z + nsed);
  m.if_x[4].assign(moho_x, moho_x + nmoho); m.if_z[4].assign(moho_z, moho_z + nmoho);
  m.if_x[5] = xs; m.if_z[5].assign(nx_s, zrel[nz_s - 1] + maxtopo);
  // ---- propagation grid: square cells of the smesh x spacing
  m.nx = nx_s;
  m.dz = xpos[1] - xpos[0];
  m.dx = m.dz;
  m.z0 = 0.0;
  const double zmax = zrel[nz_s - 1] + maxtopo + 5.0;
  m.nz = (int)lround((zmax - m.z0) / m.dz) + 1;
  m.x0 = xpos[0];
  m.tol = std::min(m.dx, m.dz) * 0.01;
  m.xg.resize(m.nx);
  m.zg.resize(m.nz);
  for (int i = 0; i < m.nx; i++) m.xg[i] = m.x0 + i * m.dx;
  m.vx.assign(xpos, xpos + nx_s);
  for (int j = 0; j < m.nz; j++) m.zg[j] = m.z0 + j * m.dz;
  const int nx = m.nx, nz = m.nz;
  // interfaces must sit on the grid columns (the cut-cell bookkeeping uses
  // the node index as the column, src/propagate2d.f90:1385-1410)
  for (int k = 1; k <= NIF; k++) {
    if ((int)m.if_x[k].size() != nx) {
      err << "interface " << k << " has " << m.if_x[k].size() << " nodes, grid has " << nx << " columns";
      return err.str();
    }
    for (int c = 0; c < nx; c++)
      if (fabs(m.if_x[k][c] - m.xg[c]) > 1e-6 * m.dx) {
        err << "interface " << k << " node " << c + 1 << " is not on grid column x=" << m.xg[c];
        return err.str();
      }
  }
  // ---- velocity resampled to the uniform dz: water above the seafloor,
  // linear in depth below it, last smesh row below the mesh (:225-321)
  m.vgrid.assign((size_t)nx * nz, 0.0);
  for (int i = 0; i < nx; i++) {
    const double zsf = topo[i];
    for (int j = 0; j < nz; j++) {
      const double pr = m.zg[j] - zsf;
      double v;
      if (pr < 0.0) v = vwater;
      else if (pr >= zrel[nz_s - 1]) v = vel[(size_t)i * nz_s + nz_s - 1];
      else {
        int jo;
        for (jo = 1; jo <= nz_s - 1; jo++)
          if (pr >= zrel[jo - 1] && pr < zrel[jo]) break;
        if (jo >= nz_s) jo = nz_s - 1;
        const double w = (pr - zrel[jo - 1]) / (zrel[jo] - zrel[jo - 1]);
        v = (1.0 - w) * vel[(size_t)i * nz_s + jo - 1] + w * vel[(size_t)i * nz_s + jo];
      }
      m.vgrid[(size_t)j * nx + i] = v;
    }
  }
  return rebuild_geometry(m);
}

// interface_normal2d (rays2d.f90:1264-1306) at node c's own x: the cell is
// found by floor on the grid spacing, and the slope is a central difference
// on that cell's nodes (one-sided at the first cell).
static void program_normal(const HostInterface& is, double dx, int c, double& nx_, double& nz_) {
  const int n = is.n;
  if (n < 2) { nx_ = 0.0; nz_ = 1.0; return; }
  int ix = (int)floor((is.x[c] - is.x[0]) / dx) + 1;
  if (ix < 1) ix = 1;
  if (ix >= n) ix = n - 1;
  double dzdx;
  if (ix > 1 && ix < n - 1) dzdx = (is.z[ix] - is.z[ix - 2]) / (2.0 * dx);
  else if (ix == 1 && n >= 3) dzdx = (is.z[ix] - is.z[ix - 1]) / dx;
  else if (ix == n - 1 && n >= 3) dzdx = (is.z[ix] - is.z[ix - 2]) / (2.0 * dx);
  else dzdx = 0.0;
  const double nl = sqrt(1.0 + dzdx * dzdx);
  nx_ = -dzdx / nl;
  nz_ = 1.0 / nl;
}

std::string rebuild_geometry(HostModel& m) {
  const int nx = m.nx, nz = m.nz;
  // ---- intersections: normals, cut cells (:1335-1500)
  m.assoc.assign((size_t)nx * nz, 0);
  for (int k = 1; k <= NIF; k++) {
    HostInterface& is = m.isec[k];
    is.n = nx;
    is.x = m.if_x[k];
    is.z = m.if_z[k];
    is.nrm_x.resize(nx); is.nrm_z.resize(nx);
    is.pnrm_x.resize(nx); is.pnrm_z.resize(nx);
    is.cc.assign(nx, -1); is.ok.assign(nx, 0);
    for (int c = 0; c < nx; c++) {
      double dzdx;
      if (c > 0 && c < nx - 1) dzdx = (is.z[c + 1] - is.z[c - 1]) / (is.x[c + 1] - is.x[c - 1]);
      else if (c == 0) dzdx = (is.z[1] - is.z[0]) / (is.x[1] - is.x[0]);
      else dzdx = (is.z[c] - is.z[c - 1]) / (is.x[c] - is.x[c - 1]);
      const double nl = sqrt(1.0 + dzdx * dzdx);
      is.nrm_x[c] = -dzdx / nl;
      is.nrm_z[c] = 1.0 / nl;
      program_normal(is, m.dx, c, is.pnrm_x[c], is.pnrm_z[c]);
      const double zv = is.z[c];
      int izc = -1;
      for (int j = 0; j < nz - 1; j++)
        if (m.zg[j] <= zv && zv < m.zg[j + 1]) { izc = j; break; }
      if (izc < 0) {
        if (zv <= m.zg[0]) izc = 0;
        if (zv >= m.zg[nz - 1]) izc = nz - 2;
      }
      is.cc[c] = izc;
      if (izc >= 0) {
        is.ok[c] = (m.zg[izc] <= zv && zv <= m.zg[izc + 1]) ? 1 : 0;
        if (c < nx - 1 && izc < nz - 1) m.assoc[(size_t)izc * nx + c] = 1;
      }
    }
  }
  for (int k = 1; k <= NIF; k++) intersection_velocities(m, m.isec[k]);
  // ---- regions (:1503-1610): water, sediment, crust, mantle, seafloor->bottom,
  // seafloor->Moho
  static const int RTOP[NREG + 1] = {0, 1, 2, 3, 4, 2, 2};
  static const int RBOT[NREG + 1] = {0, 2, 3, 4, 5, 5, 4};
  // Every per-node loop below runs row by row (i fastest), in the arrays'
  // storage order: the nodes are independent of each other, so the order of
  // the loops does not change any value.
  // Regions are independent of each other: each one's membership, and later
  // its flags, velocities and node order, is built on a host thread of its
  // own.
  auto per_region = [](const std::function<void(int)>& fn) {
    std::vector<std::thread> th;
    for (int r = 1; r <= NREG; r++) th.emplace_back(fn, r);
    for (auto& t : th) t.join();
  };
  per_region([&](int r) {
    HostRegion& R = m.reg[r];
    R.id = r; R.itop = RTOP[r]; R.ibot = RBOT[r];
    R.flags.assign((size_t)nx * nz, 0);
    const HostInterface& t = m.isec[R.itop];
    const HostInterface& b = m.isec[R.ibot];
    std::vector<double> zt(nx), zb(nx);
    for (int i = 0; i < nx; i++) { zt[i] = iface_z_grid(m, t, m.xg[i]); zb[i] = iface_z_grid(m, b, m.xg[i]); }
    for (int j = 0; j < nz; j++)
      for (int i = 0; i < nx; i++)
        if (zt[i] <= m.zg[j] && m.zg[j] <= zb[i]) R.flags[(size_t)j * nx + i] |= RF_IN;
  });
  // the last region that holds a node names it (regions are written in
  // order, :1576-1577); -1 for a node of none
  m.node_region.assign((size_t)nx * nz, -1);
  for (int r = 1; r <= NREG; r++) {
    const std::vector<uint16_t>& F = m.reg[r].flags;
    for (size_t g = 0; g < F.size(); g++)
      if (F[g] & RF_IN) m.node_region[g] = r;
  }
  // fully regular nodes (:1932-1979)
  m.fully_regular.assign((size_t)nx * nz, 0);
  for (int j = 0; j < nz; j++)
    for (int i = 0; i < nx; i++) {
      const size_t g = (size_t)j * nx + i;
      if (m.node_region[g] <= 0) continue;
      bool reg = true;
      for (int di = 0; di <= 1; di++)
        for (int dj = 0; dj <= 1; dj++) {
          const int cx = i - di, cz = j - dj;
          if (cx >= 0 && cx < nx - 1 && cz >= 0 && cz < nz - 1 && m.assoc[(size_t)cz * nx + cx]) reg = false;
        }
      if (i == 0 || i == nx - 1 || j == 0 || j == nz - 1) reg = false;
      m.fully_regular[g] = reg;
    }
  per_region([&](int r) {
    HostRegion& R = m.reg[r];
    const HostInterface& t = m.isec[R.itop];
    const HostInterface& b = m.isec[R.ibot];
    auto in = [&](int i, int j) { return (R.flags[(size_t)j * nx + i] & RF_IN) != 0; };
    for (int j = 0; j < nz; j++)
      for (int i = 0; i < nx; i++) {
        const size_t g = (size_t)j * nx + i;
        if (!in(i, j)) continue;
        uint16_t f = R.flags[g];
        for (int k = 0; k < 4; k++) {
          const int cx = i + cell_dx(k), cz = j + cell_dz(k);
          if (cx < 0 || cx > nx - 2 || cz < 0 || cz > nz - 2) continue;
          f |= (uint16_t)(RF_CELL0 << k);
          if (m.assoc[(size_t)cz * nx + cx]) f |= (uint16_t)(RF_ASSOC0 << k);
          if (t.cc[cx] == cz || b.cc[cx] == cz) f |= (uint16_t)(RF_IRR0 << k);
        }
        if (m.fully_regular[g]) f |= RF_FULLREG;
        // coincidence with the region's own interface node of this column
        // (which must be a member of one of this node's cells)
        const HostInterface* sides[2] = {&t, &b};
        for (int s = 0; s < 2; s++) {
          const HostInterface& is = *sides[s];
          if (i > nx - 2 || is.cc[i] < 0 || !is.ok[i]) continue;
          if (!(is.cc[i] == j || is.cc[i] == j - 1) || is.cc[i] > nz - 2) continue;
          const double ddx = m.xg[i] - is.x[i], ddz = m.zg[j] - is.z[i];
          if (sqrt(ddx * ddx + ddz * ddz) < 0.01 * m.tol) { f |= (uint16_t)(RF_COI0 << s); break; }
        }
        R.flags[g] = f;
      }
    // velocities_to_region2d (:1788-1929)
    R.vel.assign((size_t)(nz + 2) * nx, 0.0);
    for (int i = 0; i < nx; i++) {
      R.vel[(size_t)nz * nx + i] = t.vel_bot[i];
      R.vel[(size_t)(nz + 1) * nx + i] = b.vel_top[i];
    }
    for (int j = 0; j < nz; j++)
      for (int i = 0; i < nx; i++) {
        if (!in(i, j)) continue;
        double v = m.vgrid[(size_t)j * nx + i];
        if (j < nz - 1 && !in(i, j + 1)) {
          int i3 = (int)lround((m.xg[i] - b.x[0]) / m.dx);
          i3 = std::min(std::max(i3, 0), b.n - 1);
          v = b.vel_top[i3];
        }
        if (j > 0 && !in(i, j - 1)) {
          int i3 = (int)lround((m.xg[i] - t.x[0]) / m.dx);
          i3 = std::min(std::max(i3, 0), t.n - 1);
          v = t.vel_bot[i3];
        }
        R.vel[(size_t)j * nx + i] = v;
      }
    // reference node order: grid nodes ix-outer / iz-inner, then top, then
    // bottom; membership is gathered column by column from a transposed copy
    std::vector<uint8_t> cm((size_t)nx * nz);
    for (int j = 0; j < nz; j++)
      for (int i = 0; i < nx; i++) cm[(size_t)i * nz + j] = in(i, j) ? 1 : 0;
    R.ref_order.clear();
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < nz; j++)
        if (cm[(size_t)i * nz + j]) R.ref_order.push_back(j * nx + i);
    R.ngnode = (int)R.ref_order.size();
    for (int c = 0; c < t.n; c++) R.ref_order.push_back(nz * nx + c);
    for (int c = 0; c < b.n; c++) R.ref_order.push_back((nz + 1) * nx + c);
    R.nnode = (int)R.ref_order.size();
  });
  return "";
}

}  // namespace fmt2d
