// Host-side model build of the MI355X eikonal path: propagation grid,
// interface/grid intersections, regions 1-6 and their velocities, in the
// dense layout the kernels use (see eik_common.h).
#pragma once
#include <stdint.h>
#include <string>
#include <vector>

namespace fmt2d {

constexpr int NIF = 5;   // sea surface, seafloor, sediment base, Moho, bottom
constexpr int NREG = 6;  // water, sediment, crust, mantle, seafloor->bottom, seafloor->Moho

struct HostInterface {      // one intersection of an interface with the grid
  int n = 0;
  std::vector<double> x, z, nrm_x, nrm_z, vel_top, vel_bot;
  // normals as the forward program's reflection takes them: interface_normal2d
  // evaluated at each node's x (rays2d.f90:1264-1306, fmt2d_main.f90:1563)
  std::vector<double> pnrm_x, pnrm_z;
  std::vector<int> cc;      // 0-based cut-cell row of node c (cell column c), -1 none
  std::vector<uint8_t> ok;  // node inside its own cut cell
};

struct HostRegion {
  int id = 0, itop = 0, ibot = 0;          // interface ids 1..5
  std::vector<uint16_t> flags;             // (nz, nx)
  std::vector<double> vel;                 // (nz+2, nx)
  int ngnode = 0, nnode = 0;
  std::vector<int> ref_order;              // reference node order -> layout index
};

struct HostModel {
  int nx = 0, nz = 0;
  double dx = 0, dz = 0, x0 = 0, z0 = 0, tol = 0;
  double vwater = 1.5;
  std::vector<double> xg, zg;
  std::vector<double> vx;  // velocity-grid x: the smesh node positions (fmt2d_main.f90:181)
  std::vector<double> vgrid;               // (nz, nx) z-resampled velocity
  std::vector<int> node_region;            // last writer (reference semantics)
  std::vector<uint8_t> assoc;              // (nz, nx) cell cut by any interface
  std::vector<uint8_t> fully_regular;      // (nz, nx)
  HostInterface isec[NIF + 1];
  HostRegion reg[NREG + 1];
  // interfaces as read (for placing sources/receivers on the seafloor)
  std::vector<double> if_x[NIF + 1], if_z[NIF + 1];
  // seafloor depths as read at set-up: sources and the in-water test of
  // receivers stay on it when the seafloor is later inverted
  // (invert2d_main.f90:517-531, 544-546)
  std::vector<double> sf_setup;
  double iface_z_at(int k, double px) const;
  double setup_seafloor_at(double px) const;
};

// Builds the model of src/fmt2d_main.f90:91-418 from the .smesh content and
// the sediment-base / Moho interface nodes.  vel is (nx_s, nz_s), iz fastest.
// Returns an empty string on success, else an error message.
std::string build_model(HostModel& m, int nx_s, int nz_s, const double* xpos, const double* topo,
                        const double* zrel, const double* vel, double vwater, int nsed,
                        const double* sed_x, const double* sed_z, int nmoho, const double* moho_x,
                        const double* moho_z);

// Intersections, regions, cut cells and region velocities from the model's
// velocity grid and interface nodes (rebuild_geometry2d,
// src/inversion2d.f90:3661-3816); build_model ends with it, a model step
// (new velocities or interface depths) calls it again.
std::string rebuild_geometry(HostModel& m);

}  // namespace fmt2d
