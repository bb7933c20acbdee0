// Shared definitions of the MI355X eikonal path (host and device).
//
// Layout of one traveltime field in HBM (FP64, SoA):
//   rows 0..nz-1   : grid nodes, row-major (ix fastest)   index iz*nx + ix
//   row  nz        : nodes of the region's top interface, one per column
//   row  nz+1      : nodes of the region's bottom interface, one per column
// so a field is (nz+2)*nx doubles per component (T, dT/dx, dT/dz).  The
// reference keeps the same nodes as a compacted list (grid ix-outer/iz-inner,
// then top, then bottom interface nodes; src/propagate2d.f90:1564-1609);
// the host maps between the two orders.
#pragma once
#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define FMT_HD __host__ __device__ inline
#else
// host-only compilation of the shared operator (unit tests of the update math)
#define FMT_HD inline
#endif

#define FMT_HUGE_TIME 1.0e20f
#define FMT_HUGE_HALF 0.5e20f

// per-region grid-node flag bits
enum {
  RF_IN = 1u << 0,         // node belongs to the region
  RF_IRR0 = 1u << 1,       // cell k (k=0..3) cut by the region's top/bottom interface
  RF_ASSOC0 = 1u << 5,     // cell k cut by any interface (fit-gradient path)
  RF_CELL0 = 1u << 9,      // cell k exists inside the grid
  RF_FULLREG = 1u << 13,   // fully regular grid node (src/propagate2d.f90:1932-1979)
  RF_COI0 = 1u << 14,      // coincides with the region's top (bit 14) / bottom (bit 15)
                           // interface node of its column, which is a member of its cells
};
// cell k of grid node (ix,iz), in the reference's enumeration order
// (find_connected_cells2d, src/propagate2d.f90:360-371): k=0 (ix,iz),
// k=1 (ix,iz-1), k=2 (ix-1,iz), k=3 (ix-1,iz-1)  (cells named by top-left node)
FMT_HD int cell_dx(int k) { return (k >> 1) ? -1 : 0; }
FMT_HD int cell_dz(int k) { return (k & 1) ? -1 : 0; }

// Static description of one region on the device (all arrays device pointers).
struct RegionDev {
  int nx, nz;
  double dx, dz;
  const double* xg;       // grid x (nx)
  const double* zg;       // grid z (nz)
  const uint16_t* flags;  // (nz, nx)
  const double* vel;      // (nz+2, nx): region velocity incl. interface rows
  // interface nodes: side 0 = top, 1 = bottom
  const double* ifx[2];   // x of interface node per column
  const double* ifz[2];   // z
  const int* ifcc[2];     // cut-cell row (0-based) of the node's own cell, -1 none
  const uint8_t* ifok[2]; // node lies inside its cut cell (inodes list)
  double tol;             // 0.01*min(dx,dz)
};
