// Ray backtracing and Frechet rows on the MI355X (device side).
//
// One thread per (receiver, phase) ray walks the steepest-descent path of the
// resident stage fields from the receiver back to the source, section by
// section, as trace_ray_from_receiver2d / trace_ray_section2d do
// (src/rays2d.f90:7-762), after placing an in-water receiver at its optimal
// seafloor entry point (find_optimal_entry_inv, src/inversion2d.f90:6344-6438).
// Points stay in HBM; the Frechet rows (compute_frechet_for_ray2d,
// src/inversion2d.f90:4548-4787) are built from them as CSR.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fmt2d {

constexpr int RAY_MAXP = 5000;      // maxpoints of one section (rays2d.f90:205)
constexpr int RAY_MAXSEC = 4;

struct RayRegionDev {
  int id, nnode, itop, ibot;
  const int* order;        // reference node k (0-based) -> field slot
  const uint16_t* flags;   // (nz, nx)
  const double* vel;       // (nz+2, nx) region node velocities
  // column index of the nearest-node search (nearest_node): rows of the
  // region's grid nodes per column (jlo > jhi: none), and bounds on the x of
  // every node (grid and interface) of the columns from c on (xlo_from) and
  // up to c (xhi_upto); indexed = 0 when some column's rows are not one run
  int indexed;
  const int* jlo;
  const int* jhi;
  const double* xlo_from;
  const double* xhi_upto;
};

struct RayGeomDev {
  int nx, nz;
  double dx, dz, x0, z0;
  const double* xg;
  const double* zg;
  const double* isx[6];    // intersections 1..5: node x, z (one per grid column)
  const double* isz[6];
  int isn[6];
  RayRegionDev reg[7];
  const double* sfx;       // seafloor interface as read (intrface(2)%x, %z_nodes)
  const double* sfz;
  const double* sfz_setup; // seafloor depths at set-up (in-water test of receivers)
  int nsf;
  double sfdx;
  double vwater;
  const double* pvel;      // propagation-grid velocity (nz, nx)
  // inversion velocity grid (resampled smesh grid) and its parameter numbers
  int vnx, vnz;
  double vdx, vdz;
  const double* vx;
  const double* vz;
  const int* inv_idx;      // (vnz, vnx): 1-based parameter, 0 = not inverted
  int n_vel_par;
  int n_par;               // all parameters (n_inv_parms)
  // inverted interfaces: parameter block of interface k (input nodes), -1 none
  int if_start[6];
  int if_nnode[6];
  const double* ifx_in[6];
  double ifdx_in[6];
};

// one stage field of one source
struct RayField {
  const double* T;
  const double* GX;
  const double* GZ;
  int region;
};

struct RayJob {
  int src;       // source index of the last forward
  int path;      // 1..5: Pg, PsedP, PmP, Pg_ss, Pg_ss2
  double rx, rz; // receiver position as given
};

struct RayOut {
  double t;            // receiver time incl. the water leg
  double water_time;
  double ex, ez;       // receiver position the ray starts from (entry point)
  int valid;
  int nsec;
  int npts[RAY_MAXSEC];
  int headwave, diffracted;
  int direct;          // Pg taken as the direct water wave (no sections)
  int work;            // gradient interpolations the walk took (rounds of the trace loop)
};

hipError_t launch_seafloor_times(const RayGeomDev* G, const RayField* fields, int nsrc, int nsf, double* tsf,
                                 hipStream_t st);
// adjoint sources of the receiver times (weights g per ray) in the
// adjoint fields lam[field_of[src*12 + itf-1]][slot]; active[f] set for
// fields that receive one
hipError_t launch_adjoint_sources(const RayGeomDev* G, const RayField* fields, const RayJob* jobs, const RayOut* out,
                                  const double* g, int nray, const int* field_of, long long nslot, double* lam,
                                  int* active, hipStream_t st);
// order: the rays in the order they are traced (a permutation of 0..nray-1:
// the longest first, so that no long wave starts late; rays of one source
// and path side by side)
hipError_t launch_trace(const RayGeomDev* G, const RayField* fields, const double* src_x, const double* src_z,
                        const double* tsf, const RayJob* jobs, const int* order, int nray, const long long* pt_off,
                        double2* pts, RayOut* out, int program, hipStream_t st);
// Frechet contributions of each ray in the reference's accumulation order:
// pass 1 (keys == nullptr) counts them per ray, pass 2 writes
// key = ray << 32 | (parameter - 1), value = weight * length (or the
// interface geometry factor).
hipError_t launch_frechet_contrib(const RayGeomDev* G, const RayField* fields, const RayJob* jobs, int nray,
                                  int ray0, const long long* pt_off, const double2* pts, const RayOut* out,
                                  const long long* coff, long long* count, unsigned long long* keys, double* vals,
                                  hipStream_t st);

// nearest node of region `region` for n points, by the column index (idx;
// -2 when the region has none) and by the scan (scan)
hipError_t launch_nearest_check(const RayGeomDev* G, int region, int n, const double* px, const double* pz, int* idx,
                                int* scan, hipStream_t st);
// count of nearest-node searches (fallbacks of the interpolators)
hipError_t nearest_calls_reset(hipStream_t st);
hipError_t nearest_calls_read(unsigned long long* n, hipStream_t st);

}  // namespace fmt2d
