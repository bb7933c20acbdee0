/*
 * ORACLE -- CPU restatement of the reference FMT2D eikonal path, used only as
 * a checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline.
 * It is never linked into, called by, or shipped with the product library.
 *
 * Restated (FP64, same operation order) from the reference sources:
 *   geometry       src/propagate2d.f90:1335-1979  (find_intersection2d,
 *                  define_region2d, velocities_to_*, tag_regular_nodes2d)
 *   model build    src/fmt2d_main.f90:91-418 (interfaces, resampling, regions)
 *   FMM            src/propagate2d.f90:33-946     (propagate2d,
 *                  find_connected_nodes2d/cells2d, update_time2d,
 *                  regular_update2d, time_from_1_node2d, time_from_2_nodes2d)
 *   gradient fit   src/propagate2d.f90:949-1196   (fit_gradient_at_node2d)
 *   heap           src/propagate2d.f90:1199-1332
 *   seeding        src/propagate2d.f90:1982-2741  (sweep_region_from_interface2d,
 *                  initialize_region_from_source2d, transfer_times_to_intersection,
 *                  reflect_gradient_at_interface)
 *   stage schedule src/inversion2d.f90:4184-4416  (run_forward2d, incl. the
 *                  water top-row gradient fix and the PmP head-wave post-pass)
 *
 * Pinned against the reference's own routines (oracle/ref_fields.f90 driving
 * the reference modules built by oracle/build_ref.sh): geometry, velocities and
 * every stage field are compared node by node in tests/test_oracle_pin.py.
 *
 * Indexing follows the reference: node ids and grid indices are 1-based.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define HUGE_TIME 1.0e20
#define NIF 5
#define NREG 6

typedef struct {
  int nnode, id;
  double *x, *z, *nx_, *nz_;          /* normal(1:2,n) as nx_, nz_ */
  double *vel_top, *vel_bot;
  double *atime, *gx, *gz;            /* arrivaltime / time_gradient */
  int n_ccells;
  int *cc_ix, *cc_iz;                 /* ccells(icell)%ix/iz, 1-based icell */
  int *n_inodes, *inodes;             /* inodes(4, icell) */
  int *ccell_from_inode;              /* (4, nnode) */
  int *rleft_node_id, *rright_node_id;
  int regleft, regright;              /* region ids, 0 = none */
} Isec;

typedef struct {
  int id, itop, ibot;                 /* intersection ids */
  int nnode, ngnode;
  int *i1, *i2, *i3;
  double *x, *z, *vel;
  int *status;
  double *T, *gx, *gz;
  int *rnode;                         /* region_rnode_id(nx,nz) */
  int *itop_rnode, *ibot_rnode;
} Region;

typedef struct {
  int nx, nz;
  double dx, dz, x0, z0, tol;
  double *x, *z;
  double *vel;                        /* pgrid%velocity(1,ix,iz) */
  int *node_region, *rnode_id;
  int *fully_regular;
  int *ccind;                         /* (NIF, nx, nz): 0 = none; -1 marks "p associated" via ccassoc */
  int *ccassoc;                       /* (nx, nz) */
  /* resampled velocity grid (vgrid) */
  int vnx, vnz;
  double vdx, vdz;
  double *vx, *vz, *vvel;
  Isec isec[NIF + 1];
  Region reg[NREG + 1];
  double vwater;
  /* interfaces (Tinterface2d) for source placement */
  int if_n[NIF + 1];
  double if_dx[NIF + 1];
  double *if_x[NIF + 1], *if_z[NIF + 1];
} Model;

#define G2(a, ix, iz) ((a)[((iz) - 1) * M->nx + ((ix) - 1)])
#define CCIND(ix, iz, f) (M->ccind[(((iz) - 1) * M->nx + ((ix) - 1)) * NIF + ((f) - 1)])

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

/* ------------------------------------------------------------ geometry */

static void find_intersection(Model *M, Isec *is, int id, int n, const double *x, const double *z) {
  is->id = id;
  is->nnode = n;
  is->x = xcalloc(n + 1, sizeof(double));
  is->z = xcalloc(n + 1, sizeof(double));
  is->nx_ = xcalloc(n + 1, sizeof(double));
  is->nz_ = xcalloc(n + 1, sizeof(double));
  for (int i = 1; i <= n; i++) { is->x[i] = x[i - 1]; is->z[i] = z[i - 1]; }
  for (int ix = 1; ix <= n; ix++) {
    double dzdx;
    if (ix > 1 && ix < n) dzdx = (is->z[ix + 1] - is->z[ix - 1]) / (is->x[ix + 1] - is->x[ix - 1]);
    else if (ix == 1 && n >= 2) dzdx = (is->z[ix + 1] - is->z[ix]) / (is->x[ix + 1] - is->x[ix]);
    else if (ix == n && n >= 2) dzdx = (is->z[ix] - is->z[ix - 1]) / (is->x[ix] - is->x[ix - 1]);
    else dzdx = 0.0;
    double nl = sqrt(1.0 + dzdx * dzdx);
    is->nx_[ix] = -dzdx / nl;
    is->nz_[ix] = 1.0 / nl;
  }
  is->cc_ix = xcalloc(n + 1, sizeof(int));
  is->cc_iz = xcalloc(n + 1, sizeof(int));
  int ncc = 0;
  for (int ix = 1; ix <= n; ix++) {
    double zv = is->z[ix];
    int izc = 0;
    for (int iz = 1; iz <= M->nz - 1; iz++)
      if (M->z[iz] <= zv && zv < M->z[iz + 1]) { izc = iz; break; }
    if (izc == 0) {
      if (zv <= M->z[1]) izc = 1;
      if (zv >= M->z[M->nz]) izc = M->nz - 1;
    }
    if (izc > 0) { ncc++; is->cc_ix[ncc] = ix; is->cc_iz[ncc] = izc; }
  }
  is->n_ccells = ncc;
  is->n_inodes = xcalloc(ncc + 2, sizeof(int));
  is->inodes = xcalloc(4 * (ncc + 2), sizeof(int));
  /* every cut cell belongs to the node of its own column (cc_ix == node) */
  for (int ix = 1; ix <= n; ix++) {
    double zv = is->z[ix];
    for (int ic = 1; ic <= ncc; ic++) {
      if (is->cc_ix[ic] != ix) continue;
      int izc = is->cc_iz[ic];
      if (M->z[izc] <= zv && zv <= M->z[izc + 1]) {
        is->n_inodes[ic]++;
        if (is->n_inodes[ic] > 4) { fprintf(stderr, "too many inodes\n"); abort(); }
        is->inodes[(ic - 1) * 4 + is->n_inodes[ic] - 1] = ix;
      }
    }
  }
  is->ccell_from_inode = xcalloc(4 * (n + 1), sizeof(int));
  for (int ix = 1; ix <= n; ix++) {
    int cnt = 0;
    for (int ic = 1; ic <= ncc; ic++)
      if (is->cc_ix[ic] == ix) { cnt++; if (cnt <= 4) is->ccell_from_inode[(ix - 1) * 4 + cnt - 1] = ic; }
  }
  for (int ic = 1; ic <= ncc; ic++) {
    int ix = is->cc_ix[ic], iz = is->cc_iz[ic];
    if (ix >= 1 && ix < M->nx && iz >= 1 && iz < M->nz) {
      G2(M->ccassoc, ix, iz) = 1;
      CCIND(ix, iz, id) = ic;
    }
  }
}

static double iface_z_grid(const Model *M, const Isec *is, double px) {
  /* interpolate_iface_z2d (propagate2d.f90:1615-1645): uses the grid dx */
  if (is->nnode < 2) return is->z[1];
  int ixx = (int)floor((px - is->x[1]) / M->dx) + 1;
  if (ixx < 1) return is->z[1];
  if (ixx >= is->nnode) return is->z[is->nnode];
  double w = (px - is->x[ixx]) / M->dx;
  if (w < 0.0) w = 0.0;
  if (w > 1.0) w = 1.0;
  return (1.0 - w) * is->z[ixx] + w * is->z[ixx + 1];
}

static void define_region(Model *M, Region *R, int id, int itop, int ibot) {
  Isec *t = &M->isec[itop], *b = &M->isec[ibot];
  R->id = id; R->itop = itop; R->ibot = ibot;
  int ng = 0;
  for (int ix = 1; ix <= M->nx; ix++)
    for (int iz = 1; iz <= M->nz; iz++) {
      double zt = iface_z_grid(M, t, M->x[ix]), zb = iface_z_grid(M, b, M->x[ix]);
      if (zt <= M->z[iz] && M->z[iz] <= zb) ng++;
    }
  R->ngnode = ng;
  R->nnode = ng + t->nnode + b->nnode;
  int nn = R->nnode + 1;
  R->i1 = xcalloc(nn, sizeof(int)); R->i2 = xcalloc(nn, sizeof(int)); R->i3 = xcalloc(nn, sizeof(int));
  R->x = xcalloc(nn, sizeof(double)); R->z = xcalloc(nn, sizeof(double));
  R->vel = xcalloc(nn, sizeof(double));
  R->status = xcalloc(nn, sizeof(int));
  R->T = xcalloc(nn, sizeof(double)); R->gx = xcalloc(nn, sizeof(double)); R->gz = xcalloc(nn, sizeof(double));
  R->rnode = xcalloc((size_t)M->nx * M->nz, sizeof(int));
  for (int i = 1; i <= R->nnode; i++) { R->status[i] = -1; R->T[i] = HUGE_TIME; }
  int n = 0;
  for (int ix = 1; ix <= M->nx; ix++)
    for (int iz = 1; iz <= M->nz; iz++) {
      double zt = iface_z_grid(M, t, M->x[ix]), zb = iface_z_grid(M, b, M->x[ix]);
      if (zt <= M->z[iz] && M->z[iz] <= zb) {
        n++;
        R->i1[n] = ix; R->i2[n] = iz; R->i3[n] = 0;
        R->x[n] = M->x[ix]; R->z[n] = M->z[iz];
        G2(M->rnode_id, ix, iz) = n;
        G2(M->node_region, ix, iz) = id;
        G2(R->rnode, ix, iz) = n;
      } else if (G2(M->node_region, ix, iz) == 0) {
        G2(M->node_region, ix, iz) = -1;
      }
    }
  free(t->rright_node_id);
  t->rright_node_id = xcalloc(t->nnode + 1, sizeof(int));
  R->itop_rnode = xcalloc(t->nnode + 1, sizeof(int));
  for (int m = 1; m <= t->nnode; m++) {
    n++;
    R->i1[n] = 0; R->i2[n] = t->id; R->i3[n] = m;
    R->x[n] = t->x[m]; R->z[n] = t->z[m];
    t->rright_node_id[m] = n; R->itop_rnode[m] = n;
  }
  free(b->rleft_node_id);
  b->rleft_node_id = xcalloc(b->nnode + 1, sizeof(int));
  R->ibot_rnode = xcalloc(b->nnode + 1, sizeof(int));
  for (int m = 1; m <= b->nnode; m++) {
    n++;
    R->i1[n] = 0; R->i2[n] = b->id; R->i3[n] = m;
    R->x[n] = b->x[m]; R->z[n] = b->z[m];
    b->rleft_node_id[m] = n; R->ibot_rnode[m] = n;
  }
}

static void tag_regular(Model *M) {
  for (int ix = 1; ix <= M->nx; ix++)
    for (int iz = 1; iz <= M->nz; iz++) {
      if (G2(M->node_region, ix, iz) <= 0) { G2(M->fully_regular, ix, iz) = 0; continue; }
      int reg = 1;
      for (int i = 0; i <= 1; i++)
        for (int j = 0; j <= 1; j++)
          if (ix - i > 0 && ix - i < M->nx && iz - j > 0 && iz - j < M->nz)
            if (G2(M->ccassoc, ix - i, iz - j)) reg = 0;
      if (ix == 1 || ix == M->nx) reg = 0;
      if (iz == 1 || iz == M->nz) reg = 0;
      G2(M->fully_regular, ix, iz) = reg;
    }
}

#define VV(ix, iz) (M->vvel[((iz) - 1) * M->vnx + ((ix) - 1)])

static void vel_to_intersection(Model *M, Isec *is) {
  is->vel_top = xcalloc(is->nnode + 1, sizeof(double));
  is->vel_bot = xcalloc(is->nnode + 1, sizeof(double));
  double dz = M->vdz;
  for (int m = 1; m <= is->nnode; m++) {
    double px = is->x[m], pz = is->z[m];
    int ivx = (int)floor((px - M->vx[1]) / M->vdx) + 1;
    if (ivx < 1) ivx = 1;
    if (ivx >= M->vnx) ivx = M->vnx - 1;
    double wx = (px - M->vx[ivx]) / M->vdx;
    if (wx < 0.0) wx = 0.0;
    if (wx > 1.0) wx = 1.0;
    for (int side = 0; side < 2; side++) {
      double po = side == 0 ? pz - dz : pz + dz, wz;
      int ivz;
      if (po < M->vz[1]) { ivz = 1; wz = 0.0; }
      else if (po >= M->vz[M->vnz]) { ivz = M->vnz - 1; wz = 1.0; }
      else {
        for (ivz = 1; ivz <= M->vnz - 1; ivz++)
          if (po >= M->vz[ivz] && po < M->vz[ivz + 1]) break;
        if (ivz >= M->vnz) ivz = M->vnz - 1;
        double lo = po - M->vz[ivz], up = M->vz[ivz + 1] - M->vz[ivz];
        wz = up > 0.0 ? lo / up : 0.0;
      }
      if (wz < 0.0) wz = 0.0;
      if (wz > 1.0) wz = 1.0;
      double v = (1.0 - wx) * (1.0 - wz) * VV(ivx, ivz) + wx * (1.0 - wz) * VV(ivx + 1, ivz) +
                 (1.0 - wx) * wz * VV(ivx, ivz + 1) + wx * wz * VV(ivx + 1, ivz + 1);
      if (side == 0) is->vel_top[m] = v; else is->vel_bot[m] = v;
    }
  }
}

static void vel_to_region(Model *M, Region *R) {
  Isec *t = &M->isec[R->itop], *b = &M->isec[R->ibot];
  for (int n = 1; n <= R->nnode; n++) {
    if (R->i1[n] != 0) R->vel[n] = G2(M->vel, R->i1[n], R->i2[n]);
    else if (R->i2[n] == t->id) R->vel[n] = t->vel_bot[R->i3[n]];
    else if (R->i2[n] == b->id) R->vel[n] = b->vel_top[R->i3[n]];
  }
  for (int n = 1; n <= R->nnode; n++) {
    int i1 = R->i1[n], i2 = R->i2[n];
    if (i1 == 0) continue;
    if (i2 < M->nz && G2(R->rnode, i1, i2 + 1) <= 0) {
      int i3 = (int)lround((R->x[n] - b->x[1]) / M->dx) + 1;
      if (i3 < 1) i3 = 1;
      if (i3 > b->nnode) i3 = b->nnode;
      R->vel[n] = b->vel_top[i3];
    }
    if (i2 > 1 && G2(R->rnode, i1, i2 - 1) <= 0) {
      int i3 = (int)lround((R->x[n] - t->x[1]) / M->dx) + 1;
      if (i3 < 1) i3 = 1;
      if (i3 > t->nnode) i3 = t->nnode;
      R->vel[n] = t->vel_bot[i3];
    }
  }
}

/* Model build of fmt2d_main.f90:91-418 from the smesh, Moho and sediment-base
 * inputs.  vel is (nx_s, nz_s) with iz fastest (the file's row order). */
Model *orc_model_create(int nx_s, int nz_s, const double *xpos, const double *topo,
                        const double *zrel, const double *vel, double vwater,
                        int nsed, const double *sed_x, const double *sed_z,
                        int nmoho, const double *moho_x, const double *moho_z) {
  Model *M = xcalloc(1, sizeof(Model));
  M->vwater = vwater;
  double maxtopo = topo[0];
  for (int i = 1; i < nx_s; i++) if (topo[i] > maxtopo) maxtopo = topo[i];
  /* interfaces */
  double *zero = xcalloc(nx_s, sizeof(double)), *bot = xcalloc(nx_s, sizeof(double));
  for (int i = 0; i < nx_s; i++) bot[i] = zrel[nz_s - 1] + maxtopo;
  const double *ifx[NIF + 1] = {0, xpos, xpos, sed_x, moho_x, xpos};
  const double *ifz[NIF + 1] = {0, zero, topo, sed_z, moho_z, bot};
  int ifn[NIF + 1] = {0, nx_s, nx_s, nsed, nmoho, nx_s};
  for (int k = 1; k <= NIF; k++) {
    M->if_n[k] = ifn[k];
    M->if_dx[k] = ifx[k][1] - ifx[k][0];
    M->if_x[k] = xcalloc(ifn[k], sizeof(double));
    M->if_z[k] = xcalloc(ifn[k], sizeof(double));
    memcpy(M->if_x[k], ifx[k], ifn[k] * sizeof(double));
    memcpy(M->if_z[k], ifz[k], ifn[k] * sizeof(double));
  }
  /* propagation grid */
  M->nx = nx_s;
  M->dz = xpos[1] - xpos[0];
  M->dx = M->dz;
  M->z0 = 0.0;
  double zmax = zrel[nz_s - 1] + maxtopo + 5.0;
  M->nz = (int)lround((zmax - M->z0) / M->dz) + 1;
  M->x0 = xpos[0];
  M->tol = (M->dx < M->dz ? M->dx : M->dz) * 0.01;
  M->x = xcalloc(M->nx + 1, sizeof(double));
  M->z = xcalloc(M->nz + 1, sizeof(double));
  for (int n = 1; n <= M->nx; n++) M->x[n] = M->x0 + (n - 1) * M->dx;
  for (int n = 1; n <= M->nz; n++) M->z[n] = M->z0 + (n - 1) * M->dz;
  /* z-resampled velocity grid (fmt2d_main.f90:225-321) */
  M->vnx = nx_s; M->vnz = M->nz; M->vdx = xpos[1] - xpos[0]; M->vdz = M->dz;
  M->vx = xcalloc(M->vnx + 1, sizeof(double));
  M->vz = xcalloc(M->vnz + 1, sizeof(double));
  M->vvel = xcalloc((size_t)M->vnx * M->vnz, sizeof(double));
  for (int i = 1; i <= nx_s; i++) M->vx[i] = xpos[i - 1];
  for (int j = 1; j <= M->vnz; j++) M->vz[j] = (j - 1) * M->dz;
  for (int i = 1; i <= nx_s; i++) {
    int ivx = i;
    if (ivx >= nx_s) ivx = nx_s - 1;
    double wx = (double)(i - ivx);
    if (wx < 0.0) wx = 0.0;
    if (wx > 1.0) wx = 1.0;
    double zsf = (1.0 - wx) * topo[ivx - 1] + wx * topo[ivx];
    for (int j = 1; j <= M->vnz; j++) {
      double pr = M->vz[j] - zsf, v;
      if (pr < 0.0) v = vwater;
      else if (pr >= zrel[nz_s - 1]) v = vel[(size_t)(i - 1) * nz_s + nz_s - 1];
      else {
        int jo;
        for (jo = 1; jo <= nz_s - 1; jo++)
          if (pr >= zrel[jo - 1] && pr < zrel[jo]) break;
        if (jo >= nz_s) jo = nz_s - 1;
        double w = (pr - zrel[jo - 1]) / (zrel[jo] - zrel[jo - 1]);
        v = (1.0 - w) * vel[(size_t)(i - 1) * nz_s + jo - 1] + w * vel[(size_t)(i - 1) * nz_s + jo];
      }
      VV(i, j) = v;
    }
  }
  size_t ng = (size_t)M->nx * M->nz;
  M->node_region = xcalloc(ng, sizeof(int));
  M->rnode_id = xcalloc(ng, sizeof(int));
  M->fully_regular = xcalloc(ng, sizeof(int));
  M->ccind = xcalloc(ng * NIF, sizeof(int));
  M->ccassoc = xcalloc(ng, sizeof(int));
  M->vel = xcalloc(ng, sizeof(double));
  for (int k = 1; k <= NIF; k++) find_intersection(M, &M->isec[k], k, ifn[k], ifx[k], ifz[k]);
  for (int r = 1; r <= 4; r++) define_region(M, &M->reg[r], r, r, r + 1);
  free(M->isec[2].rright_node_id); M->isec[2].rright_node_id = NULL;
  free(M->isec[5].rleft_node_id); M->isec[5].rleft_node_id = NULL;
  define_region(M, &M->reg[5], 5, 2, 5);
  M->isec[2].regright = 5;
  M->isec[5].regleft = 5;
  define_region(M, &M->reg[6], 6, 2, 4);
  tag_regular(M);
  /* velocities_to_grid2d: the resampled grid has the pgrid shape */
  for (int ix = 1; ix <= M->nx; ix++)
    for (int iz = 1; iz <= M->nz; iz++) G2(M->vel, ix, iz) = VV(ix, iz);
  for (int k = 1; k <= NIF; k++) vel_to_intersection(M, &M->isec[k]);
  for (int r = 1; r <= NREG; r++) vel_to_region(M, &M->reg[r]);
  M->isec[1].regleft = 0; M->isec[1].regright = 1;
  M->isec[3].regleft = 2; M->isec[3].regright = 3;
  M->isec[4].regleft = 3; M->isec[4].regright = 4;
  M->isec[5].regright = 0;
  for (int k = 1; k <= NIF; k++) {
    Isec *is = &M->isec[k];
    is->atime = xcalloc(is->nnode + 1, sizeof(double));
    is->gx = xcalloc(is->nnode + 1, sizeof(double));
    is->gz = xcalloc(is->nnode + 1, sizeof(double));
  }
  free(zero); free(bot);
  return M;
}

/* ------------------------------------------------------------ FMM state */

typedef struct {
  Model *M;
  Region *R;
  double *vel;
  int *tree;            /* node_from_tree_ind, 1-based */
  int ntree;
  int *counted;         /* node_is_counted */
  char *suspect;
  int newalive;
  int connode[64], n_connode;
  int cc_ix[4], cc_iz[4], n_concell;
  int cc_nnode[4], cc_nbase[4], cc_regular[4];
  long n_det_negative;
} FMM;

static void heap_add(FMM *F, int nn) {
  Region *R = F->R;
  F->ntree++;
  R->status[nn] = F->ntree;
  F->tree[F->ntree] = nn;
  int child = F->ntree, parent = F->ntree / 2;
  while (parent > 0) {
    if (R->T[nn] < R->T[F->tree[parent]]) {
      R->status[nn] = parent;
      R->status[F->tree[parent]] = child;
      int tmp = F->tree[child]; F->tree[child] = F->tree[parent]; F->tree[parent] = tmp;
      child = parent; parent = child / 2;
    } else parent = 0;
  }
}

static void heap_remove_root(FMM *F) {
  Region *R = F->R;
  if (F->ntree == 1) { F->ntree = 0; return; }
  R->status[F->tree[F->ntree]] = 1;
  F->tree[1] = F->tree[F->ntree];
  F->ntree--;
  int parent = 1, child = 2;
  while (child < F->ntree) {
    if (R->T[F->tree[child]] > R->T[F->tree[child + 1]]) child++;
    if (R->T[F->tree[child]] < R->T[F->tree[parent]]) {
      R->status[F->tree[parent]] = child;
      R->status[F->tree[child]] = parent;
      int tmp = F->tree[child]; F->tree[child] = F->tree[parent]; F->tree[parent] = tmp;
      parent = child; child = 2 * parent;
    } else child = F->ntree + 1;
  }
  if (child == F->ntree) {
    if (R->T[F->tree[child]] < R->T[F->tree[parent]]) {
      R->status[F->tree[parent]] = child;
      R->status[F->tree[child]] = parent;
      int tmp = F->tree[child]; F->tree[child] = F->tree[parent]; F->tree[parent] = tmp;
    }
  }
}

static void heap_update(FMM *F, int nn) {
  Region *R = F->R;
  int child = R->status[nn], parent = child / 2;
  while (parent > 0) {
    if (R->T[nn] < R->T[F->tree[parent]]) {
      R->status[nn] = parent;
      R->status[F->tree[parent]] = child;
      int tmp = F->tree[child]; F->tree[child] = F->tree[parent]; F->tree[parent] = tmp;
      child = parent; parent = child / 2;
    } else parent = 0;
  }
}

static Isec *node_isec(FMM *F, int node) {
  Model *M = F->M;
  Region *R = F->R;
  return R->i2[node] == M->isec[R->itop].id ? &M->isec[R->itop] : &M->isec[R->ibot];
}

static void gather_concells(FMM *F, int node) {
  Model *M = F->M;
  Region *R = F->R;
  int i1 = R->i1[node], i2 = R->i2[node], i3 = R->i3[node];
  F->n_concell = 0;
  if (i1 != 0) {
    for (int i = 0; i <= 1; i++)
      if (i1 - i > 0 && i1 - i < M->nx)
        for (int j = 0; j <= 1; j++)
          if (i2 - j > 0 && i2 - j < M->nz) {
            F->cc_ix[F->n_concell] = i1 - i; F->cc_iz[F->n_concell] = i2 - j; F->n_concell++;
          }
  } else {
    Isec *is = node_isec(F, node);
    for (int n = 0; n < 4; n++) {
      int ic = is->ccell_from_inode[(i3 - 1) * 4 + n];
      if (ic > 0) { F->cc_ix[F->n_concell] = is->cc_ix[ic]; F->cc_iz[F->n_concell] = is->cc_iz[ic]; F->n_concell++; }
    }
  }
}

static int cell_assoc(Model *M, int ix, int iz) {
  if (ix < 1 || ix > M->nx || iz < 1 || iz > M->nz) return 0;
  return G2(M->ccassoc, ix, iz);
}
static int cell_ccind(Model *M, int ix, int iz, int f) { return CCIND(ix, iz, f); }

static void add_counted(FMM *F, int m, int center) {
  if (F->counted[m] != center) { F->connode[F->n_connode++] = m; F->counted[m] = center; }
}

static void find_connected_nodes(FMM *F, int center) {
  Model *M = F->M;
  Region *R = F->R;
  int i1 = R->i1[center], i2 = R->i2[center];
  if (i1 != 0 && G2(M->fully_regular, i1, i2)) {
    F->n_connode = 0;
    int m;
    if ((m = G2(R->rnode, i1 - 1, i2)) > 0) F->connode[F->n_connode++] = m;
    if ((m = G2(R->rnode, i1 + 1, i2)) > 0) F->connode[F->n_connode++] = m;
    if ((m = G2(R->rnode, i1, i2 - 1)) > 0) F->connode[F->n_connode++] = m;
    if ((m = G2(R->rnode, i1, i2 + 1)) > 0) F->connode[F->n_connode++] = m;
    return;
  }
  gather_concells(F, center);
  F->counted[center] = center;
  F->n_connode = 0;
  Isec *top = &M->isec[R->itop], *bot = &M->isec[R->ibot];
  for (int n = 0; n < F->n_concell; n++) {
    int cx = F->cc_ix[n], cz = F->cc_iz[n];
    int regular = (i1 != 0);
    if (cell_assoc(M, cx, cz)) {
      int ic = cell_ccind(M, cx, cz, top->id);
      if (ic != 0) {
        regular = 0;
        for (int jj = 0; jj < top->n_inodes[ic]; jj++) add_counted(F, R->itop_rnode[top->inodes[(ic - 1) * 4 + jj]], center);
      }
      ic = cell_ccind(M, cx, cz, bot->id);
      if (ic != 0) {
        regular = 0;
        for (int jj = 0; jj < bot->n_inodes[ic]; jj++) add_counted(F, R->ibot_rnode[bot->inodes[(ic - 1) * 4 + jj]], center);
      }
    }
    for (int i = 0; i <= 1; i++) {
      int ii = cx + i;
      for (int j = 0; j <= 1; j++) {
        int jj = cz + j;
        if (ii >= 1 && ii <= M->nx && jj >= 1 && jj <= M->nz) {
          if ((i1 != 0 && abs(i1 - ii) + abs(i2 - jj) == 1) || !regular) {
            int m = G2(R->rnode, ii, jj);
            if (m > 0) add_counted(F, m, center);
          }
        }
      }
    }
  }
  if (i1 == 0) {
    int ix = (int)lround((R->x[center] - M->x0) / M->dx) + 1;
    if (ix < 1) ix = 1;
    if (ix > M->nx) ix = M->nx;
    int izs = (int)lround((R->z[center] - M->z0) / M->dz) + 1;
    if (izs < 1) izs = 1;
    if (izs > M->nz) izs = M->nz;
    int lo = ix - 1 < 1 ? 1 : ix - 1, hi = ix + 1 > M->nx ? M->nx : ix + 1;
    if (i2 == M->isec[R->ibot].id) {
      for (int ixx = lo; ixx <= hi; ixx++)
        for (int izz = izs; izz >= 1; izz--) {
          int m = G2(R->rnode, ixx, izz);
          if (m > 0) { add_counted(F, m, center); break; }
        }
    } else {
      for (int ixx = lo; ixx <= hi; ixx++)
        for (int izz = izs; izz <= M->nz; izz++) {
          int m = G2(R->rnode, ixx, izz);
          if (m > 0) { add_counted(F, m, center); break; }
        }
    }
  }
}

static void find_connected_cells(FMM *F, int center) {
  Model *M = F->M;
  Region *R = F->R;
  int i1 = R->i1[center], i2 = R->i2[center];
  gather_concells(F, center);
  F->n_connode = 0;
  Isec *top = &M->isec[R->itop], *bot = &M->isec[R->ibot];
  for (int n = 0; n < F->n_concell; n++) {
    int cx = F->cc_ix[n], cz = F->cc_iz[n];
    F->cc_regular[n] = (i1 != 0);
    if (cell_assoc(M, cx, cz)) {
      int ic = cell_ccind(M, cx, cz, top->id);
      if (ic != 0) {
        F->cc_regular[n] = 0;
        for (int jj = 0; jj < top->n_inodes[ic]; jj++) {
          int m = R->itop_rnode[top->inodes[(ic - 1) * 4 + jj]];
          if (R->status[m] == 0) F->connode[F->n_connode++] = m;
        }
      }
      ic = cell_ccind(M, cx, cz, bot->id);
      if (ic != 0) {
        F->cc_regular[n] = 0;
        for (int jj = 0; jj < bot->n_inodes[ic]; jj++) {
          int m = R->ibot_rnode[bot->inodes[(ic - 1) * 4 + jj]];
          if (R->status[m] == 0) F->connode[F->n_connode++] = m;
        }
      }
    }
    for (int i = 0; i <= 1; i++) {
      int ii = cx + i;
      for (int j = 0; j <= 1; j++) {
        int jj = cz + j;
        if (ii < 1 || ii > M->nx || jj < 1 || jj > M->nz) continue;
        if ((i1 != 0 && abs(i1 - ii) + abs(i2 - jj) == 1) || !F->cc_regular[n]) {
          int m = G2(R->rnode, ii, jj);
          if (m > 0 && m <= R->nnode && R->status[m] == 0) F->connode[F->n_connode++] = m;
        }
      }
    }
    F->cc_nnode[n] = F->n_connode;
  }
  F->cc_nbase[0] = 1;
  for (int n = 0; n < F->n_concell; n++) {
    F->cc_nnode[n] = F->cc_nnode[n] - F->cc_nbase[n] + 1;
    if (n < F->n_concell - 1) F->cc_nbase[n + 1] = F->cc_nbase[n] + F->cc_nnode[n];
  }
}

static void regular_update(FMM *F, int unode, const int *anode, int na) {
  Model *M = F->M;
  Region *R = F->R;
  const double *vel = F->vel;
  if (na > 2) { fprintf(stderr, "too many nodes for regular update2d\n"); abort(); }
  int iu = R->i1[unode], ju = R->i2[unode];
  double a = 0, b = 0, c = 0, t2x = 0, t2z = 0, tx = 0, tz = 0, tmax = 0;
  int x_info = 0, z_info = 0, ioff = 0, joff = 0;
  for (int n = 0; n < na; n++) {
    int an = anode[n];
    if (R->T[an] > tmax) tmax = R->T[an];
    int ia = R->i1[an], ja = R->i2[an];
    if (ia == 0) continue;
    if (ia != iu) {
      x_info = 1; ioff = ia;
      int di = ia - iu, ias = iu + 2 * di;
      double velx = 0.5 * (vel[unode] + vel[an]);
      if (ias >= 1 && ias <= M->nx) {
        int sec = G2(R->rnode, ias, ja);
        if (sec > 0 && sec <= R->nnode) {
          if (R->status[sec] == 0 && R->T[an] >= R->T[sec]) {
            if (vel[an] * vel[sec] > 2.2250738585072014e-308 &&
                R->gx[an] * R->gx[sec] + R->gz[an] * R->gz[sec] >= 0.71 / (vel[an] * vel[sec])) {
              tx = R->T[an]; t2x = R->T[sec];
              double d2 = (velx / M->dx) * (velx / M->dx);
              a = a + 2.25 * d2;
              b = b + (-6.0 * tx + 1.5 * t2x) * d2;
              c = c + (4.0 * tx * tx - 2.0 * tx * t2x + 0.25 * t2x * t2x) * d2;
              continue;
            }
          }
        }
      }
      tx = R->T[an];
      double d2 = (velx / M->dx) * (velx / M->dx);
      a = a + d2;
      b = b + (-2.0 * tx) * d2;
      c = c + (tx * tx) * d2;
    } else if (ja != ju) {
      z_info = 1; joff = ja;
      int dj = ja - ju, jas = ju + 2 * dj;
      double velz = 0.5 * (vel[unode] + vel[an]);
      if (jas >= 1 && jas <= M->nz) {
        int sec = G2(R->rnode, ia, jas);
        if (sec > 0 && sec <= R->nnode) {
          if (R->status[sec] == 0 && R->T[an] >= R->T[sec]) {
            if (vel[an] * vel[sec] > 2.2250738585072014e-308 &&
                R->gx[an] * R->gx[sec] + R->gz[an] * R->gz[sec] >= 0.71 / (vel[an] * vel[sec])) {
              tz = R->T[an]; t2z = R->T[sec];
              double d2 = (velz / M->dz) * (velz / M->dz);
              a = a + 2.25 * d2;
              b = b + (-6.0 * tz + 1.5 * t2z) * d2;
              c = c + (4.0 * tz * tz - 2.0 * tz * t2z + 0.25 * t2z * t2z) * d2;
              continue;
            }
          }
        }
      }
      tz = R->T[an];
      double d2 = (velz / M->dz) * (velz / M->dz);
      a = a + d2;
      b = b + (-2.0 * tz) * d2;
      c = c + (tz * tz) * d2;
    }
  }
  c = c - 1.0;
  if (fabs(a) < 2.2250738585072014e-308) return;
  double det = b * b - 4.0 * a * c;
  if (det < 0.0) {
    if (0.5 * sqrt(fabs(det)) / fabs(a) < 1.0e20) { det = 0.0; F->n_det_negative++; }
    else { fprintf(stderr, "negative determinant too large\n"); abort(); }
  }
  double t1 = 0.5 * (-b + sqrt(det)) / a;
  if (tmax > t1) t1 = tmax;
  if (t1 < R->T[unode]) {
    R->T[unode] = t1;
    R->gx[unode] = 0.0; R->gz[unode] = 0.0;
    if (x_info) {
      if (t2x == 0.0) R->gx[unode] = (t1 - tx) / (M->x[iu] - M->x[ioff]);
      else R->gx[unode] = 0.5 * (3.0 * t1 - 4.0 * tx + t2x) / (M->x[iu] - M->x[ioff]);
    }
    if (z_info) {
      if (t2z == 0.0) R->gz[unode] = (t1 - tz) / (M->z[ju] - M->z[joff]);
      else R->gz[unode] = 0.5 * (3.0 * t1 - 4.0 * tz + t2z) / (M->z[ju] - M->z[joff]);
    }
    F->suspect[unode] = !(x_info && z_info);
  }
}

static void time_from_1(FMM *F, int u, int n1) {
  Region *R = F->R;
  double dx = R->x[u] - R->x[n1], dz = R->z[u] - R->z[n1];
  double dist = sqrt(dx * dx + dz * dz);
  if (dist < 0.01 * F->M->tol) {
    R->T[u] = R->T[n1]; R->gx[u] = R->gx[n1]; R->gz[u] = R->gz[n1];
    return;
  }
  double tt = dist / (0.5 * (F->vel[u] + F->vel[n1]));
  double at = R->T[n1] + tt;
  if (at < R->T[u]) {
    R->T[u] = at;
    R->gx[u] = dx / (dist * F->vel[u]);
    R->gz[u] = dz / (dist * F->vel[u]);
  }
}

static void time_from_2(FMM *F, int u, int a1, int a2) {
  Region *R = F->R;
  const double *vel = F->vel;
  double tol = 0.01 * F->M->tol;
  int n1 = a1, n2 = a2;
  if (R->T[n1] < R->T[n2]) { int t = n1; n1 = n2; n2 = t; }
  double xu0 = R->x[u], xu1 = R->z[u];
  double x10 = R->x[n1], x11 = R->z[n1], x20 = R->x[n2], x21 = R->z[n2];
  double r10 = x10 - xu0, r11 = x11 - xu1, r20 = x20 - xu0, r21 = x21 - xu1;
  if (sqrt(r10 * r10 + r11 * r11) < tol) { R->T[u] = R->T[n1]; R->gx[u] = R->gx[n1]; R->gz[u] = R->gz[n1]; return; }
  if (sqrt(r20 * r20 + r21 * r21) < tol) { R->T[u] = R->T[n2]; R->gx[u] = R->gx[n2]; R->gz[u] = R->gz[n2]; return; }
  double e0 = x20 - x10, e1 = x21 - x11;
  if (sqrt(e0 * e0 + e1 * e1) < tol) { time_from_1(F, u, n1); return; }
  double v12 = 0.5 * (vel[n1] + vel[n2]);
  double d0 = (x10 - x20) / v12, d1 = (x11 - x21) / v12;
  double sdd = d0 * d0 + d1 * d1;
  double p0 = (R->T[n1] - R->T[n2]) / sdd * d0, p1 = (R->T[n1] - R->T[n2]) / sdd * d1;
  double q0 = -d1, q1 = d0;
  double ae = q0 * q0 + q1 * q1;
  double be = 2.0 * (p0 * q0 + p1 * q1);
  double ce = (p0 * p0 + p1 * p1) - 1.0;
  double det = be * be - 4.0 * ae * ce;
  if (det < 0.0) return;
  double u1 = 0.5 * (-be + sqrt(det)) / ae, u2 = 0.5 * (-be - sqrt(det)) / ae;
  double w10 = p0 + u1 * q0, w11 = p1 + u1 * q1, w20 = p0 + u2 * q0, w21 = p1 + u2 * q1;
  double n10 = -r11, n11 = r10, l1 = sqrt(n10 * n10 + n11 * n11);
  n10 = n10 / l1; n11 = n11 / l1;
  if (n10 * r20 + n11 * r21 < 0.0) { n10 = -n10; n11 = -n11; }
  double n20 = -r21, n21 = r20, l2 = sqrt(n20 * n20 + n21 * n21);
  n20 = n20 / l2; n21 = n21 / l2;
  if (n20 * r10 + n21 * r11 < 0.0) { n20 = -n20; n21 = -n21; }
  if (w10 * n10 + w11 * n11 <= 0.0 && w10 * n20 + w11 * n21 <= 0.0) {
    double at = R->T[n1] - (w10 * r10 + w11 * r11) / (0.5 * (vel[u] + vel[n1]));
    if (at >= R->T[n1] && at < R->T[u]) {
      R->T[u] = at;
      double wl = sqrt(w10 * w10 + w11 * w11);
      R->gx[u] = w10 / (wl * vel[n1]); R->gz[u] = w11 / (wl * vel[n1]);
    }
  }
  if (w20 * n10 + w21 * n11 <= 0.0 && w20 * n20 + w21 * n21 <= 0.0) {
    double at = R->T[n1] - (w20 * r10 + w21 * r11) / (0.5 * (vel[u] + vel[n1]));
    if (at >= R->T[n1] && at < R->T[u]) {
      R->T[u] = at;
      double wl = sqrt(w20 * w20 + w21 * w21);
      R->gx[u] = w20 / (wl * vel[n1]); R->gz[u] = w21 / (wl * vel[n1]);
    }
  }
}

static int trace_node = -2;
static void update_time_impl(FMM *F, int u);
static void update_time(FMM *F, int u) {
  /* ORC_TRACE=<region node id> (diagnostics): report every change of that node */
  if (trace_node == -2) { const char *e = getenv("ORC_TRACE"); trace_node = e ? atoi(e) : -1; }
  double old = F->R->T[u];
  update_time_impl(F, u);
  if (u == trace_node && F->R->T[u] != old)
    fprintf(stderr, "trace node %d: newalive %d (i1=%d i2=%d i3=%d T=%.17g) T %.17g -> %.17g g=(%.17g,%.17g)\n", u,
            F->newalive, F->R->i1[F->newalive], F->R->i2[F->newalive], F->R->i3[F->newalive], F->R->T[F->newalive],
            old, F->R->T[u], F->R->gx[u], F->R->gz[u]);
}
static void update_time_impl(FMM *F, int u) {
  Model *M = F->M;
  Region *R = F->R;
  int i1 = R->i1[u], i2 = R->i2[u];
  int anode[32], na;
  if (i1 != 0 && G2(M->fully_regular, i1, i2)) {
    for (int j = -1; j <= 1; j += 2)
      for (int i = -1; i <= 1; i += 2) {
        na = 0;
        int id = G2(R->rnode, i1 + i, i2);
        if (id > 0 && id <= R->nnode && R->status[id] == 0) anode[na++] = id;
        id = G2(R->rnode, i1, i2 + j);
        if (id > 0 && id <= R->nnode && R->status[id] == 0) anode[na++] = id;
        if (na > 0) {
          for (int n = 0; n < na; n++)
            if (anode[n] == F->newalive) { anode[n] = anode[0]; anode[0] = F->newalive; }
          if (anode[0] == F->newalive) regular_update(F, u, anode, na);
        }
      }
    return;
  }
  find_connected_cells(F, u);
  for (int n = 0; n < F->n_concell; n++) {
    if (F->cc_nnode[n] <= 0) continue;
    na = F->cc_nnode[n];
    if (na > 30) { fprintf(stderr, "nanode > 30\n"); abort(); }
    for (int k = 0; k < na; k++) anode[k] = F->connode[F->cc_nbase[n] - 1 + k];
    for (int i = 0; i < na; i++)
      if (anode[i] == F->newalive) { int t = anode[0]; anode[0] = anode[i]; anode[i] = t; }
    if (anode[0] != F->newalive) continue;
    if (F->cc_regular[n]) regular_update(F, u, anode, na);
    else {
      time_from_1(F, u, anode[0]);
      for (int ii = 1; ii < na; ii++) time_from_2(F, u, anode[0], anode[ii]);
    }
  }
}

void orc_fit_gradient(Model *M, Region *R, int center, double *dtdx, double *dtdz) {
  enum { MAXFIT = 30 };
  int i1 = R->i1[center], i2 = R->i2[center], i3 = R->i3[center];
  int ncx[4], ncz[4], nc = 0, irr[4];
  if (i1 != 0) {
    for (int i = 0; i <= 1; i++)
      if (i1 - i > 0 && i1 - i < M->nx)
        for (int j = 0; j <= 1; j++)
          if (i2 - j > 0 && i2 - j < M->nz) { ncx[nc] = i1 - i; ncz[nc] = i2 - j; nc++; }
  } else {
    Isec *is = (i2 == M->isec[R->itop].id) ? &M->isec[R->itop] : &M->isec[R->ibot];
    for (int n = 0; n < 4; n++) {
      int ic = is->ccell_from_inode[(i3 - 1) * 4 + n];
      if (ic > 0) { ncx[nc] = is->cc_ix[ic]; ncz[nc] = is->cc_iz[ic]; nc++; }
    }
  }
  int nirr = 0;
  for (int n = 0; n < nc; n++) { irr[n] = cell_assoc(M, ncx[n], ncz[n]); nirr += irr[n]; }
  if (nirr == 0) {
    if (i1 != 0) {
      int m = G2(R->rnode, i1 - 1 < 1 ? 1 : i1 - 1, i2), n = G2(R->rnode, i1 + 1 > M->nx ? M->nx : i1 + 1, i2);
      if (m > 0 && m <= R->nnode && n > 0 && n <= R->nnode) *dtdx = (R->T[n] - R->T[m]) / (R->x[n] - R->x[m]);
      else if (m > 0 && m <= R->nnode) *dtdx = (R->T[center] - R->T[m]) / (R->x[center] - R->x[m]);
      else if (n > 0 && n <= R->nnode) *dtdx = (R->T[n] - R->T[center]) / (R->x[n] - R->x[center]);
      else *dtdx = 0.0;
      m = G2(R->rnode, i1, i2 - 1 < 1 ? 1 : i2 - 1); n = G2(R->rnode, i1, i2 + 1 > M->nz ? M->nz : i2 + 1);
      if (m > 0 && m <= R->nnode && n > 0 && n <= R->nnode) *dtdz = (R->T[n] - R->T[m]) / (R->z[n] - R->z[m]);
      else if (m > 0 && m <= R->nnode) *dtdz = (R->T[center] - R->T[m]) / (R->z[center] - R->z[m]);
      else if (n > 0 && n <= R->nnode) *dtdz = (R->T[n] - R->T[center]) / (R->z[n] - R->z[center]);
      else *dtdz = 0.0;
      return;
    }
    if (nc == 0) { *dtdx = 0.0; *dtdz = 0.0; return; }
  }
  int mstore[MAXFIT], ncn = 0;
  double rx[MAXFIT], rz[MAXFIT], at[MAXFIT];
#define ADDFIT(mm)                                                              \
  do {                                                                          \
    int _m = (mm), _dup = 0;                                                    \
    for (int _k = 0; _k < ncn; _k++) if (mstore[_k] == _m) _dup = 1;            \
    if (!_dup && ncn < MAXFIT) {                                                \
      mstore[ncn] = _m; rx[ncn] = R->x[_m]; rz[ncn] = R->z[_m]; at[ncn] = R->T[_m]; ncn++; \
    }                                                                           \
  } while (0)
  Isec *top = &M->isec[R->itop], *bot = &M->isec[R->ibot];
  for (int n = 0; n < nc; n++) {
    if (irr[n]) {
      int ic = CCIND(ncx[n], ncz[n], top->id);
      if (ic != 0)
        for (int jj = 0; jj < top->n_inodes[ic]; jj++) {
          int m = R->itop_rnode[top->inodes[(ic - 1) * 4 + jj]];
          if (m == center) continue;
          ADDFIT(m);
        }
      ic = CCIND(ncx[n], ncz[n], bot->id);
      if (ic != 0)
        for (int jj = 0; jj < bot->n_inodes[ic]; jj++) {
          int m = R->ibot_rnode[bot->inodes[(ic - 1) * 4 + jj]];
          if (m == center) continue;
          ADDFIT(m);
        }
    }
    for (int i = 0; i <= 1; i++) {
      int ii = ncx[n] + i;
      if (ii < 1 || ii > M->nx) continue;
      for (int j = 0; j <= 1; j++) {
        int jj = ncz[n] + j;
        if (jj < 1 || jj > M->nz) continue;
        int m = G2(R->rnode, ii, jj);
        if (m <= 0 || m > R->nnode || m == center) continue;
        ADDFIT(m);
      }
    }
  }
#undef ADDFIT
  double a11 = 0, a12 = 0, a22 = 0, b1 = 0, b2 = 0;
  for (int i = 0; i < ncn; i++) {
    double dx = rx[i] - R->x[center], dz = rz[i] - R->z[center];
    double dw = 1.0 / (sqrt(dx * dx + dz * dz) + M->tol);
    a11 = a11 + dx * dx * dw * dw;
    a12 = a12 + dx * dz * dw * dw;
    a22 = a22 + dz * dz * dw * dw;
    b1 = b1 + (at[i] - R->T[center]) * dx * dw * dw;
    b2 = b2 + (at[i] - R->T[center]) * dz * dw * dw;
  }
  double a21 = a12;
  double det = a11 * a22 - a12 * a21;
  if (fabs(det) < 1.0e-30) { *dtdx = 0.0; *dtdz = 0.0; return; }
  *dtdx = (a22 * b1 - a12 * b2) / det;
  *dtdz = (a11 * b2 - a21 * b1) / det;
}

/* propagate2d (propagate2d.f90:33-127) on region R from its current status/T. */
void orc_propagate(Model *M, Region *R) {
  FMM F;
  memset(&F, 0, sizeof(F));
  F.M = M; F.R = R; F.vel = R->vel;
  F.tree = xcalloc(R->nnode + 2, sizeof(int));
  F.counted = xcalloc(R->nnode + 2, sizeof(int));
  F.suspect = xcalloc(R->nnode + 2, 1);
  for (int n = 1; n <= R->nnode; n++)
    if (R->status[n] > 0) heap_add(&F, n);
  int first = 1;
  while (F.ntree > 0) {
    F.newalive = F.tree[1];
    R->status[F.newalive] = 0;
    heap_remove_root(&F);
    if (F.ntree == 0 && !first) {
      if (getenv("ORC_DEBUG_EXIT"))
        fprintf(stderr, "early exit: region %d node %d (i1=%d i2=%d i3=%d) T=%.9f\n", R->id, F.newalive,
                R->i1[F.newalive], R->i2[F.newalive], R->i3[F.newalive], R->T[F.newalive]);
      break;
    }
    first = 0;
    find_connected_nodes(&F, F.newalive);
    int ntrial = F.n_connode, trial[64];
    memcpy(trial, F.connode, ntrial * sizeof(int));
    for (int nt = 0; nt < ntrial; nt++) {
      int u = trial[nt];
      if (R->status[u] > 0) {
        double old = R->T[u];
        update_time(&F, u);
        if (R->T[u] < old) heap_update(&F, u);
      }
      if (R->status[u] == -1) {
        R->T[u] = HUGE_TIME;
        double old = R->T[u];
        update_time(&F, u);
        if (R->T[u] < old) heap_add(&F, u);
      }
    }
  }
  /* ORC_NO_REFIT (diagnostics): keep the raw marching gradients */
  if (!getenv("ORC_NO_REFIT"))
    for (int n = 1; n <= R->nnode; n++)
      if (F.suspect[n]) orc_fit_gradient(M, R, n, &R->gx[n], &R->gz[n]);
  free(F.tree); free(F.counted); free(F.suspect);
}

/* ------------------------------------------------------------ seeding */

static void reset_region(Region *R) {
  for (int n = 1; n <= R->nnode; n++) { R->status[n] = -1; R->T[n] = HUGE_TIME; R->gx[n] = 0.0; R->gz[n] = 0.0; }
}

void orc_init_from_source(Model *M, Region *R, double sx, double sz, double t0) {
  reset_region(R);
  int ix0 = (int)floor((sx - M->x0) / M->dx) + 1;
  int iz0 = (int)floor((sz - M->z0) / M->dz) + 1;
  if (ix0 < 1) ix0 = 1;
  if (ix0 >= M->nx) ix0 = M->nx - 1;
  if (iz0 < 1) iz0 = 1;
  if (iz0 >= M->nz) iz0 = M->nz - 1;
  for (int ix = ix0; ix <= ix0 + 1; ix++)
    for (int iz = iz0; iz <= iz0 + 1; iz++) {
      if (ix < 1 || ix > M->nx || iz < 1 || iz > M->nz) continue;
      int m = G2(R->rnode, ix, iz);
      if (m <= 0 || m > R->nnode) continue;
      double ddx = M->x[ix] - sx, ddz = M->z[iz] - sz;
      double dist = sqrt(ddx * ddx + ddz * ddz);
      double va = R->vel[m];
      if (va < 0.01) va = 0.01;
      if (dist < 0.01 * M->tol) { R->T[m] = t0; R->status[m] = 0; R->gx[m] = 0.0; R->gz[m] = 0.0; }
      else {
        R->T[m] = t0 + dist / va; R->status[m] = 1;
        R->gx[m] = (M->x[ix] - sx) / (dist * va);
        R->gz[m] = (M->z[iz] - sz) / (dist * va);
      }
    }
  for (int n = 1; n <= R->nnode; n++) {
    if (R->status[n] != 0) continue;
    int i1 = R->i1[n], i2 = R->i2[n];
    if (i1 == 0) continue;
    for (int dx = -1; dx <= 1; dx += 2)
      for (int dz = -1; dz <= 1; dz += 2) {
        int a = i1 + dx, b = i2 + dz;
        if (a < 1 || a > M->nx || b < 1 || b > M->nz) continue;
        int m = G2(R->rnode, a, b);
        if (m > 0 && m <= R->nnode && R->status[m] == -1) R->status[m] = 1;
      }
  }
}

/* transfer_times_to_intersection (propagate2d.f90:2414-2554); the source
 * field's T/grad are the arrays passed in (tf%arrivaltime/time_gradient). */
static void transfer_times(Model *M, Isec *is, Region *Rs, const double *T, const double *gx, const double *gz) {
  for (int n = 1; n <= is->nnode; n++) { is->atime[n] = HUGE_TIME; is->gx[n] = 0.0; is->gz[n] = 0.0; }
  const int *map = NULL;
  if (is->id == M->isec[Rs->ibot].id) map = Rs->ibot_rnode;
  else if (is->id == M->isec[Rs->itop].id) map = Rs->itop_rnode;
  if (map) {
    for (int n = 1; n <= is->nnode; n++) {
      int m = map[n];
      if (m > 0 && m <= Rs->nnode) { is->atime[n] = T[m]; is->gx[n] = gx[m]; is->gz[n] = gz[m]; }
    }
    return;
  }
  for (int n = 1; n <= is->nnode; n++) {
    double px = is->x[n], pz = is->z[n];
    int ix = (int)floor((px - M->x0) / M->dx) + 1, iz = (int)floor((pz - M->z0) / M->dz) + 1;
    if (ix < 1) ix = 1;
    if (ix >= M->nx) ix = M->nx - 1;
    if (iz < 1) iz = 1;
    if (iz >= M->nz) iz = M->nz - 1;
    double w1 = (px - M->x[ix]) / M->dx, w2 = (pz - M->z[iz]) / M->dz;
    if (w1 < 0.0) w1 = 0.0;
    if (w1 > 1.0) w1 = 1.0;
    if (w2 < 0.0) w2 = 0.0;
    if (w2 > 1.0) w2 = 1.0;
    int n1 = G2(Rs->rnode, ix, iz), n2 = G2(Rs->rnode, ix + 1, iz), n3 = G2(Rs->rnode, ix, iz + 1), n4 = G2(Rs->rnode, ix + 1, iz + 1);
    if (n1 > 0 && n2 > 0 && n3 > 0 && n4 > 0) {
      /* region_rnode fields are read from the region's loaded arrays */
      is->atime[n] = (1.0 - w1) * (1.0 - w2) * Rs->T[n1] + w1 * (1.0 - w2) * Rs->T[n2] + (1.0 - w1) * w2 * Rs->T[n3] + w1 * w2 * Rs->T[n4];
      is->gx[n] = (1.0 - w1) * (1.0 - w2) * Rs->gx[n1] + w1 * (1.0 - w2) * Rs->gx[n2] + (1.0 - w1) * w2 * Rs->gx[n3] + w1 * w2 * Rs->gx[n4];
      is->gz[n] = (1.0 - w1) * (1.0 - w2) * Rs->gz[n1] + w1 * (1.0 - w2) * Rs->gz[n2] + (1.0 - w1) * w2 * Rs->gz[n3] + w1 * w2 * Rs->gz[n4];
    } else {
      int nn = 0;
      double dmin = 1.0e100;
      for (int i = 1; i <= Rs->nnode; i++) {
        double d = (px - Rs->x[i]) * (px - Rs->x[i]) + (pz - Rs->z[i]) * (pz - Rs->z[i]);
        if (d < dmin) { dmin = d; nn = i; }
      }
      if (nn > 0) { is->atime[n] = Rs->T[nn]; is->gx[n] = Rs->gx[nn]; is->gz[n] = Rs->gz[nn]; }
    }
  }
}

/* reflect_gradient_at_interface (propagate2d.f90:2632-2741) */
static void reflect_gradient(Model *M, Isec *is, Region *Rt) {
  for (int n = 1; n <= is->nnode; n++) {
    double n0 = is->nx_[n], n1 = is->nz_[n];
    if (is->atime[n] >= HUGE_TIME * 0.5) { is->gx[n] = 0.0; is->gz[n] = 0.0; continue; }
    double gp = is->gx[n] * n0 + is->gz[n] * n1;
    double p0 = is->gx[n] - gp * n0, p1 = is->gz[n] - gp * n1;
    int side;
    if (M->isec[Rt->itop].id == is->id) side = 1;
    else if (M->isec[Rt->ibot].id == is->id) side = -1;
    else side = gp < 0.0 ? -1 : 1;
    double vt = side > 0 ? is->vel_bot[n] : is->vel_top[n];
    if (!(vt > 0.01)) {
      int m = 0;
      if (M->isec[Rt->itop].id == is->id) m = Rt->itop_rnode[n];
      if (m == 0 && M->isec[Rt->ibot].id == is->id) m = Rt->ibot_rnode[n];
      if (m > 0 && m <= Rt->nnode) vt = Rt->vel[m];
    }
    if (vt <= 0.01) { is->atime[n] = HUGE_TIME; is->gx[n] = 0.0; is->gz[n] = 0.0; continue; }
    double pt = 1.0 / vt, pp2 = p0 * p0 + p1 * p1, pn2 = pt * pt - pp2;
    if (pn2 < -1.0e-10) { is->atime[n] = HUGE_TIME; is->gx[n] = 0.0; is->gz[n] = 0.0; }
    else {
      if (pn2 < 0.0) pn2 = 0.0;
      is->gx[n] = p0 + side * sqrt(pn2) * n0;
      is->gz[n] = p1 + side * sqrt(pn2) * n1;
    }
  }
}

/* sweep_region_from_interface2d (propagate2d.f90:1982-2304) */
static void sweep_from_interface(Model *M, Region *R, Isec *is) {
  reset_region(R);
  int nn = is->nnode;
  int *map = xcalloc(nn + 1, sizeof(int));
  int mapped = 0;
  if (is->id == M->isec[R->itop].id) { memcpy(map, R->itop_rnode, (nn + 1) * sizeof(int)); mapped = 1; }
  else if (is->id == M->isec[R->ibot].id) { memcpy(map, R->ibot_rnode, (nn + 1) * sizeof(int)); mapped = 1; }
  else if (is->regright && is->rright_node_id) { if (is->regright == R->id) { memcpy(map, is->rright_node_id, (nn + 1) * sizeof(int)); mapped = 1; } }
  else if (is->regleft && is->rleft_node_id) { if (is->regleft == R->id) { memcpy(map, is->rleft_node_id, (nn + 1) * sizeof(int)); mapped = 1; } }
  if (!mapped) {
    for (int n = 1; n <= nn; n++) {
      double it = is->atime[n];
      if (it >= HUGE_TIME * 0.5) continue;
      double g0 = is->gx[n], g1 = is->gz[n];
      int side = (g0 * is->nx_[n] + g1 * is->nz_[n] < 0.0) ? -1 : 1;
      int ix = (int)lround((is->x[n] - M->x0) / M->dx) + 1;
      if (ix < 1) ix = 1;
      if (ix > M->nx) ix = M->nx;
      int izf = 0;
      double zi = is->z[n];
      for (int iz = 1; iz <= M->nz - 1; iz++) if (M->z[iz] <= zi && zi < M->z[iz + 1]) { izf = iz; break; }
      if (izf == 0) { if (zi <= M->z[1]) izf = 1; if (zi >= M->z[M->nz]) izf = M->nz - 1; }
      if (izf == 0) continue;
      int s0, s1, st;
      if (side > 0) { s0 = izf + 1; s1 = izf + 3 < M->nz ? izf + 3 : M->nz; st = 1; }
      else { s0 = izf; s1 = izf - 2 > 1 ? izf - 2 : 1; st = -1; }
      for (int iz = s0; st > 0 ? iz <= s1 : iz >= s1; iz += st)
        for (int ixx = (ix - 1 > 1 ? ix - 1 : 1); ixx <= (ix + 1 < M->nx ? ix + 1 : M->nx); ixx++) {
          int nt = G2(R->rnode, ixx, iz);
          if (nt <= 0 || nt > R->nnode) continue;
          double dv0 = R->x[nt] - is->x[n], dv1 = R->z[nt] - is->z[n];
          double dist = sqrt(dv0 * dv0 + dv1 * dv1);
          if (dist <= 1.0e-10) continue;
          double vl = R->vel[nt];
          if (vl < 0.01) vl = 0.01;
          double dt;
          if (sqrt(g0 * g0 + g1 * g1) > 1.0e-10) { dt = dv0 * g0 + dv1 * g1; if (dt <= 0.0) dt = dist / vl; }
          else { g0 = dv0 / (dist * vl); g1 = dv1 / (dist * vl); dt = dist / vl; }
          if (R->status[nt] == -1 || it + dt < R->T[nt]) { R->status[nt] = 1; R->T[nt] = it + dt; R->gx[nt] = g0; R->gz[nt] = g1; }
        }
    }
    free(map);
    return;
  }
  for (int n = 1; n <= nn; n++) {
    int m = map[n];
    if (m > 0 && m <= R->nnode) { R->status[m] = 1; R->T[m] = is->atime[n]; R->gx[m] = is->gx[n]; R->gz[m] = is->gz[n]; }
  }
  int is_bottom = (is->id == M->isec[R->ibot].id);
  for (int n = 1; n <= nn; n++) {
    int m = map[n];
    if (m <= 0 || m > R->nnode) continue;
    if (R->status[m] != 1) continue;
    if (R->i1[m] != 0) continue;
    int ix = (int)lround((R->x[m] - M->x0) / M->dx) + 1;
    if (ix < 1) ix = 1;
    if (ix > M->nx) ix = M->nx;
    int izf = 0;
    for (int iz = 1; iz <= M->nz - 1; iz++) if (M->z[iz] <= R->z[m] && R->z[m] < M->z[iz + 1]) { izf = iz; break; }
    if (izf == 0) { if (R->z[m] <= M->z[1]) izf = 1; if (R->z[m] >= M->z[M->nz]) izf = M->nz - 1; }
    if (izf == 0) continue;
    int s0, s1, st;
    if (is_bottom) { s0 = izf; s1 = izf - 2 > 1 ? izf - 2 : 1; st = -1; }
    else { s0 = izf + 1; s1 = izf + 3 < M->nz ? izf + 3 : M->nz; st = 1; }
    for (int iz = s0; st > 0 ? iz <= s1 : iz >= s1; iz += st)
      for (int ixx = (ix - 1 > 1 ? ix - 1 : 1); ixx <= (ix + 1 < M->nx ? ix + 1 : M->nx); ixx++) {
        int nt = G2(R->rnode, ixx, iz);
        if (nt <= 0 || nt > R->nnode) continue;
        if (R->status[nt] != -1) continue;
        R->status[nt] = 1;
        double ddx = R->x[nt] - R->x[m], ddz = R->z[nt] - R->z[m];
        double dist = sqrt(ddx * ddx + ddz * ddz);
        double vl = R->vel[nt];
        if (vl < 0.01) vl = 0.01;
        R->T[nt] = R->T[m] + dist / vl;
        R->gx[nt] = (R->x[nt] - R->x[m]) / (dist * vl);
        R->gz[nt] = (R->z[nt] - R->z[m]) / (dist * vl);
      }
  }
  free(map);
}

/* ------------------------------------------------------------ stage schedule */

static const int PATH_NTF[5] = {1, 2, 2, 4, 4};
static const int PATH_TF[5][4] = {{1, 0, 0, 0}, {2, 3, 0, 0}, {4, 5, 0, 0}, {1, 6, 7, 8}, {9, 10, 11, 12}};
static const int TF_REGION[13] = {0, 5, 2, 2, 6, 6, 1, 1, 1, 1, 1, 5, 1};
static const int TF_ISTART[13] = {0, 2, 2, 3, 2, 4, 2, 1, 2, 2, 1, 2, 2};

int orc_tf_region(int itf) { return (itf >= 1 && itf <= 12) ? TF_REGION[itf] : 0; }
int orc_region_nnode(Model *M, int r) { return M->reg[r].nnode; }
int orc_grid_nx(Model *M) { return M->nx; }
int orc_grid_nz(Model *M) { return M->nz; }

double orc_iface_z_at(Model *M, int k, double px) {
  /* interpolate_iface_z on the interface's own nodes (fmt2d_main.f90:1389-1415) */
  int n = M->if_n[k];
  const double *x = M->if_x[k], *z = M->if_z[k];
  if (n < 2) return z[0];
  int i = (int)floor((px - x[0]) / M->if_dx[k]) + 1;
  if (i < 1) return z[0];
  if (i >= n) return z[n - 1];
  double w = (px - x[i - 1]) / M->if_dx[k];
  if (w < 0.0) w = 0.0;
  if (w > 1.0) w = 1.0;
  return (1.0 - w) * z[i - 1] + w * z[i];
}

static void water_top_fix(Model *M, Region *R, int up) {
  for (int ix = 1; ix <= M->nx; ix++) {
    int mt = G2(R->rnode, ix, 1), mb = G2(R->rnode, ix, 2);
    if (mt > 0 && mb > 0) {
      double g = R->gz[mt];
      int bad = up ? (g > 0.0 || fabs(g) < 0.01) : (g < 0.0 || fabs(g) < 0.01);
      if (bad) { R->gx[mt] = R->gx[mb]; R->gz[mt] = R->gz[mb]; }
    }
  }
}

/* PmP head-wave post-pass on tf5 (inversion2d.f90:4294-4409) */
static void pmp_headwave(Model *M, Region *R, Isec *is) {
  int nn = is->nnode;
  double *thw = xcalloc(nn + 2, sizeof(double)), *vhw = xcalloc(nn + 2, sizeof(double));
  for (int n = 1; n <= nn; n++) {
    int m = 0;
    if (is->rright_node_id) m = is->rright_node_id[n];
    if (m == 0 && is->rleft_node_id) m = is->rleft_node_id[n];
    thw[n] = (m > 0 && m <= R->nnode) ? R->T[m] : HUGE_TIME;
    vhw[n] = 0.0;
    if (m > 0 && m <= R->nnode) {
      if (R->i1[m] == 0) {
        int i3 = R->i3[m];
        for (int k = 0; k < 4; k++) {
          int ic = is->ccell_from_inode[(i3 - 1) * 4 + k];
          if (ic <= 0) continue;
          for (int ii = 0; ii <= 1; ii++)
            for (int jj = 0; jj <= 1; jj++) {
              int ixx = is->cc_ix[ic] + ii, izz = is->cc_iz[ic] + jj;
              if (ixx >= 1 && ixx <= M->nx && izz >= 1 && izz <= M->nz && G2(M->node_region, ixx, izz) == R->id) {
                int mm = G2(M->rnode_id, ixx, izz);
                if (mm > 0 && mm <= R->nnode && R->i1[mm] > 0 && R->z[mm] > is->z[n] + 0.5 * M->dz)
                  if (R->vel[mm] > vhw[n]) vhw[n] = R->vel[mm];
              }
            }
        }
      } else vhw[n] = R->vel[m];
    }
  }
  double dxif = is->x[2] - is->x[1];
  for (int n = 2; n <= nn; n++) { double v = vhw[n]; if (v > 0.0) { double c = thw[n - 1] + dxif / v; if (c < thw[n]) thw[n] = c; } }
  for (int n = nn - 1; n >= 1; n--) { double v = vhw[n]; if (v > 0.0) { double c = thw[n + 1] + dxif / v; if (c < thw[n]) thw[n] = c; } }
  for (int n = 1; n <= nn; n++) {
    int m = 0;
    if (is->rright_node_id) m = is->rright_node_id[n];
    if (m == 0 && is->rleft_node_id) m = is->rleft_node_id[n];
    if (m > 0 && m <= R->nnode && thw[n] < R->T[m] - 1.0e-10) R->T[m] = thw[n];
  }
  for (int n = 1; n <= nn; n++) {
    int m = 0;
    if (is->rright_node_id) m = is->rright_node_id[n];
    if (m == 0 && is->rleft_node_id) m = is->rleft_node_id[n];
    if (!(m > 0 && m <= R->nnode) || R->status[m] != 0) continue;
    if (R->i1[m] > 0) orc_fit_gradient(M, R, m, &R->gx[m], &R->gz[m]);
    else {
      double dtdx, dtdz;
      if (n > 1 && n < nn) dtdx = (thw[n + 1] - thw[n - 1]) / (2.0 * dxif);
      else if (n == 1 && nn > 1) dtdx = (thw[2] - thw[1]) / dxif;
      else if (n == nn && nn > 1) dtdx = (thw[nn] - thw[nn - 1]) / dxif;
      else dtdx = 0.0;
      double v = vhw[n];
      if (v > 0.0) { double s = 1.0 / (v * v) - dtdx * dtdx; dtdz = s > 0.0 ? sqrt(s) : 0.0; }
      else dtdz = 0.0;
      R->gx[m] = dtdx; R->gz[m] = dtdz;
    }
  }
  free(thw); free(vhw);
}

/* Runs the run_forward2d stage schedule for one source at x=sx (placed on the
 * seafloor).  out_T/out_gx/out_gz[itf-1] receive nnode values of tf itf's
 * region (NULL entries are skipped).  headwave != 0 applies the PmP
 * post-pass to tf5.  Stages with itf > nstage are skipped. */
void orc_forward_source(Model *M, double sx, int nstage, int headwave,
                        double **out_T, double **out_gx, double **out_gz) {
  double sz = orc_iface_z_at(M, 2, sx);
  double *tfT[13] = {0}, *tfx[13] = {0}, *tfz[13] = {0};
  for (int t = 1; t <= 12; t++) {
    int nn = M->reg[TF_REGION[t]].nnode + 1;
    tfT[t] = xcalloc(nn, sizeof(double)); tfx[t] = xcalloc(nn, sizeof(double)); tfz[t] = xcalloc(nn, sizeof(double));
  }
  for (int ip = 0; ip < 5; ip++)
    for (int k = 0; k < PATH_NTF[ip]; k++) {
      int itf = PATH_TF[ip][k];
      if (itf > nstage) continue;
      Region *R = &M->reg[TF_REGION[itf]];
      reset_region(R);
      if (k == 0) orc_init_from_source(M, R, sx, sz, 0.0);
      else {
        int ip_ = PATH_TF[ip][k - 1];
        Region *Rp = &M->reg[TF_REGION[ip_]];
        int nn = Rp->nnode + 1;
        memcpy(Rp->T, tfT[ip_], nn * sizeof(double));
        memcpy(Rp->gx, tfx[ip_], nn * sizeof(double));
        memcpy(Rp->gz, tfz[ip_], nn * sizeof(double));
        Isec *is = &M->isec[TF_ISTART[itf]];
        transfer_times(M, is, Rp, tfT[ip_], tfx[ip_], tfz[ip_]);
        int refl = (ip == 1 && k == 1) || (ip == 2 && k == 1) || (ip == 3 && k >= 1) || (ip == 4 && k >= 1);
        if (refl) reflect_gradient(M, is, R);
        sweep_from_interface(M, R, is);
      }
      orc_propagate(M, R);
      if (R->id == 1) {
        if ((ip == 3 && (k == 1 || k == 3)) || (ip == 4 && (k == 0 || k == 3))) water_top_fix(M, R, 1);
        else if ((ip == 3 && k == 2) || (ip == 4 && k == 1)) water_top_fix(M, R, 0);
      }
      if (headwave && ip == 2 && k == 1) pmp_headwave(M, R, &M->isec[TF_ISTART[itf]]);
      int nn = R->nnode + 1;
      memcpy(tfT[itf], R->T, nn * sizeof(double));
      memcpy(tfx[itf], R->gx, nn * sizeof(double));
      memcpy(tfz[itf], R->gz, nn * sizeof(double));
    }
  for (int t = 1; t <= 12; t++) {
    int nn = M->reg[TF_REGION[t]].nnode;
    if (out_T && out_T[t - 1]) memcpy(out_T[t - 1], tfT[t] + 1, nn * sizeof(double));
    if (out_gx && out_gx[t - 1]) memcpy(out_gx[t - 1], tfx[t] + 1, nn * sizeof(double));
    if (out_gz && out_gz[t - 1]) memcpy(out_gz[t - 1], tfz[t] + 1, nn * sizeof(double));
    free(tfT[t]); free(tfx[t]); free(tfz[t]);
  }
}

/* Geometry export for tests: region node lists and maps (0-based copies). */
void orc_region_nodes(Model *M, int r, int *i1, int *i2, int *i3, double *x, double *z, double *vel) {
  Region *R = &M->reg[r];
  for (int n = 1; n <= R->nnode; n++) {
    if (i1) i1[n - 1] = R->i1[n];
    if (i2) i2[n - 1] = R->i2[n];
    if (i3) i3[n - 1] = R->i3[n];
    if (x) x[n - 1] = R->x[n];
    if (z) z[n - 1] = R->z[n];
    if (vel) vel[n - 1] = R->vel[n];
  }
}
void orc_region_rnode(Model *M, int r, int *out) { memcpy(out, M->reg[r].rnode, sizeof(int) * M->nx * M->nz); }
void orc_fully_regular(Model *M, int *out) { memcpy(out, M->fully_regular, sizeof(int) * M->nx * M->nz); }
void orc_node_region(Model *M, int *out) { memcpy(out, M->node_region, sizeof(int) * M->nx * M->nz); }
void orc_isec(Model *M, int k, double *x, double *z, double *nx, double *nz, double *vt, double *vb) {
  Isec *is = &M->isec[k];
  for (int n = 1; n <= is->nnode; n++) {
    x[n - 1] = is->x[n]; z[n - 1] = is->z[n]; nx[n - 1] = is->nx_[n]; nz[n - 1] = is->nz_[n];
    vt[n - 1] = is->vel_top[n]; vb[n - 1] = is->vel_bot[n];
  }
}
int orc_isec_nnode(Model *M, int k) { return M->isec[k].nnode; }
double orc_grid_dx(Model *M) { return M->dx; }
