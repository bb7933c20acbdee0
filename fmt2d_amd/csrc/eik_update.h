// Node update operator of the multistage eikonal solver (host + device).
//
// The reference solves each stage with a binary-heap fast marching method
// (src/propagate2d.f90:33-127).  Its result at every node is the minimum of
// the local updates evaluated while the node was still trial, each update
// using the neighbours that were alive at that moment.  Because every local
// update is >= the largest time it uses (the clamp t1=max(t1,tmax) at
// propagate2d.f90:760, the acceptance test atime>=T(node1) at :916-942, and
// T+dist/v at :813), that set of evaluations depends only on the final times,
// so the FMM result is the fixed point of the order-free operator below:
//
//   * regular cell (quadrant with an x- and a z-neighbour, :495-531, :605-790):
//       min( update from the earlier of the two neighbours alone,
//            update from both )          -- the one-neighbour update of the
//       later neighbour is never evaluated because the earlier one is
//       already alive when the later one pops;
//   * cell cut by the region's top/bottom interface (:533-602, :793-946):
//       min over every member of time_from_1_node and over every pair of
//       members of time_from_2_nodes;
//   * second-order terms (:651-680, :697-726) are used when the node two
//     steps away is alive, i.e. reached no later than the neighbour itself;
//   * seeds enter as an upper bound, nodes seeded alive (an on-node point
//     source, :2369-2380) never pop, so they contribute only through pairs;
//   * interface nodes of the last column are never connected to any node
//     (their cut cell is not registered, :1468-1480) and keep their seed.
//
// The GPU iterates this operator (block fast-iterative method) until no
// node changes.  All arithmetic is FP64 in the reference's operation order
// (compiled without contraction into fused multiply-adds), so the fixed
// point reproduces the reference's times and gradients bit for bit, and
// with them the exact ties that decide its gradient refits.
#pragma once
#include "eik_common.h"

namespace fmt2d {

template <typename Real>
struct Best {
  Real T, gx, gz;
  int susp;       // gradient is one-sided: refit after convergence (:784)
  int reg_set;    // a regular update produced the current value
  int tie;        // directions of equal-time members the value rests on (TIE_*)
  int indep;      // some candidate reaches the value from an earlier time
};

// Equal-time support.  A regular update clamped to its trigger member's time
// (t1 = tmax, src/propagate2d.f90:770) gives a node the same time as that
// member; in the FMM the member popped first, so two nodes can never rest on
// each other this way.  A node records the directions of the members whose
// clamped updates alone give its value; a member that rests only on the node
// itself is not used by it.  This keeps an iteration from locking a pair of
// equal-time neighbours at a transient value.
enum { TIE_XP = 1, TIE_XM = 2, TIE_ZP = 4, TIE_ZM = 8 };
FMT_HD int tie_dir(int di, int dj) {
  return di > 0 ? TIE_XP : (di < 0 ? TIE_XM : (dj > 0 ? TIE_ZP : (dj < 0 ? TIE_ZM : 0)));
}

// Candidate visitors.  Every local update is announced with its evaluation
// time (the time of the member whose pop triggers it = the largest member
// time) before it is computed, so a visitor can skip it cheaply.
enum { CK_REGULAR = 1, CK_POINT = 2 };
// What an offered update is, announced (note) just before its evaluation
// time (want): for visitors that need to know which update a value came
// from (the adjoint); the others ignore it.
enum { CN_REG1 = 1, CN_REG2 = 2, CN_C1 = 3, CN_C2 = 4, CN_PARTNER = 5 };

template <typename Real>
struct MinVisitor {       // fixed-point iteration: keep the minimum
  Best<Real> b;
  // every local update is >= its evaluation time (the clamp to the largest
  // member time, T1 + dist/v, the acceptance atime >= T(node1)), so an update
  // evaluated no earlier than the current minimum cannot lower it
  FMT_HD bool want(Real evt) const { return evt < b.T; }
  template <class M>
  FMT_HD void note(int, const M&, const M&) const {}
  FMT_HD void add(Real T, Real gx, Real gz, int kind, int susp, int tie) {
    if (T < b.T || T != T) {  // a NaN candidate (regular_cand's stop case) sticks
      b.T = T; b.gx = gx; b.gz = gz; b.reg_set = (kind == CK_REGULAR); if (kind == CK_REGULAR) b.susp = susp;
      b.tie = tie; b.indep = tie == 0;
    } else if (T == b.T) {
      if (tie) b.tie |= tie; else b.indep = 1;
    }
  }
  // support mask to store with the node
  FMT_HD int tie_mask() const { return b.indep ? 0 : b.tie; }
};

// ---------------------------------------------------------------- regular

// regular_update2d for the members (x and/or z neighbours) of one quadrant,
// in the reference's operation order (src/propagate2d.f90:605-790) so that
// FP64 results, and the ties they decide, are reproduced bit for bit.
// Returns false when the update is undefined (a == 0).
template <typename Real, class V>
FMT_HD bool regular_cand(const V& v, int iu, int ju, const int* mx, const int* mz, int nm,
                         Real& T, Real& gx, Real& gz, int& susp) {
  const RegionDev& R = v.reg();
  Real tmax = 0;
  const Real vu = v.vel(iu, ju);
  Real a = 0, b = 0, c = 0;
  Real tx = 0, tz = 0, t2x = 0, t2z = 0;
  bool xi = false, zi = false;
  int ioff = 0, joff = 0;
  for (int n = 0; n < nm; n++) {
    const int ia = mx[n], ja = mz[n];
    const Real ta = v.T(ia, ja), va = v.vel(ia, ja);
    if (ta > tmax) tmax = ta;
    const bool isx = (ia != iu);
    const int is = isx ? iu + 2 * (ia - iu) : ia;
    const int js = isx ? ja : ju + 2 * (ja - ju);
    const Real h = isx ? (Real)R.dx : (Real)R.dz;
    const Real vbar = (Real)0.5 * (vu + va);
    if (isx) { xi = true; ioff = ia; } else { zi = true; joff = ja; }
    bool second = false;
    if (is >= 0 && is < R.nx && js >= 0 && js < R.nz && v.in(is, js)) {
      const Real ts = v.T(is, js);
      // the second node is alive when it was reached no later than this one
      if (ts < (Real)FMT_HUGE_HALF && ta >= ts) {
        const Real vs = v.vel(is, js);
        if (va * vs > (Real)2.2250738585072014e-308 &&
            (Real)v.GX(ia, ja) * (Real)v.GX(is, js) + (Real)v.GZ(ia, ja) * (Real)v.GZ(is, js) >=
                (Real)0.71 / (va * vs)) {
          second = true;
          const Real q = vbar / h;
          const Real d2 = q * q;
          if (isx) { tx = ta; t2x = ts; } else { tz = ta; t2z = ts; }
          a = a + (Real)2.25 * d2;
          b = b + ((Real)-6.0 * ta + (Real)1.5 * ts) * d2;
          c = c + ((Real)4.0 * ta * ta - (Real)2.0 * ta * ts + (Real)0.25 * ts * ts) * d2;
        }
      }
    }
    if (!second) {
      const Real q = vbar / h;
      const Real d2 = q * q;
      if (isx) tx = ta; else tz = ta;
      a = a + d2;
      b = b + ((Real)-2.0 * ta) * d2;
      c = c + (ta * ta) * d2;
    }
  }
  c = c - (Real)1.0;
  if (!(fabs((double)a) >= 2.2250738585072014e-308)) return false;
  Real det = b * b - (Real)4.0 * a * c;
  // a negative discriminant is taken as 0, unless the root it would move is
  // beyond 1e20 s, where the reference stops (:745-756): the time becomes
  // NaN, which the forward reports as an error
  if (det < 0) det = (Real)0.5 * sqrt(-det) / fabs(a) < (Real)1.0e20 ? (Real)0 : (Real)NAN;
  Real t1 = (Real)0.5 * (-b + sqrt(det)) / a;
  if (tmax > t1) t1 = tmax;
  T = t1;
  gx = 0; gz = 0;
  // t2x==0 selects the one-sided first-order formula (second order unused,
  // or the second node is the source itself at time 0)
  if (xi) {
    const Real dxu = (Real)v.xg(iu) - (Real)v.xg(ioff);
    gx = (t2x == (Real)0) ? (t1 - tx) / dxu : (Real)0.5 * ((Real)3.0 * t1 - (Real)4.0 * tx + t2x) / dxu;
  }
  if (zi) {
    const Real dzu = (Real)v.zg(ju) - (Real)v.zg(joff);
    gz = (t2z == (Real)0) ? (t1 - tz) / dzu : (Real)0.5 * ((Real)3.0 * t1 - (Real)4.0 * tz + t2z) / dzu;
  }
  susp = !(xi && zi);
  return true;
}

// ---------------------------------------------------------------- point geometry

// time_from_1_node2d (src/propagate2d.f90:793-828)
template <typename Real, class Vis>
FMT_HD void cand1(Real xu, Real zu, Real vu, Real x1, Real z1, Real t1, Real g1x, Real g1z, Real v1,
                  Real tol, Vis& vis) {
  const Real dx = xu - x1, dz = zu - z1;
  const Real dist = sqrt(dx * dx + dz * dz);
  if (dist < (Real)0.01 * tol) { vis.add(t1, g1x, g1z, CK_POINT, 0, 0); return; }
  const Real at = t1 + dist / ((Real)0.5 * (vu + v1));
  vis.add(at, dx / (dist * vu), dz / (dist * vu), CK_POINT, 0, 0);
}

// One root of time_from_2_nodes2d: the slowness vector p + u q, kept when it
// points into the wedge of the two members and gives a time no earlier than
// the later member's (src/propagate2d.f90:903-942).
template <typename Real, class Vis>
FMT_HD void cand2_root(const Real& uu, const Real& p0, const Real& p1, const Real& q0, const Real& q1,
                       const Real& n10, const Real& n11, const Real& n20, const Real& n21, const Real& r10,
                       const Real& r11, const Real& ta, const Real& va, const Real& vm, Vis& vis) {
  const Real w0 = p0 + uu * q0, w1 = p1 + uu * q1;
  if (w0 * n10 + w1 * n11 <= 0 && w0 * n20 + w1 * n21 <= 0) {
    const Real at = ta - (w0 * r10 + w1 * r11) / vm;
    if (at >= ta) {
      const Real wl = sqrt(w0 * w0 + w1 * w1);
      vis.add(at, w0 / (wl * va), w1 / (wl * va), CK_POINT, 0, 0);
    }
  }
}

// time_from_2_nodes2d (src/propagate2d.f90:831-946)
template <typename Real, class Vis>
FMT_HD void cand2(Real xu, Real zu, Real vu,
                  Real xa, Real za, Real ta, Real gax, Real gaz, Real va,
                  Real xb, Real zb, Real tb, Real gbx, Real gbz, Real vb, Real tol, Vis& vis) {
  // node1 is the later of the two
  if (ta < tb) {
    Real t;
    t = xa; xa = xb; xb = t; t = za; za = zb; zb = t; t = ta; ta = tb; tb = t;
    t = gax; gax = gbx; gbx = t; t = gaz; gaz = gbz; gbz = t; t = va; va = vb; vb = t;
  }
  const Real ctol = (Real)0.01 * tol;
  const Real r10 = xa - xu, r11 = za - zu, r20 = xb - xu, r21 = zb - zu;
  if (sqrt(r10 * r10 + r11 * r11) < ctol) { vis.add(ta, gax, gaz, CK_POINT, 0, 0); return; }
  if (sqrt(r20 * r20 + r21 * r21) < ctol) { vis.add(tb, gbx, gbz, CK_POINT, 0, 0); return; }
  const Real e0 = xb - xa, e1 = zb - za;
  if (sqrt(e0 * e0 + e1 * e1) < ctol) {
    cand1<Real>(xu, zu, vu, xa, za, ta, gax, gaz, va, tol, vis);
    return;
  }
  const Real v12 = (Real)0.5 * (va + vb);
  const Real d0 = (xa - xb) / v12, d1 = (za - zb) / v12;
  const Real sdd = d0 * d0 + d1 * d1;
  const Real f = (ta - tb) / sdd;
  const Real p0 = f * d0, p1 = f * d1;
  const Real q0 = -d1, q1 = d0;
  const Real ae = q0 * q0 + q1 * q1;
  const Real be = (Real)2.0 * (p0 * q0 + p1 * q1);
  const Real ce = (p0 * p0 + p1 * p1) - (Real)1.0;
  const Real det = be * be - (Real)4.0 * ae * ce;
  if (det < 0) return;
  const Real sq = sqrt(det);
  const Real u1 = (Real)0.5 * (-be + sq) / ae, u2 = (Real)0.5 * (-be - sq) / ae;
  Real n10 = -r11, n11 = r10;
  const Real l1 = sqrt(n10 * n10 + n11 * n11);
  n10 /= l1; n11 /= l1;
  if (n10 * r20 + n11 * r21 < 0) { n10 = -n10; n11 = -n11; }
  Real n20 = -r21, n21 = r20;
  const Real l2 = sqrt(n20 * n20 + n21 * n21);
  n20 /= l2; n21 /= l2;
  if (n20 * r10 + n21 * r11 < 0) { n20 = -n20; n21 = -n21; }
  const Real vm = (Real)0.5 * (vu + va);
  // the two roots in order, each by a call of its own (no selection between
  // two local values: the device compiler has mis-selected such a pair for
  // class-typed Real)
  cand2_root<Real>(u1, p0, p1, q0, q1, n10, n11, n20, n21, r10, r11, ta, va, vm, vis);
  cand2_root<Real>(u2, p0, p1, q0, q1, n10, n11, n20, n21, r10, r11, ta, va, vm, vis);
}

// member of an irregular cell: grid node (kind 0) or interface node (kind 1+s)
struct Member {
  int kind, i, j;  // grid: (ix,iz); interface: side=kind-1, column i
};

template <typename Real, class V>
FMT_HD void member_data(const V& v, const Member& m, Real& x, Real& z, Real& t, Real& gx, Real& gz, Real& vel) {
  if (m.kind == 0) {
    x = v.xg(m.i); z = v.zg(m.j); t = v.T(m.i, m.j); gx = v.GX(m.i, m.j); gz = v.GZ(m.i, m.j); vel = v.vel(m.i, m.j);
  } else {
    const int s = m.kind - 1;
    x = v.ifx(s, m.i); z = v.ifz(s, m.i); t = v.iT(s, m.i); gx = v.iGX(s, m.i); gz = v.iGZ(s, m.i); vel = v.ivel(s, m.i);
  }
}

// Members of the irregular cell (cx,cz) for updating node `self`: the
// region's top then bottom interface nodes in the cell, then the corners in
// the region (src/propagate2d.f90:392-441).  `self` and `excl` (the node's
// coincident partner, see update_grid_node) are left out, and members not yet
// reached are skipped (they cannot lower the node).  The members are returned
// as a mask over six slots in that order: 0/1 top/bottom interface node,
// 2..5 corners (cx+i, cz+j) with i outer, j inner (slot_member), so callers
// walk them with bit scans and keep everything in registers.
FMT_HD Member slot_member(int cx, int cz, int k) {
  Member m;
  if (k < 2) { m.kind = 1 + k; m.i = cx; m.j = 0; }
  else { m.kind = 0; m.i = cx + ((k - 2) >> 1); m.j = cz + ((k - 2) & 1); }
  return m;
}

FMT_HD int lowest_slot(unsigned m) { return __builtin_ctz(m); }

template <class V>
FMT_HD unsigned irregular_members(const V& v, int cx, int cz, const Member& self, const Member& excl) {
  const RegionDev& R = v.reg();
  unsigned mask = 0;
  for (int s = 0; s < 2; s++) {
    if (v.ifcc(s, cx) == cz && v.ifok(s, cx)) {
      const bool me = (self.kind == 1 + s && self.i == cx) || (excl.kind == 1 + s && excl.i == cx);
      if (!me && v.iT(s, cx) < FMT_HUGE_HALF) mask |= 1u << s;
    }
  }
  for (int i = 0; i <= 1; i++)
    for (int j = 0; j <= 1; j++) {
      const int ii = cx + i, jj = cz + j;
      if (ii < 0 || ii >= R.nx || jj < 0 || jj >= R.nz) continue;
      if (self.kind == 0 && self.i == ii && self.j == jj) continue;
      if (excl.kind == 0 && excl.i == ii && excl.j == jj) continue;
      if (!v.in(ii, jj)) continue;
      if (v.T(ii, jj) >= FMT_HUGE_HALF) continue;
      mask |= 1u << (2 + 2 * i + j);
    }
  return mask;
}

template <typename Real, class V, class Vis>
FMT_HD void irregular_cell(const V& v, int cx, int cz, const Member& self, const Member& excl, Real xu,
                           Real zu, Real vu, Vis& vis) {
  const Real tol = (Real)v.reg().tol;
  // every member alone, then every later member paired with it, in slot order
  for (unsigned ma = irregular_members(v, cx, cz, self, excl); ma; ma &= ma - 1) {
    const Member A = slot_member(cx, cz, lowest_slot(ma));
    Real xa, za, ta, gax, gaz, va;
    member_data<Real>(v, A, xa, za, ta, gax, gaz, va);
    vis.note(CN_C1, A, A);
    if (!v.fixed(A) && vis.want(ta)) cand1<Real>(xu, zu, vu, xa, za, ta, gax, gaz, va, tol, vis);
    for (unsigned mc = ma & (ma - 1); mc; mc &= mc - 1) {
      const Member Cm = slot_member(cx, cz, lowest_slot(mc));
      Real xc, zc, tc, gcx, gcz, vc;
      member_data<Real>(v, Cm, xc, zc, tc, gcx, gcz, vc);
      vis.note(CN_C2, A, Cm);
      if (vis.want(ta > tc ? ta : tc))
        cand2<Real>(xu, zu, vu, xa, za, ta, gax, gaz, va, xc, zc, tc, gcx, gcz, vc, tol, vis);
    }
  }
}

// Equal-time bookkeeping of a regular update of value T: `tie` gets the
// directions of the members whose time equals T; returns false when one of
// them rests only on node (iu,ju) (it popped after the node, so the FMM
// never evaluates this update).
template <typename Real, class V>
FMT_HD bool regular_tie(const V& v, int iu, int ju, const int* mx, const int* mz, int nm, Real T, int& tie) {
  tie = 0;
  for (int n = 0; n < nm; n++) {
    if ((Real)v.T(mx[n], mz[n]) != T) continue;
    if (v.tie(mx[n], mz[n]) == tie_dir(iu - mx[n], ju - mz[n])) return false;
    tie |= tie_dir(mx[n] - iu, mz[n] - ju);
  }
  return true;
}

// quadrant / regular cell (cx,cz) of grid node (iu,ju)
template <typename Real, class V, class Vis>
FMT_HD void regular_cell(const V& v, int iu, int ju, int cx, int cz, Vis& vis) {
  const int xn = (cx == iu) ? iu + 1 : iu - 1;   // x-neighbour inside the cell
  const int zn = (cz == ju) ? ju + 1 : ju - 1;   // z-neighbour inside the cell
  const bool hx = v.in(xn, ju) && v.T(xn, ju) < FMT_HUGE_HALF;
  const bool hz = v.in(iu, zn) && v.T(iu, zn) < FMT_HUGE_HALF;
  int mx[2], mz[2];
  Real T, gx, gz;
  int susp;
  if (hx || hz) {
    // one-neighbour update from the neighbour that is alive first
    bool xfirst = hx && (!hz || v.T(xn, ju) <= v.T(iu, zn));
    const int fi = xfirst ? xn : iu, fj = xfirst ? ju : zn;
    Member fm; fm.kind = 0; fm.i = fi; fm.j = fj;
    vis.note(CN_REG1, fm, fm);
    if (!v.fixed(fm) && vis.want((Real)v.T(fi, fj))) {
      mx[0] = fi; mz[0] = fj;
      if (regular_cand<Real>(v, iu, ju, mx, mz, 1, T, gx, gz, susp)) {
        int tie;
        if (regular_tie<Real>(v, iu, ju, mx, mz, 1, T, tie)) vis.add(T, gx, gz, CK_REGULAR, susp, tie);
      }
    }
  }
  if (hx && hz) {
    // both alive: evaluated when the later one pops, with the later first
    const bool zlater = v.T(iu, zn) >= v.T(xn, ju);
    mx[0] = zlater ? iu : xn; mz[0] = zlater ? zn : ju;
    mx[1] = zlater ? xn : iu; mz[1] = zlater ? ju : zn;
    Member la; la.kind = 0; la.i = mx[0]; la.j = mz[0];
    Member ea; ea.kind = 0; ea.i = mx[1]; ea.j = mz[1];
    vis.note(CN_REG2, la, ea);
    if (vis.want((Real)v.T(mx[0], mz[0])) && regular_cand<Real>(v, iu, ju, mx, mz, 2, T, gx, gz, susp)) {
      int tie;
      if (regular_tie<Real>(v, iu, ju, mx, mz, 2, T, tie)) vis.add(T, gx, gz, CK_REGULAR, susp, tie);
    }
  }
}

template <typename Real, class V, class Vis>
FMT_HD void own_grid_candidates(const V& v, int iu, int ju, const Member& excl, Vis& vis) {
  const uint16_t fl = v.flags(iu, ju);
  Member self; self.kind = 0; self.i = iu; self.j = ju;
  const Real xu = v.xg(iu), zu = v.zg(ju), vu = v.vel(iu, ju);
  for (int k = 0; k < 4; k++) {
    if (!(fl & (RF_CELL0 << k))) continue;
    const int cx = iu + cell_dx(k), cz = ju + cell_dz(k);
    if (fl & (RF_IRR0 << k)) irregular_cell<Real>(v, cx, cz, self, excl, xu, zu, vu, vis);
    else regular_cell<Real>(v, iu, ju, cx, cz, vis);
  }
}

template <typename Real, class V, class Vis>
FMT_HD void own_iface_candidates(const V& v, int s, int c, const Member& excl, Vis& vis) {
  Member self; self.kind = 1 + s; self.i = c; self.j = 0;
  irregular_cell<Real>(v, c, v.ifcc(s, c), self, excl, (Real)v.ifx(s, c), (Real)v.ifz(s, c),
                       (Real)v.ivel(s, c), vis);
}

// Coincident pairs.  A grid node lying exactly on its region's interface node
// (the sea surface and the flat model bottom coincide with grid rows) is a
// member of that node's cut cell and vice versa; each copies the other's time
// and gradient when the other pops first (the dist<0.01*tolerance branches,
// :807-811, :863-873).  The pair therefore ends with the smaller of the two
// nodes' own minima.  Each node evaluates its partner's own minimum (partner's
// seed and candidates, itself excluded) as one copy candidate, so the
// iteration cannot lock the pair onto a transient value through the mutual
// copy.

template <class V>
FMT_HD int grid_partner_side(const V& v, int iu, int ju) {
  const uint16_t fl = v.flags(iu, ju);
  return (fl & RF_COI0) ? 0 : ((fl & (RF_COI0 << 1)) ? 1 : -1);
}

template <class V>
FMT_HD int iface_partner_row(const V& v, int s, int c) {
  const RegionDev& R = v.reg();
  if (c >= R.nx - 1) return -1;
  const int cz = v.ifcc(s, c);
  if (cz < 0) return -1;
  for (int j = cz; j <= cz + 1 && j < R.nz; j++)
    if (v.in(c, j) && (v.flags(c, j) & (RF_COI0 << s))) return j;
  return -1;
}

// Row of the tile schedule that interface node (s,c) with cut-cell row cz is
// solved in: cz, or cz+1 when it coincides with the grid node below.  An
// interface node reads its cell (rows cz, cz+1) and the candidates of its
// coincident grid node, which reach two rows beyond that node; homing the
// interface node at its partner's row keeps every read and every dependency
// within two rows of the node's home.
template <class V>
FMT_HD int iface_home_row(const V& v, int s, int c, int cz) {
  const int pj = iface_partner_row(v, s, c);
  return pj == cz + 1 ? pj : cz;
}

// own minimum of interface node (s,c) leaving out grid node (pi,pj)
template <typename Real, class V>
FMT_HD MinVisitor<Real> iface_own_min(const V& v, int s, int c, int pi, int pj) {
  MinVisitor<Real> m;
  m.b.T = v.iT0(s, c); m.b.gx = v.iGX0(s, c); m.b.gz = v.iGZ0(s, c); m.b.susp = 0; m.b.reg_set = 0; m.b.tie = 0; m.b.indep = 1;
  Member ex; ex.kind = 0; ex.i = pi; ex.j = pj;
  own_iface_candidates<Real>(v, s, c, ex, m);
  return m;
}

// own minimum of grid node (c,j) leaving out interface node (s,c)
template <typename Real, class V>
FMT_HD MinVisitor<Real> grid_own_min(const V& v, int c, int j, int s) {
  MinVisitor<Real> m;
  m.b.T = v.T0(c, j); m.b.gx = v.GX0(c, j); m.b.gz = v.GZ0(c, j); m.b.susp = 0; m.b.reg_set = 0; m.b.tie = 0; m.b.indep = 1;
  Member ex; ex.kind = 1 + s; ex.i = c; ex.j = 0;
  Member gm; gm.kind = 0; gm.i = c; gm.j = j;
  if (!v.fixed(gm)) own_grid_candidates<Real>(v, c, j, ex, m);
  return m;
}

// All candidates of grid node (iu,ju) except its seed.
template <typename Real, class V, class Vis>
FMT_HD void grid_candidates(const V& v, int iu, int ju, Vis& vis) {
  const int ps = grid_partner_side(v, iu, ju);
  Member excl; excl.kind = -1; excl.i = -1; excl.j = -1;
  if (ps >= 0) { excl.kind = 1 + ps; excl.i = iu; }
  own_grid_candidates<Real>(v, iu, ju, excl, vis);
  if (ps >= 0) {
    MinVisitor<Real> p = iface_own_min<Real>(v, ps, iu, iu, ju);
    Member pm; pm.kind = 1 + ps; pm.i = iu; pm.j = 0;
    vis.note(CN_PARTNER, pm, pm);
    if (p.b.T < (Real)FMT_HUGE_HALF && vis.want(p.b.T)) vis.add(p.b.T, p.b.gx, p.b.gz, CK_POINT, 0, 0);
  }
}

// All candidates of interface node (s,c) except its seed.  The last column
// is never reached (see header).
template <typename Real, class V, class Vis>
FMT_HD void iface_candidates(const V& v, int s, int c, Vis& vis) {
  const RegionDev& R = v.reg();
  if (c >= R.nx - 1 || v.ifcc(s, c) < 0) return;
  const int pj = iface_partner_row(v, s, c);
  Member excl; excl.kind = -1; excl.i = -1; excl.j = -1;
  if (pj >= 0) { excl.kind = 0; excl.i = c; excl.j = pj; }
  own_iface_candidates<Real>(v, s, c, excl, vis);
  if (pj >= 0) {
    MinVisitor<Real> p = grid_own_min<Real>(v, c, pj, s);
    Member pm; pm.kind = 0; pm.i = c; pm.j = pj;
    vis.note(CN_PARTNER, pm, pm);
    if (p.b.T < (Real)FMT_HUGE_HALF && vis.want(p.b.T)) vis.add(p.b.T, p.b.gx, p.b.gz, CK_POINT, 0, 0);
  }
}

// Fixed-point update of grid node (iu,ju); b holds the node's seed on entry,
// on return b.tie is the node's equal-time support mask.
template <typename Real, class V>
FMT_HD void update_grid_node(const V& v, int iu, int ju, Best<Real>& b) {
  MinVisitor<Real> m; m.b = b;
  grid_candidates<Real>(v, iu, ju, m);
  b = m.b;
  b.tie = m.tie_mask();
}

template <typename Real, class V>
FMT_HD void update_iface_node(const V& v, int s, int c, Best<Real>& b) {
  MinVisitor<Real> m; m.b = b;
  iface_candidates<Real>(v, s, c, m);
  b = m.b;
}

// ---------------------------------------------------------------- heap exhaustion

// The reference stops marching as soon as its heap is empty right after a pop
// (propagate2d.f90:80): the popped node's neighbours are then never updated
// and every node not yet alive keeps huge_time.  In a thin layer the front can
// shrink to that one node.  A node j is on the front from its own pop until
// its last connected node pops, so the exit time is the first pop time t
// (after the first pop, and after every seed has popped) at which no other
// node has T_j < t < (largest time among j's connected nodes).  This returns
// that largest time for grid node (i,j) or interface node (s,c): the members
// of its cells, which are exactly the nodes its pop updates.
template <typename Real, class V>
FMT_HD Real max_member_time_grid(const V& v, int iu, int ju) {
  const uint16_t fl = v.flags(iu, ju);
  Real mx = -1;
  Member self; self.kind = 0; self.i = iu; self.j = ju;
  Member none; none.kind = -1; none.i = -1; none.j = -1;
  for (int k = 0; k < 4; k++) {
    if (!(fl & (RF_CELL0 << k))) continue;
    const int cx = iu + cell_dx(k), cz = ju + cell_dz(k);
    if (fl & (RF_IRR0 << k)) {
      for (unsigned ma = irregular_members(v, cx, cz, self, none); ma; ma &= ma - 1) {
        const Member A = slot_member(cx, cz, lowest_slot(ma));
        if (v.fixed(A)) continue;
        const Real t = A.kind == 0 ? (Real)v.T(A.i, A.j) : (Real)v.iT(A.kind - 1, A.i);
        mx = t > mx ? t : mx;
      }
    } else {
      const int xn = (cx == iu) ? iu + 1 : iu - 1, zn = (cz == ju) ? ju + 1 : ju - 1;
      Member a; a.kind = 0; a.i = xn; a.j = ju;
      if (v.in(xn, ju) && !v.fixed(a)) { const Real t = v.T(xn, ju); mx = t > mx ? t : mx; }
      a.i = iu; a.j = zn;
      if (v.in(iu, zn) && !v.fixed(a)) { const Real t = v.T(iu, zn); mx = t > mx ? t : mx; }
    }
  }
  return mx;
}

template <typename Real, class V>
FMT_HD Real max_member_time_iface(const V& v, int s, int c) {
  const RegionDev& R = v.reg();
  if (c >= R.nx - 1 || v.ifcc(s, c) < 0) return -1;
  Member self; self.kind = 1 + s; self.i = c; self.j = 0;
  Member none; none.kind = -1; none.i = -1; none.j = -1;
  const int cz = v.ifcc(s, c);
  Real mx = -1;
  for (unsigned ma = irregular_members(v, c, cz, self, none); ma; ma &= ma - 1) {
    const Member A = slot_member(c, cz, lowest_slot(ma));
    if (v.fixed(A)) continue;
    const Real t = A.kind == 0 ? (Real)v.T(A.i, A.j) : (Real)v.iT(A.kind - 1, A.i);
    mx = t > mx ? t : mx;
  }
  return mx;
}

// ---------------------------------------------------------------- suspect flag

// Evaluation times of a node's candidates (distinct, ascending, at most 16).
template <typename Real>
struct EvtCollector {
  Real e[16];
  int n = 0;
  FMT_HD bool want(Real t) {
    int k = 0;
    while (k < n && e[k] < t) k++;
    if (k < n && e[k] == t) return false;
    if (n == 16) return false;
    for (int q = n; q > k; q--) e[q] = e[q - 1];
    e[k] = t; n++;
    return false;
  }
  template <class M>
  FMT_HD void note(int, const M&, const M&) const {}
  FMT_HD void add(Real, Real, Real, int, int, int) {}
};

// Replays the updates of one evaluation time in the FMM's order.
template <typename Real>
struct ReplayVisitor {
  Real evt, T;
  int susp;
  FMT_HD bool want(Real t) const { return t == evt; }
  template <class M>
  FMT_HD void note(int, const M&, const M&) const {}
  FMT_HD void add(Real t, Real, Real, int kind, int s, int) {
    if (t < T) { T = t; if (kind == CK_REGULAR) susp = s; }
  }
};

// The reference refits the gradient of a node when its last improving
// regular update was one-sided (propagate2d.f90:784, :119-123).  Updates are
// replayed in the order of their triggering pops (ascending evaluation time),
// starting from the seed.  Fully regular nodes only have regular updates, for
// them the last improving one is the minimum itself.
template <typename Real, class V>
FMT_HD int grid_suspect(const V& v, int iu, int ju, Real seedT) {
  const uint16_t fl = v.flags(iu, ju);
  if (!(fl & (RF_IRR0 * 15)) && grid_partner_side(v, iu, ju) < 0) {
    MinVisitor<Real> m;
    m.b.T = seedT; m.b.gx = 0; m.b.gz = 0; m.b.susp = 0; m.b.reg_set = 0; m.b.tie = 0; m.b.indep = 1;
    grid_candidates<Real>(v, iu, ju, m);
    return m.b.susp;
  }
  EvtCollector<Real> ec;
  grid_candidates<Real>(v, iu, ju, ec);
  ReplayVisitor<Real> rp;
  rp.T = seedT; rp.susp = 0;
  for (int k = 0; k < ec.n; k++) {
    rp.evt = ec.e[k];
    grid_candidates<Real>(v, iu, ju, rp);
  }
  return rp.susp;
}

// ---------------------------------------------------------------- gradient refit

// fit_gradient_at_node2d (src/propagate2d.f90:949-1196) for a grid node
template <typename Real, class V>
FMT_HD void fit_gradient_grid(const V& v, int iu, int ju, Real& dtdx, Real& dtdz) {
  const RegionDev& R = v.reg();
  const uint16_t fl = v.flags(iu, ju);
  int nassoc = 0, ncell = 0;
  for (int k = 0; k < 4; k++)
    if (fl & (RF_CELL0 << k)) { ncell++; if (fl & (RF_ASSOC0 << k)) nassoc++; }
  const Real tc = v.T(iu, ju);
  if (nassoc == 0) {
    const int ml = iu > 0 ? iu - 1 : 0, mr = iu < R.nx - 1 ? iu + 1 : R.nx - 1;
    const bool hl = v.in(ml, ju), hr = v.in(mr, ju);
    if (hl && hr) dtdx = (v.T(mr, ju) - v.T(ml, ju)) / (v.xg(mr) - v.xg(ml));
    else if (hl) dtdx = (tc - v.T(ml, ju)) / (v.xg(iu) - v.xg(ml));
    else if (hr) dtdx = (v.T(mr, ju) - tc) / (v.xg(mr) - v.xg(iu));
    else dtdx = 0;
    const int mu = ju > 0 ? ju - 1 : 0, md = ju < R.nz - 1 ? ju + 1 : R.nz - 1;
    const bool hu = v.in(iu, mu), hd = v.in(iu, md);
    if (hu && hd) dtdz = (v.T(iu, md) - v.T(iu, mu)) / (v.zg(md) - v.zg(mu));
    else if (hu) dtdz = (tc - v.T(iu, mu)) / (v.zg(ju) - v.zg(mu));
    else if (hd) dtdz = (v.T(iu, md) - tc) / (v.zg(md) - v.zg(ju));
    else dtdz = 0;
    return;
  }
  // weighted least squares over the connected cells' nodes (<=30, no repeats)
  Real a11 = 0, a12 = 0, a22 = 0, b1 = 0, b2 = 0;
  const Real xu = v.xg(iu), zu = v.zg(ju);
  int seen_g[16]; int nseen_g = 0;
  int seen_i[8]; int nseen_i = 0;
  for (int k = 0; k < 4; k++) {
    if (!(fl & (RF_CELL0 << k))) continue;
    const int cx = iu + cell_dx(k), cz = ju + cell_dz(k);
    if (fl & (RF_ASSOC0 << k)) {
      for (int s = 0; s < 2; s++) {
        if (v.ifcc(s, cx) == cz && v.ifok(s, cx)) {
          const int key = s * 65536 + cx;
          bool dup = false;
          for (int q = 0; q < nseen_i; q++) dup |= (seen_i[q] == key);
          if (dup || nseen_i >= 8) continue;
          seen_i[nseen_i++] = key;
          const Real dx = (Real)v.ifx(s, cx) - xu, dz = (Real)v.ifz(s, cx) - zu;
          const Real dw = (Real)1.0 / (sqrt(dx * dx + dz * dz) + (Real)R.tol);
          const Real dt = v.iT(s, cx) - tc;
          a11 += dx * dx * dw * dw; a12 += dx * dz * dw * dw; a22 += dz * dz * dw * dw;
          b1 += dt * dx * dw * dw; b2 += dt * dz * dw * dw;
        }
      }
    }
    for (int i = 0; i <= 1; i++)
      for (int j = 0; j <= 1; j++) {
        const int ii = cx + i, jj = cz + j;
        if (ii < 0 || ii >= R.nx || jj < 0 || jj >= R.nz) continue;
        if (ii == iu && jj == ju) continue;
        if (!v.in(ii, jj)) continue;
        const int key = jj * R.nx + ii;
        bool dup = false;
        for (int q = 0; q < nseen_g; q++) dup |= (seen_g[q] == key);
        if (dup || nseen_g >= 16) continue;
        seen_g[nseen_g++] = key;
        const Real dx = v.xg(ii) - xu, dz = v.zg(jj) - zu;
        const Real dw = (Real)1.0 / (sqrt(dx * dx + dz * dz) + (Real)R.tol);
        const Real dt = v.T(ii, jj) - tc;
        a11 += dx * dx * dw * dw; a12 += dx * dz * dw * dw; a22 += dz * dz * dw * dw;
        b1 += dt * dx * dw * dw; b2 += dt * dz * dw * dw;
      }
  }
  const Real det = a11 * a22 - a12 * a12;
  if (fabs((double)det) < 1.0e-30) { dtdx = 0; dtdz = 0; return; }
  dtdx = (a22 * b1 - a12 * b2) / det;
  dtdz = (a11 * b2 - a12 * b1) / det;
}

}  // namespace fmt2d
