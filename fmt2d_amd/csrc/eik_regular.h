// Fast evaluation of the node update for a grid node whose four cells are all
// regular and which has no coincident interface partner (host + device).
//
// update_grid_node (eik_update.h) walks the node's quadrants and, per
// quadrant, evaluates regular_update2d (src/propagate2d.f90:605-790) for the
// neighbour that is alive first and for both neighbours, reading every
// member and second node again each time.  The same result follows from
// less work:
//
//  * a member's contribution to the quadratic (first- or second-order term,
//    :651-745) depends only on the member, its second node and the node's
//    velocity, not on the quadrant: the four members' terms are formed once
//    and summed per candidate in the order regular_update2d sums them
//    (a = 0 + later + earlier), so every candidate time has the same bits;
//  * the one-member update of a neighbour is the same in both quadrants that
//    contain it; evaluating it again after the visitor has seen it changes
//    nothing (a repeated candidate is either no lower than the running
//    minimum or equal to it with the same equal-time bits), so each is
//    formed once and offered in every position the quadrant walk offers it;
//  * the gradient of a candidate matters only for the candidate that the
//    walk keeps (the first one reaching the final minimum): candidate times
//    are formed for all, the gradient of the kept one at the end;
//  * what depends only on velocities and node positions (v_bar/h squared,
//    0.71/(v_a v_s), the member offsets) is formed once per tile visit.
//
// Directions d: 0 x+ (i+1), 1 x- (i-1), 2 z+ (j+1), 3 z- (j-1).  Quadrant k
// is cell k of find_connected_cells2d (:360-371): members (x+, z+), (x+, z-),
// (x-, z+), (x-, z-) for k = 0..3.
#pragma once
#include "eik_update.h"

namespace fmt2d {

constexpr double REG_TINY = 2.2250738585072014e-308;

// What the update of a fast-path node reads that does not change while a
// field is solved.
struct RegStatic {
  double d2[4];    // ((0.5 (v_u + v_a)) / h)^2 of the member in direction d
  double thr[4];   // 0.71 / (v_a v_s) of the member and its second node
  double off[4];   // x_u - x_a (d = 0, 1) or z_u - z_a (d = 2, 3)
  unsigned mem;    // bit d: member in the region
  unsigned sec;    // bit d: second node in the grid and the region, v_a v_s > tiny
  unsigned cells;  // bit k: cell k exists
  unsigned fix;    // bit d: member is the on-node source (its own update never runs)
};

FMT_HD int reg_dir_di(int d) { return d == 0 ? 1 : (d == 1 ? -1 : 0); }
FMT_HD int reg_dir_dj(int d) { return d == 2 ? 1 : (d == 3 ? -1 : 0); }

// True for grid nodes the fast path evaluates: every cell regular, no
// coincident interface node.
FMT_HD bool reg_fast_node(uint16_t fl) { return (fl & ((RF_IRR0 * 15) | (RF_COI0 * 3))) == 0; }

template <class V>
FMT_HD RegStatic reg_static(const V& v, int iu, int ju) {
  const RegionDev& R = v.reg();
  RegStatic s;
  s.mem = 0; s.sec = 0; s.fix = 0;
  s.cells = (v.flags(iu, ju) >> 9) & 15u;  // RF_CELL0 << k
  const double vu = v.vel(iu, ju);
  for (int d = 0; d < 4; d++) {
    s.d2[d] = 0.0; s.thr[d] = 0.0; s.off[d] = 0.0;
    const int ia = iu + reg_dir_di(d), ja = ju + reg_dir_dj(d);
    if (!v.in(ia, ja)) continue;
    s.mem |= 1u << d;
    Member m; m.kind = 0; m.i = ia; m.j = ja;
    if (v.fixed(m)) s.fix |= 1u << d;
    const double va = v.vel(ia, ja);
    const bool isx = d < 2;
    const double h = isx ? R.dx : R.dz;
    const double vbar = 0.5 * (vu + va);
    const double q = vbar / h;
    s.d2[d] = q * q;
    s.off[d] = isx ? v.xg(iu) - v.xg(ia) : v.zg(ju) - v.zg(ja);
    const int is = iu + 2 * reg_dir_di(d), js = ju + 2 * reg_dir_dj(d);
    if (is >= 0 && is < R.nx && js >= 0 && js < R.nz && v.in(is, js)) {
      const double vs = v.vel(is, js);
      if (va * vs > REG_TINY) {
        s.sec |= 1u << d;
        s.thr[d] = 0.71 / (va * vs);
      }
    }
  }
  return s;
}

// Field values the update reads: members (m*) and second nodes (s*) per
// direction.
struct RegDyn {
  double mT[4], mGX[4], mGZ[4];
  int mTie[4];
  double sT[4], sGX[4], sGZ[4];
};

// Quadratic terms of the member in direction d (regular_update2d's loop body).
struct RegTerm {
  double a, b, c, ta, t2;
  bool ok;  // member reached
};

FMT_HD RegTerm reg_term(const RegStatic& s, const RegDyn& y, int d) {
  RegTerm t;
  const double ta = y.mT[d];
  t.ta = ta;
  t.ok = ((s.mem >> d) & 1u) && ta < (double)FMT_HUGE_HALF;
  const double d2 = s.d2[d];
  bool second = false;
  if ((s.sec >> d) & 1u) {
    const double ts = y.sT[d];
    if (ts < (double)FMT_HUGE_HALF && ta >= ts && y.mGX[d] * y.sGX[d] + y.mGZ[d] * y.sGZ[d] >= s.thr[d]) {
      second = true;
      t.t2 = ts;
      t.a = 2.25 * d2;
      t.b = (-6.0 * ta + 1.5 * ts) * d2;
      t.c = (4.0 * ta * ta - 2.0 * ta * ts + 0.25 * ts * ts) * d2;
    }
  }
  if (!second) {
    t.t2 = 0.0;
    t.a = d2;
    t.b = (-2.0 * ta) * d2;
    t.c = (ta * ta) * d2;
  }
  return t;
}

// Root of the candidate's quadratic with the clamp to its latest member
// (:745-760); false when the quadratic degenerates (a == 0).
FMT_HD bool reg_root(double a, double b, double c, double tmax, double& t1) {
  c = c - 1.0;
  if (!(fabs(a) >= REG_TINY)) return false;
  double det = b * b - 4.0 * a * c;
  // as regular_cand: beyond 1e20 s the reference stops, here the time is NaN
  if (det < 0) det = 0.5 * sqrt(-det) / fabs(a) < 1.0e20 ? 0.0 : (double)NAN;
  t1 = 0.5 * (-b + sqrt(det)) / a;
  if (tmax > t1) t1 = tmax;
  return true;
}

// One-sided gradient component along a member (:768-782).
FMT_HD double reg_grad(double t1, double ta, double t2, double off) {
  return (t2 == 0.0) ? (t1 - ta) / off : 0.5 * (3.0 * t1 - 4.0 * ta + t2) / off;
}

// Equal-time bookkeeping of regular_tie for one member: false when the
// member rests only on the node (its support mask is the direction back to
// the node).
FMT_HD bool reg_tie_member(const RegDyn& y, int d, double T, int& tie) {
  if (y.mT[d] != T) return true;
  const int back = d == 0 ? TIE_XM : (d == 1 ? TIE_XP : (d == 2 ? TIE_ZM : TIE_ZP));
  const int own = d == 0 ? TIE_XP : (d == 1 ? TIE_XM : (d == 2 ? TIE_ZP : TIE_ZM));
  if (y.mTie[d] == back) return false;
  tie |= own;
  return true;
}

// Running minimum of the quadrant walk (MinVisitor's rule for regular
// candidates) and the gradient inputs of the candidate it keeps.
struct RegBest {
  double T;
  int tie, indep;
  int gsel;                  // 0 seed, 1 x member only, 2 z member only, 3 both
  double xta, xt2, xoff, zta, zt2, zoff;
};

FMT_HD void reg_add(RegBest& b, double T, int tie, int gsel, const RegTerm& tx, double xoff, const RegTerm& tz,
                    double zoff) {
  if (T < b.T || T != T) {  // a NaN candidate (reg_root's stop case) sticks
    b.T = T; b.tie = tie; b.indep = tie == 0; b.gsel = gsel;
    b.xta = tx.ta; b.xt2 = tx.t2; b.xoff = xoff; b.zta = tz.ta; b.zt2 = tz.t2; b.zoff = zoff;
  } else if (T == b.T) {
    if (tie) b.tie |= tie; else b.indep = 1;
  }
}

// One-member candidate of a member (regular_update2d with that member alone).
FMT_HD bool reg_one(const RegTerm& t, const RegDyn& y, int d, double& r, int& tie) {
  if (!t.ok) return false;
  double a = 0.0, b = 0.0, c = 0.0;
  a = a + t.a; b = b + t.b; c = c + t.c;
  const double tmax = t.ta > 0.0 ? t.ta : 0.0;
  if (!reg_root(a, b, c, tmax, r)) return false;
  tie = 0;
  return reg_tie_member(y, d, r, tie);
}

// Quadrant with x member X and z member Z (directions dx, dz).
FMT_HD void reg_quadrant(RegBest& b, const RegStatic& s, const RegDyn& y, const RegTerm& X, int dx, bool vX, double cX,
                         int tX, const RegTerm& Z, int dz, bool vZ, double cZ, int tZ) {
  const bool hx = X.ok, hz = Z.ok;
  if (hx || hz) {
    // one-member update of the neighbour alive first
    const bool xf = hx && (!hz || X.ta <= Z.ta);
    const int dF = xf ? dx : dz;
    const double taF = xf ? X.ta : Z.ta;
    if (!((s.fix >> dF) & 1u) && taF < b.T) {
      if (xf) { if (vX) reg_add(b, cX, tX, 1, X, s.off[dx], Z, 0.0); }
      else if (vZ) reg_add(b, cZ, tZ, 2, X, 0.0, Z, s.off[dz]);
    }
  }
  if (hx && hz) {
    // both, the later one first
    const bool zlater = Z.ta >= X.ta;
    const RegTerm& Lm = zlater ? Z : X;
    const RegTerm& Em = zlater ? X : Z;
    if (Lm.ta < b.T) {
      double a = 0.0, bb = 0.0, c = 0.0;
      a = a + Lm.a; bb = bb + Lm.b; c = c + Lm.c;
      a = a + Em.a; bb = bb + Em.b; c = c + Em.c;
      double tmax = 0.0;
      tmax = Lm.ta > tmax ? Lm.ta : tmax;
      tmax = Em.ta > tmax ? Em.ta : tmax;
      double r;
      if (reg_root(a, bb, c, tmax, r)) {
        int tie = 0;
        const bool ok = zlater ? (reg_tie_member(y, dz, r, tie) && reg_tie_member(y, dx, r, tie))
                               : (reg_tie_member(y, dx, r, tie) && reg_tie_member(y, dz, r, tie));
        if (ok) reg_add(b, r, tie, 3, X, s.off[dx], Z, s.off[dz]);
      }
    }
  }
}

// The quadrant walk of a fast-path node seeded with time T0: the running
// minimum and the candidate it keeps.
FMT_HD RegBest reg_walk(const RegStatic& s, const RegDyn& y, double T0) {
  const RegTerm t0 = reg_term(s, y, 0), t1 = reg_term(s, y, 1), t2 = reg_term(s, y, 2), t3 = reg_term(s, y, 3);
  double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
  int e0 = 0, e1 = 0, e2 = 0, e3 = 0;
  const bool v0 = reg_one(t0, y, 0, c0, e0), v1 = reg_one(t1, y, 1, c1, e1);
  const bool v2 = reg_one(t2, y, 2, c2, e2), v3 = reg_one(t3, y, 3, c3, e3);
  RegBest b;
  b.T = T0; b.tie = 0; b.indep = 1; b.gsel = 0;
  b.xta = b.xt2 = b.xoff = b.zta = b.zt2 = b.zoff = 0.0;
  if (s.cells & 1u) reg_quadrant(b, s, y, t0, 0, v0, c0, e0, t2, 2, v2, c2, e2);
  if (s.cells & 2u) reg_quadrant(b, s, y, t0, 0, v0, c0, e0, t3, 3, v3, c3, e3);
  if (s.cells & 4u) reg_quadrant(b, s, y, t1, 1, v1, c1, e1, t2, 2, v2, c2, e2);
  if (s.cells & 8u) reg_quadrant(b, s, y, t1, 1, v1, c1, e1, t3, 3, v3, c3, e3);
  return b;
}

// Suspect flag of a fast-path node (grid_suspect): its value comes from a
// one-member regular update (x or z member alone, :784), which for a node
// with regular cells only is the candidate the walk keeps.
FMT_HD bool reg_suspect(const RegStatic& s, const RegDyn& y, double T0) {
  const int g = reg_walk(s, y, T0).gsel;
  return g == 1 || g == 2;
}

// The update of a fast-path node: the same (T, dT/dx, dT/dz, equal-time
// mask) as update_grid_node for a node seeded with (T0, gx0, gz0).
FMT_HD void reg_update(const RegStatic& s, const RegDyn& y, double T0, double gx0, double gz0, double& T,
                       double& gx, double& gz, int& tie_out) {
  const RegBest b = reg_walk(s, y, T0);
  T = b.T;
  tie_out = b.indep ? 0 : b.tie;
  if (b.gsel == 0) { gx = gx0; gz = gz0; return; }
  gx = (b.gsel & 1) ? reg_grad(b.T, b.xta, b.xt2, b.xoff) : 0.0;
  gz = (b.gsel & 2) ? reg_grad(b.T, b.zta, b.zt2, b.zoff) : 0.0;
}

// Reads RegDyn for node (iu, ju) through a view (host checks, post-passes).
template <class V>
FMT_HD RegDyn reg_dyn(const V& v, const RegStatic& s, int iu, int ju) {
  RegDyn y;
  for (int d = 0; d < 4; d++) {
    y.mT[d] = FMT_HUGE_TIME; y.mGX[d] = 0.0; y.mGZ[d] = 0.0; y.mTie[d] = 0;
    y.sT[d] = FMT_HUGE_TIME; y.sGX[d] = 0.0; y.sGZ[d] = 0.0;
    if (!((s.mem >> d) & 1u)) continue;
    const int ia = iu + reg_dir_di(d), ja = ju + reg_dir_dj(d);
    y.mT[d] = v.T(ia, ja); y.mGX[d] = v.GX(ia, ja); y.mGZ[d] = v.GZ(ia, ja); y.mTie[d] = v.tie(ia, ja);
    if (!((s.sec >> d) & 1u)) continue;
    const int is = iu + 2 * reg_dir_di(d), js = ju + 2 * reg_dir_dj(d);
    y.sT[d] = v.T(is, js); y.sGX[d] = v.GX(is, js); y.sGZ[d] = v.GZ(is, js);
  }
  return y;
}

}  // namespace fmt2d
