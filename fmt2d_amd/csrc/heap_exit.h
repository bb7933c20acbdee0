// Heap-exhaustion stop of the reference FMM, evaluated on the device.
//
// The reference marches a binary heap and leaves its loop when a pop other
// than the first empties the heap, before the popped node updates its
// neighbours (src/propagate2d.f90:74-80): the nodes that would have been
// reached through that node keep huge_time.  On a converged order-free field
// the stop time is
//
//   t* = min { T_c : lo < T_c < hi, T_c >= smax, no other node has time T_c,
//                    max{ maxnb_i : T_i < T_c } <= T_c }        (else +inf)
//
// over the field's reached, non-fixed nodes, where lo/hi are their smallest /
// largest time, smax the largest time of the nodes seeded before marching,
// and maxnb_i the largest time among the nodes connected to node i.  A field
// with a seed that is never reached runs to the end (+inf).
//
// The condition says that T_c lies in no open interval (T_i, maxnb_i).  The
// device evaluates it without sorting: the field's time range [lo, hi] is cut
// into EXIT_NBIN bins; a difference array marks bins that one interval covers
// completely (no candidate there), a prefix maximum of the per-bin maxnb
// gives max{maxnb_i} over all earlier bins, and only the nodes of the few
// uncovered bins are checked exactly against the nodes of their own bin.  The
// bin index is a monotone function of the time, so "bin(a) < bin(b)" implies
// "a < b" and every pruning step is exact.
#pragma once
#include "eik_common.h"

namespace fmt2d {

constexpr int EXIT_NBIN = 8192;

// per-field accumulators (order-preserving keys, see ord_key)
struct ExitField {
  unsigned long long lo, hi;   // smallest / largest reached time
  unsigned long long smax;     // largest seed time
  unsigned long long texit;    // stop time (key of +inf: runs to the end)
  int count;                   // reached, non-fixed nodes
  int bad;                     // a seed is never reached
  int ncand;                   // nodes in uncovered bins
  int pad;
};

// a node of an uncovered bin
struct ExitEntry {
  double t, m;  // time, max time of its connected nodes
  int f, b;     // field, bin
};

// order-preserving map of doubles onto unsigned 64-bit keys
FMT_HD unsigned long long ord_key(double d) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, d);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
FMT_HD double ord_val(unsigned long long k) {
  const unsigned long long u = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  return __builtin_bit_cast(double, u);
}

FMT_HD double exit_scale(double lo, double hi) { return hi > lo ? (double)EXIT_NBIN / (hi - lo) : 0.0; }

// bin of time t; non-decreasing in t
FMT_HD int exit_bin(double t, double lo, double scale) {
  if (!(t > lo)) return 0;
  const double x = (t - lo) * scale;
  return x >= (double)(EXIT_NBIN - 1) ? EXIT_NBIN - 1 : (int)x;
}

// bins [exit_bin(t)+1, exit_cover_end) lie strictly inside (t, m)
FMT_HD int exit_cover_end(double m, double lo, double hi, double scale) {
  return m > hi ? EXIT_NBIN : exit_bin(m, lo, scale);
}

// A bin can hold the stop node only if no single interval covers it, its
// latest node is at or after every seed, and no earlier bin's connection
// reaches past that latest node.
FMT_HD bool exit_bin_open(int cnt, int cover, double mprev, double bmax, double smax) {
  return cnt > 0 && cover == 0 && bmax >= smax && mprev <= bmax;
}

}  // namespace fmt2d
