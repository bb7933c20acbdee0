// Front-to-back FP64 sums of non-negative terms, s_k = fl(s_{k-1} + x_k),
// in bulk and bit for bit.
//
// Write the running sum as s = m * 2^e with an integer m < 2^53 (e the
// exponent of s's last mantissa bit, -1074 for zero and subnormal sums).
// While the exact sum stays below 2^53 * 2^e, every representable value the
// addition can round to is a multiple of 2^e, so fl(s + x) = (m + d) * 2^e
// with d = floor(x / 2^e) rounded up when the remainder is over half of
// 2^e: an increment that depends on x and e only, not on m.  Increments of
// consecutive terms therefore add as integers (in any order, exactly) until
// m + (their sum) would reach 2^53 (the sum moves to the next binade), or a
// term's remainder is exactly half (round-half-even then depends on m's
// parity), or a term is not a finite non-negative number.  That term is
// added the ordinary way and the bulk step starts again from it.
#pragma once
#include "eik_common.h"

namespace fmt2d {

constexpr uint64_t SUM_TOP = 1ull << 53;  // m must stay below this

FMT_HD uint64_t sum_bits(double x) { return __builtin_bit_cast(uint64_t, x); }

struct SumBase {
  uint64_t m;
  int e;
};

// s (finite, >= +0) as m * 2^e
FMT_HD SumBase sum_base(double s) {
  const uint64_t b = sum_bits(s);
  const int be = (int)(b >> 52);
  const uint64_t f = b & ((1ull << 52) - 1);
  if (be == 0) return SumBase{f, -1074};
  return SumBase{f | (1ull << 52), be - 1075};
}

// m * 2^e back to a double: m < 2^53, and m >= 2^52 unless e is -1074
// (a base from sum_base plus increments satisfies both)
FMT_HD double sum_value(uint64_t m, int e) {
  if (e == -1074) return __builtin_bit_cast(double, m);  // biased exponent 0 or 1
  return __builtin_bit_cast(double, ((uint64_t)(e + 1075) << 52) | (m - (1ull << 52)));
}

// The increment of m when x is added at exponent e, as a double (an integer
// below 2^53: sums of increments stay exact while they are below 2^53, and
// a sum that passes the limit compares as passed).  x * 2^-e is exact
// whenever it is 0.5 or more; below that the increment is 0 and no tie is
// possible.  False when the addition must be made the ordinary way: x
// negative, not finite, an exact half remainder, or an increment of
// 2^53 or more.
FMT_HD bool sum_incr(double x, int e, double* d) {
  // y < 0 only for x < 0; a negative x small enough for y to round to -0
  // is below half of 2^(e - 1074) and leaves s unchanged, like +0
  const double y = ldexp(x, -e);
  if (!(y >= 0.0 && y < 9007199254740992.0)) return false;
  const double fl = floor(y);
  const double f = y - fl;
  if (f == 0.5) return false;
  *d = f > 0.5 ? fl + 1.0 : fl;
  return true;
}

// 2^53 - m as a double (exact): the bulk sum of increments from base b must
// stay below it
FMT_HD double sum_room(const SumBase& b) { return (double)(SUM_TOP - b.m); }

}  // namespace fmt2d
