// The bulk front-to-back sum of fmt2d_amd/csrc/ordered_sum.h run on the CPU
// the way seqnorm_kernel runs it (chunks, increments summed as integers,
// the first term that breaks added the ordinary way, restart after it), for
// comparison with the plain loop s = s + x[k].
#include <math.h>
#include <stdint.h>

#include "ordered_sum.h"

using namespace fmt2d;

extern "C" double serial_sum(const double* x, long long n) {
  double s = 0.0;
  for (long long k = 0; k < n; k++) s = s + x[k];
  return s;
}

// *steps: additions made the ordinary way.  plant != 0 plants a mistake
// (for the test that the cases see it): 1 leaves out the binade check, 2
// rounds the terms that must be added the ordinary way as if their
// increment were 1 (right for some ties, wrong for the rest).
extern "C" double bulk_sum(const double* x, long long n, int chunk, long long* steps, int plant) {
  double acc = 0.0;
  long long serial = 0;
  for (long long c0 = 0; c0 < n; c0 += chunk) {
    const long long c1 = c0 + chunk < n ? c0 + chunk : n;
    long long start = c0;
    while (start < c1) {
      if (!isfinite(acc)) {
        for (long long k = start; k < n; k++) acc = acc + x[k];
        *steps = serial;
        return acc;
      }
      const SumBase B = sum_base(acc);
      const double room = sum_room(B);
      double P = 0.0;
      long long kb = -1;
      for (long long k = start; k < c1; k++) {
        double d;
        if (!sum_incr(x[k], B.e, &d)) {
          if (plant == 2 && x[k] > 0 && isfinite(x[k])) {
            d = 1.0;
          } else {
            kb = k;
            break;
          }
        }
        const double Q = P + d;
        if (plant != 1 && Q >= room) {
          kb = k;
          break;
        }
        P = Q;
      }
      if (kb < 0) {
        acc = sum_value(B.m + (uint64_t)P, B.e);
        break;
      }
      acc = sum_value(B.m + (uint64_t)P, B.e) + x[kb];
      serial++;
      start = kb + 1;
    }
  }
  *steps = serial;
  return acc;
}
