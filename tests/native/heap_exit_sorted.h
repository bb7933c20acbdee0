// Heap-exhaustion stop time by the plain sort-and-scan statement of the rule
// (pops in time order; the reference leaves its loop when a pop other than
// the first empties the heap, src/propagate2d.f90:74-80).  Test
// infrastructure: the checker of the binned device evaluation in
// fmt2d_amd/csrc/heap_exit.h and the stop rule of the CPU harness.
#pragma once
#include <algorithm>
#include <limits>
#include <vector>

namespace fmt2d_test {

// T: final times (>= 0.5e20: never reached), maxnb: largest time among each
// node's connected nodes (-1 none), seed: in the heap from the start, fixed:
// alive from the start, never pops.  Returns the time of the pop after which
// the heap is empty, or +inf when the march runs to the last node.
inline double heap_exit_time(const std::vector<double>& T, const std::vector<double>& maxnb,
                             const std::vector<char>& seed, const std::vector<char>& fixed) {
  const double inf = std::numeric_limits<double>::infinity();
  double smax = -inf;
  std::vector<int> idx;
  idx.reserve(T.size());
  for (size_t k = 0; k < T.size(); k++) {
    if (fixed[k]) continue;
    if (seed[k]) {
      if (!(T[k] < 0.5e20)) return inf;  // a seed that is never improved pops last
      smax = std::max(smax, T[k]);
    }
    if (T[k] < 0.5e20) idx.push_back((int)k);
  }
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return T[a] < T[b] || (T[a] == T[b] && a < b); });
  double h = -inf;  // largest connected-node time over the nodes popped so far
  for (size_t k = 0; k < idx.size(); k++) {
    const double t = T[idx[k]];
    const bool distinct = (k == 0 || T[idx[k - 1]] < t) && (k + 1 == idx.size() || T[idx[k + 1]] > t);
    if (k >= 1 && distinct && t >= smax && h <= t) return (k + 1 == idx.size()) ? inf : t;
    h = std::max(h, maxnb[idx[k]]);
  }
  return inf;
}

}  // namespace fmt2d_test
