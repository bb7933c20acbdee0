// Host-side evaluation of the reference FMM's heap-exhaustion stop
// (src/propagate2d.f90:80) on a converged field; see max_member_time_grid in
// eik_update.h for the rule.
#pragma once
#include <vector>

namespace fmt2d {

// T: final times of the field's nodes (huge >= 0.5e20 = never reached),
// maxnb: largest time among each node's connected nodes (-1 none),
// seed: node is in the heap from the start, fixed: node is alive from the
// start and never pops.  Returns the pop time after which the reference stops,
// or +infinity when it runs to the end.
double heap_exit_time(const std::vector<double>& T, const std::vector<double>& maxnb,
                      const std::vector<char>& seed, const std::vector<char>& fixed);

}  // namespace fmt2d
