// Serial walk through the binned heap-exhaustion evaluation of the device
// kernels (exit_stats / exit_bins / exit_scan / exit_gather / exit_check in
// fmt2d_amd/csrc/eik_kernels.hip, using the same helpers of heap_exit.h),
// checked against the sort-and-scan statement of the rule on random fields.
// Test infrastructure; prints "ok <ncases> <nfinite>" or the first mismatch.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <limits>
#include <random>
#include <vector>

#include "../../fmt2d_amd/csrc/heap_exit.h"
#include "heap_exit_sorted.h"

using namespace fmt2d;

static double binned(const std::vector<double>& T, const std::vector<double>& M, const std::vector<char>& seed,
                     const std::vector<char>& fixed) {
  const double inf = std::numeric_limits<double>::infinity();
  // exit_stats
  unsigned long long lo = ~0ull, hi = 0, smax = ord_key(-inf);
  int count = 0, bad = 0;
  for (size_t k = 0; k < T.size(); k++) {
    if (fixed[k]) continue;
    const bool listed = T[k] < 0.5e20;
    if (seed[k]) {
      if (!listed) bad = 1;
      else smax = std::max(smax, ord_key(T[k]));
    }
    if (listed) { lo = std::min(lo, ord_key(T[k])); hi = std::max(hi, ord_key(T[k])); count++; }
  }
  if (bad || count < 3) return inf;
  const double dlo = ord_val(lo), dhi = ord_val(hi), dsmax = ord_val(smax), sc = exit_scale(dlo, dhi);
  // exit_bins
  std::vector<int> cnt(EXIT_NBIN, 0), diff(EXIT_NBIN + 1, 0);
  std::vector<unsigned long long> bmax(EXIT_NBIN, ord_key(-inf)), bm(EXIT_NBIN, ord_key(-inf));
  for (size_t k = 0; k < T.size(); k++) {
    if (fixed[k] || !(T[k] < 0.5e20)) continue;
    const int b = exit_bin(T[k], dlo, sc);
    cnt[b]++;
    bmax[b] = std::max(bmax[b], ord_key(T[k]));
    bm[b] = std::max(bm[b], ord_key(M[k]));
    const int e = exit_cover_end(M[k], dlo, dhi, sc);
    if (e > b + 1) { diff[b + 1]++; diff[e]--; }
  }
  // exit_scan
  std::vector<char> open(EXIT_NBIN, 0);
  std::vector<double> mprev(EXIT_NBIN);
  int cover = 0;
  double run = -inf;
  for (int b = 0; b < EXIT_NBIN; b++) {
    cover += diff[b];
    mprev[b] = run;
    open[b] = exit_bin_open(cnt[b], cover, run, ord_val(bmax[b]), dsmax);
    run = std::max(run, ord_val(bm[b]));
  }
  // exit_gather + exit_check
  std::vector<std::vector<int>> members(EXIT_NBIN);
  for (size_t k = 0; k < T.size(); k++) {
    if (fixed[k] || !(T[k] < 0.5e20)) continue;
    const int b = exit_bin(T[k], dlo, sc);
    if (open[b]) members[b].push_back((int)k);
  }
  double te = inf;
  for (int b = 0; b < EXIT_NBIN; b++)
    for (int c : members[b]) {
      const double t = T[c];
      if (!(t >= dsmax && t > dlo && t < dhi)) continue;
      double mx = mprev[b];
      bool dup = false;
      for (int j : members[b]) {
        if (j == c) continue;
        if (T[j] == t) dup = true;
        if (T[j] < t) mx = std::max(mx, M[j]);
      }
      if (!dup && mx <= t) te = std::min(te, t);
    }
  return te;
}

int main(int argc, char** argv) {
  const int ncase = argc > 1 ? atoi(argv[1]) : 3000;
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  int nfinite = 0;
  for (int c = 0; c < ncase; c++) {
    const int n = 1 + (int)(U(rng) * (c % 7 == 0 ? 20000 : 300));
    const int mode = c % 5;
    std::vector<double> T(n), M(n);
    std::vector<char> seed(n, 0), fixed(n, 0);
    const double quant = (c % 3 == 0) ? 0.01 : 0.0;  // ties
    for (int k = 0; k < n; k++) {
      double t = U(rng) * 50.0;
      if (quant > 0) t = floor(t / quant) * quant;
      T[k] = t;
    }
    if (mode == 1) {  // a march along a chain: every node reaches the next one
      std::sort(T.begin(), T.end());
      for (int k = 0; k < n; k++) M[k] = (k + 1 < n) ? T[k + 1] : -1.0;
    } else if (mode == 2) {  // wide front with a narrow neck around t = 30
      for (int k = 0; k < n; k++) {
        const double reach = (T[k] > 28.0 && T[k] < 30.0) ? 0.0 : 3.0 * U(rng);
        M[k] = T[k] + reach;
      }
    } else {
      for (int k = 0; k < n; k++) {
        const double r = U(rng);
        M[k] = r < 0.05 ? -1.0 : (r < 0.08 ? 1.0e20 : T[k] + (mode == 3 ? 0.2 : 1.0) * U(rng) * U(rng));
        if (r > 0.97) M[k] = T[k] - U(rng);  // connections that are all earlier
      }
    }
    for (int k = 0; k < n; k++) {
      const double r = U(rng);
      if (r < 0.02) T[k] = 1.0e20;  // unreached
      if (r > 0.995) seed[k] = 1;
    }
    if (n > 3 && U(rng) < 0.5) fixed[(int)(U(rng) * n)] = 1;
    if (c % 4 == 0)  // seeds that are all reached
      for (int k = 0; k < n; k++) if (seed[k] && !(T[k] < 0.5e20)) seed[k] = 0;
    const double a = fmt2d_test::heap_exit_time(T, M, seed, fixed), b = binned(T, M, seed, fixed);
    if (!(a == b || (std::isinf(a) && std::isinf(b)))) {
      printf("mismatch case %d (n=%d mode=%d): sorted %.17g binned %.17g\n", c, n, mode, a, b);
      return 1;
    }
    if (!std::isinf(a)) nfinite++;
  }
  printf("ok %d %d\n", ncase, nfinite);
  return 0;
}
