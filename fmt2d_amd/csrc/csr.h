// CSR assembly of Frechet rows (device side), see csr.hip.
#pragma once
#include <hip/hip_runtime.h>

namespace fmt2d {

// Device scratch of the sorts below, kept between calls and only grown:
// slot k holds the largest request made of it so far (repeated assemblies of
// similar size then allocate nothing).  Held until release().
struct CsrWork {
  static constexpr int NSLOT = 10;
  void* p[NSLOT] = {};
  size_t bytes[NSLOT] = {};
  CsrWork() = default;
  CsrWork(const CsrWork&) = delete;
  CsrWork& operator=(const CsrWork&) = delete;
  template <typename T>
  hipError_t get(int slot, size_t n, T** out) {
    const size_t b = (n ? n : 1) * sizeof(T);
    if (bytes[slot] < b) {
      if (p[slot]) (void)hipFree(p[slot]);
      p[slot] = nullptr;
      bytes[slot] = 0;
      const hipError_t e = hipMalloc(&p[slot], b);
      if (e != hipSuccess) return e;
      bytes[slot] = b;
    }
    *out = static_cast<T*>(p[slot]);
    return hipSuccess;
  }
  void release() {
    for (int k = 0; k < NSLOT; k++) {
      if (p[k]) (void)hipFree(p[k]);
      p[k] = nullptr;
      bytes[k] = 0;
    }
  }
  ~CsrWork() { release(); }
};

struct CsrOut {
  int* rowptr = nullptr;   // nray + 1 (0-based offsets)
  int* col = nullptr;      // parameter index + base
  double* val = nullptr;
  int nnz = 0;
};

// keys: ray << 32 | (parameter - 1) in accumulation order, n pairs; sorted
// over key_bits bits.  Allocates out's arrays (caller frees with hipFree);
// its scratch comes from w.
hipError_t build_csr(unsigned long long* keys, double* vals, long long n, int nray, int key_bits, int base,
                     CsrOut& out, CsrWork& w, hipStream_t st);

// Products of CSR rows (0-based columns) with a sparse P (row per column):
// with keys == nullptr, cnt[r] = number of products of row r; otherwise row
// r writes (r << 32 | P column, value * weight) from off[r] on, its entries
// in stored order and each entry's P row in order.
hipError_t launch_prolong_rows(const int* rp, const int* col, const double* val, int nrow, const int* p_rp,
                               const int* p_col, const double* p_w, const long long* off, long long* cnt,
                               unsigned long long* keys, double* vals, hipStream_t st);

}  // namespace fmt2d
