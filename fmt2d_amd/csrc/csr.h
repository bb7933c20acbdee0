// CSR assembly of Frechet rows (device side), see csr.hip.
#pragma once
#include <hip/hip_runtime.h>

namespace fmt2d {

struct CsrOut {
  int* rowptr = nullptr;   // nray + 1 (0-based offsets)
  int* col = nullptr;      // parameter index + base
  double* val = nullptr;
  int nnz = 0;
};

// keys: ray << 32 | (parameter - 1) in accumulation order, n pairs; sorted
// over key_bits bits.  Allocates out's arrays (caller frees with hipFree).
hipError_t build_csr(unsigned long long* keys, double* vals, long long n, int nray, int key_bits, int base,
                     CsrOut& out, hipStream_t st);

// Products of CSR rows (0-based columns) with a sparse P (row per column):
// with keys == nullptr, cnt[r] = number of products of row r; otherwise row
// r writes (r << 32 | P column, value * weight) from off[r] on, its entries
// in stored order and each entry's P row in order.
hipError_t launch_prolong_rows(const int* rp, const int* col, const double* val, int nrow, const int* p_rp,
                               const int* p_col, const double* p_w, const long long* off, long long* cnt,
                               unsigned long long* keys, double* vals, hipStream_t st);

}  // namespace fmt2d
