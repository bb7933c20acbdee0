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

}  // namespace fmt2d
