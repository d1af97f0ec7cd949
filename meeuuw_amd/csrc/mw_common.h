// Shared device helpers for the meeuuw HIP kernels (gfx950).
//
// Array layouts follow the Python interface exactly:
//   icon (m, nel) row-major: icon[k * nel + iel]
//   qpoint fields (nel, 9):  a[iel * 9 + iq]
//   reference tables (9, m): N[iq * m + k]
// The FE/PIC kernels are built with -ffp-contract=off so that every product
// and sum rounds exactly as written; together with IEEE division this makes
// them bit-identical to a sequential FP64 evaluation in the same order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "meeuuw_hip.h"

namespace mw {

// status codes (MW_OK, MW_EINVAL, ...) come from the public C header


void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define MW_HIP(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return ::mw::fail(MW_EHIP, std::string(#call) + ": " +          \
                                           hipGetErrorString(e_));          \
  } while (0)

#define MW_CHECK_LAUNCH() MW_HIP(hipGetLastError())

inline unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ---------------------------------------------------------------------------
// Q2 / Q1 basis functions, same written expressions as basis_functions.py:20-138
// ---------------------------------------------------------------------------
__device__ __forceinline__ void basis_V(double r, double s, double* N) {
  N[0] = 0.5 * r * (r - 1.0) * 0.5 * s * (s - 1.0);
  N[1] = 0.5 * r * (r + 1.0) * 0.5 * s * (s - 1.0);
  N[2] = 0.5 * r * (r + 1.0) * 0.5 * s * (s + 1.0);
  N[3] = 0.5 * r * (r - 1.0) * 0.5 * s * (s + 1.0);
  N[4] = (1.0 - r * r) * 0.5 * s * (s - 1.0);
  N[5] = 0.5 * r * (r + 1.0) * (1.0 - s * s);
  N[6] = (1.0 - r * r) * 0.5 * s * (s + 1.0);
  N[7] = 0.5 * r * (r - 1.0) * (1.0 - s * s);
  N[8] = (1.0 - r * r) * (1.0 - s * s);
}

__device__ __forceinline__ double dot9(const double* a, const double* b) {
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 9; k++) s += a[k] * b[k];
  return s;
}

}  // namespace mw
