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

// propagate a non-zero status
#define RC(x)            \
  do {                   \
    int rc_ = (x);       \
    if (rc_) return rc_; \
  } while (0)

inline unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// XCD-aware tile order.  The dispatcher hands consecutive workgroups to the 8
// XCDs in turn, and each XCD has its own L2; read in launch order, neighbouring
// tiles land on different XCDs and every XCD fetches the shared halo rows from
// HBM.  Remapping so that each XCD owns one contiguous band of tiles lets the
// halo (and partially used cache lines) be served from that XCD's L2.
constexpr int NXCD = 8;
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
  const int q = nblk / NXCD, r = nblk % NXCD, x = bid % NXCD, k = bid / NXCD;
  return (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

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

__device__ __forceinline__ void basis_V_dr(double r, double s, double* d) {
  d[0] = 0.5 * (2.0 * r - 1.0) * 0.5 * s * (s - 1);
  d[1] = 0.5 * (2.0 * r + 1.0) * 0.5 * s * (s - 1);
  d[2] = 0.5 * (2.0 * r + 1.0) * 0.5 * s * (s + 1);
  d[3] = 0.5 * (2.0 * r - 1.0) * 0.5 * s * (s + 1);
  d[4] = (-2.0 * r) * 0.5 * s * (s - 1);
  d[5] = 0.5 * (2.0 * r + 1.0) * (1.0 - s * s);
  d[6] = (-2.0 * r) * 0.5 * s * (s + 1);
  d[7] = 0.5 * (2.0 * r - 1.0) * (1.0 - s * s);
  d[8] = (-2.0 * r) * (1.0 - s * s);
}

__device__ __forceinline__ void basis_V_dt(double r, double s, double* d) {
  d[0] = 0.5 * r * (r - 1.0) * 0.5 * (2.0 * s - 1.0);
  d[1] = 0.5 * r * (r + 1.0) * 0.5 * (2.0 * s - 1.0);
  d[2] = 0.5 * r * (r + 1.0) * 0.5 * (2.0 * s + 1.0);
  d[3] = 0.5 * r * (r - 1.0) * 0.5 * (2.0 * s + 1.0);
  d[4] = (1.0 - r * r) * 0.5 * (2.0 * s - 1.0);
  d[5] = 0.5 * r * (r + 1.0) * (-2.0 * s);
  d[6] = (1.0 - r * r) * 0.5 * (2.0 * s + 1.0);
  d[7] = 0.5 * r * (r - 1.0) * (-2.0 * s);
  d[8] = (1.0 - r * r) * (-2.0 * s);
}

__device__ __forceinline__ void basis_P(double r, double s, double* N) {
  N[0] = 0.25 * (1 - r) * (1 - s);
  N[1] = 0.25 * (1 + r) * (1 - s);
  N[2] = 0.25 * (1 + r) * (1 + s);
  N[3] = 0.25 * (1 - r) * (1 + s);
}

// ---------------------------------------------------------------------------
// structured box helpers: Q2 node grid (2*nelx+1) x (2*nelz+1), element
// e = ez*nelx + ex, local node numbering 3-6-2 / 7-8-5 / 0-4-1
// ---------------------------------------------------------------------------
__device__ __forceinline__ int q2_local(int lx, int lz) {
  // lz = 0: 0 4 1 ; lz = 1: 7 8 5 ; lz = 2: 3 6 2
  const int t[9] = {0, 4, 1, 7, 8, 5, 3, 6, 2};
  return t[lz * 3 + lx];
}

// reference local node b -> (lx, lz) on the element's 3x3 node grid
__device__ __forceinline__ void q2_pos(int b, int& lx, int& lz) {
  const int X[9] = {0, 2, 2, 0, 1, 2, 1, 0, 1};
  const int Z[9] = {0, 0, 2, 2, 0, 1, 2, 1, 1};
  lx = X[b];
  lz = Z[b];
}

// elements containing grid line I (0..2*nel) and the node's local position
// in each, in increasing element order; returns the count (1 or 2)
// Node k (reference order) of element iel on the row-major (2 nelx + 1)-wide
// node grid of a box mesh (build_velocity_connectivity.py:67-75)
__device__ __forceinline__ int grid_node(int iel, int nelx, int nnx, int k) {
  const int ez = iel / nelx, ex = iel - ez * nelx;
  const int base = 2 * ex + 2 * ez * nnx;
  constexpr int di[9] = {0, 2, 2, 0, 1, 2, 1, 0, 1}, dj[9] = {0, 0, 2, 2, 0, 1, 2, 1, 1};
  return base + di[k] + dj[k] * nnx;
}

__device__ __forceinline__ int q2_adjacent(int I, int nel, int* e, int* l) {
  if (I & 1) {
    e[0] = I >> 1;
    l[0] = 1;
    return 1;
  }
  int c = 0;
  int h = I >> 1;
  if (h - 1 >= 0) { e[c] = h - 1; l[c] = 2; c++; }
  if (h < nel) { e[c] = h; l[c] = 0; c++; }
  return c;
}

// The first INC4 entries of node n's incidence list (element << 5 | local
// node), read together, and the list's length: on a box mesh a node has at
// most four incident elements, so gathers over the list can issue every load
// before the first add (longer lists are left to the caller's loop).
constexpr int INC4 = 4;
__device__ __forceinline__ int inc_entries4(const int* __restrict__ ptr,
                                            const int* __restrict__ list, long n, int* ent) {
  const int i0 = ptr[n], c = ptr[n + 1] - i0;
#pragma unroll
  for (int q = 0; q < INC4; q++) ent[q] = q < c ? list[i0 + q] : 0;
  return c;
}

__device__ __forceinline__ double dot9(const double* a, const double* b) {
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 9; k++) s += a[k] * b[k];
  return s;
}

}  // namespace mw
