// Geometric multigrid for the Q2 velocity block on a structured box mesh.
//
// Level l halves the element counts of level l-1.  The operator of level l is
// applied matrix-free: an element contributes eta_e * Khat_l, where Khat_l is
// the 18x18 stiffness of a unit-viscosity rectangle of that level (computed on
// the device with the same quadrature as the assembly).  Coarse viscosities are
// the means of their four children (rediscretisation).  Transfers are the
// exact Q2 interpolation between nested meshes (tensor product of 1-D
// quadratic interpolation) and its transpose.  Smoother: Chebyshev-Jacobi with
// the largest eigenvalue of D^-1 A estimated by power iteration on every
// update; coarsest level: dense inverse formed on a side stream (Gauss-Jordan
// in one workgroup's LDS up to 64 unknowns, blocked over the whole GPU above)
// and applied as a matrix-vector product.
// A V-cycle level runs four kernels: pre-smoothing steps 0 and 1 fused,
// residual fused with restriction, prolongation fused with post-smoothing
// step 0, and post-smoothing step 1 (degree 2); below the finest level each
// kernel is bound by launch and memory latency, not bandwidth, so kernel count
// and the number of dependent memory round trips per kernel set its cost.
//
// Every node-parallel kernel gathers from the (at most 4) elements around its
// node, in increasing element order: no atomics, deterministic results.
#include "mw_common.h"
#include "mw_solver.h"

namespace mw {

template <class T>
struct LV {  // kernel-side view of a level
  int nelx, nelz, nnx, nnz;
  const T* eta;
  const uint8_t* fix;
  const T* Kh;
  const T* diag;
  const T* invdiag;
};

template <class T>
static LV<T> view(const Level<T>& L) {
  return LV<T>{L.nelx, L.nelz, L.nnx, L.nnz, L.eta.p, L.fix.p, L.Khat.p, L.diag.p, L.invdiag.p};
}

// Tiled node-parallel operator kernels.  A 256-thread block owns a 16x16 tile
// of nodes and stages in LDS the 20x20 node patch its elements touch (masked
// x) and the viscosities of the covering elements.
constexpr int TT = 16;           // tile edge (nodes)
constexpr int TP = TT + 4;       // patch edge
constexpr int TE = TT / 2 + 3;   // covering elements per direction

enum { OP_APPLY = 0, OP_RESID = 1, OP_CHEB = 2 };

// XCD-aware tile order (xcd_remap, in the common header) for 2-D and 1-D grids
__device__ __forceinline__ void xcd_tile(int& tx, int& tz) {
  const int nbx = gridDim.x;
  const int lin = xcd_remap(blockIdx.y * nbx + blockIdx.x, gridDim.x * gridDim.y);
  tx = lin % nbx;
  tz = lin / nbx;
}
// 1-D grids: a contiguous range of blocks (rows of nodes) per XCD
__device__ __forceinline__ long xcd_index() {
  return (long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
}

// Two-component vector type of a node's (u, w) pair.
template <class T> struct V2;
template <> struct V2<float> { using t = float2; };
template <> struct V2<double> { using t = double2; };
// pair accessors: the vectors may be Krylov columns at any 4-byte offset
template <class T>
__device__ __forceinline__ typename V2<T>::t ld2(const T* p, long n) {
  return typename V2<T>::t{p[2 * n], p[2 * n + 1]};
}
template <class T>
__device__ __forceinline__ void st2(T* p, long n, T a, T b) {
  p[2 * n] = a;
  p[2 * n + 1] = b;
}
// Read-only tables (Khat, Gh) read through the constant address space: a
// wave-uniform index then always becomes a scalar load, whatever else the
// kernel writes.
template <class T>
using cptr = const __attribute__((address_space(4))) T*;
template <class T>
__device__ __forceinline__ cptr<T> as_const(const T* p) {
  return (cptr<T>)p;
}

// a * b + c with one rounding (the operator gathers' tap sums: half the VALU
// operations of a multiply and an add, the same in every kernel that forms a row)
__device__ __forceinline__ float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }

__device__ __forceinline__ double wave_sum64(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ unsigned fix2(const uint8_t* f, long n) {
  return (unsigned)f[2 * n] | ((unsigned)f[2 * n + 1] << 8);
}

// Operator row pair of node (I, J) of parity class (PX, PZ) from a staged
// patch: sx holds the masked (u, w) of node lines xo.., zo.. (PW wide), se the
// viscosities of elements ex0.., ez0.. (EW wide).  An even grid line touches the
// elements h-1 (local 2) and h (local 0), an odd one the element (I-1)/2 (local
// 1); elements outside the mesh are skipped.  With the parity fixed at compile
// time every Khat offset is a constant and the rows are scalar loads.  The
// fixed-dof override of the node's own row is left to the caller.
template <class T, int PX, int PZ, int PW, int EW>
__device__ __forceinline__ void gather_row(const LV<T>& L, const typename V2<T>::t (*sx)[PW],
                                           const T (*se)[EW], int I, int J, int xo, int zo,
                                           int ex0, int ez0, T& y0, T& y1) {
  using T2 = typename V2<T>::t;
  constexpr int NX = PX ? 1 : 2, NZ = PZ ? 1 : 2;
  // every patch and viscosity read first, so they wait together; then the
  // element sums in the reference order
  T2 v[NZ][NX][9];
  T et[NZ][NX];
  bool inside[NZ][NX];
#pragma unroll
  for (int iz = 0; iz < NZ; iz++) {
    const int ez = PZ ? (J - 1) >> 1 : (J >> 1) - 1 + iz;
#pragma unroll
    for (int ix = 0; ix < NX; ix++) {
      const int ex = PX ? (I - 1) >> 1 : (I >> 1) - 1 + ix;
      const int bx = 2 * ex - xo, bz = 2 * ez - zo;
#pragma unroll
      for (int mz = 0; mz < 3; mz++)
#pragma unroll
        for (int mx = 0; mx < 3; mx++) v[iz][ix][mz * 3 + mx] = sx[bz + mz][bx + mx];
      // an element outside the mesh is formed (its patch entries exist) but
      // not added: selects, not branches, so the Khat loads stay scalar
      inside[iz][ix] = ex >= 0 && ex < L.nelx && ez >= 0 && ez < L.nelz;
      et[iz][ix] = se[inside[iz][ix] ? ez - ez0 : 0][inside[iz][ix] ? ex - ex0 : 0];
    }
  }
  y0 = 0;
  y1 = 0;
#pragma unroll
  for (int iz = 0; iz < NZ; iz++) {
    const int lza = PZ ? 1 : (iz ? 0 : 2);
#pragma unroll
    for (int ix = 0; ix < NX; ix++) {
      const int lxa = PX ? 1 : (ix ? 0 : 2);
      const int a = q2_local(lxa, lza);
      const cptr<T> r0 = as_const(L.Kh) + (2 * a) * 18;
      const cptr<T> r1 = r0 + 18;
      T s0 = 0, s1 = 0;
#pragma unroll
      for (int mz = 0; mz < 3; mz++)
#pragma unroll
        for (int mx = 0; mx < 3; mx++) {
          const int bb = q2_local(mx, mz);
          const T2 w = v[iz][ix][mz * 3 + mx];
          s0 = fmad(r0[2 * bb + 1], w.y, fmad(r0[2 * bb], w.x, s0));
          s1 = fmad(r1[2 * bb + 1], w.y, fmad(r1[2 * bb], w.x, s1));
        }
      const T c0 = et[iz][ix] * s0, c1 = et[iz][ix] * s1;
      y0 = inside[iz][ix] ? y0 + c0 : y0;
      y1 = inside[iz][ix] ? y1 + c1 : y1;
    }
  }
}

// gather_row with the parity class chosen at run time (wave-uniform w)
template <class T, int PW, int EW>
__device__ __forceinline__ void gather_row_w(int w, const LV<T>& L,
                                             const typename V2<T>::t (*sx)[PW], const T (*se)[EW],
                                             int I, int J, int xo, int zo, int ex0, int ez0,
                                             T& y0, T& y1) {
  switch (w) {
    case 0: gather_row<T, 0, 0, PW, EW>(L, sx, se, I, J, xo, zo, ex0, ez0, y0, y1); break;
    case 1: gather_row<T, 1, 0, PW, EW>(L, sx, se, I, J, xo, zo, ex0, ez0, y0, y1); break;
    case 2: gather_row<T, 0, 1, PW, EW>(L, sx, se, I, J, xo, zo, ex0, ez0, y0, y1); break;
    default: gather_row<T, 1, 1, PW, EW>(L, sx, se, I, J, xo, zo, ex0, ez0, y0, y1); break;
  }
}

// One 16x16 node tile per 256-thread block.  Wave w of the block owns the 64
// nodes of one parity class (I - I0, J - J0) = (2*lx + (w & 1), 2*lz + (w >> 1)),
// lane = lz*8 + lx, so every lane of a wave has the same element slots around
// its node (2 or 1 per direction), the same local node number in each of them
// and therefore the same Khat rows: those are wave-uniform (scalar) loads and
// the only LDS reads of the gather are the 9 masked (u, w) pairs of each
// element.  Every global load of the block (own operands, x patch, flags,
// viscosities) is issued before the first wait.  The arithmetic of each row is
// the element-ordered sum of the reference formulation: elements in increasing
// order, taps in local (mz, mx) order, masked dofs contributing x = 0, so the
// result does not depend on the tiling.  This is the body for node tile
// (tX, tZ) of level L.
template <class T, int OP>
__device__ __forceinline__ void level_op_tile(const LV<T>& L, const T* __restrict__ x,
                                              const T* __restrict__ b, T* __restrict__ y,
                                              T* __restrict__ d, T* __restrict__ xout,
                                              const T* __restrict__ coef, int step, int zero_x,
                                              int tX, int tZ) {
  using T2 = typename V2<T>::t;
  __shared__ T2 sx[TP][TP];  // masked (u, w) of the patch
  __shared__ T se[TE][TE];
  const int tid = threadIdx.x;
  const int I0 = tX * TT, J0 = tZ * TT;
  const int ex0 = I0 > 0 ? (I0 - 1) >> 1 : 0, ez0 = J0 > 0 ? (J0 - 1) >> 1 : 0;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int px = w & 1, pz = w >> 1, lane = tid & 63;
  const int I = I0 + 2 * (lane & 7) + px, J = J0 + 2 * (lane >> 3) + pz;
  const bool valid = I < L.nnx && J < L.nnz;
  const long n = valid ? (long)J * L.nnx + I : 0;
  const bool need_x = !(OP == OP_CHEB && zero_x);
  // own operands
  T2 pb = {0, 0}, pi = {0, 0}, pd = {0, 0}, px0 = {0, 0};
  unsigned pf = 0;
  if (valid) {
    if (OP != OP_APPLY) pb = ld2(b, n);
    if (OP == OP_CHEB) {
      pi = ld2(L.invdiag, n);
      if (step > 0) pd = ld2<T>(d, n);
    }
    if (need_x) {
      px0 = ld2(x, n);
      pf = fix2(L.fix, n);
    }
  }
  // patch: entries tid and tid + 256 of the 20x20 node square around the tile
  T2 pv[2];
  unsigned pfl[2];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int i = tid + 256 * k;
    const int gI = I0 - 2 + i % TP, gJ = J0 - 2 + i / TP;
    pv[k] = T2{0, 0};
    pfl[k] = 0x0101;
    if (need_x && i < TP * TP && gI >= 0 && gI < L.nnx && gJ >= 0 && gJ < L.nnz) {
      const long m = (long)gJ * L.nnx + gI;
      pfl[k] = fix2(L.fix, m);
      pv[k] = ld2(x, m);
    }
  }
  T pe = 0;
  {
    const int eX = ex0 + tid % TE, eZ = ez0 + tid / TE;
    if (need_x && tid < TE * TE && eX < L.nelx && eZ < L.nelz) pe = L.eta[(long)eZ * L.nelx + eX];
  }
  if (need_x) {
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const int i = tid + 256 * k;
      if (i < TP * TP) {
        T2 v = pv[k];
        if (pfl[k] & 0xff) v.x = 0;
        if (pfl[k] >> 8) v.y = 0;
        sx[i / TP][i % TP] = v;
      }
    }
    if (tid < TE * TE) se[tid / TE][tid % TE] = pe;
  }
  __syncthreads();
  if (!valid) return;
  T y0 = 0, y1 = 0;
  if (need_x) {
    gather_row_w<T, TP, TE>(w, L, sx, se, I, J, I0 - 2, J0 - 2, ex0, ez0, y0, y1);
    if (pf & 0xff) y0 = L.diag[2 * n] * px0.x;
    if (pf >> 8) y1 = L.diag[2 * n + 1] * px0.y;
  }
  if (OP == OP_APPLY) {
    st2(y, n, y0, y1);
  } else if (OP == OP_RESID) {
    st2(y, n, pb.x - y0, pb.y - y1);
  } else {
    const T c1 = coef[2 * step], c2 = coef[2 * step + 1];
    T d0 = c2 * pi.x * (pb.x - y0);
    T d1 = c2 * pi.y * (pb.y - y1);
    if (step > 0) {
      d0 += c1 * pd.x;
      d1 += c1 * pd.y;
    }
    st2(d, n, d0, d1);
    st2(xout, n, px0.x + d0, px0.y + d1);
  }
}

// FINE only names the finest-level instance (profiles list it on its own)
template <class T, int OP, bool FINE = false>
__global__ void __launch_bounds__(256)
k_level_op(LV<T> L, const T* __restrict__ x, const T* __restrict__ b, T* __restrict__ y,
           T* __restrict__ d, T* __restrict__ xout, const T* __restrict__ coef, int step,
           int zero_x) {
  int tX, tZ;
  xcd_tile(tX, tZ);
  level_op_tile<T, OP>(L, x, b, y, d, xout, coef, step, zero_x, tX, tZ);
}

template <class T>
static dim3 tiles(const Level<T>& L) {
  return dim3((L.nnx + TT - 1) / TT, (L.nnz + TT - 1) / TT);
}

// Khat of a unit-viscosity hx*hz rectangle: sum_q B^T C B JxW_q (C = diag(2,2,1))
template <class T>
__global__ void k_khat(double hx, double hz, T* __restrict__ Kh) {
  __shared__ double dx[81], dz[81], jw[9];
  int t = threadIdx.x;
  if (t < 9) {
    const double q[3] = {-sqrt(3.0 / 5.0), 0.0, sqrt(3.0 / 5.0)};
    const double w[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    int iq = t / 3, jq = t % 3;
    double dr[9], ds[9];
    basis_V_dr(q[iq], q[jq], dr);
    basis_V_dt(q[iq], q[jq], ds);
    for (int k = 0; k < 9; k++) {
      dx[t * 9 + k] = (2.0 / hx) * dr[k];
      dz[t * 9 + k] = (2.0 / hz) * ds[k];
    }
    jw[t] = (hx / 2) * (hz / 2) * (w[iq] * w[jq]);
  }
  __syncthreads();
  for (int s = t; s < 324; s += blockDim.x) {
    int i = s / 18, j = s % 18;
    int a = i >> 1, ca = i & 1, b = j >> 1, cb = j & 1;
    double sum = 0.0;
    for (int qq = 0; qq < 9; qq++) {
      double xa = dx[qq * 9 + a], za = dz[qq * 9 + a], xb = dx[qq * 9 + b], zb = dz[qq * 9 + b];
      double v;
      if (ca == 0 && cb == 0) v = 2.0 * xa * xb + za * zb;
      else if (ca == 0 && cb == 1) v = za * xb;
      else if (ca == 1 && cb == 0) v = xa * zb;
      else v = 2.0 * za * zb + xa * xb;
      sum += v * jw[qq];
    }
    Kh[s] = (T)sum;
  }
}

// coarse viscosity from the four children: arithmetic (mode 0), geometric (1)
// or harmonic (2) mean; the arithmetic mean is the default, the others follow
// the weak phase across a sharp viscosity jump more closely
template <class T>
__global__ void k_coarsen_eta(int nelx_c, int nelz_c, const T* __restrict__ ef,
                              T* __restrict__ ec, int mode) {
  int E = blockIdx.x * blockDim.x + threadIdx.x;
  if (E >= nelx_c * nelz_c) return;
  int X = E % nelx_c, Z = E / nelx_c, nfx = 2 * nelx_c;
  long f = (long)(2 * Z) * nfx + 2 * X;
  const T a = ef[f], b = ef[f + 1], c = ef[f + nfx], d = ef[f + nfx + 1];
  if (mode == 1)
    ec[E] = (T)exp(0.25 * (log((double)a) + log((double)b) + log((double)c) + log((double)d)));
  else if (mode == 2)
    ec[E] = (T)(4.0 / (1.0 / (double)a + 1.0 / (double)b + 1.0 / (double)c + 1.0 / (double)d));
  else
    ec[E] = (T)0.25 * (a + b + c + d);
}

__global__ void k_eta_from_double(long n, const double* __restrict__ a, float* __restrict__ b) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// coarse fixed mask by injection from the coinciding fine node
__global__ void k_coarsen_fix(int nnx_c, int nnz_c, int nnx_f, const uint8_t* __restrict__ ff,
                              uint8_t* __restrict__ fc) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= (long)nnx_c * nnz_c) return;
  int i = (int)(n % nnx_c), j = (int)(n / nnx_c);
  long m = (long)(2 * j) * nnx_f + 2 * i;
  fc[2 * n] = ff[2 * m];
  fc[2 * n + 1] = ff[2 * m + 1];
}

// 1-D quadratic interpolation weights between nested Q2 lines.  Fine line
// index I lies in coarse element E = I>>2 at r = (I&3)/2 - 1; its weights on
// the coarse nodes 2E, 2E+1, 2E+2:
//   I & 3 = 0 (r = -1):   1,     0,    0
//   I & 3 = 1 (r = -1/2): 3/8,   3/4, -1/8
//   I & 3 = 2 (r =  0):   0,     1,    0
//   I & 3 = 3 (r = +1/2): -1/8,  3/4,  3/8
// (selects, not branches: lanes of a wavefront hold different I & 3)
__device__ __forceinline__ double pw(int I, int k) {
  const int r = I & 3;
  const double w0 = r == 0 ? 1.0 : (r == 1 ? 0.375 : (r == 2 ? 0.0 : -0.125));
  const double w1 = (r & 1) ? 0.75 : (r == 2 ? 1.0 : 0.0);
  const double w2 = r == 1 ? -0.125 : (r == 3 ? 0.375 : 0.0);
  return k == 0 ? w0 : (k == 1 ? w1 : w2);
}

// xf += P xc on free fine dofs (all 9 candidate loads issued together)
template <class T>
__global__ void k_prolong_add(int nnx_f, int nnz_f, int nnx_c, int nnz_c,
                              const uint8_t* __restrict__ fixf, const T* __restrict__ xc,
                              T* __restrict__ xf) {
  long n = xcd_index();
  if (n >= (long)nnx_f * nnz_f) return;
  int I = (int)(n % nnx_f), J = (int)(n / nnx_f);
  int bi = (I >> 2) * 2, bj = (J >> 2) * 2;
  double s0 = 0, s1 = 0;
#pragma unroll
  for (int kb = 0; kb < 3; kb++)
#pragma unroll
    for (int ka = 0; ka < 3; ka++) {
      double w = pw(I, ka) * pw(J, kb);
      int ci = bi + ka, cj = bj + kb;
      if (w != 0.0 && ci < nnx_c && cj < nnz_c) {
        long m = (long)cj * nnx_c + ci;
        s0 += w * (double)xc[2 * m];
        s1 += w * (double)xc[2 * m + 1];
      }
    }
  if (!fixf[2 * n]) xf[2 * n] += (T)s0;
  if (!fixf[2 * n + 1]) xf[2 * n + 1] += (T)s1;
}

// restriction weight of fine line index I = 2*i + o onto coarse index i
__device__ __forceinline__ double rw(int i, int o) {
  if (i & 1) return (o == 0) ? 1.0 : ((o == -1 || o == 1) ? 0.75 : 0.0);
  switch (o) {
    case 0: return 1.0;
    case -1: case 1: return 0.375;
    case -3: case 3: return -0.125;
    default: return 0.0;
  }
}

// 1-D restriction taps of a coarse line of parity P (offsets in increasing
// order and their weights; the same values as rw())
template <int P>
__device__ __forceinline__ void rtap(int k, int& o, double& w) {
  if (P) {
    o = k - 1;
    w = k == 1 ? 1.0 : 0.75;
  } else {
    const int off[5] = {-3, -1, 0, 1, 3};
    const double ww[5] = {-0.125, 0.375, 1.0, 0.375, -0.125};
    o = off[k];
    w = ww[k];
  }
}

constexpr int RT = 16;          // coarse tile edge (nodes)
constexpr int RP = 2 * RT + 5;  // fine patch edge: fine lines 2*i0-3 .. 2*(i0+RT-1)+3

// bc = P^T rf of coarse node (i, j), parity class (PX, PZ), from the masked
// fine patch (fixed and out-of-mesh entries hold 0, which adds nothing);
// taps in increasing offset order, j outer, as the untiled sum
template <class T, int PX, int PZ, int P = RP>
__device__ __forceinline__ void restrict_row(const typename V2<T>::t (*sr)[P], int li, int lj,
                                             double& s0, double& s1) {
  s0 = 0;
  s1 = 0;
#pragma unroll
  for (int b = 0; b < (PZ ? 3 : 5); b++) {
    int ob;
    double wb;
    rtap<PZ>(b, ob, wb);
#pragma unroll
    for (int a = 0; a < (PX ? 3 : 5); a++) {
      int oa;
      double wa;
      rtap<PX>(a, oa, wa);
      const double w = wa * wb;
      const auto v = sr[2 * lj + 3 + ob][2 * li + 3 + oa];
      s0 += w * (double)v.x;
      s1 += w * (double)v.y;
    }
  }
}

// Tiled restriction: one 16x16 coarse tile per block, the 37x37 fine patch it
// reads staged (masked) in LDS in one round trip of loads; wave w owns the
// coarse nodes of parity class (w & 1, w >> 1).
template <class T>
__global__ void __launch_bounds__(256)
k_restrict_tiled(int nnx_c, int nnz_c, int nnx_f, int nnz_f, const uint8_t* __restrict__ fixf,
                 const uint8_t* __restrict__ fixc, const T* __restrict__ rf, T* __restrict__ bc) {
  using T2 = typename V2<T>::t;
  __shared__ T2 sr[RP][RP];
  const int tid = threadIdx.x;
  int tX, tZ;
  xcd_tile(tX, tZ);
  const int i0 = tX * RT, j0 = tZ * RT;
  const int fi0 = 2 * i0 - 3, fj0 = 2 * j0 - 3;
  constexpr int NLD = (RP * RP + 255) / 256;
  T2 v[NLD];
  unsigned f[NLD];
#pragma unroll
  for (int k = 0; k < NLD; k++) {
    const int e = tid + 256 * k;
    const int gI = fi0 + e % RP, gJ = fj0 + e / RP;
    v[k] = T2{0, 0};
    f[k] = 0x0101;
    if (e < RP * RP && gI >= 0 && gI < nnx_f && gJ >= 0 && gJ < nnz_f) {
      const long m = (long)gJ * nnx_f + gI;
      f[k] = fix2(fixf, m);
      v[k] = ld2(rf, m);
    }
  }
#pragma unroll
  for (int k = 0; k < NLD; k++) {
    const int e = tid + 256 * k;
    if (e < RP * RP) {
      T2 u = v[k];
      if (f[k] & 0xff) u.x = 0;
      if (f[k] >> 8) u.y = 0;
      sr[e / RP][e % RP] = u;
    }
  }
  __syncthreads();
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int li = 2 * (lane & 7) + (w & 1), lj = 2 * (lane >> 3) + (w >> 1);
  const int i = i0 + li, j = j0 + lj;
  if (i >= nnx_c || j >= nnz_c) return;
  double s0, s1;
  switch (w) {
    case 0: restrict_row<T, 0, 0>(sr, li, lj, s0, s1); break;
    case 1: restrict_row<T, 1, 0>(sr, li, lj, s0, s1); break;
    case 2: restrict_row<T, 0, 1>(sr, li, lj, s0, s1); break;
    default: restrict_row<T, 1, 1>(sr, li, lj, s0, s1); break;
  }
  const long n = (long)j * nnx_c + i;
  const unsigned fc = fix2(fixc, n);
  st2(bc, n, (fc & 0xff) ? (T)0 : (T)s0, (fc >> 8) ? (T)0 : (T)s1);
}

template <class T>
static void launch_restrict(const Level<T>& L, const Level<T>& C, const T* rf, T* bc,
                            hipStream_t st) {
  const dim3 grid((C.nnx + RT - 1) / RT, (C.nnz + RT - 1) / RT);
  k_restrict_tiled<T><<<grid, 256, 0, st>>>(C.nnx, C.nnz, L.nnx, L.nnz, L.fix.p, C.fix.p, rf, bc);
}

// ------------------------------------------------------------------------
// Fused V-cycle kernels.  Each replaces two or three of the kernels above
// with the same arithmetic, so a V-cycle gives the same bits either way; they
// exist because below the finest level every kernel is pure launch and
// memory latency.

// Pre-smoothing from x = 0: Chebyshev steps 0 and 1 in one pass.  Step 0 is
// pointwise (d0 = c2_0 D^-1 b, x0 = 0 + d0), so the x0 patch that step 1
// gathers is formed from b and D^-1 while staging.
template <class T, bool FINE = false>
__global__ void __launch_bounds__(256)
k_cheb01(LV<T> L, const T* __restrict__ b, T* __restrict__ d, T* __restrict__ xout,
         const T* __restrict__ coef) {
  using T2 = typename V2<T>::t;
  __shared__ T2 sx[TP][TP];
  __shared__ T se[TE][TE];
  const int tid = threadIdx.x;
  int tX, tZ;
  xcd_tile(tX, tZ);
  const int I0 = tX * TT, J0 = tZ * TT;
  const int ex0 = I0 > 0 ? (I0 - 1) >> 1 : 0, ez0 = J0 > 0 ? (J0 - 1) >> 1 : 0;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int I = I0 + 2 * (lane & 7) + (w & 1), J = J0 + 2 * (lane >> 3) + (w >> 1);
  const bool valid = I < L.nnx && J < L.nnz;
  const long n = valid ? (long)J * L.nnx + I : 0;
  T2 pb = {0, 0}, pi = {0, 0};
  unsigned pf = 0;
  if (valid) {
    pb = ld2(b, n);
    pi = ld2(L.invdiag, n);
    pf = fix2(L.fix, n);
  }
  T2 vb[2], vi[2];
  unsigned vf[2];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int e = tid + 256 * k;
    const int gI = I0 - 2 + e % TP, gJ = J0 - 2 + e / TP;
    vb[k] = T2{0, 0};
    vi[k] = T2{0, 0};
    vf[k] = 0x0101;
    if (e < TP * TP && gI >= 0 && gI < L.nnx && gJ >= 0 && gJ < L.nnz) {
      const long m = (long)gJ * L.nnx + gI;
      vf[k] = fix2(L.fix, m);
      vb[k] = ld2(b, m);
      vi[k] = ld2(L.invdiag, m);
    }
  }
  T pe = 0;
  {
    const int eX = ex0 + tid % TE, eZ = ez0 + tid / TE;
    if (tid < TE * TE && eX < L.nelx && eZ < L.nelz) pe = L.eta[(long)eZ * L.nelx + eX];
  }
  const T c20 = coef[1], c11 = coef[2], c21 = coef[3];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int e = tid + 256 * k;
    if (e < TP * TP) {
      T2 v;
      v.x = (vf[k] & 0xff) ? (T)0 : (T)0 + c20 * vi[k].x * (vb[k].x - (T)0);
      v.y = (vf[k] >> 8) ? (T)0 : (T)0 + c20 * vi[k].y * (vb[k].y - (T)0);
      sx[e / TP][e % TP] = v;
    }
  }
  if (tid < TE * TE) se[tid / TE][tid % TE] = pe;
  __syncthreads();
  if (!valid) return;
  const T d00 = c20 * pi.x * (pb.x - (T)0), d01 = c20 * pi.y * (pb.y - (T)0);
  const T x00 = (T)0 + d00, x01 = (T)0 + d01;
  T y0, y1;
  gather_row_w<T, TP, TE>(w, L, sx, se, I, J, I0 - 2, J0 - 2, ex0, ez0, y0, y1);
  if (pf & 0xff) y0 = L.diag[2 * n] * x00;
  if (pf >> 8) y1 = L.diag[2 * n + 1] * x01;
  T d0 = c21 * pi.x * (pb.x - y0);
  T d1 = c21 * pi.y * (pb.y - y1);
  d0 += c11 * d00;
  d1 += c11 * d01;
  st2(d, n, d0, d1);
  st2(xout, n, x00 + d0, x01 + d1);
}

constexpr int CP = TT / 2 + 5;  // coarse patch edge behind a node tile's patch

// xf += P xc (free dofs) at fine node (I, J) from the staged coarse patch sc
// (origin ci0, cj0): the taps, their order and the skipped ones as k_prolong_add
template <class T, int P = CP>
__device__ __forceinline__ void prolong_at(const typename V2<T>::t (*sc)[P], int I, int J,
                                           int ci0, int cj0, int nnx_c, int nnz_c, double& s0,
                                           double& s1) {
  const int bi = (I >> 2) * 2, bj = (J >> 2) * 2;
  s0 = 0;
  s1 = 0;
#pragma unroll
  for (int kb = 0; kb < 3; kb++)
#pragma unroll
    for (int ka = 0; ka < 3; ka++) {
      const double w = pw(I, ka) * pw(J, kb);
      const int ci = bi + ka, cj = bj + kb;
      // every patch entry's taps lie inside the staged coarse patch, so the
      // read is unconditional and only the add is skipped
      const bool use = w != 0.0 && ci < nnx_c && cj < nnz_c;
      const auto v = sc[cj - cj0][ci - ci0];
      const double t0 = w * (double)v.x, t1 = w * (double)v.y;
      s0 = use ? s0 + t0 : s0;
      s1 = use ? s1 + t1 : s1;
    }
}

// Post-smoothing entry: xf += P xc on free dofs, then Chebyshev step 0 on the
// result.  The prolonged x is only ever read by that step, so it is formed in
// LDS (from a staged coarse patch) and never stored.
template <class T, bool FINE = false, int BS = 256>
__global__ void __launch_bounds__(BS)
k_prolong_cheb0(LV<T> L, int nnx_c, int nnz_c, const T* __restrict__ xc, const T* __restrict__ x,
                const T* __restrict__ b, T* __restrict__ d, T* __restrict__ xout,
                const T* __restrict__ coef) {
  using T2 = typename V2<T>::t;
  __shared__ T2 sx[TP][TP];
  __shared__ T2 sxu[TT][TT];  // unmasked prolonged x of the tile's own nodes
  __shared__ T2 sc[CP][CP];
  __shared__ T se[TE][TE];
  const int tid = threadIdx.x;
  int tX, tZ;
  xcd_tile(tX, tZ);
  const int I0 = tX * TT, J0 = tZ * TT;
  const int ex0 = I0 > 0 ? (I0 - 1) >> 1 : 0, ez0 = J0 > 0 ? (J0 - 1) >> 1 : 0;
  const int ci0 = I0 / 2 - 2, cj0 = J0 / 2 - 2;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int I = I0 + 2 * (lane & 7) + (w & 1), J = J0 + 2 * (lane >> 3) + (w >> 1);
  const bool valid = tid < 256 && I < L.nnx && J < L.nnz;  // waves 0-3 own the nodes
  const long n = valid ? (long)J * L.nnx + I : 0;
  T2 pb = {0, 0}, pi = {0, 0};
  unsigned pf = 0;
  if (valid) {
    pb = ld2(b, n);
    pi = ld2(L.invdiag, n);
    pf = fix2(L.fix, n);
  }
  constexpr int NXP = (TP * TP + BS - 1) / BS;
  T2 vx[NXP];
  unsigned vf[NXP];
#pragma unroll
  for (int k = 0; k < NXP; k++) {
    const int e = tid + BS * k;
    const int gI = I0 - 2 + e % TP, gJ = J0 - 2 + e / TP;
    vx[k] = T2{0, 0};
    vf[k] = 0x0101;
    if (e < TP * TP && gI >= 0 && gI < L.nnx && gJ >= 0 && gJ < L.nnz) {
      const long m = (long)gJ * L.nnx + gI;
      vf[k] = fix2(L.fix, m);
      vx[k] = ld2(x, m);
    }
  }
  T2 vc = {0, 0};
  {
    const int ci = ci0 + tid % CP, cj = cj0 + tid / CP;
    if (tid < CP * CP && ci >= 0 && ci < nnx_c && cj >= 0 && cj < nnz_c)
      vc = ld2(xc, (long)cj * nnx_c + ci);
  }
  T pe = 0;
  {
    const int eX = ex0 + tid % TE, eZ = ez0 + tid / TE;
    if (tid < TE * TE && eX < L.nelx && eZ < L.nelz) pe = L.eta[(long)eZ * L.nelx + eX];
  }
  if (tid < CP * CP) sc[tid / CP][tid % CP] = vc;
  if (tid < TE * TE) se[tid / TE][tid % TE] = pe;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NXP; k++) {
    const int e = tid + BS * k;
    const int gI = I0 - 2 + e % TP, gJ = J0 - 2 + e / TP;
    if (e < TP * TP) {
      T2 v = {0, 0};
      if (gI >= 0 && gI < L.nnx && gJ >= 0 && gJ < L.nnz) {
        double s0, s1;
        prolong_at<T>(sc, gI, gJ, ci0, cj0, nnx_c, nnz_c, s0, s1);
        v = vx[k];
        if (!(vf[k] & 0xff)) v.x += (T)s0;
        if (!(vf[k] >> 8)) v.y += (T)s1;
        const int pi_ = e % TP - 2, pj_ = e / TP - 2;
        if (pi_ >= 0 && pi_ < TT && pj_ >= 0 && pj_ < TT) sxu[pj_][pi_] = v;
        if (vf[k] & 0xff) v.x = 0;
        if (vf[k] >> 8) v.y = 0;
      }
      sx[e / TP][e % TP] = v;
    }
  }
  __syncthreads();
  if (!valid) return;
  const T x00 = sxu[J - J0][I - I0].x, x01 = sxu[J - J0][I - I0].y;
  T y0, y1;
  gather_row_w<T, TP, TE>(w, L, sx, se, I, J, I0 - 2, J0 - 2, ex0, ez0, y0, y1);
  if (pf & 0xff) y0 = L.diag[2 * n] * x00;
  if (pf >> 8) y1 = L.diag[2 * n + 1] * x01;
  const T c20 = coef[1];
  const T d0 = c20 * pi.x * (pb.x - y0);
  const T d1 = c20 * pi.y * (pb.y - y1);
  st2(d, n, d0, d1);
  st2(xout, n, x00 + d0, x01 + d1);
}

#ifndef PC_BS
#define PC_BS 512  // threads of k_prolong_cheb0: one patch entry per thread while staging
#endif
#ifndef PC_BS_FINE
#define PC_BS_FINE 256  // on the finest level, where 1089 tiles fill the GPU anyway
#endif

constexpr int RXW = RP + 4;  // x patch edge behind a residual patch
constexpr int REW = RT + 4;  // element patch edge behind a residual patch

// The residual rows of a 37x37 fine patch fall into four parity classes with
// 4, 2, 2 and 1 element slots per row; the 512-thread residual kernel gives
// them 3, 2, 2 and 1 waves so that no lane has more than 8 slots to gather.
constexpr int RR_BS = 512;
__device__ __forceinline__ void rr_class(int w, int& cls, int& first, int& nw) {
  if (w < 3) { cls = 0; first = 0; nw = 3; }
  else if (w < 5) { cls = 1; first = 3; nw = 2; }
  else if (w < 7) { cls = 2; first = 5; nw = 2; }
  else { cls = 3; first = 7; nw = 1; }
}
// residual rows of parity class (PX, PZ) given to lane ls of NW waves: rows
// q = ls + 64*NW*t of the class, t < NR (masked into sr: fixed and
// out-of-mesh entries 0); rb holds their b
template <class T, int PX, int PZ, int NW, int P = RP, int XW = RXW, int EW = REW>
struct ResidRows {
  static constexpr int NCX = PX ? (P + 1) / 2 : P / 2, NCZ = PZ ? (P + 1) / 2 : P / 2;
  static constexpr int NR = (NCX * NCZ + 64 * NW - 1) / (64 * NW);
  __device__ static __forceinline__ bool row(int ls, int t, int fi0, int fj0, int& ri, int& rj) {
    const int q = ls + 64 * NW * t;
    ri = 2 * (q % NCX) + (1 - PX);
    rj = 2 * (q / NCX) + (1 - PZ);
    return q < NCX * NCZ;
  }
  __device__ static __forceinline__ void rhs(const LV<T>& L, const T* __restrict__ b, int ls,
                                             int fi0, int fj0, typename V2<T>::t* rb) {
#pragma unroll
    for (int t = 0; t < NR; t++) {
      int ri, rj;
      rb[t] = typename V2<T>::t{0, 0};
      if (row(ls, t, fi0, fj0, ri, rj)) {
        const int I = fi0 + ri, J = fj0 + rj;
        if (I >= 0 && I < L.nnx && J >= 0 && J < L.nnz) rb[t] = ld2(b, (long)J * L.nnx + I);
      }
    }
  }
  __device__ static __forceinline__ void rows(const LV<T>& L, const typename V2<T>::t (*sx)[XW],
                                              const uint16_t (*sf)[XW], const T (*se)[EW],
                                              typename V2<T>::t (*sr)[P], int ls, int fi0,
                                              int fj0, const typename V2<T>::t* rb) {
#pragma unroll
    for (int t = 0; t < NR; t++) {
      int ri, rj;
      if (row(ls, t, fi0, fj0, ri, rj)) {
        const int I = fi0 + ri, J = fj0 + rj;
        typename V2<T>::t r = {0, 0};
        if (I >= 0 && I < L.nnx && J >= 0 && J < L.nnz) {
          T y0, y1;
          gather_row<T, PX, PZ, XW, EW>(L, sx, se, I, J, fi0 - 2, fj0 - 2, (fi0 >> 1) - 1,
                                        (fj0 >> 1) - 1, y0, y1);
          const unsigned f = sf[rj + 2][ri + 2];
          r.x = (f & 0xff) ? (T)0 : rb[t].x - y0;
          r.y = (f >> 8) ? (T)0 : rb[t].y - y1;
        }
        sr[rj][ri] = r;
      }
    }
  }
};

// Residual and restriction in one pass: bc = P^T (b - A x) over a 16x16
// coarse tile.  The 37x37 fine residual patch that the tile's restriction
// reads is computed in LDS from a 41x41 x patch (about a third more rows than
// the fine nodes the tile covers, in exchange for one kernel and one round
// trip of the fine residual through memory fewer).  Residual rows: 8 waves
// by parity class (rr_class); restriction: waves 0-3, one coarse parity class
// each.
template <class T, bool FINE = false>
__global__ void __launch_bounds__(RR_BS)
k_resid_restrict(LV<T> L, int nnx_c, int nnz_c, const uint8_t* __restrict__ fixc,
                 const T* __restrict__ x, const T* __restrict__ b, T* __restrict__ bc) {
  using T2 = typename V2<T>::t;
  __shared__ T2 sx[RXW][RXW];
  __shared__ uint16_t sf[RXW][RXW];
  __shared__ T se[REW][REW];
  __shared__ T2 sr[RP][RP];
  const int tid = threadIdx.x;
  int tX, tZ;
  xcd_tile(tX, tZ);
  const int i0 = tX * RT, j0 = tZ * RT;
  const int fi0 = 2 * i0 - 3, fj0 = 2 * j0 - 3;
  const int xo = fi0 - 2, zo = fj0 - 2, ex0 = (fi0 >> 1) - 1, ez0 = (fj0 >> 1) - 1;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  int cls, first, nw;
  rr_class(w, cls, first, nw);
  const int ls = (w - first) * 64 + lane;
  constexpr int NX = (RXW * RXW + RR_BS - 1) / RR_BS;
  T2 vx[NX];
  unsigned vf[NX];
#pragma unroll
  for (int k = 0; k < NX; k++) {
    const int e = tid + RR_BS * k;
    const int gI = xo + e % RXW, gJ = zo + e / RXW;
    vx[k] = T2{0, 0};
    vf[k] = 0x0101;
    if (e < RXW * RXW && gI >= 0 && gI < L.nnx && gJ >= 0 && gJ < L.nnz) {
      const long m = (long)gJ * L.nnx + gI;
      vf[k] = fix2(L.fix, m);
      vx[k] = ld2(x, m);
    }
  }
  T ve = 0;
  {
    const int eX = ex0 + tid % REW, eZ = ez0 + tid / REW;
    if (tid < REW * REW && eX >= 0 && eX < L.nelx && eZ >= 0 && eZ < L.nelz)
      ve = L.eta[(long)eZ * L.nelx + eX];
  }
  T2 rb[6];
  switch (cls) {
    case 0: ResidRows<T, 0, 0, 3>::rhs(L, b, ls, fi0, fj0, rb); break;
    case 1: ResidRows<T, 1, 0, 2>::rhs(L, b, ls, fi0, fj0, rb); break;
    case 2: ResidRows<T, 0, 1, 2>::rhs(L, b, ls, fi0, fj0, rb); break;
    default: ResidRows<T, 1, 1, 1>::rhs(L, b, ls, fi0, fj0, rb); break;
  }
#pragma unroll
  for (int k = 0; k < NX; k++) {
    const int e = tid + RR_BS * k;
    if (e < RXW * RXW) {
      T2 v = vx[k];
      if (vf[k] & 0xff) v.x = 0;
      if (vf[k] >> 8) v.y = 0;
      sx[e / RXW][e % RXW] = v;
      sf[e / RXW][e % RXW] = (uint16_t)vf[k];
    }
  }
  if (tid < REW * REW) se[tid / REW][tid % REW] = ve;
  __syncthreads();
  switch (cls) {
    case 0: ResidRows<T, 0, 0, 3>::rows(L, sx, sf, se, sr, ls, fi0, fj0, rb); break;
    case 1: ResidRows<T, 1, 0, 2>::rows(L, sx, sf, se, sr, ls, fi0, fj0, rb); break;
    case 2: ResidRows<T, 0, 1, 2>::rows(L, sx, sf, se, sr, ls, fi0, fj0, rb); break;
    default: ResidRows<T, 1, 1, 1>::rows(L, sx, sf, se, sr, ls, fi0, fj0, rb); break;
  }
  __syncthreads();
  if (w >= 4) return;
  const int li = 2 * (lane & 7) + (w & 1), lj = 2 * (lane >> 3) + (w >> 1);
  const int i = i0 + li, j = j0 + lj;
  if (i >= nnx_c || j >= nnz_c) return;
  double s0, s1;
  switch (w) {
    case 0: restrict_row<T, 0, 0>(sr, li, lj, s0, s1); break;
    case 1: restrict_row<T, 1, 0>(sr, li, lj, s0, s1); break;
    case 2: restrict_row<T, 0, 1>(sr, li, lj, s0, s1); break;
    default: restrict_row<T, 1, 1>(sr, li, lj, s0, s1); break;
  }
  const long nc = (long)j * nnx_c + i;
  const unsigned fc = fix2(fixc, nc);
  st2(bc, nc, (fc & 0xff) ? (T)0 : (T)s0, (fc >> 8) ? (T)0 : (T)s1);
}

// The same residual and restriction over an 8x8 coarse tile (a 21x21 fine
// residual patch from a 25x25 x patch): four times the tiles with 512 threads
// each, two waves per parity class, so no lane forms more than one residual
// row.  On the latency-bound levels a tile's chain is what the launch waits
// for; the rows, their order and the restriction sums are those of
// k_resid_restrict.
constexpr int R8T = 8, R8P = 2 * R8T + 5, R8XW = R8P + 4, R8EW = R8T + 4;
#ifndef RR8_LEVELS
#define RR8_LEVELS 2  // 8x8 tiles on: 0 no level, 1 the levels below the finest, 2 every level
#endif
template <class T, bool FINE = false>
__global__ void __launch_bounds__(RR_BS)
k_resid_restrict8(LV<T> L, int nnx_c, int nnz_c, const uint8_t* __restrict__ fixc,
                  const T* __restrict__ x, const T* __restrict__ b, T* __restrict__ bc) {
  using T2 = typename V2<T>::t;
  __shared__ T2 sx[R8XW][R8XW];
  __shared__ uint16_t sf[R8XW][R8XW];
  __shared__ T se[R8EW][R8EW];
  __shared__ T2 sr[R8P][R8P];
  const int tid = threadIdx.x;
  int tX, tZ;
  xcd_tile(tX, tZ);
  const int i0 = tX * R8T, j0 = tZ * R8T;
  const int fi0 = 2 * i0 - 3, fj0 = 2 * j0 - 3;
  const int xo = fi0 - 2, zo = fj0 - 2, ex0 = (fi0 >> 1) - 1, ez0 = (fj0 >> 1) - 1;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int cls = w >> 1, ls = (w & 1) * 64 + lane;
  constexpr int NX = (R8XW * R8XW + RR_BS - 1) / RR_BS;
  T2 vx[NX];
  unsigned vf[NX];
#pragma unroll
  for (int k = 0; k < NX; k++) {
    const int e = tid + RR_BS * k;
    const int gI = xo + e % R8XW, gJ = zo + e / R8XW;
    vx[k] = T2{0, 0};
    vf[k] = 0x0101;
    if (e < R8XW * R8XW && gI >= 0 && gI < L.nnx && gJ >= 0 && gJ < L.nnz) {
      const long m = (long)gJ * L.nnx + gI;
      vf[k] = fix2(L.fix, m);
      vx[k] = ld2(x, m);
    }
  }
  T ve = 0;
  {
    const int eX = ex0 + tid % R8EW, eZ = ez0 + tid / R8EW;
    if (tid < R8EW * R8EW && eX >= 0 && eX < L.nelx && eZ >= 0 && eZ < L.nelz)
      ve = L.eta[(long)eZ * L.nelx + eX];
  }
  T2 rb[1];
  switch (cls) {
    case 0: ResidRows<T, 0, 0, 2, R8P, R8XW, R8EW>::rhs(L, b, ls, fi0, fj0, rb); break;
    case 1: ResidRows<T, 1, 0, 2, R8P, R8XW, R8EW>::rhs(L, b, ls, fi0, fj0, rb); break;
    case 2: ResidRows<T, 0, 1, 2, R8P, R8XW, R8EW>::rhs(L, b, ls, fi0, fj0, rb); break;
    default: ResidRows<T, 1, 1, 2, R8P, R8XW, R8EW>::rhs(L, b, ls, fi0, fj0, rb); break;
  }
#pragma unroll
  for (int k = 0; k < NX; k++) {
    const int e = tid + RR_BS * k;
    if (e < R8XW * R8XW) {
      T2 v = vx[k];
      if (vf[k] & 0xff) v.x = 0;
      if (vf[k] >> 8) v.y = 0;
      sx[e / R8XW][e % R8XW] = v;
      sf[e / R8XW][e % R8XW] = (uint16_t)vf[k];
    }
  }
  if (tid < R8EW * R8EW) se[tid / R8EW][tid % R8EW] = ve;
  __syncthreads();
  switch (cls) {
    case 0: ResidRows<T, 0, 0, 2, R8P, R8XW, R8EW>::rows(L, sx, sf, se, sr, ls, fi0, fj0, rb); break;
    case 1: ResidRows<T, 1, 0, 2, R8P, R8XW, R8EW>::rows(L, sx, sf, se, sr, ls, fi0, fj0, rb); break;
    case 2: ResidRows<T, 0, 1, 2, R8P, R8XW, R8EW>::rows(L, sx, sf, se, sr, ls, fi0, fj0, rb); break;
    default: ResidRows<T, 1, 1, 2, R8P, R8XW, R8EW>::rows(L, sx, sf, se, sr, ls, fi0, fj0, rb); break;
  }
  __syncthreads();
  if (w >= 4 || lane >= 16) return;
  const int li = 2 * (lane & 3) + (w & 1), lj = 2 * (lane >> 2) + (w >> 1);
  const int i = i0 + li, j = j0 + lj;
  if (i >= nnx_c || j >= nnz_c) return;
  double s0, s1;
  switch (w) {
    case 0: restrict_row<T, 0, 0, R8P>(sr, li, lj, s0, s1); break;
    case 1: restrict_row<T, 1, 0, R8P>(sr, li, lj, s0, s1); break;
    case 2: restrict_row<T, 0, 1, R8P>(sr, li, lj, s0, s1); break;
    default: restrict_row<T, 1, 1, R8P>(sr, li, lj, s0, s1); break;
  }
  const long nc = (long)j * nnx_c + i;
  const unsigned fc = fix2(fixc, nc);
  st2(bc, nc, (fc & 0xff) ? (T)0 : (T)s0, (fc >> 8) ? (T)0 : (T)s1);
}

// ------------------------------------------------------------------------
// The bottom of the V-cycle in one workgroup: the last smoothed level (at
// most 16x16 nodes, one tile) and the dense coarsest solve.  Pre-smoothing
// steps 0 and 1, residual, restriction, x_c = A_c^-1 b_c, prolongation and
// post-smoothing steps 0 and 1 run as LDS phases of one kernel, each with
// the arithmetic of the kernel it replaces (k_cheb01, k_resid_restrict,
// k_dense_mv, k_prolong_cheb0, k_level_op), so the V-cycle is bit-identical;
// it saves four launches per V-cycle where each of them is pure latency.
// Wave w owns the nodes of parity class w (one node per lane: a class of a
// 16x16 grid has at most 64 nodes).  Chebyshev degree 2 only.
constexpr int BOT_MAX_NC = 64;  // coarsest unknowns held in LDS

template <class T, int PX, int PZ>
struct BotRows {
  // node of lane q in class (PX, PZ) of an nnx x nnz grid; false past the class
  __device__ static __forceinline__ bool node(int q, int nnx, int nnz, int& I, int& J) {
    const int ncx = (nnx - PX + 1) / 2, ncz = (nnz - PZ + 1) / 2;
    if (q >= ncx * ncz) return false;
    I = 2 * (q % ncx) + PX;
    J = 2 * (q / ncx) + PZ;
    return true;
  }
};

template <class T>
__global__ void __launch_bounds__(256)
k_vcycle_bottom(LV<T> L, int nnx_c, int nnz_c, const uint8_t* __restrict__ fixc, int nc,
                const T* __restrict__ cinv, const T* __restrict__ b, T* __restrict__ xout,
                const T* __restrict__ coef) {
  using T2 = typename V2<T>::t;
  __shared__ T2 sx[TP][TP];          // masked x patch (lines -2 ..), reused per phase
  __shared__ T se[TE][TE];
  __shared__ T2 sb[TT][TT], si[TT][TT], sdg[TT][TT];  // b, 1/diag, diag per node
  __shared__ uint16_t sfx[TT][TT];
  __shared__ T2 su[TT][TT], sd[TT][TT];  // unmasked x and d of the current step per node
  __shared__ T2 sr[RP][RP];          // masked residual (lines -3 ..)
  __shared__ T2 sc[CP][CP];          // coarse x (lines -2 ..)
  __shared__ T sbc[BOT_MAX_NC];      // coarse b, dof order
  __shared__ T sinv[BOT_MAX_NC * BOT_MAX_NC];
  __shared__ double smv[BOT_MAX_NC][64];  // per-row lane terms of the dense product
  const int tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  // ---- phase 1: everything into LDS
  for (int i = tid; i < RP * RP; i += 256) sr[i / RP][i % RP] = T2{0, 0};
  for (int i = tid; i < CP * CP; i += 256) sc[i / CP][i % CP] = T2{0, 0};
  {  // all loads of the inverse in flight at once (indices past nc*nc re-read the last entry)
    constexpr int NI = BOT_MAX_NC * BOT_MAX_NC / 256;
    const int last = nc * nc - 1;
    T vi[NI];
#pragma unroll
    for (int k = 0; k < NI; k++) vi[k] = cinv[min(tid + 256 * k, last)];
#pragma unroll
    for (int k = 0; k < NI; k++)
      if (tid + 256 * k <= last) sinv[tid + 256 * k] = vi[k];
  }
  {
    const int eX = tid % TE, eZ = tid / TE;
    T pe = 0;
    if (tid < TE * TE && eX < L.nelx && eZ < L.nelz) pe = L.eta[(long)eZ * L.nelx + eX];
    if (tid < TE * TE) se[eZ][eX] = pe;
  }
  {
    const int I = tid % TT, J = tid / TT;
    T2 vb = {0, 0}, vi = {0, 0}, vd = {0, 0};
    unsigned vf = 0x0101;
    if (I < L.nnx && J < L.nnz) {
      const long n = (long)J * L.nnx + I;
      vb = ld2(b, n);
      vi = ld2(L.invdiag, n);
      vd = ld2(L.diag, n);
      vf = fix2(L.fix, n);
    }
    sb[J][I] = vb;
    si[J][I] = vi;
    sdg[J][I] = vd;
    sfx[J][I] = (uint16_t)vf;
  }
  const T c20 = coef[1], c11 = coef[2], c21 = coef[3];
  __syncthreads();
  // ---- phase 2: pre-smoothing step 0 (pointwise) into the patch, then step 1
  for (int i = tid; i < TP * TP; i += 256) {
    const int I = i % TP - 2, J = i / TP - 2;
    T2 v = {0, 0};
    if (I >= 0 && I < L.nnx && J >= 0 && J < L.nnz) {
      const unsigned f = sfx[J][I];
      const T2 pb = sb[J][I], pi = si[J][I];
      v.x = (f & 0xff) ? (T)0 : (T)0 + c20 * pi.x * (pb.x - (T)0);
      v.y = (f >> 8) ? (T)0 : (T)0 + c20 * pi.y * (pb.y - (T)0);
    }
    sx[i / TP][i % TP] = v;
  }
  __syncthreads();
  int I = 0, J = 0;
  bool mine = false;
  switch (w) {
    case 0: mine = BotRows<T, 0, 0>::node(lane, L.nnx, L.nnz, I, J); break;
    case 1: mine = BotRows<T, 1, 0>::node(lane, L.nnx, L.nnz, I, J); break;
    case 2: mine = BotRows<T, 0, 1>::node(lane, L.nnx, L.nnz, I, J); break;
    default: mine = BotRows<T, 1, 1>::node(lane, L.nnx, L.nnz, I, J); break;
  }
  unsigned pf = 0;
  T2 pb = {0, 0}, pi = {0, 0}, pdg = {0, 0};
  if (mine) {
    pf = sfx[J][I];
    pb = sb[J][I];
    pi = si[J][I];
    pdg = sdg[J][I];
  }
  T2 x1 = {0, 0};
  if (mine) {
    const T d00 = c20 * pi.x * (pb.x - (T)0), d01 = c20 * pi.y * (pb.y - (T)0);
    const T x00 = (T)0 + d00, x01 = (T)0 + d01;
    T y0, y1;
    gather_row_w<T, TP, TE>(w, L, sx, se, I, J, -2, -2, 0, 0, y0, y1);
    if (pf & 0xff) y0 = pdg.x * x00;
    if (pf >> 8) y1 = pdg.y * x01;
    T d0 = c21 * pi.x * (pb.x - y0);
    T d1 = c21 * pi.y * (pb.y - y1);
    d0 += c11 * d00;
    d1 += c11 * d01;
    x1 = T2{x00 + d0, x01 + d1};
  }
  __syncthreads();  // every step-1 gather has read sx
  // ---- phase 3: residual of x1 (masked x1 in the patch)
  for (int i = tid; i < TP * TP; i += 256) sx[i / TP][i % TP] = T2{0, 0};
  __syncthreads();
  if (mine) {
    T2 v = x1;
    if (pf & 0xff) v.x = 0;
    if (pf >> 8) v.y = 0;
    sx[J + 2][I + 2] = v;
    su[J][I] = x1;
  }
  __syncthreads();
  if (mine) {
    T y0, y1;
    gather_row_w<T, TP, TE>(w, L, sx, se, I, J, -2, -2, 0, 0, y0, y1);
    T2 r;
    r.x = (pf & 0xff) ? (T)0 : pb.x - y0;
    r.y = (pf >> 8) ? (T)0 : pb.y - y1;
    sr[J + 3][I + 3] = r;
  }
  __syncthreads();
  // ---- phase 4: restriction to the coarsest level
  {
    int i = 0, j = 0;
    bool cm = false;
    switch (w) {
      case 0: cm = BotRows<T, 0, 0>::node(lane, nnx_c, nnz_c, i, j); break;
      case 1: cm = BotRows<T, 1, 0>::node(lane, nnx_c, nnz_c, i, j); break;
      case 2: cm = BotRows<T, 0, 1>::node(lane, nnx_c, nnz_c, i, j); break;
      default: cm = BotRows<T, 1, 1>::node(lane, nnx_c, nnz_c, i, j); break;
    }
    if (cm) {
      double s0, s1;
      switch (w) {
        case 0: restrict_row<T, 0, 0>(sr, i, j, s0, s1); break;
        case 1: restrict_row<T, 1, 0>(sr, i, j, s0, s1); break;
        case 2: restrict_row<T, 0, 1>(sr, i, j, s0, s1); break;
        default: restrict_row<T, 1, 1>(sr, i, j, s0, s1); break;
      }
      const long nn = (long)j * nnx_c + i;
      const unsigned fc = fix2(fixc, nn);
      sbc[2 * nn] = (fc & 0xff) ? (T)0 : (T)s0;
      sbc[2 * nn + 1] = (fc >> 8) ? (T)0 : (T)s1;
    }
  }
  __syncthreads();
  // ---- phase 5: x_c = A_c^-1 b_c with the sums of k_dense_mv (lane l of a
  // row takes the terms j = l, l+64, ..; then the shuffle-down tree, here
  // level by level in LDS over all rows at once)
  for (int i = tid; i < nc * 64; i += 256) {
    const int row = i >> 6, l = i & 63;
    double s = 0.0;
    for (int j = l; j < nc; j += 64) s += (double)sinv[row * nc + j] * (double)sbc[j];
    smv[row][l] = s;
  }
  __syncthreads();
  for (int o = 32; o > 0; o >>= 1) {
    for (int i = tid; i < nc * o; i += 256) {
      const int row = i / o, l = i % o;
      smv[row][l] += smv[row][l + o];
    }
    __syncthreads();
  }
  for (int row = tid; row < nc; row += 256) {
    const double s = smv[row][0];
    const int cn = row >> 1, ci = cn % nnx_c, cj = cn / nnx_c;
    if (row & 1) sc[cj + 2][ci + 2].y = (T)s;
    else sc[cj + 2][ci + 2].x = (T)s;
  }
  __syncthreads();
  // ---- phase 6: prolongation into the patch (x1 + P x_c on free dofs), step 0
  for (int i = tid; i < TP * TP; i += 256) {
    const int gI = i % TP - 2, gJ = i / TP - 2;
    T2 v = {0, 0};
    if (gI >= 0 && gI < L.nnx && gJ >= 0 && gJ < L.nnz) {
      double p0, p1;
      prolong_at<T>(sc, gI, gJ, -2, -2, nnx_c, nnz_c, p0, p1);
      const unsigned f = sfx[gJ][gI];
      v = su[gJ][gI];
      if (!(f & 0xff)) v.x += (T)p0;
      if (!(f >> 8)) v.y += (T)p1;
      sd[gJ][gI] = v;  // unmasked prolonged x (sd is free until step 0 below)
      if (f & 0xff) v.x = 0;
      if (f >> 8) v.y = 0;
    }
    sx[i / TP][i % TP] = v;
  }
  __syncthreads();
  T2 x2 = {0, 0}, d0v = {0, 0};
  if (mine) {
    const T2 xp = sd[J][I];
    T y0, y1;
    gather_row_w<T, TP, TE>(w, L, sx, se, I, J, -2, -2, 0, 0, y0, y1);
    if (pf & 0xff) y0 = pdg.x * xp.x;
    if (pf >> 8) y1 = pdg.y * xp.y;
    const T d0 = c20 * pi.x * (pb.x - y0);
    const T d1 = c20 * pi.y * (pb.y - y1);
    d0v = T2{d0, d1};
    x2 = T2{xp.x + d0, xp.y + d1};
  }
  __syncthreads();
  // ---- phase 7: post-smoothing step 1 from the masked x2 patch
  for (int i = tid; i < TP * TP; i += 256) sx[i / TP][i % TP] = T2{0, 0};
  __syncthreads();
  if (mine) {
    T2 v = x2;
    if (pf & 0xff) v.x = 0;
    if (pf >> 8) v.y = 0;
    sx[J + 2][I + 2] = v;
  }
  __syncthreads();
  if (!mine) return;
  T y0, y1;
  gather_row_w<T, TP, TE>(w, L, sx, se, I, J, -2, -2, 0, 0, y0, y1);
  if (pf & 0xff) y0 = pdg.x * x2.x;
  if (pf >> 8) y1 = pdg.y * x2.y;
  T d0 = c21 * pi.x * (pb.x - y0);
  T d1 = c21 * pi.y * (pb.y - y1);
  d0 += c11 * d0v.x;
  d1 += c11 * d0v.y;
  st2(xout, (long)J * L.nnx + I, x2.x + d0, x2.y + d1);
}

// ------------------------------------------------------------------------
// Multi-level bottom of the V-cycle: the last (up to three) smoothed levels
// and the dense coarsest solve as one 1024-thread workgroup, every vector in
// LDS.  The fused levels fill fixed slots of at most 33x33, 17x17 and 9x9
// nodes (the 16x16, 8x8 and 4x4 element grids of a halving hierarchy), the
// last fused level sits in the last slot.  Each sweep is a phase of the
// one-level bottom kernel above, generalised to a whole level per phase:
// every wave owns nodes of one parity class (ml_class), and the arithmetic of
// every sweep is the per-level kernels' (pre-smoothing
// steps 0+1, residual, restriction, prolongation + step 0, step 1): the
// V-cycle is bit-identical either way.  The levels share one row pitch and
// run through one copy of each phase (a run-time loop), which keeps the code
// inside the instruction cache.
constexpr int ML_BS = 1024;
constexpr int ML_PW = 37;  // masked x patch pitch: node lines -2 .. 34
constexpr int ML_RW = 39;  // residual patch pitch: fine lines -3 .. 35
constexpr int ML_CW = 21;  // coarse x patch pitch: coarse lines -2 .. 18
constexpr int ML_EW = 16;  // element row pitch
constexpr int ML_NP = 33;  // node row pitch of every slot
constexpr int ML_N0 = 33, ML_N1 = 17, ML_N2 = 9;  // slot node edges
constexpr int ML_MAXLEV = 3;
constexpr int ML_ROWS = ML_N0 + ML_N1 + ML_N2;             // node rows of all slots
constexpr int ML_EROWS = (ML_N0 + ML_N1 + ML_N2 - 3) / 2;  // element rows of all slots

template <class T>
struct MlArgs {
  int nlev;                 // fused smoothed levels (1..3), in slots 3-nlev .. 2
  LV<T> L[ML_MAXLEV];       // by slot
  const T* coef[ML_MAXLEV];
  const T* b;               // right-hand side of the first fused level
  T* x;                     // its result
  int nnx_c, nnz_c, nc;     // coarsest level
  const uint8_t* fixc;
  const T* cinv;
};

template <class T>
struct MlView {  // one level's LDS arrays (rows of the shared slot arrays)
  typename V2<T>::t (*b)[ML_NP];
  typename V2<T>::t (*id)[ML_NP];
  typename V2<T>::t (*dg)[ML_NP];
  typename V2<T>::t (*x1)[ML_NP];  // x after pre-smoothing
  uint16_t (*fx)[ML_NP];
  T (*e)[ML_EW];
  LV<T> L;  // the level (Khat read through the scalar cache)
  int nelx, nelz, nnx, nnz;
  T c20, c11, c21;
};

// Waves per parity class, in proportion to the work of a node's row: a corner
// node (class 0) sums 4 elements, an edge node 2, a centre node 1.
__device__ __forceinline__ void ml_class(int w, int& c, int& s, int& nw) {
  if (w < 7) { c = 0; s = w; nw = 7; }
  else if (w < 11) { c = 1; s = w - 7; nw = 4; }
  else if (w < 14) { c = 2; s = w - 11; nw = 3; }
  else { c = 3; s = w - 14; nw = 2; }
}

// f(I, J) for the nodes of parity class c owned by wave s of the nw waves of
// that class (at most 289 nodes per class: three passes of 2 waves suffice)
template <class F>
__device__ __forceinline__ void ml_nodes(int c, int s, int nw, int lane, int nnx, int nnz, F f) {
#pragma unroll 1
  for (int k = 0; k < 3; k++) {
    const int q = s * 64 + lane + nw * 64 * k;
    int I = 0, J = 0;
    bool mine;
    switch (c) {
      case 0: mine = BotRows<float, 0, 0>::node(q, nnx, nnz, I, J); break;
      case 1: mine = BotRows<float, 1, 0>::node(q, nnx, nnz, I, J); break;
      case 2: mine = BotRows<float, 0, 1>::node(q, nnx, nnz, I, J); break;
      default: mine = BotRows<float, 1, 1>::node(q, nnx, nnz, I, J); break;
    }
    if (mine) f(I, J);
  }
}

// P^T r at coarse node (i, j) of class c
template <class T>
__device__ __forceinline__ void ml_restrict_at(int c, const typename V2<T>::t (*sr)[ML_RW], int i,
                                               int j, double& s0, double& s1) {
  switch (c) {
    case 0: restrict_row<T, 0, 0, ML_RW>(sr, i, j, s0, s1); break;
    case 1: restrict_row<T, 1, 0, ML_RW>(sr, i, j, s0, s1); break;
    case 2: restrict_row<T, 0, 1, ML_RW>(sr, i, j, s0, s1); break;
    default: restrict_row<T, 1, 1, ML_RW>(sr, i, j, s0, s1); break;
  }
}

template <class T>
__global__ void __launch_bounds__(ML_BS) k_vcycle_bottom_ml(MlArgs<T> A) {
  using T2 = typename V2<T>::t;
  __shared__ T2 sb[ML_ROWS][ML_NP], si[ML_ROWS][ML_NP], sg[ML_ROWS][ML_NP], su[ML_ROWS][ML_NP];
  __shared__ uint16_t sf[ML_ROWS][ML_NP];
  __shared__ T se[ML_EROWS][ML_EW];
  // two masked x patches: a sweep reads one and writes the next iterate, masked,
  // into the other, so the masking needs no phase of its own
  __shared__ T2 sx[ML_PW][ML_PW], sx2[ML_PW][ML_PW];
  __shared__ T2 sr[ML_RW][ML_RW];  // residual patch; the coarse x patch between levels
  __shared__ T2 sp[ML_N0][ML_N0], sd[ML_N0][ML_N0];
  __shared__ T sbc[BOT_MAX_NC];
  __shared__ uint16_t sfc[BOT_MAX_NC / 2];
  __shared__ T sinv[BOT_MAX_NC * BOT_MAX_NC];
  T2(*sc)[ML_CW] = reinterpret_cast<T2(*)[ML_CW]>(&sr[0][0]);
  const int tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  int c, sg_, nw;
  ml_class(w, c, sg_, nw);
  const int first = ML_MAXLEV - A.nlev;  // slot of the first fused level
  const int row0[ML_MAXLEV] = {0, ML_N0, ML_N0 + ML_N1};
  const int erow0[ML_MAXLEV] = {0, ML_N0 / 2, ML_N0 / 2 + ML_N1 / 2};
  // every global load of the kernel in flight at once: the slots' operands
  // (slot 0 needs two passes of the block), viscosities, Chebyshev
  // coefficients, the coarse inverse and mask
  T cf[ML_MAXLEV][3];
#pragma unroll
  for (int l = 0; l < ML_MAXLEV; l++) {
    const T* cp = A.coef[l >= first ? l : ML_MAXLEV - 1];
#pragma unroll
    for (int k = 0; k < 3; k++) cf[l][k] = cp[k + 1];
  }
  auto view = [&](int l) {
    MlView<T> M;
    M.b = &sb[row0[l]];
    M.id = &si[row0[l]];
    M.dg = &sg[row0[l]];
    M.x1 = &su[row0[l]];
    M.fx = &sf[row0[l]];
    M.e = &se[erow0[l]];
    M.L = A.L[l];
    M.nelx = A.L[l].nelx;
    M.nelz = A.L[l].nelz;
    M.nnx = A.L[l].nnx;
    M.nnz = A.L[l].nnz;
    M.c20 = l == 0 ? cf[0][0] : (l == 1 ? cf[1][0] : cf[2][0]);
    M.c11 = l == 0 ? cf[0][1] : (l == 1 ? cf[1][1] : cf[2][1]);
    M.c21 = l == 0 ? cf[0][2] : (l == 1 ? cf[1][2] : cf[2][2]);
    return M;
  };
  {
    constexpr int NP = 4;  // (slot, node) entries of a thread: slot 0 twice, slots 1 and 2
    const int pl[NP] = {0, 0, 1, 2};
    const int pi_[NP] = {tid, tid + ML_BS, tid, tid};
    T2 vb[NP], vi[NP], vg[NP];
    unsigned vf[NP];
    bool ok[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) {
      const int l = pl[k];
      const LV<T>& L = A.L[l >= first ? l : ML_MAXLEV - 1];
      const int nn = L.nnx * L.nnz;
      ok[k] = l >= first && pi_[k] < nn;
      const int i = ok[k] ? pi_[k] : 0;
      vb[k] = ld2(A.b, l == first ? i : 0);
      vi[k] = ld2(L.invdiag, i);
      vg[k] = ld2(L.diag, i);
      vf[k] = fix2(L.fix, i);
    }
    T ve[ML_MAXLEV];
    bool eok[ML_MAXLEV];
#pragma unroll
    for (int l = 0; l < ML_MAXLEV; l++) {
      const LV<T>& L = A.L[l >= first ? l : ML_MAXLEV - 1];
      eok[l] = l >= first && tid < L.nelx * L.nelz;
      ve[l] = L.eta[eok[l] ? tid : 0];
    }
    constexpr int NI = BOT_MAX_NC * BOT_MAX_NC / ML_BS;
    const int last = A.nc * A.nc - 1;
    T vc[NI];
#pragma unroll
    for (int k = 0; k < NI; k++) vc[k] = A.cinv[min(tid + ML_BS * k, last)];
    const bool cok = tid < A.nnx_c * A.nnz_c;
    const unsigned vfc = fix2(A.fixc, cok ? tid : 0);
    // into LDS
#pragma unroll
    for (int k = 0; k < NI; k++)
      if (tid + ML_BS * k <= last) sinv[tid + ML_BS * k] = vc[k];
    if (cok) sfc[tid] = (uint16_t)vfc;
#pragma unroll
    for (int k = 0; k < NP; k++) {
      if (!ok[k]) continue;
      const int l = pl[k], nnx = A.L[l].nnx, r0 = row0[l] + pi_[k] / nnx, I = pi_[k] % nnx;
      if (l == first) sb[r0][I] = vb[k];
      si[r0][I] = vi[k];
      sg[r0][I] = vg[k];
      sf[r0][I] = (uint16_t)vf[k];
    }
#pragma unroll
    for (int l = 0; l < ML_MAXLEV; l++)
      if (eok[l]) se[erow0[l] + tid / A.L[l].nelx][tid % A.L[l].nelx] = ve[l];
  }
  __syncthreads();
  // ---- down: pre-smoothing (steps 0+1), residual, restriction
#pragma unroll 1
  for (int l = first; l < ML_MAXLEV; l++) {
    const MlView<T> M = view(l);
    const int nnx = M.nnx, nnz = M.nnz, pw_ = nnx + 4;
    // step 0 (pointwise from x = 0) into the masked patch; the second patch's
    // ring outside the level and the residual patch cleared
    for (int i = tid; i < pw_ * (nnz + 4); i += ML_BS) {
      const int I = i % pw_ - 2, J = i / pw_ - 2;
      T2 v = {0, 0};
      if (I >= 0 && I < nnx && J >= 0 && J < nnz) {
        const unsigned f = M.fx[J][I];
        const T2 pb = M.b[J][I], pi = M.id[J][I];
        v.x = (f & 0xff) ? (T)0 : (T)0 + M.c20 * pi.x * (pb.x - (T)0);
        v.y = (f >> 8) ? (T)0 : (T)0 + M.c20 * pi.y * (pb.y - (T)0);
      } else {
        sx2[J + 2][I + 2] = T2{0, 0};
      }
      sx[J + 2][I + 2] = v;
    }
    for (int i = tid; i < ML_RW * ML_RW; i += ML_BS) sr[i / ML_RW][i % ML_RW] = T2{0, 0};
    __syncthreads();
    ml_nodes(c, sg_, nw, lane, nnx, nnz, [&](int I, int J) {  // step 1
      const unsigned pf = M.fx[J][I];
      const T2 pb = M.b[J][I], pi = M.id[J][I], pdg = M.dg[J][I];
      const T d00 = M.c20 * pi.x * (pb.x - (T)0), d01 = M.c20 * pi.y * (pb.y - (T)0);
      const T x00 = (T)0 + d00, x01 = (T)0 + d01;
      T y0, y1;
      gather_row_w<T, ML_PW, ML_EW>(c, M.L, sx, M.e, I, J, -2, -2, 0, 0, y0, y1);
      if (pf & 0xff) y0 = pdg.x * x00;
      if (pf >> 8) y1 = pdg.y * x01;
      T d0 = M.c21 * pi.x * (pb.x - y0);
      T d1 = M.c21 * pi.y * (pb.y - y1);
      d0 += M.c11 * d00;
      d1 += M.c11 * d01;
      const T2 x1 = T2{x00 + d0, x01 + d1};
      M.x1[J][I] = x1;
      sx2[J + 2][I + 2] = T2{(pf & 0xff) ? (T)0 : x1.x, (pf >> 8) ? (T)0 : x1.y};
    });
    __syncthreads();
    // residual of x1 (masked x1 in the second patch) into sr, zero elsewhere
    ml_nodes(c, sg_, nw, lane, nnx, nnz, [&](int I, int J) {
      const unsigned pf = M.fx[J][I];
      const T2 pb = M.b[J][I];
      T y0, y1;
      gather_row_w<T, ML_PW, ML_EW>(c, M.L, sx2, M.e, I, J, -2, -2, 0, 0, y0, y1);
      T2 r;
      r.x = (pf & 0xff) ? (T)0 : pb.x - y0;
      r.y = (pf >> 8) ? (T)0 : pb.y - y1;
      sr[J + 3][I + 3] = r;
    });
    __syncthreads();
    // restriction into the next level's b (masked by its fix) or the coarsest b
    if (l + 1 < ML_MAXLEV) {
      const MlView<T> C = view(l + 1);
      ml_nodes(c, sg_, nw, lane, C.nnx, C.nnz, [&](int i, int j) {
        double s0, s1;
        ml_restrict_at<T>(c, sr, i, j, s0, s1);
        const unsigned fc = C.fx[j][i];
        C.b[j][i] = T2{(fc & 0xff) ? (T)0 : (T)s0, (fc >> 8) ? (T)0 : (T)s1};
      });
    } else {
      ml_nodes(c, sg_, nw, lane, A.nnx_c, A.nnz_c, [&](int i, int j) {
        double s0, s1;
        ml_restrict_at<T>(c, sr, i, j, s0, s1);
        const int nn = j * A.nnx_c + i;
        const unsigned fc = sfc[nn];
        sbc[2 * nn] = (fc & 0xff) ? (T)0 : (T)s0;
        sbc[2 * nn + 1] = (fc >> 8) ? (T)0 : (T)s1;
      });
    }
    __syncthreads();
  }
  // ---- coarsest: x_c = A_c^-1 b_c with the sums of k_dense_mv, into the coarse patch
  for (int row = w; row < A.nc; row += ML_BS / 64) {
    double s = 0.0;
    for (int j = lane; j < A.nc; j += 64) s += (double)sinv[row * A.nc + j] * (double)sbc[j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) {
      const int cn = row >> 1, ci = cn % A.nnx_c, cj = cn / A.nnx_c;
      if (row & 1) sc[cj + 2][ci + 2].y = (T)s;
      else sc[cj + 2][ci + 2].x = (T)s;
    }
  }
  __syncthreads();
  // ---- up: prolongation of the coarse patch onto x1, post-smoothing steps 0
  // and 1; the result into the coarse patch of the level above (or to x)
#pragma unroll 1
  for (int l = ML_MAXLEV - 1; l >= first; l--) {
    const MlView<T> M = view(l);
    const int nnx = M.nnx, nnz = M.nnz, pw_ = nnx + 4;
    const int nnx_c = l + 1 < ML_MAXLEV ? A.L[l + 1].nnx : A.nnx_c;
    const int nnz_c = l + 1 < ML_MAXLEV ? A.L[l + 1].nnz : A.nnz_c;
    for (int i = tid; i < pw_ * (nnz + 4); i += ML_BS) {
      const int gI = i % pw_ - 2, gJ = i / pw_ - 2;
      T2 v = {0, 0};
      if (gI >= 0 && gI < nnx && gJ >= 0 && gJ < nnz) {
        double p0, p1;
        prolong_at<T, ML_CW>(sc, gI, gJ, -2, -2, nnx_c, nnz_c, p0, p1);
        const unsigned f = M.fx[gJ][gI];
        v = M.x1[gJ][gI];
        if (!(f & 0xff)) v.x += (T)p0;
        if (!(f >> 8)) v.y += (T)p1;
        sp[gJ][gI] = v;  // unmasked prolonged x
        if (f & 0xff) v.x = 0;
        if (f >> 8) v.y = 0;
      } else {
        sx2[gJ + 2][gI + 2] = T2{0, 0};
      }
      sx[gJ + 2][gI + 2] = v;
    }
    __syncthreads();
    ml_nodes(c, sg_, nw, lane, nnx, nnz, [&](int I, int J) {  // step 0
      const unsigned pf = M.fx[J][I];
      const T2 pb = M.b[J][I], pi = M.id[J][I], pdg = M.dg[J][I], xp = sp[J][I];
      T y0, y1;
      gather_row_w<T, ML_PW, ML_EW>(c, M.L, sx, M.e, I, J, -2, -2, 0, 0, y0, y1);
      if (pf & 0xff) y0 = pdg.x * xp.x;
      if (pf >> 8) y1 = pdg.y * xp.y;
      const T d0 = M.c20 * pi.x * (pb.x - y0);
      const T d1 = M.c20 * pi.y * (pb.y - y1);
      sd[J][I] = T2{d0, d1};
      const T2 x2 = T2{xp.x + d0, xp.y + d1};
      sp[J][I] = x2;  // only this node reads its entry
      sx2[J + 2][I + 2] = T2{(pf & 0xff) ? (T)0 : x2.x, (pf >> 8) ? (T)0 : x2.y};
    });
    __syncthreads();
    T* xout = l == first ? A.x : nullptr;
    ml_nodes(c, sg_, nw, lane, nnx, nnz, [&](int I, int J) {  // step 1
      const unsigned pf = M.fx[J][I];
      const T2 pb = M.b[J][I], pi = M.id[J][I], pdg = M.dg[J][I], x2 = sp[J][I], d0v = sd[J][I];
      T y0, y1;
      gather_row_w<T, ML_PW, ML_EW>(c, M.L, sx2, M.e, I, J, -2, -2, 0, 0, y0, y1);
      if (pf & 0xff) y0 = pdg.x * x2.x;
      if (pf >> 8) y1 = pdg.y * x2.y;
      T d0 = M.c21 * pi.x * (pb.x - y0);
      T d1 = M.c21 * pi.y * (pb.y - y1);
      d0 += M.c11 * d0v.x;
      d1 += M.c11 * d0v.y;
      if (xout)
        st2(xout, (long)J * nnx + I, x2.x + d0, x2.y + d1);
      else
        sc[J + 2][I + 2] = T2{x2.x + d0, x2.y + d1};
    });
    __syncthreads();
  }
}

// dense coarsest operator (FP64), row i = dof, rows ld apart; zero: clear the
// row first (else the row holds only its diagonal entry, which is reset)
template <class T>
__global__ void k_coarse_dense(LV<T> L, double* __restrict__ A, long ld, bool zero) {
  long ndof = 2L * L.nnx * L.nnz;
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= ndof) return;
  double* row = A + i * ld;
  if (zero)
    for (long j = 0; j < ndof; j++) row[j] = 0.0;
  else
    row[i] = 0.0;
  long n = i >> 1;
  int c = (int)(i & 1);
  if (L.fix[i]) {
    row[i] = (double)L.diag[i];
    return;
  }
  int I = (int)(n % L.nnx), J = (int)(n / L.nnx);
  int ex[2], lx[2], ez[2], lz[2];
  int cx = q2_adjacent(I, L.nelx, ex, lx), cz = q2_adjacent(J, L.nelz, ez, lz);
  for (int iz = 0; iz < cz; iz++)
    for (int ix = 0; ix < cx; ix++) {
      int e = ez[iz] * L.nelx + ex[ix];
      int a = q2_local(lx[ix], lz[iz]);
      double et = (double)L.eta[e];
      int base = (2 * ez[iz]) * L.nnx + 2 * ex[ix];
      for (int mz = 0; mz < 3; mz++)
        for (int mx = 0; mx < 3; mx++) {
          int b = q2_local(mx, mz);
          long m = base + (long)mz * L.nnx + mx;
          for (int cb = 0; cb < 2; cb++) {
            long j = 2 * m + cb;
            if (L.fix[j]) continue;
            row[j] += et * (double)L.Kh[(2 * a + c) * 18 + 2 * b + cb];
          }
        }
    }
}

// In-place Gauss-Jordan inverse of a small SPD coarsest operator with the n x 2n
// work matrix [A | I] in LDS (one workgroup), writing the inverse to Ainv.
constexpr int GJ_LDS_MAX_N = 64;
template <class T>
__global__ void __launch_bounds__(1024) k_gauss_jordan_lds(int n, const double* __restrict__ A,
                                                          T* __restrict__ Ainv) {
  __shared__ double W[GJ_LDS_MAX_N * 2 * GJ_LDS_MAX_N];
  const int n2 = 2 * n;
  for (int k = threadIdx.x; k < n * n2; k += blockDim.x) {
    int i = k / n2, j = k % n2;
    W[k] = j < n ? A[i * n + j] : (j - n == i ? 1.0 : 0.0);
  }
  __syncthreads();
  // thread (column j, row group i0) owns rows i0, i0 + RG, ..: per pivot it
  // reads the pivot row entry of its column and the pivot column entries of
  // its rows, then (after one barrier) scales or eliminates its own entries
  // with the arithmetic of the textbook sweep (row p divided by the pivot,
  // W[i][j] -= W[i][p] * W[p][j] with the scaled pivot row)
  constexpr int NC = 2 * GJ_LDS_MAX_N, RG = 1024 / NC, NR = GJ_LDS_MAX_N / RG;
  const int j = threadIdx.x % NC, i0 = threadIdx.x / NC;
  for (int p = 0; p < n; p++) {
    const double piv = W[p * n2 + p];
    const double wp = j < n2 ? W[p * n2 + j] : 0.0;
    double f[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const int i = i0 + RG * r;
      f[r] = i < n ? W[i * n2 + p] : 0.0;
    }
    __syncthreads();
    if (j < n2) {
      const double np = wp / piv;
#pragma unroll
      for (int r = 0; r < NR; r++) {
        const int i = i0 + RG * r;
        if (i < n) W[i * n2 + j] = (i == p) ? np : W[i * n2 + j] - f[r] * np;
      }
    }
    __syncthreads();
  }
  for (int k = threadIdx.x; k < n * n; k += blockDim.x) {
    int i = k / n, j = k % n;
    Ainv[k] = (T)W[i * n2 + n + j];
  }
}

// Blocked Gauss-Jordan inverse of an SPD matrix over the whole GPU, for
// coarsest levels of a few hundred to ~2000 unknowns.  The work matrix is
// np x np (np a multiple of GJB_TILE, the identity outside the leading n x n)
// and is swept one block K of GJB pivots at a time; each sweep is one launch
// over 64 x 64 tiles that reads S and writes D (two buffers, so no tile reads
// what another one writes).  Every workgroup forms P = inv(S_KK) itself in its
// first wavefront (lane l holds column l of [S_KK | I]), then writes
//   D_KK = P,  D_Kj = P S_Kj,  D_iK = -S_iK P,  D_ij = S_ij - S_iK (P S_Kj)
// for i, j outside K.  Once every block has been swept the work matrix holds
// the inverse.  SPD: every pivot is a diagonal entry of a Schur complement of
// an SPD matrix, so no pivoting is needed.
constexpr int GJB = 16, GJB_TILE = 64;
constexpr int GJ_MAX_N = 2304;  // coarsest unknowns at most (dense solve)
static inline int gj_pad(int n) { return (n + GJB_TILE - 1) / GJB_TILE * GJB_TILE; }

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__global__ void __launch_bounds__(256) k_gj_sweep(int np, int k, const double* __restrict__ S,
                                                  double* __restrict__ D) {
  __shared__ double sP[GJB][GJB + 1];
  __shared__ double sC[GJB_TILE][GJB + 1];  // S_iK of the tile's rows
  __shared__ double sR[GJB][GJB_TILE + 1];  // S_Kj of the tile's columns, then P S_Kj
  constexpr int NQ = GJB * GJB_TILE / 256, NT = GJB_TILE / 4;
  const int tid = threadIdx.x;
  const int i0 = blockIdx.y * GJB_TILE, j0 = blockIdx.x * GJB_TILE;
  const int cc = tid & 63, rg = tid >> 6;
  const int j = j0 + cc;
  // every global load first: the tile entries of this thread (column j, rows
  // rg + 4t), its share of S_iK and S_Kj, and (first wavefront) S_KK by columns
  double sv[NT], cv[NQ], rv[NQ], w[GJB];
#pragma unroll
  for (int t = 0; t < NT; t++) sv[t] = S[(long)(i0 + rg + 4 * t) * np + j];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int e = tid + 256 * q;
    cv[q] = S[(long)(i0 + e / GJB) * np + k + e % GJB];
    rv[q] = S[(long)(k + e / GJB_TILE) * np + j0 + e % GJB_TILE];
  }
  const int l = tid;  // wavefront 0: lane l < 16 holds column l of S_KK, 16..31 of I
  if (tid < 64)
#pragma unroll
    for (int r = 0; r < GJB; r++)
      w[r] = l < GJB ? S[(long)(k + r) * np + k + l] : (l - GJB == r ? 1.0 : 0.0);
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int e = tid + 256 * q;
    sC[e / GJB][e % GJB] = cv[q];
    sR[e / GJB_TILE][e % GJB_TILE] = rv[q];
  }
  if (tid < 64) {
#pragma unroll
    for (int p = 0; p < GJB; p++) {
      double f[GJB];
#pragma unroll
      for (int r = 0; r < GJB; r++) f[r] = readlane_f64(w[r], p);  // column p
      w[p] = w[p] / f[p];
#pragma unroll
      for (int r = 0; r < GJB; r++)
        if (r != p) w[r] -= f[r] * w[p];
    }
    if (l >= GJB && l < 2 * GJB)
#pragma unroll
      for (int r = 0; r < GJB; r++) sP[r][l - GJB] = w[r];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int e = tid + 256 * q, rr = e / GJB_TILE, c = e % GJB_TILE;
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < GJB; m++) acc += sP[rr][m] * sR[m][c];
    rv[q] = acc;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int e = tid + 256 * q;
    sR[e / GJB_TILE][e % GJB_TILE] = rv[q];
  }
  __syncthreads();
  const bool jK = j >= k && j < k + GJB;
  double pr[GJB];  // this column of P S_Kj, or of P when j is in K
#pragma unroll
  for (int m = 0; m < GJB; m++) pr[m] = jK ? sP[m][j - k] : sR[m][cc];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int r = rg + 4 * t, i = i0 + r;
    double v;
    if (i >= k && i < k + GJB) {
      v = jK ? sP[i - k][j - k] : sR[i - k][cc];
    } else {
      v = jK ? 0.0 : sv[t];
#pragma unroll
      for (int m = 0; m < GJB; m++) v -= sC[r][m] * pr[m];
    }
    D[(long)i * np + j] = v;
  }
}

// np x np identity
__global__ void k_identity(int np, double* __restrict__ A) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < (long)np * np) A[e] = (e / np == e % np) ? 1.0 : 0.0;
}

// leading n x n block of the np-strided work matrix -> Ainv (n x n, T)
template <class T>
__global__ void k_extract_padded(int n, int np, const double* __restrict__ W,
                                 T* __restrict__ Ainv) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  Ainv[e] = (T)W[(e / n) * np + e % n];
}

// x = Ainv b, one wavefront per row
template <class T>
__global__ void k_dense_mv(int n, const T* __restrict__ Ainv, const T* __restrict__ b,
                           T* __restrict__ x) {
  int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= n) return;
  double s = 0.0;
  for (int j = lane; j < n; j += 64) s += (double)Ainv[(long)row * n + j] * (double)b[j];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if (lane == 0) x[row] = (T)s;
}

template <class T>
__global__ void k_copy(long n, const T* __restrict__ a, T* __restrict__ b) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// ------------------------------------------------------------------------
// Per-update work on all levels at once.  The levels are independent during
// the power iterations (and for the diagonals), so each step runs as one
// launch over the blocks of every level instead of one launch per level; each
// level keeps its own block decomposition, so every value is the one the
// per-level kernels give.
template <class T>
struct LevelSet {
  int nlev;                   // levels in the set
  int tile0[MAX_LEVELS + 1];  // first block of each level's 16x16 node tiles
  int ntx[MAX_LEVELS];        // node tiles per row
  int vblk0[MAX_LEVELS + 1];  // first block of each level's 256-wide dof blocks
  int nblk0[MAX_LEVELS + 1];  // first block of each level's 256-wide node blocks
  int dblk0[MAX_LEVELS + 1];  // first partial of each level's dot (dot_blocks per level)
  long ndof[MAX_LEVELS];
  LV<T> L[MAX_LEVELS];
  T* x[MAX_LEVELS];
  T* r[MAX_LEVELS];
  T* diag[MAX_LEVELS];
  T* invdiag[MAX_LEVELS];
  T* coef[MAX_LEVELS];
  int steps[MAX_LEVELS];      // Chebyshev degree per level
};

// level of block b given the levels' first blocks (nlev levels)
__device__ __forceinline__ int set_level(const int* first, int nlev, int b) {
  int l = 0;
  while (l + 1 < nlev && b >= first[l + 1]) l++;
  return l;
}

// diagonal of the unmasked operator: sum of eta_e * Khat[ii] over the
// elements around the node
template <class T>
__global__ void __launch_bounds__(256) k_set_diag(LevelSet<T> S) {
  const int l = set_level(S.nblk0, S.nlev, blockIdx.x);
  const LV<T>& L = S.L[l];
  const long n = (long)(blockIdx.x - S.nblk0[l]) * 256 + threadIdx.x;
  if (n >= (long)L.nnx * L.nnz) return;
  int I = (int)(n % L.nnx), J = (int)(n / L.nnx);
  int ex[2], lx[2], ez[2], lz[2];
  int cx = q2_adjacent(I, L.nelx, ex, lx), cz = q2_adjacent(J, L.nelz, ez, lz);
  // the (up to) four elements' loads issued together, then added in the
  // order of the element loop (iz outer, ix inner)
  T et[4], k0[4], k1[4];
#pragma unroll
  for (int iz = 0; iz < 2; iz++)
#pragma unroll
    for (int ix = 0; ix < 2; ix++) {
      const int q = iz * 2 + ix;
      const bool ok = iz < cz && ix < cx;
      const int a = ok ? q2_local(lx[ix], lz[iz]) : 0;
      et[q] = ok ? L.eta[ez[iz] * L.nelx + ex[ix]] : (T)0;
      k0[q] = L.Kh[(2 * a) * 18 + 2 * a];
      k1[q] = L.Kh[(2 * a + 1) * 18 + 2 * a + 1];
    }
  T d0 = 0, d1 = 0;
#pragma unroll
  for (int iz = 0; iz < 2; iz++)
#pragma unroll
    for (int ix = 0; ix < 2; ix++) {
      const int q = iz * 2 + ix;
      if (iz < cz && ix < cx) {
        d0 += et[q] * k0[q];
        d1 += et[q] * k1[q];
      }
    }
  S.diag[l][2 * n] = d0;
  S.diag[l][2 * n + 1] = d1;
  S.invdiag[l][2 * n] = (T)1 / d0;
  S.invdiag[l][2 * n + 1] = (T)1 / d1;
}

template <class T>
__global__ void __launch_bounds__(256) k_set_hash_fill(LevelSet<T> S) {
  const int l = set_level(S.vblk0, S.nlev, blockIdx.x);
  const long i = (long)(blockIdx.x - S.vblk0[l]) * 256 + threadIdx.x;
  if (i >= S.ndof[l]) return;
  unsigned h = (unsigned)(i * 2654435761u) ^ 0x9e3779b9u;
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  S.x[l][i] = (T)(0.5 + (h & 0xffff) / 65536.0);
}

// r = A x on every level of the set
template <class T>
__global__ void __launch_bounds__(256) k_set_apply(LevelSet<T> S) {
  const int l = set_level(S.tile0, S.nlev, blockIdx.x);
  const int t = blockIdx.x - S.tile0[l];
  level_op_tile<T, OP_APPLY>(S.L[l], S.x[l], nullptr, S.r[l], nullptr, nullptr, nullptr, 0, 0,
                             t % S.ntx[l], t / S.ntx[l]);
}

// partial sums of <x, x> per level (as k_dot_partial over that level's blocks)
template <class T>
__global__ void __launch_bounds__(DOT_BS) k_set_dot_partial(LevelSet<T> S, double* partial) {
  const int l = set_level(S.dblk0, S.nlev, blockIdx.x);
  const int lb = blockIdx.x - S.dblk0[l], nb = S.dblk0[l + 1] - S.dblk0[l];
  const T* a = S.x[l];
  const long n = S.ndof[l];
  double s = 0.0;
  for (long i = lb * (long)DOT_BS + threadIdx.x; i < n; i += (long)nb * DOT_BS)
    s += (double)a[i] * (double)a[i];
  __shared__ double sh[DOT_BS / 64];
  s = wave_sum64(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < DOT_BS / 64; w++) t += sh[w];
    partial[blockIdx.x] = t;
  }
}

// out[l] = sum of level l's partials (one wavefront per level, as k_dot_final)
template <class T>
__global__ void k_set_dot_final(LevelSet<T> S, const double* __restrict__ partial,
                                double* __restrict__ out) {
  const int l = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (l >= S.nlev) return;
  double s = 0.0;
  for (int b = S.dblk0[l] + lane; b < S.dblk0[l + 1]; b += 64) s += partial[b];
  s = wave_sum64(s);
  if (lane == 0) out[l] = s;
}

// x = D^-1 r / ||x|| per level (||x||^2 in nrm2[l])
template <class T>
__global__ void __launch_bounds__(256) k_set_scale(LevelSet<T> S, const double* __restrict__ nrm2) {
  const int l = set_level(S.vblk0, S.nlev, blockIdx.x);
  const long i = (long)(blockIdx.x - S.vblk0[l]) * 256 + threadIdx.x;
  if (i >= S.ndof[l]) return;
  double s = 1.0 / sqrt(nrm2[l] > 0 ? nrm2[l] : 1.0);
  S.x[l][i] = (T)((double)S.invdiag[l][i] * (double)S.r[l][i] * s);
}

template <class T>
__global__ void k_set_cheb_coef(LevelSet<T> S, const double* __restrict__ nrm2) {
  const int l = threadIdx.x;
  if (l >= S.nlev) return;
  const int steps = S.steps[l];
  double lm = sqrt(nrm2[l]);
  const double lmax = 1.1 * lm, lmin = 0.1 * lm;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
  double rho = 1.0 / sigma;
  for (int k = 0; k < steps; k++) {
    double c1, c2;
    if (k == 0) {
      c1 = 0.0;
      c2 = 1.0 / theta;
    } else {
      double rn = 1.0 / (2.0 * sigma - rho);
      c1 = rn * rho;
      c2 = 2.0 * rn / delta;
      rho = rn;
    }
    S.coef[l][2 * k] = (T)c1;
    S.coef[l][2 * k + 1] = (T)c2;
  }
}

static int dot_blocks_of(long n) {
  long b = (n + DOT_BS - 1) / DOT_BS;
  return (int)(b < DOT_BLOCKS ? (b < 1 ? 1 : b) : DOT_BLOCKS);
}

template <class T>
static LevelSet<T> level_set(const std::vector<Level<T>*>& lev, int nlev) {
  LevelSet<T> S{};
  S.nlev = nlev;
  S.tile0[0] = S.vblk0[0] = S.nblk0[0] = S.dblk0[0] = 0;
  for (int l = 0; l < nlev; l++) {
    Level<T>& L = *lev[l];
    const dim3 t = tiles(L);
    S.ntx[l] = (int)t.x;
    S.tile0[l + 1] = S.tile0[l] + (int)(t.x * t.y);
    S.vblk0[l + 1] = S.vblk0[l] + (int)((L.ndof + 255) / 256);
    S.nblk0[l + 1] = S.nblk0[l] + (int)((L.nn + 255) / 256);
    S.dblk0[l + 1] = S.dblk0[l] + dot_blocks_of(L.ndof);
    S.ndof[l] = L.ndof;
    S.L[l] = view(L);
    S.x[l] = L.x.p;
    S.r[l] = L.r.p;
    S.diag[l] = L.diag.p;
    S.invdiag[l] = L.invdiag.p;
    S.coef[l] = L.coef.p;
  }
  return S;
}

// ------------------------------------------------------------------------
// Fused tail of the V-cycle: one 1024-thread workgroup runs every level from
// the first one with at most TAIL_MAX_DOF unknowns down to the coarsest and
// back, entirely in LDS (operator data, work vectors and the coarse inverse
// are staged in at the start; the first tail level's x is written back at the
// end), with block barriers between sweeps instead of kernel boundaries.  At
// these sizes the per-level kernels are pure launch latency.  The arithmetic
// of every sweep is that of the per-level kernels, in the same order, so the
// V-cycle is bit-identical either way.
constexpr int TAIL_MAX_DOF = 600;
constexpr int TAIL_BS = 256;
constexpr size_t TAIL_MAX_LDS = 150 * 1024;

template <class T>
struct TailLevel {
  LV<T> v;  // global operator data
  T *x, *b;  // global x (written back for the first level) and b (read for the first)
  const T* coef;
  int nsteps;
};

template <class T>
struct TailArgs {
  int nlev, nsmooth, ncoarse;
  const T* coarse_inv;
  TailLevel<T> L[MAX_LEVELS];
};

// bytes of LDS a tail level needs: Khat, coef, eta, diag, invdiag, x, b, r, d, fix
template <class T>
__host__ __device__ inline size_t tail_level_bytes(int nelx, int nelz, long ndof, int nsmooth) {
  size_t e = 324 + 2 * (size_t)nsmooth + (size_t)nelx * nelz + 6 * (size_t)ndof;
  return ((e * sizeof(T) + ndof + 15) / 16) * 16;
}

template <class T>
struct TailLds {  // a level's arrays in LDS
  LV<T> v;
  T *x, *b, *r, *d, *coef;
};

// masked operator row pair of node (I, J), x read from LDS
template <class T>
__device__ __forceinline__ void tail_row(const LV<T>& L, const T* x, int I, int J, T& y0, T& y1) {
  int exa[2], lxa[2], eza[2], lza[2];
  int cx = q2_adjacent(I, L.nelx, exa, lxa), cz = q2_adjacent(J, L.nelz, eza, lza);
  y0 = 0;
  y1 = 0;
  for (int iz = 0; iz < cz; iz++)
    for (int ix = 0; ix < cx; ix++) {
      int a = q2_local(lxa[ix], lza[iz]);
      const T* r0 = L.Kh + (2 * a) * 18;
      const T* r1 = r0 + 18;
      T s0 = 0, s1 = 0;
#pragma unroll
      for (int mz = 0; mz < 3; mz++)
#pragma unroll
        for (int mx = 0; mx < 3; mx++) {
          int bb = q2_local(mx, mz);
          int m = (2 * eza[iz] + mz) * L.nnx + 2 * exa[ix] + mx;
          T xu = L.fix[2 * m] ? (T)0 : x[2 * m];
          T xw = L.fix[2 * m + 1] ? (T)0 : x[2 * m + 1];
          s0 = fmad(r0[2 * bb + 1], xw, fmad(r0[2 * bb], xu, s0));
          s1 = fmad(r1[2 * bb + 1], xw, fmad(r1[2 * bb], xu, s1));
        }
      T et = L.eta[eza[iz] * L.nelx + exa[ix]];
      y0 += et * s0;
      y1 += et * s1;
    }
  int n = J * L.nnx + I;
  if (L.fix[2 * n]) y0 = L.diag[2 * n] * x[2 * n];
  if (L.fix[2 * n + 1]) y1 = L.diag[2 * n + 1] * x[2 * n + 1];
}

template <class T>
__device__ void tail_cheb(const TailLds<T>& P, int steps, bool zero_x) {
  const LV<T>& L = P.v;
  const int nn = L.nnx * L.nnz;
  for (int k = 0; k < steps; k++) {
    const bool zx = zero_x && k == 0;
    const T c1 = P.coef[2 * k], c2 = P.coef[2 * k + 1];
    for (int n = threadIdx.x; n < nn; n += TAIL_BS) {
      T y0 = 0, y1 = 0;
      if (!zx) tail_row(L, P.x, n % L.nnx, n / L.nnx, y0, y1);
      T d0 = c2 * L.invdiag[2 * n] * (P.b[2 * n] - y0);
      T d1 = c2 * L.invdiag[2 * n + 1] * (P.b[2 * n + 1] - y1);
      if (k > 0) {
        d0 += c1 * P.d[2 * n];
        d1 += c1 * P.d[2 * n + 1];
      }
      P.d[2 * n] = d0;
      P.d[2 * n + 1] = d1;
    }
    __syncthreads();
    for (int n = threadIdx.x; n < nn; n += TAIL_BS) {
      T x0 = zx ? (T)0 : P.x[2 * n], x1 = zx ? (T)0 : P.x[2 * n + 1];
      P.x[2 * n] = x0 + P.d[2 * n];
      P.x[2 * n + 1] = x1 + P.d[2 * n + 1];
    }
    __syncthreads();
  }
}

template <class T>
__device__ __forceinline__ void lds_copy(T* dst, const T* __restrict__ src, long n) {
  for (long i = threadIdx.x; i < n; i += TAIL_BS) dst[i] = src[i];
}

// LDS arrays of tail level l (offsets recomputed on the fly: no private arrays)
template <class T>
__device__ __forceinline__ TailLds<T> tail_lds(const TailArgs<T>& A, unsigned char* base, int l) {
  unsigned char* p = base;
  for (int k = 0; k < l; k++) {
    const TailLevel<T>& G = A.L[k];
    p += tail_level_bytes<T>(G.v.nelx, G.v.nelz, 2L * G.v.nnx * G.v.nnz, A.nsmooth);
  }
  const TailLevel<T>& G = A.L[l];
  const long ndof = 2L * G.v.nnx * G.v.nnz;
  const long nel = (long)G.v.nelx * G.v.nelz;
  T* t = reinterpret_cast<T*>(p);
  TailLds<T> S;
  T* Kh = t; t += 324;
  S.coef = t; t += 2 * A.nsmooth;
  T* eta = t; t += nel;
  T* diag = t; t += ndof;
  T* invdiag = t; t += ndof;
  S.x = t; t += ndof;
  S.b = t; t += ndof;
  S.r = t; t += ndof;
  S.d = t; t += ndof;
  S.v = LV<T>{G.v.nelx, G.v.nelz, G.v.nnx, G.v.nnz, eta, reinterpret_cast<uint8_t*>(t), Kh, diag,
              invdiag};
  return S;
}

template <class T>
__global__ void __launch_bounds__(TAIL_BS) k_vcycle_tail(TailArgs<T> A) {
  extern __shared__ __align__(16) unsigned char tail_smem[];
  const int last = A.nlev - 1;
  // stage every level in LDS
  unsigned char* p = tail_smem;
  for (int l = 0; l < A.nlev; l++) {
    const TailLevel<T>& G = A.L[l];
    const long ndof = 2L * G.v.nnx * G.v.nnz;
    const long nel = (long)G.v.nelx * G.v.nelz;
    TailLds<T> S = tail_lds(A, tail_smem, l);
    lds_copy(const_cast<T*>(S.v.Kh), G.v.Kh, 324);
    lds_copy(S.coef, G.coef, 2L * G.nsteps);
    lds_copy(const_cast<T*>(S.v.eta), G.v.eta, nel);
    lds_copy(const_cast<T*>(S.v.diag), G.v.diag, ndof);
    lds_copy(const_cast<T*>(S.v.invdiag), G.v.invdiag, ndof);
    lds_copy(const_cast<uint8_t*>(S.v.fix), G.v.fix, ndof);
    if (l == 0) lds_copy(S.b, G.b, ndof);
    p += tail_level_bytes<T>(G.v.nelx, G.v.nelz, ndof, A.nsmooth);
  }
  T* cinv = reinterpret_cast<T*>(p);
  lds_copy(cinv, A.coarse_inv, (long)A.ncoarse * A.ncoarse);
  __syncthreads();
  for (int l = 0; l < last; l++) {
    const TailLds<T> P = tail_lds(A, tail_smem, l);
    const TailLds<T> C = tail_lds(A, tail_smem, l + 1);
    tail_cheb(P, A.nsmooth, true);
    const int nn = P.v.nnx * P.v.nnz;
    for (int n = threadIdx.x; n < nn; n += TAIL_BS) {
      T y0, y1;
      tail_row(P.v, P.x, n % P.v.nnx, n / P.v.nnx, y0, y1);
      P.r[2 * n] = P.b[2 * n] - y0;
      P.r[2 * n + 1] = P.b[2 * n + 1] - y1;
    }
    __syncthreads();
    const int nc = C.v.nnx * C.v.nnz;
    for (int n = threadIdx.x; n < nc; n += TAIL_BS) {  // bc = P^T r (as k_restrict)
      int i = n % C.v.nnx, j = n / C.v.nnx;
      double s0 = 0, s1 = 0;
#pragma unroll
      for (int ob = -3; ob <= 3; ob++)
#pragma unroll
        for (int oa = -3; oa <= 3; oa++) {
          double w = rw(i, oa) * rw(j, ob);
          int fi = 2 * i + oa, fj = 2 * j + ob;
          if (w != 0.0 && fi >= 0 && fi < P.v.nnx && fj >= 0 && fj < P.v.nnz) {
            int m = fj * P.v.nnx + fi;
            if (!P.v.fix[2 * m]) s0 += w * (double)P.r[2 * m];
            if (!P.v.fix[2 * m + 1]) s1 += w * (double)P.r[2 * m + 1];
          }
        }
      C.b[2 * n] = C.v.fix[2 * n] ? (T)0 : (T)s0;
      C.b[2 * n + 1] = C.v.fix[2 * n + 1] ? (T)0 : (T)s1;
    }
    __syncthreads();
  }
  {  // coarsest: x = Ainv b (as k_dense_mv: lane-strided terms, tree sum per wavefront)
    const TailLds<T> C = tail_lds(A, tail_smem, last);
    const int n = A.ncoarse, lane = threadIdx.x & 63;
    for (int row = threadIdx.x >> 6; row < n; row += TAIL_BS / 64) {
      double s = 0.0;
      for (int j = lane; j < n; j += 64) s += (double)cinv[row * n + j] * (double)C.b[j];
      for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
      if (lane == 0) C.x[row] = (T)s;
    }
    __syncthreads();
  }
  for (int l = last - 1; l >= 0; l--) {
    const TailLds<T> P = tail_lds(A, tail_smem, l);
    const TailLds<T> C = tail_lds(A, tail_smem, l + 1);
    const int nn = P.v.nnx * P.v.nnz;
    for (int n = threadIdx.x; n < nn; n += TAIL_BS) {  // xf += P xc (as k_prolong_add)
      int I = n % P.v.nnx, J = n / P.v.nnx;
      int bi = (I >> 2) * 2, bj = (J >> 2) * 2;
      double s0 = 0, s1 = 0;
#pragma unroll
      for (int kb = 0; kb < 3; kb++)
#pragma unroll
        for (int ka = 0; ka < 3; ka++) {
          double w = pw(I, ka) * pw(J, kb);
          int ci = bi + ka, cj = bj + kb;
          if (w != 0.0 && ci < C.v.nnx && cj < C.v.nnz) {
            int m = cj * C.v.nnx + ci;
            s0 += w * (double)C.x[2 * m];
            s1 += w * (double)C.x[2 * m + 1];
          }
        }
      if (!P.v.fix[2 * n]) P.x[2 * n] += (T)s0;
      if (!P.v.fix[2 * n + 1]) P.x[2 * n + 1] += (T)s1;
    }
    __syncthreads();
    tail_cheb(P, A.nsmooth, false);
  }
  const long nd0 = 2L * A.L[0].v.nnx * A.L[0].v.nnz;
  const TailLds<T> S0 = tail_lds(A, tail_smem, 0);
  for (long i = threadIdx.x; i < nd0; i += TAIL_BS) A.L[0].x[i] = S0.x[i];
}

// ------------------------------------------------------------------------
// The FP32 Stokes operator on the finest level in one launch, and the front
// of the block preconditioner.  G is applied from its 18x4 template Gh (FP64,
// uniform over the box), with the arithmetic of the stand-alone template
// kernels: y_u = K z_u + (float)(G z_p) on free rows (K z_u rounded first),
// y_p = (float)(G^T z_u) over free velocity dofs.

// reference local node b -> (lx, lz) on the element's 3x3 node grid
__device__ __forceinline__ void q2_pos_ref(int b, int& lx, int& lz) {
  const int X[9] = {0, 2, 2, 0, 1, 2, 1, 0, 1};
  const int Z[9] = {0, 0, 2, 2, 0, 1, 2, 1, 1};
  lx = X[b];
  lz = Z[b];
}

// (G p)(I, J) in FP64: elements around the node in increasing order, the
// element's pressure nodes in reference order; pv(e_x, e_z, k) gives p
template <class PV>
__device__ __forceinline__ void g_row(const double* __restrict__ Gh, int I, int J, int nelx,
                                      int nelz, PV pv, double& a0, double& a1) {
  int ex[2], lx[2], ez[2], lz[2];
  const int cx = q2_adjacent(I, nelx, ex, lx), cz = q2_adjacent(J, nelz, ez, lz);
  a0 = 0.0;
  a1 = 0.0;
#pragma unroll
  for (int iz = 0; iz < 2; iz++)
#pragma unroll
    for (int ix = 0; ix < 2; ix++) {
      if (iz < cz && ix < cx) {
        const int a = q2_local(lx[ix], lz[iz]);
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const double p = pv(ex[ix], ez[iz], k);
          a0 += Gh[(2 * a) * 4 + k] * p;
          a1 += Gh[(2 * a + 1) * 4 + k] * p;
        }
      }
    }
}

// g_row for a node of parity class (PX, PZ) (the element slots and local node
// numbers of gather_row, so the Gh entries are wave-uniform)
template <int PX, int PZ, class PV>
__device__ __forceinline__ void g_row_cls(const double* __restrict__ Gh, int I, int J, int nelx,
                                          int nelz, PV pv, double& a0, double& a1) {
  a0 = 0.0;
  a1 = 0.0;
#pragma unroll
  for (int iz = 0; iz < (PZ ? 1 : 2); iz++) {
    const int lza = PZ ? 1 : (iz ? 0 : 2);
    const int ez = PZ ? (J - 1) >> 1 : (J >> 1) - 1 + iz;
#pragma unroll
    for (int ix = 0; ix < (PX ? 1 : 2); ix++) {
      const int lxa = PX ? 1 : (ix ? 0 : 2);
      const int ex = PX ? (I - 1) >> 1 : (I >> 1) - 1 + ix;
      const int a = q2_local(lxa, lza);
      const bool inside = ex >= 0 && ex < nelx && ez >= 0 && ez < nelz;
      // the taps of a skipped element are formed from a clamped element and
      // not added, so the Gh loads stay unconditional (scalar)
      const int cx = inside ? ex : 0, cz = inside ? ez : 0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const double p = pv(cx, cz, k);
        const double t0 = as_const(Gh)[(2 * a) * 4 + k] * p;
        const double t1 = as_const(Gh)[(2 * a + 1) * 4 + k] * p;
        a0 = inside ? a0 + t0 : a0;
        a1 = inside ? a1 + t1 : a1;
      }
    }
  }
}

// pressure node of element (ex, ez) in reference order k (nelx+1 per row)
__device__ __forceinline__ long pnode(int ex, int ez, int k, int npx) {
  const long pb = (long)ez * npx + ex;
  return k == 0 ? pb : k == 1 ? pb + 1 : k == 2 ? pb + npx + 1 : pb + npx;
}

constexpr int PT = 16;          // pressure tile edge
constexpr int PV = 2 * PT + 3;  // velocity patch edge behind a pressure tile

// yp = alpha * G^T u on a 16x16 pressure tile (fixed velocity dofs excluded):
// the 35x35 velocity patch is staged masked in LDS, the Gh entries of each
// element slot are wave-uniform
template <class T>
__device__ __forceinline__ void gt_tile(int nelx, int nelz, int nnx, int nnz,
                                        const uint8_t* __restrict__ fix,
                                        const double* __restrict__ Gh, const T* __restrict__ u,
                                        T* __restrict__ yp, double alpha, int tX, int tZ) {
  using T2 = typename V2<T>::t;
  __shared__ T2 su[PV][PV];
  const int tid = threadIdx.x;
  const int i0 = tX * PT, j0 = tZ * PT;
  const int uo = 2 * i0 - 2, vo = 2 * j0 - 2;
  constexpr int NL = (PV * PV + 255) / 256;
  T2 v[NL];
  unsigned f[NL];
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const int e = tid + 256 * k;
    const int gI = uo + e % PV, gJ = vo + e / PV;
    v[k] = T2{0, 0};
    f[k] = 0x0101;
    if (e < PV * PV && gI >= 0 && gI < nnx && gJ >= 0 && gJ < nnz) {
      const long m = (long)gJ * nnx + gI;
      f[k] = fix2(fix, m);
      v[k] = ld2(u, m);
    }
  }
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const int e = tid + 256 * k;
    if (e < PV * PV) {
      T2 w = v[k];
      if (f[k] & 0xff) w.x = 0;
      if (f[k] >> 8) w.y = 0;
      su[e / PV][e % PV] = w;
    }
  }
  __syncthreads();
  const int li = tid % PT, lj = tid / PT;
  const int i = i0 + li, j = j0 + lj;
  if (i > nelx || j > nelz) return;
  double s = 0.0;
#pragma unroll
  for (int dz = 0; dz < 2; dz++)
#pragma unroll
    for (int dx = 0; dx < 2; dx++) {
      const int ez = j - 1 + dz, ex = i - 1 + dx;
      const int k = dx ? (dz ? 0 : 3) : (dz ? 1 : 2);
      // elements outside the mesh are formed (their patch entries exist) but
      // not added, so the Gh loads stay unconditional (scalar)
      const bool inside = ez >= 0 && ez < nelz && ex >= 0 && ex < nelx;
      const int bx0 = 2 * ex - uo, bz0 = 2 * ez - vo;
#pragma unroll
      for (int b = 0; b < 9; b++) {
        int bx, bz;
        q2_pos_ref(b, bx, bz);
        const T2 w = su[bz0 + bz][bx0 + bx];
        const double t0 = as_const(Gh)[(2 * b) * 4 + k] * (double)w.x;
        const double t1 = as_const(Gh)[(2 * b + 1) * 4 + k] * (double)w.y;
        s = inside ? s + t0 : s;
        s = inside ? s + t1 : s;
      }
    }
  yp[(long)j * (nelx + 1) + i] = (T)(alpha * s);
}

// y_u = K z_u + G z_p, y_p = G^T z_u on the finest level: blocks [0, nvt) are
// the velocity node tiles, the rest the pressure tiles
template <class T>
__global__ void __launch_bounds__(256)
k_stokes_fine(LV<T> L, const double* __restrict__ Gh, const T* __restrict__ zu,
              const T* __restrict__ zp, T* __restrict__ yu, T* __restrict__ yp, int nvt,
              int ntxp) {
  if ((int)blockIdx.x >= nvt) {
    const int t = blockIdx.x - nvt;
    gt_tile<T>(L.nelx, L.nelz, L.nnx, L.nnz, L.fix, Gh, zu, yp, 1.0, t % ntxp, t / ntxp);
    return;
  }
  using T2 = typename V2<T>::t;
  __shared__ T2 sx[TP][TP];
  __shared__ T se[TE][TE];
  const int tid = threadIdx.x;
  const int ntx = (L.nnx + TT - 1) / TT;
  const int lin = xcd_remap(blockIdx.x, nvt);
  const int I0 = (lin % ntx) * TT, J0 = (lin / ntx) * TT;
  const int ex0 = I0 > 0 ? (I0 - 1) >> 1 : 0, ez0 = J0 > 0 ? (J0 - 1) >> 1 : 0;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int I = I0 + 2 * (lane & 7) + (w & 1), J = J0 + 2 * (lane >> 3) + (w >> 1);
  const bool valid = I < L.nnx && J < L.nnz;
  const long n = valid ? (long)J * L.nnx + I : 0;
  const int npx = L.nelx + 1;
  T2 px0 = {0, 0};
  unsigned pf = 0;
  double a0 = 0.0, a1 = 0.0;
  if (valid) {
    px0 = ld2(zu, n);
    pf = fix2(L.fix, n);
    auto pv = [&](int ex, int ez, int k) { return (double)zp[pnode(ex, ez, k, npx)]; };
    switch (w) {
      case 0: g_row_cls<0, 0>(Gh, I, J, L.nelx, L.nelz, pv, a0, a1); break;
      case 1: g_row_cls<1, 0>(Gh, I, J, L.nelx, L.nelz, pv, a0, a1); break;
      case 2: g_row_cls<0, 1>(Gh, I, J, L.nelx, L.nelz, pv, a0, a1); break;
      default: g_row_cls<1, 1>(Gh, I, J, L.nelx, L.nelz, pv, a0, a1); break;
    }
  }
  T2 pv[2];
  unsigned pfl[2];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int e = tid + 256 * k;
    const int gI = I0 - 2 + e % TP, gJ = J0 - 2 + e / TP;
    pv[k] = T2{0, 0};
    pfl[k] = 0x0101;
    if (e < TP * TP && gI >= 0 && gI < L.nnx && gJ >= 0 && gJ < L.nnz) {
      const long m = (long)gJ * L.nnx + gI;
      pfl[k] = fix2(L.fix, m);
      pv[k] = ld2(zu, m);
    }
  }
  T pe = 0;
  {
    const int eX = ex0 + tid % TE, eZ = ez0 + tid / TE;
    if (tid < TE * TE && eX < L.nelx && eZ < L.nelz) pe = L.eta[(long)eZ * L.nelx + eX];
  }
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int e = tid + 256 * k;
    if (e < TP * TP) {
      T2 v = pv[k];
      if (pfl[k] & 0xff) v.x = 0;
      if (pfl[k] >> 8) v.y = 0;
      sx[e / TP][e % TP] = v;
    }
  }
  if (tid < TE * TE) se[tid / TE][tid % TE] = pe;
  __syncthreads();
  if (!valid) return;
  T y0, y1;
  gather_row_w<T, TP, TE>(w, L, sx, se, I, J, I0 - 2, J0 - 2, ex0, ez0, y0, y1);
  if (pf & 0xff) {
    y0 = L.diag[2 * n] * px0.x;
    a0 = 0.0;
  }
  if (pf >> 8) {
    y1 = L.diag[2 * n + 1] * px0.y;
    a1 = 0.0;
  }
  st2(yu, n, y0 + (T)a0, y1 + (T)a1);
}

// Front of the block-triangular preconditioner: z_p = -S^-1 v_p (S^-1 the
// diagonal minv) and t = v_u - G z_p (free rows; fixed rows v_u), one launch:
// blocks [0, nbu) the velocity nodes (z_p formed on the fly where G reads it),
// the rest the pressure nodes
// (xs given: z_p = -xs, an approximate solve of S xs = v_p done beforehand)
template <class T>
__global__ void __launch_bounds__(256)
k_schur_front(int nelx, int nelz, int nnx, int nnz, const uint8_t* __restrict__ fix,
              const double* __restrict__ Gh, const double* __restrict__ minv,
              const T* __restrict__ vu, const T* __restrict__ vp, const T* __restrict__ xs,
              T* __restrict__ zp, T* __restrict__ t, int nbu, double scale) {
  const int npx = nelx + 1;
  auto zval = [&](long m) -> T {
    return xs ? (T)(-scale * (double)xs[m]) : (T)(-scale * minv[m] * (double)vp[m]);
  };
  if ((int)blockIdx.x >= nbu) {
    const long m = (long)(blockIdx.x - nbu) * 256 + threadIdx.x;
    if (m < (long)npx * (nelz + 1)) zp[m] = zval(m);
    return;
  }
  const long n = (long)blockIdx.x * 256 + threadIdx.x;
  if (n >= (long)nnx * nnz) return;
  const int I = (int)(n % nnx), J = (int)(n / nnx);
  double a0, a1;
  g_row(Gh, I, J, nelx, nelz,
        [&](int ex, int ez, int k) { return (double)zval(pnode(ex, ez, k, npx)); }, a0, a1);
  const unsigned f = fix2(fix, n);
  const T g0 = (f & 0xff) ? (T)0 : (T)a0, g1 = (f >> 8) ? (T)0 : (T)a1;
  const typename V2<T>::t v = ld2(vu, n);
  st2(t, n, (T)(1.0 * (double)v.x + -1.0 * (double)g0), (T)(1.0 * (double)v.y + -1.0 * (double)g1));
}

template <class T>
void stokes_apply_fine(const Level<T>& L, const double* Gh, const T* z, T* y, long nv,
                       hipStream_t st) {
  const dim3 vt = tiles(L);
  const int nvt = (int)(vt.x * vt.y);
  const int ntxp = (L.nelx + 1 + PT - 1) / PT, ntzp = (L.nelz + 1 + PT - 1) / PT;
  k_stokes_fine<T><<<nvt + ntxp * ntzp, 256, 0, st>>>(view(L), Gh, z, z + nv, y, y + nv, nvt,
                                                       ntxp);
}

template <class T>
void schur_front(const Level<T>& L, const double* Gh, const double* minv, const T* v, T* zp,
                 T* t, long nv, hipStream_t st, const T* xs, double scale) {
  const int nbu = (int)((L.nn + 255) / 256);
  const int nbp = (int)(((long)(L.nelx + 1) * (L.nelz + 1) + 255) / 256);
  k_schur_front<T><<<nbu + nbp, 256, 0, st>>>(L.nelx, L.nelz, L.nnx, L.nnz, L.fix.p, Gh, minv,
                                              v, v + nv, xs, zp, t, nbu, scale);
}

// ---------------------------------------------------------------------------
// Chebyshev-Jacobi on the viscosity-weighted Q1 pressure mass matrix
// M_eta = sum_e (area / eta_e) Mhat (Mhat: the unit bilinear element mass,
// entries 4, 2, 1 / 36 for the node itself, an edge neighbour, the opposite
// corner).  Element by element, the Rayleigh quotient of M_e against its
// diagonal lies in [1/4, 9/4], so D^-1 M_eta has its spectrum there whatever
// eta is and the Chebyshev bounds are fixed.
__device__ __forceinline__ double pm_weight(int a, int b) {  // 36 * Mhat[a][b], reference order
  const int d = (a - b) & 3;
  return d == 0 ? 4.0 : (d == 2 ? 1.0 : 2.0);
}

// Element weights w_e = area / (36 eta_e) and the diagonal of M_eta with its
// inverse, once per viscosity field (the Chebyshev steps only read them).
__global__ void k_pmass_prepare(int nelx, int nelz, double area, const double* __restrict__ eta,
                                double* __restrict__ w, double* __restrict__ dg,
                                double* __restrict__ dinv) {
  const long m = (long)blockIdx.x * 256 + threadIdx.x;
  const int npx = nelx + 1;
  if (m < (long)nelx * nelz) w[m] = area / (36.0 * eta[m]);
  if (m >= (long)npx * (nelz + 1)) return;
  const int i = (int)(m % npx), j = (int)(m / npx);
  double d = 0.0;
  for (int dz = 0; dz < 2; dz++)
    for (int dxx = 0; dxx < 2; dxx++) {
      const int ex = i - 1 + dxx, ez = j - 1 + dz;
      if (ex < 0 || ex >= nelx || ez < 0 || ez >= nelz) continue;
      d += (area / (36.0 * eta[(long)ez * nelx + ex])) * 4.0;
    }
  dg[m] = d;
  dinv[m] = 1.0 / d;
}

// (M_eta x)(i, j) from the element weights; x(k, l) read through px
template <class PX>
__device__ __forceinline__ double pmass_row(int i, int j, int nelx, int nelz,
                                            const double* __restrict__ w, PX px) {
  double y = 0.0;
  for (int dz = 0; dz < 2; dz++)
    for (int dxx = 0; dxx < 2; dxx++) {
      const int ex = i - 1 + dxx, ez = j - 1 + dz;
      if (ex < 0 || ex >= nelx || ez < 0 || ez >= nelz) continue;
      const double we = w[(long)ez * nelx + ex];
      const int a = dxx ? (dz ? 0 : 3) : (dz ? 1 : 2);  // (i, j) in the element
      double s = 0.0;
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const int bi = ex + ((b == 1 || b == 2) ? 1 : 0), bj = ez + (b >= 2 ? 1 : 0);
        s += pm_weight(a, b) * px(bi, bj);
      }
      y += we * s;
    }
  return y;
}

// One launch per Chebyshev step; fused = 1: steps 0 and 1 (step 0 is
// pointwise from x = 0, so its result is formed where step 1 reads it)
template <class T>
__global__ void __launch_bounds__(256)
k_pmass_cheb(int nelx, int nelz, const double* __restrict__ w, const double* __restrict__ dgv,
             const double* __restrict__ dinv, const T* __restrict__ b, const T* __restrict__ xin,
             T* __restrict__ d, T* __restrict__ xout, double c20, double c1, double c2,
             int fused) {
  const int npx = nelx + 1;
  const long m = (long)blockIdx.x * 256 + threadIdx.x;
  if (m >= (long)npx * (nelz + 1)) return;
  const int i = (int)(m % npx), j = (int)(m / npx);
  const double dg = dgv[m];
  if (fused) {
    auto x1 = [&](int k, int l) {
      const long q = (long)l * npx + k;
      return (double)(T)(c20 * dinv[q] * (double)b[q]);
    };
    const double y = pmass_row(i, j, nelx, nelz, w, x1);
    const double x0 = x1(i, j);
    const double dn = c1 * x0 + c2 * ((double)b[m] - y) / dg;
    d[m] = (T)dn;
    xout[m] = (T)(x0 + dn);
  } else {
    const double y = pmass_row(i, j, nelx, nelz, w,
                               [&](int k, int l) { return (double)xin[(long)l * npx + k]; });
    const double dn = c1 * (double)d[m] + c2 * ((double)b[m] - y) / dg;
    d[m] = (T)dn;
    xout[m] = (T)((double)xin[m] + dn);
  }
}

int pmass_prepare(int nelx, int nelz, double area, const double* eta, double* w, double* dg,
                  double* dinv, hipStream_t st) {
  if (!eta || !w || !dg || !dinv) return fail(MW_EINVAL, "pmass_prepare: null pointer");
  const long n = std::max((long)nelx * nelz, (long)(nelx + 1) * (nelz + 1));
  k_pmass_prepare<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(nelx, nelz, area, eta, w, dg,
                                                               dinv);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

template <class T>
int pmass_cheb_solve(int nelx, int nelz, const double* w, const double* dg, const double* dinv,
                     const T* b, T* xa, T* xb, T* d, int steps, T** x, hipStream_t st) {
  if (steps < 2) return fail(MW_EINVAL, "pmass_cheb_solve: at least 2 steps");
  const double lmin = 0.25, lmax = 2.25;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
  double rho = 1.0 / sigma;
  double c1[32], c2[32];
  for (int k = 0; k < steps && k < 32; k++) {
    if (k == 0) {
      c1[k] = 0.0;
      c2[k] = 1.0 / theta;
    } else {
      const double rn = 1.0 / (2.0 * sigma - rho);
      c1[k] = rn * rho;
      c2[k] = 2.0 * rn / delta;
      rho = rn;
    }
  }
  const long np = (long)(nelx + 1) * (nelz + 1);
  const int nbk = (int)((np + 255) / 256);
  k_pmass_cheb<T><<<nbk, 256, 0, st>>>(nelx, nelz, w, dg, dinv, b, nullptr, d, xa, c2[0], c1[1],
                                       c2[1], 1);
  T* cur = xa;
  T* nxt = xb;
  for (int k = 2; k < steps && k < 32; k++) {
    k_pmass_cheb<T><<<nbk, 256, 0, st>>>(nelx, nelz, w, dg, dinv, b, cur, d, nxt, 0.0, c1[k],
                                         c2[k], 0);
    std::swap(cur, nxt);
  }
  MW_CHECK_LAUNCH();
  *x = cur;
  return MW_OK;
}

// The Chebyshev steps on M_eta and the preconditioner front in one launch.
// A block takes a 16x16 tile of velocity nodes; G reads z_p on the 10x10
// pressure nodes of the tile's elements, step k of S reads its predecessor on
// one more ring of pressure nodes, so step 0 covers a (10 + 2 (S - 1))^2
// patch and each later step one ring less (the patch is recomputed by the
// neighbouring tiles instead of a launch per step).  Every value is formed by
// the expressions of k_pmass_cheb and k_schur_front with the same operands,
// so the results are those of the separate launches bit for bit.
constexpr int SF_T = 16;               // velocity tile edge
constexpr int SF_P = SF_T / 2 + 2;     // pressure nodes G reads per tile edge
struct SchurCoef {
  double c20, scale;
  double c1[SCHUR_FUSED_MAX], c2[SCHUR_FUSED_MAX];
};

template <class T, int S>
__global__ void __launch_bounds__(256)
k_schur_front_cheb(int nelx, int nelz, int nnx, int nnz, const uint8_t* __restrict__ fix,
                   const double* __restrict__ Gh, const double* __restrict__ w,
                   const double* __restrict__ dgv, const double* __restrict__ dinv,
                   const T* __restrict__ vu, const T* __restrict__ b, T* __restrict__ zp,
                   T* __restrict__ t, SchurCoef c) {
  constexpr int H = S - 1, R = SF_P + 2 * H;
  __shared__ T sx[2][R * R], sd[R * R];
  const int tid = threadIdx.x;
  const int npx = nelx + 1, npz = nelz + 1;
  const int I0 = blockIdx.x * SF_T, J0 = blockIdx.y * SF_T;
  const int bi = I0 / 2 - 1 - H, bj = J0 / 2 - 1 - H;  // patch origin (pressure node)
  // step 0 from x = 0: x1 = c20 D^-1 b, pointwise (d = x1)
  for (int e = tid; e < R * R; e += 256) {
    const int i = bi + e % R, j = bj + e / R;
    T x1 = 0;
    if (i >= 0 && i < npx && j >= 0 && j < npz) {
      const long q = (long)j * npx + i;
      x1 = (T)(c.c20 * dinv[q] * (double)b[q]);
    }
    sx[0][e] = x1;
    sd[e] = x1;
  }
  __syncthreads();
#pragma unroll
  for (int k = 1; k < S; k++) {
    const T* xin = sx[(k - 1) & 1];
    T* xout = sx[k & 1];
    const int h = S - 1 - k, n = SF_P + 2 * h, o = H - h;  // this step's ring and offset
    for (int e = tid; e < n * n; e += 256) {
      const int li = o + e % n, lj = o + e / n;
      const int i = bi + li, j = bj + lj;
      if (i < 0 || i >= npx || j < 0 || j >= npz) continue;
      const long m = (long)j * npx + i;
      const double dg = dgv[m];
      const double y = pmass_row(i, j, nelx, nelz, w, [&](int kk, int ll) {
        return (double)xin[(ll - bj) * R + (kk - bi)];
      });
      const double dn = c.c1[k] * (double)sd[lj * R + li] + c.c2[k] * ((double)b[m] - y) / dg;
      sd[lj * R + li] = (T)dn;
      const T xn = (T)((double)xin[lj * R + li] + dn);
      if (k < S - 1) {
        xout[lj * R + li] = xn;
      } else {
        const T z = (T)(-c.scale * (double)xn);  // z_p = -scale x
        xout[lj * R + li] = z;
        if (2 * i >= I0 && 2 * i < I0 + SF_T && 2 * j >= J0 && 2 * j < J0 + SF_T) zp[m] = z;
      }
    }
    __syncthreads();
  }
  const T* zs = sx[(S - 1) & 1];
  const int I = I0 + tid % SF_T, J = J0 + tid / SF_T;
  if (I >= nnx || J >= nnz) return;
  const long n = (long)J * nnx + I;
  double a0, a1;
  g_row(Gh, I, J, nelx, nelz,
        [&](int ex, int ez, int k) {
          const int pi = ex + (k == 1 || k == 2 ? 1 : 0), pj = ez + (k >= 2 ? 1 : 0);
          return (double)zs[(pj - bj) * R + (pi - bi)];
        },
        a0, a1);
  const unsigned f = fix2(fix, n);
  const T g0 = (f & 0xff) ? (T)0 : (T)a0, g1 = (f >> 8) ? (T)0 : (T)a1;
  const typename V2<T>::t v = ld2(vu, n);
  st2(t, n, (T)(1.0 * (double)v.x + -1.0 * (double)g0), (T)(1.0 * (double)v.y + -1.0 * (double)g1));
}

template <class T>
int schur_front_cheb(const Level<T>& L, const double* Gh, const double* w, const double* dg,
                     const double* dinv, const T* v, T* zp, T* t, long nv, int steps,
                     double scale, hipStream_t st) {
  if (steps < 2 || steps > SCHUR_FUSED_MAX)
    return fail(MW_EINVAL, "schur_front_cheb: steps outside [2, SCHUR_FUSED_MAX]");
  SchurCoef c{};
  const double lmin = 0.25, lmax = 2.25;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
  double rho = 1.0 / sigma;
  c.c20 = 1.0 / theta;
  c.scale = scale;
  for (int k = 1; k < steps; k++) {
    const double rn = 1.0 / (2.0 * sigma - rho);
    c.c1[k] = rn * rho;
    c.c2[k] = 2.0 * rn / delta;
    rho = rn;
  }
  const dim3 grid((L.nnx + SF_T - 1) / SF_T, (L.nnz + SF_T - 1) / SF_T);
  const T* b = v + nv;
#define MW_SF(S)                                                                              \
  k_schur_front_cheb<T, S><<<grid, 256, 0, st>>>(L.nelx, L.nelz, L.nnx, L.nnz, L.fix.p, Gh, w, \
                                                 dg, dinv, v, b, zp, t, c)
  switch (steps) {
    case 2: MW_SF(2); break;
    case 3: MW_SF(3); break;
    default: MW_SF(4); break;
  }
#undef MW_SF
  MW_CHECK_LAUNCH();
  return MW_OK;
}

template <class T>
void level_apply(const Level<T>& L, const T* x, T* y, hipStream_t st) {
  k_level_op<T, OP_APPLY><<<tiles(L), 256, 0, st>>>(view(L), x, nullptr, y, nullptr, nullptr,
                                                    nullptr, 0, 0);
}

template <class T>
Gmg<T>::~Gmg() {
  side_k = -1;  // nothing more is queued
  if (side) {
    (void)hipStreamSynchronize(side);
    (void)hipStreamDestroy(side);
  }
  if (ev_dense) (void)hipEventDestroy(ev_dense);
  if (ev_inv) (void)hipEventDestroy(ev_inv);
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (auto* l : lev) delete l;
}

template <class T>
int Gmg<T>::setup(int nelx, int nelz, double Lx, double Lz, const uint8_t* fine_fix,
                  hipStream_t st) {
  if (side) MW_HIP(hipStreamSynchronize(side));  // no inverse in flight into freed buffers
  inv_pending = false;
  side_k = -1;
  for (auto* l : lev) delete l;
  lev.clear();
  have_pvec = false;
  have_next = false;
  int nx = nelx, nz = nelz;
  while (true) {
    auto* L = new Level<T>();
    L->nelx = nx; L->nelz = nz; L->nnx = 2 * nx + 1; L->nnz = 2 * nz + 1;
    L->nn = (long)L->nnx * L->nnz; L->ndof = 2 * L->nn;
    L->hx = Lx / nx; L->hz = Lz / nz;
    int rc;
    if ((rc = L->eta.alloc((size_t)nx * nz)) || (rc = L->fix.alloc(L->ndof)) ||
        (rc = L->diag.alloc(L->ndof)) || (rc = L->invdiag.alloc(L->ndof)) ||
        (rc = L->Khat.alloc(324)) || (rc = L->r.alloc(L->ndof)) || (rc = L->x.alloc(L->ndof)) ||
        (rc = L->x2.alloc(L->ndof)) || (rc = L->d.alloc(L->ndof)) || (rc = L->b.alloc(L->ndof)) ||
        (rc = L->coef.alloc(2 * (size_t)std::max(std::max(nsmooth, nsmooth_coarse), 1))) ||
        (rc = L->pvec.alloc(L->ndof))) {
      delete L;
      return rc;
    }
    k_khat<T><<<1, 256, 0, st>>>(L->hx, L->hz, L->Khat.p);
    if (lev.empty()) {
      MW_HIP(hipMemcpyAsync(L->fix.p, fine_fix, L->ndof, hipMemcpyDeviceToDevice, st));
    } else {
      Level<T>* F = lev.back();
      k_coarsen_fix<<<nblocks(L->nn, 256), 256, 0, st>>>(L->nnx, L->nnz, F->nnx, F->fix.p,
                                                           L->fix.p);
    }
    MW_CHECK_LAUNCH();
    L->fine = lev.empty();
    lev.push_back(L);
    bool can = (nx % 2 == 0) && (nz % 2 == 0) && nx >= 4 && nz >= 4;
    if (!can || L->ndof <= coarse_max_dof) break;
    nx /= 2;
    nz /= 2;
  }
  Level<T>* C = lev.back();
  ncoarse = (int)C->ndof;
  if ((int)lev.size() > MAX_LEVELS) return fail(MW_EUNSUPPORTED, "gmg: too many levels");
  tail_from = 0;
  tail_lds = 0;
  if (fuse_tail)
    for (size_t l = 1; l + 1 < lev.size(); l++) {
      if (lev[l]->ndof > TAIL_MAX_DOF) continue;
      size_t bytes = (size_t)ncoarse * ncoarse * sizeof(T);
      for (size_t k = l; k < lev.size(); k++)
        bytes += tail_level_bytes<T>(lev[k]->nelx, lev[k]->nelz, lev[k]->ndof, nsmooth);
      if (bytes > TAIL_MAX_LDS) continue;
      tail_from = (int)l;
      tail_lds = bytes;
      MW_HIP(hipFuncSetAttribute((const void*)k_vcycle_tail<T>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
      break;
    }
  if (ncoarse > GJ_MAX_N) return fail(MW_EUNSUPPORTED, "gmg: coarsest level too large for a dense solve");
  // the last smoothed level and the coarsest solve run as one workgroup when
  // the level is one 16x16 tile, the coarsest fits the LDS inverse, and the
  // smoothing is the degree-2 Chebyshev the kernel implements
  bottom = false;
  bottom_ml = 0;
  if (allow_bottom > 0 && lev.size() >= 2 && tail_from == 0 && ncoarse <= BOT_MAX_NC) {
    const int nl = (int)lev.size();
    const Level<T>& B = *lev[nl - 2];
    bottom = B.nnx <= TT && B.nnz <= TT && steps(nl - 2) == 2;
    // up to three smoothed levels fused with the coarsest solve (FP32: the
    // FP64 arrays exceed the LDS): the last in the 9x9 slot, the one above in
    // the 17x17 slot, then the 33x33 slot; the finest level keeps its own kernels
    const int edge[ML_MAXLEV] = {ML_N2, ML_N1, ML_N0};
    for (int k = 1; k <= ML_MAXLEV && k <= allow_bottom && sizeof(T) == sizeof(float); k++) {
      const int l = nl - 1 - k;
      if (l < 1) break;
      const Level<T>& M = *lev[l];
      if (M.nnx > edge[k - 1] || M.nnz > edge[k - 1] || steps(l) != 2) break;
      bottom_ml = k;
    }
    if (bottom_ml < 2) bottom_ml = 0;  // one level: the 256-thread kernel
    if (bottom_ml) bottom = false;
  }
  int rc;
  if ((rc = coarse_inv.alloc((size_t)ncoarse * ncoarse)) ||
      (rc = coarse_a.alloc(ncoarse <= GJ_LDS_MAX_N ? (size_t)ncoarse * ncoarse
                                                   : 2 * (size_t)gj_pad(ncoarse) * gj_pad(ncoarse))) ||
      (ncoarse > GJ_LDS_MAX_N && (rc = coarse_next.alloc((size_t)ncoarse * ncoarse))) ||
      (rc = pnorm.alloc(2 * MAX_LEVELS)))
    return rc;
  if (nsmooth < 1) return fail(MW_EINVAL, "gmg: Chebyshev degree must be >= 1");
  if (fuse_tail && nsmooth_coarse > 0 && nsmooth_coarse != nsmooth)
    return fail(MW_EINVAL, "gmg: the fused tail needs one Chebyshev degree on all levels");
  // the levels below the finest run as one captured graph (fixed buffers,
  // coefficients read from device memory)
  if (lev.size() > 1) {
    // captured on a private stream (the caller's may be the null stream, which
    // cannot be captured); the graph is launched on the caller's stream
    MW_HIP(hipStreamSynchronize(st));
    hipStream_t cs;
    MW_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    int rc2 = e == hipSuccess ? vcycle_level(1, cs) : MW_OK;
    hipError_t e2 = e == hipSuccess ? hipStreamEndCapture(cs, &graph) : e;
    (void)hipStreamDestroy(cs);
    if (rc2) return rc2;
    if (e2 != hipSuccess) return fail(MW_EHIP, std::string("gmg capture: ") + hipGetErrorString(e2));
    MW_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
  }
  return MW_OK;
}

template <class T>
int Gmg<T>::update(const double* eta_fine, Reducer& R, hipStream_t st) {
  Level<T>* F = lev[0];
  long nel0 = (long)F->nelx * F->nelz;
  RC(pump_side(1 << 30));  // whatever the last solve left unqueued
  if (inv_pending || have_next) {  // the last inverse's side work reads the levels and
    MW_HIP(hipStreamWaitEvent(st, ev_inv, 0));  // coarse_a: wait before changing them
    inv_pending = false;
  }
  if (sizeof(T) == sizeof(double))
    MW_HIP(hipMemcpyAsync(F->eta.p, eta_fine, nel0 * sizeof(double), hipMemcpyDeviceToDevice, st));
  else
    k_eta_from_double<<<nblocks(nel0, 256), 256, 0, st>>>(nel0, eta_fine, (float*)F->eta.p);
  for (size_t l = 1; l < lev.size(); l++) {
    Level<T>* L = lev[l];
    k_coarsen_eta<T><<<nblocks((long)L->nelx * L->nelz, 256), 256, 0, st>>>(
        L->nelx, L->nelz, lev[l - 1]->eta.p, L->eta.p, eta_coarsen);
  }
  const int nl = (int)lev.size();
  LevelSet<T> S = level_set(lev, nl);
  k_set_diag<T><<<S.nblk0[nl], 256, 0, st>>>(S);
  MW_CHECK_LAUNCH();
  // coarsest: dense operator and inverse on the side stream while the power
  // iterations below run on this one
  Level<T>* C = lev.back();
  int n = ncoarse;
  double* A = coarse_a.p;
  if (!side) {
    MW_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));  // default priority
    MW_HIP(hipEventCreateWithFlags(&ev_dense, hipEventDisableTiming));
    MW_HIP(hipEventCreateWithFlags(&ev_inv, hipEventDisableTiming));
  }
  if (n <= GJ_LDS_MAX_N) {
    MW_HIP(hipEventRecord(ev_dense, st));
    MW_HIP(hipStreamWaitEvent(side, ev_dense, 0));
    k_coarse_dense<T><<<nblocks(n, 128), 128, 0, side>>>(view(*C), A, n, true);
    k_gauss_jordan_lds<T><<<1, 1024, 0, side>>>(n, A, coarse_inv.p);
    MW_CHECK_LAUNCH();
    MW_HIP(hipEventRecord(ev_inv, side));
    inv_pending = true;
  } else {
    // Larger coarsest levels: the blocked inverse takes longer than the
    // power iterations, so it runs one solve behind.  This solve's V-cycles
    // use the inverse of the previous update's operator (copied in from
    // coarse_next, whose side-stream work ended during the previous solve);
    // the new one is formed into coarse_next meanwhile.  Its sweeps are
    // queued a few at a time by pump_side() (from the solver's iteration loop,
    // while the host would wait for the device anyway), so queuing ~40
    // launches never holds back the caller's stream.  The first update waits
    // for its own.
    if (have_next)
      MW_HIP(hipMemcpyAsync(coarse_inv.p, coarse_next.p, (size_t)n * n * sizeof(T),
                            hipMemcpyDeviceToDevice, st));
    MW_HIP(hipEventRecord(ev_dense, st));
    MW_HIP(hipStreamWaitEvent(side, ev_dense, 0));
    // padded to whole tiles, then one sweep per pivot block (pump_side),
    // ping-ponging between the two work matrices
    const int np = gj_pad(n);
    k_identity<<<nblocks((long)np * np, 256), 256, 0, side>>>(np, A);
    k_coarse_dense<T><<<nblocks(n, 128), 128, 0, side>>>(view(*C), A, np, false);
    MW_CHECK_LAUNCH();
    side_k = 0;
    if (!have_next) {
      RC(pump_side(1 << 30));
      MW_HIP(hipStreamWaitEvent(st, ev_inv, 0));
      MW_HIP(hipMemcpyAsync(coarse_inv.p, coarse_next.p, (size_t)n * n * sizeof(T),
                            hipMemcpyDeviceToDevice, st));
    }
    have_next = true;
    inv_pending = false;
  }
  // largest eigenvalue of D^-1 A per smoothed level by power iteration, all
  // smoothed levels per launch; the first update starts from a fixed hash
  // vector, later ones (power_its_warm > 0) from the previous update's vectors,
  // which the slowly changing viscosity leaves close to the new ones
  if (nl > 1) {
    LevelSet<T> P = level_set(lev, nl - 1);
    const int np = nl - 1;
    for (int l = 0; l < np; l++) P.x[l] = lev[l]->pvec.p;
    double* nrm = pnorm.p;  // [0, MAX_LEVELS): ||x||^2, [MAX_LEVELS, 2*MAX_LEVELS): final
    const bool warm = have_pvec && power_its_warm > 0;
    if (!warm) k_set_hash_fill<T><<<P.vblk0[np], 256, 0, st>>>(P);
    const int its = warm ? power_its_warm : power_its;
    have_pvec = true;
    for (int it = 0; it < its; it++) {
      k_set_apply<T><<<P.tile0[np], 256, 0, st>>>(P);
      // v = D^-1 A v / ||v|| ; ||v|| measured before scaling
      k_set_dot_partial<T><<<P.dblk0[np], DOT_BS, 0, st>>>(P, R.partial.p);
      k_set_dot_final<T><<<1, 64 * MAX_LEVELS, 0, st>>>(P, R.partial.p, nrm);
      k_set_scale<T><<<P.vblk0[np], 256, 0, st>>>(P, nrm);
    }
    k_set_dot_partial<T><<<P.dblk0[np], DOT_BS, 0, st>>>(P, R.partial.p);
    k_set_dot_final<T><<<1, 64 * MAX_LEVELS, 0, st>>>(P, R.partial.p, nrm + MAX_LEVELS);
    for (int l = 0; l < np; l++) P.steps[l] = steps(l);
    k_set_cheb_coef<T><<<1, 64, 0, st>>>(P, nrm + MAX_LEVELS);
    MW_CHECK_LAUNCH();
  }
  if (!solver_pumps) RC(pump_side(1 << 30));  // queued after this stream's work
  return MW_OK;
}

// Queue up to nmax pending sweeps of the blocked coarsest inverse on the side
// stream; after the last one, the extraction into coarse_next and ev_inv.
template <class T>
int Gmg<T>::pump_side(int nmax) {
  if (side_k < 0) return MW_OK;
  const int n = ncoarse, np = gj_pad(n);
  double* W[2] = {coarse_a.p, coarse_a.p + (size_t)np * np};
  const dim3 grid(np / GJB_TILE, np / GJB_TILE);
  for (int i = 0; i < nmax && side_k < n; i++, side_k += GJB) {
    const int s = (side_k / GJB) & 1;
    k_gj_sweep<<<grid, 256, 0, side>>>(np, side_k, W[s], W[s ^ 1]);
  }
  if (side_k >= n) {
    const int s = ((n + GJB - 1) / GJB) & 1;
    k_extract_padded<T><<<nblocks((long)n * n, 256), 256, 0, side>>>(n, np, W[s], coarse_next.p);
    MW_HIP(hipEventRecord(ev_inv, side));
    side_k = -1;
  }
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// one Chebyshev step k of level L: in -> out (in == nullptr: from x = 0)
template <class T>
static void cheb_step(Level<T>& L, const T* in, T* out, int k, hipStream_t st) {
  const int zx = in ? 0 : 1;
  if (L.fine) {
    prof().begin(PROF_CHEB_FINE, st);
    k_level_op<T, OP_CHEB, true><<<tiles(L), 256, 0, st>>>(view(L), in, L.b.p, nullptr, L.d.p,
                                                           out, L.coef.p, k, zx);
    prof().end(PROF_CHEB_FINE, st);
  } else {
    k_level_op<T, OP_CHEB><<<tiles(L), 256, 0, st>>>(view(L), in, L.b.p, nullptr, L.d.p, out,
                                                     L.coef.p, k, zx);
  }
}

// Pre-smoothing from x = 0, the result in L.x: step k writes x when
// steps-1-k is even and x2 otherwise, so the last step lands in x without a
// copy.  Steps 0 and 1 run as one fused kernel.
template <class T>
static void pre_smooth(Level<T>& L, int steps, hipStream_t st) {
  auto dst = [&](int k) { return ((steps - 1 - k) % 2 == 0) ? L.x.p : L.x2.p; };
  int k0 = 1;
  if (steps >= 2) {
    if (L.fine)
      k_cheb01<T, true><<<tiles(L), 256, 0, st>>>(view(L), L.b.p, L.d.p, dst(1), L.coef.p);
    else
      k_cheb01<T><<<tiles(L), 256, 0, st>>>(view(L), L.b.p, L.d.p, dst(1), L.coef.p);
    k0 = 2;
  } else {
    cheb_step<T>(L, nullptr, dst(0), 0, st);
  }
  for (int k = k0; k < steps; k++) cheb_step<T>(L, dst(k - 1), dst(k), k, st);
}

// Post-smoothing after the coarse correction: x += P xc fused with step 0
// (x -> x2), then the remaining steps alternate; the result ends in L.x.
template <class T>
static void post_smooth(Level<T>& L, const Level<T>& C, int steps, hipStream_t st) {
  if (L.fine)
    k_prolong_cheb0<T, true, PC_BS_FINE><<<tiles(L), PC_BS_FINE, 0, st>>>(view(L), C.nnx, C.nnz, C.x.p, L.x.p,
                                                       L.b.p, L.d.p, L.x2.p, L.coef.p);
  else
    k_prolong_cheb0<T, false, PC_BS><<<tiles(L), PC_BS, 0, st>>>(view(L), C.nnx, C.nnz, C.x.p, L.x.p, L.b.p,
                                                 L.d.p, L.x2.p, L.coef.p);
  T* a = L.x2.p;
  for (int k = 1; k < steps; k++) {
    T* out = (a == L.x.p) ? L.x2.p : L.x.p;
    cheb_step<T>(L, a, out, k, st);
    a = out;
  }
  if (a != L.x.p) k_copy<T><<<nblocks(L.ndof, 256), 256, 0, st>>>(L.ndof, a, L.x.p);
}

template <class T>
int Gmg<T>::vcycle_level(int l, hipStream_t st) {
  Level<T>& L = *lev[l];
  {  // levels that read the coarsest inverse at once wait for it here
    const int last = (int)lev.size() - 1;
    const bool direct = (tail_from > 0 && l == tail_from) || l == last ||
                        (bottom_ml && l == last - bottom_ml) || (bottom && l == last - 1);
    if (direct && inv_pending) {
      MW_HIP(hipStreamWaitEvent(st, ev_inv, 0));
      inv_pending = false;
    }
  }
  if (tail_from > 0 && l == tail_from) {
    TailArgs<T> A;
    A.nlev = (int)lev.size() - l;
    A.nsmooth = nsmooth;
    A.ncoarse = ncoarse;
    A.coarse_inv = coarse_inv.p;
    for (int k = 0; k < A.nlev; k++) {
      Level<T>& M = *lev[l + k];
      A.L[k] = TailLevel<T>{view(M), M.x.p, M.b.p, M.coef.p, nsmooth};
    }
    k_vcycle_tail<T><<<1, TAIL_BS, tail_lds, st>>>(A);
    MW_CHECK_LAUNCH();
    return MW_OK;
  }
  if (l == (int)lev.size() - 1) {
    k_dense_mv<T><<<nblocks((long)ncoarse * 64, 256), 256, 0, st>>>(ncoarse, coarse_inv.p, L.b.p,
                                                                    L.x.p);
    MW_CHECK_LAUNCH();
    return MW_OK;
  }
  if (bottom_ml && l == (int)lev.size() - 1 - bottom_ml) {
    if constexpr (sizeof(T) != sizeof(float)) {
      return fail(MW_EUNSUPPORTED, "gmg: the fused multi-level bottom is FP32 only");
    } else {
    MlArgs<T> A{};
    A.nlev = bottom_ml;
    for (int k = 0; k < bottom_ml; k++) {
      Level<T>& M = *lev[l + k];
      const int slot = ML_MAXLEV - bottom_ml + k;
      A.L[slot] = view(M);
      A.coef[slot] = M.coef.p;
    }
    const Level<T>& C = *lev.back();
    A.b = L.b.p;
    A.x = L.x.p;
    A.nnx_c = C.nnx;
    A.nnz_c = C.nnz;
    A.nc = ncoarse;
    A.fixc = C.fix.p;
    A.cinv = coarse_inv.p;
    k_vcycle_bottom_ml<T><<<1, ML_BS, 0, st>>>(A);
    MW_CHECK_LAUNCH();
    return MW_OK;
    }
  }
  if (bottom && l == (int)lev.size() - 2) {
    const Level<T>& C = *lev[l + 1];
    k_vcycle_bottom<T><<<1, 256, 0, st>>>(view(L), C.nnx, C.nnz, C.fix.p, ncoarse, coarse_inv.p,
                                          L.b.p, L.x.p, L.coef.p);
    MW_CHECK_LAUNCH();
    return MW_OK;
  }
  pre_smooth(L, steps(l), st);
  if (inv_pending) {  // the levels below use the coarsest inverse
    MW_HIP(hipStreamWaitEvent(st, ev_inv, 0));
    inv_pending = false;
  }
  Level<T>& C = *lev[l + 1];
  const dim3 cgrid((C.nnx + RT - 1) / RT, (C.nnz + RT - 1) / RT);
  const dim3 cgrid8((C.nnx + R8T - 1) / R8T, (C.nnz + R8T - 1) / R8T);
  const bool small = (L.fine ? RR8_LEVELS >= 2 : RR8_LEVELS >= 1);
  if (small && L.fine)
    k_resid_restrict8<T, true><<<cgrid8, RR_BS, 0, st>>>(view(L), C.nnx, C.nnz, C.fix.p, L.x.p,
                                                       L.b.p, C.b.p);
  else if (small)
    k_resid_restrict8<T><<<cgrid8, RR_BS, 0, st>>>(view(L), C.nnx, C.nnz, C.fix.p, L.x.p, L.b.p,
                                                 C.b.p);
  else if (L.fine)
    k_resid_restrict<T, true><<<cgrid, RR_BS, 0, st>>>(view(L), C.nnx, C.nnz, C.fix.p, L.x.p,
                                                     L.b.p, C.b.p);
  else
    k_resid_restrict<T><<<cgrid, RR_BS, 0, st>>>(view(L), C.nnx, C.nnz, C.fix.p, L.x.p, L.b.p,
                                               C.b.p);
  if (l == 0 && graph_exec && !inline_coarse) {
    MW_HIP(hipGraphLaunch(graph_exec, st));
  } else {
    int rc = vcycle_level(l + 1, st);
    if (rc) return rc;
  }
  post_smooth(L, C, steps(l), st);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

template <class T>
void Gmg<T>::restrict_to(int l, const T* rf, T* bc, hipStream_t st) {
  Level<T>& L = *lev[l];
  Level<T>& C = *lev[l + 1];
  launch_restrict(L, C, rf, bc, st);
}

template <class T>
void Gmg<T>::smooth_steps(int l, int reps, hipStream_t st) {
  Level<T>& L = *lev[l];
  const int k = nsmooth > 1 ? 1 : 0;
  for (int i = 0; i < reps; i++) {
    const T* in = (i & 1) ? L.x2.p : L.x.p;
    T* out = (i & 1) ? L.x.p : L.x2.p;
    if (L.fine)
      k_level_op<T, OP_CHEB, true><<<tiles(L), 256, 0, st>>>(view(L), in, L.b.p, nullptr, L.d.p,
                                                             out, L.coef.p, k, 0);
    else
      k_level_op<T, OP_CHEB><<<tiles(L), 256, 0, st>>>(view(L), in, L.b.p, nullptr, L.d.p, out,
                                                       L.coef.p, k, 0);
  }
}

template <class T>
void Gmg<T>::prolong_add(int l, const T* xc, T* xf, hipStream_t st) {
  Level<T>& L = *lev[l];
  Level<T>& C = *lev[l + 1];
  k_prolong_add<T><<<nblocks(L.nn, 256), 256, 0, st>>>(L.nnx, L.nnz, C.nnx, C.nnz, L.fix.p, xc,
                                                      xf);
}

template <class T>
int Gmg<T>::vcycle(const T* r, T* z, hipStream_t st) {
  Level<T>& L = *lev[0];
  if (nsmooth % 2 == 0 && lev.size() > 1 && r != z) {
    // the finest level reads b and leaves its result in x (even smoothing
    // degree), so point them at the caller's vectors for this cycle instead of
    // copying; the levels below never touch level 0's b and x
    T* b0 = L.b.p;
    T* x0 = L.x.p;
    L.b.p = const_cast<T*>(r);
    L.x.p = z;
    int rc = vcycle_level(0, st);
    L.b.p = b0;
    L.x.p = x0;
    return rc;
  }
  k_copy<T><<<nblocks(L.ndof, 256), 256, 0, st>>>(L.ndof, r, L.b.p);
  int rc = vcycle_level(0, st);
  if (rc) return rc;
  k_copy<T><<<nblocks(L.ndof, 256), 256, 0, st>>>(L.ndof, L.x.p, z);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

template struct Gmg<double>;
template struct Gmg<float>;
template void level_apply<double>(const Level<double>&, const double*, double*, hipStream_t);
template void level_apply<float>(const Level<float>&, const float*, float*, hipStream_t);
template void stokes_apply_fine<float>(const Level<float>&, const double*, const float*, float*,
                                       long, hipStream_t);
template void schur_front<float>(const Level<float>&, const double*, const double*, const float*,
                                 float*, float*, long, hipStream_t, const float*, double);
template int pmass_cheb_solve<float>(int, int, const double*, const double*, const double*,
                                     const float*, float*, float*, float*, int, float**,
                                     hipStream_t);
template int schur_front_cheb<float>(const Level<float>&, const double*, const double*,
                                     const double*, const double*, const float*, float*, float*,
                                     long, int, double, hipStream_t);

}  // namespace mw
