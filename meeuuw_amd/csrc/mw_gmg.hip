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
// update; coarsest level: dense inverse (Gauss-Jordan in one workgroup).
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
// of nodes and stages in LDS the 20x20 node patch its elements touch (x and
// the fixed-dof flags), the viscosities of the covering elements and Khat, so
// every gather of the 18x18 element rows is an LDS read and the global loads
// are one coalesced round trip.
constexpr int TT = 16;           // tile edge (nodes)
constexpr int TP = TT + 4;       // patch edge
constexpr int TE = TT / 2 + 3;   // covering elements per direction

enum { OP_APPLY = 0, OP_RESID = 1, OP_CHEB = 2 };

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

// FINE only names the finest-level instance (profiles list it on its own)
template <class T, int OP, bool FINE = false>
__global__ void __launch_bounds__(256)
k_level_op(LV<T> L, const T* __restrict__ x, const T* __restrict__ b, T* __restrict__ y,
           T* __restrict__ d, T* __restrict__ xout, const T* __restrict__ coef, int step,
           int zero_x) {
  __shared__ T sK[324];
  __shared__ T sx[TP][TP][2];
  __shared__ uint8_t sf[TP][TP][2];
  __shared__ T se[TE][TE];
  const int tid = threadIdx.x;
  int tX, tZ;
  xcd_tile(tX, tZ);
  const int I0 = tX * TT, J0 = tZ * TT;
  const int ex0 = I0 > 0 ? (I0 - 1) >> 1 : 0, ez0 = J0 > 0 ? (J0 - 1) >> 1 : 0;
  // the node's own operands are requested first, so their latency overlaps the patch staging
  const int tx = tid % TT, ty = tid / TT;
  const int I = I0 + tx, J = J0 + ty;
  const bool valid = I < L.nnx && J < L.nnz;
  const long n = valid ? (long)J * L.nnx + I : 0;
  T pb0 = 0, pb1 = 0, pi0 = 0, pi1 = 0, pd0 = 0, pd1 = 0;
  if (valid && OP != OP_APPLY) {
    pb0 = b[2 * n];
    pb1 = b[2 * n + 1];
  }
  if (valid && OP == OP_CHEB) {
    pi0 = L.invdiag[2 * n];
    pi1 = L.invdiag[2 * n + 1];
    if (step > 0) {
      pd0 = d[2 * n];
      pd1 = d[2 * n + 1];
    }
  }
  for (int i = tid; i < 324; i += 256) sK[i] = L.Kh[i];
  const bool need_x = !(OP == OP_CHEB && zero_x);
  for (int i = tid; i < TP * TP; i += 256) {
    int pi = i % TP, pj = i / TP;
    int gI = I0 - 2 + pi, gJ = J0 - 2 + pj;
    T v0 = 0, v1 = 0;
    uint8_t f0 = 1, f1 = 1;
    if (gI >= 0 && gI < L.nnx && gJ >= 0 && gJ < L.nnz) {
      long m = (long)gJ * L.nnx + gI;
      f0 = L.fix[2 * m];
      f1 = L.fix[2 * m + 1];
      if (need_x) {
        v0 = x[2 * m];
        v1 = x[2 * m + 1];
      }
    }
    sx[pj][pi][0] = v0;
    sx[pj][pi][1] = v1;
    sf[pj][pi][0] = f0;
    sf[pj][pi][1] = f1;
  }
  for (int i = tid; i < TE * TE; i += 256) {
    int a = i % TE, c = i / TE;
    int eX = ex0 + a, eZ = ez0 + c;
    se[c][a] = (eX < L.nelx && eZ < L.nelz) ? L.eta[(long)eZ * L.nelx + eX] : (T)0;
  }
  __syncthreads();
  if (!valid) return;
  T y0 = 0, y1 = 0;
  if (need_x) {
    int exa[2], lxa[2], eza[2], lza[2];
    int cx = q2_adjacent(I, L.nelx, exa, lxa), cz = q2_adjacent(J, L.nelz, eza, lza);
    for (int iz = 0; iz < cz; iz++)
      for (int ix = 0; ix < cx; ix++) {
        int a = q2_local(lxa[ix], lza[iz]);
        const T* r0 = sK + (2 * a) * 18;
        const T* r1 = r0 + 18;
        int bx = 2 * exa[ix] - (I0 - 2), bz = 2 * eza[iz] - (J0 - 2);
        T s0 = 0, s1 = 0;
#pragma unroll
        for (int mz = 0; mz < 3; mz++)
#pragma unroll
          for (int mx = 0; mx < 3; mx++) {
            int bb = q2_local(mx, mz);
            T xu = sf[bz + mz][bx + mx][0] ? (T)0 : sx[bz + mz][bx + mx][0];
            T xw = sf[bz + mz][bx + mx][1] ? (T)0 : sx[bz + mz][bx + mx][1];
            s0 += r0[2 * bb] * xu + r0[2 * bb + 1] * xw;
            s1 += r1[2 * bb] * xu + r1[2 * bb + 1] * xw;
          }
        T et = se[eza[iz] - ez0][exa[ix] - ex0];
        y0 += et * s0;
        y1 += et * s1;
      }
    T x0 = sx[ty + 2][tx + 2][0], x1 = sx[ty + 2][tx + 2][1];
    if (sf[ty + 2][tx + 2][0]) y0 = L.diag[2 * n] * x0;
    if (sf[ty + 2][tx + 2][1]) y1 = L.diag[2 * n + 1] * x1;
  }
  if (OP == OP_APPLY) {
    y[2 * n] = y0;
    y[2 * n + 1] = y1;
  } else if (OP == OP_RESID) {
    y[2 * n] = pb0 - y0;
    y[2 * n + 1] = pb1 - y1;
  } else {
    const T c1 = coef[2 * step], c2 = coef[2 * step + 1];
    T x0 = 0, x1 = 0;
    if (need_x) {
      x0 = sx[ty + 2][tx + 2][0];
      x1 = sx[ty + 2][tx + 2][1];
    }
    T d0 = c2 * pi0 * (pb0 - y0);
    T d1 = c2 * pi1 * (pb1 - y1);
    if (step > 0) {
      d0 += c1 * pd0;
      d1 += c1 * pd1;
    }
    d[2 * n] = d0;
    d[2 * n + 1] = d1;
    xout[2 * n] = x0 + d0;
    xout[2 * n + 1] = x1 + d1;
  }
}

template <class T>
static dim3 tiles(const Level<T>& L) {
  return dim3((L.nnx + TT - 1) / TT, (L.nnz + TT - 1) / TT);
}

// diagonal of the unmasked operator: sum of eta_e * Khat[ii] over elements
template <class T>
__global__ void k_level_diag(LV<T> L, T* __restrict__ diag, T* __restrict__ invdiag) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= (long)L.nnx * L.nnz) return;
  int I = (int)(n % L.nnx), J = (int)(n / L.nnx);
  int ex[2], lx[2], ez[2], lz[2];
  int cx = q2_adjacent(I, L.nelx, ex, lx), cz = q2_adjacent(J, L.nelz, ez, lz);
  T d0 = 0, d1 = 0;
  for (int iz = 0; iz < cz; iz++)
    for (int ix = 0; ix < cx; ix++) {
      int a = q2_local(lx[ix], lz[iz]);
      T et = L.eta[ez[iz] * L.nelx + ex[ix]];
      d0 += et * L.Kh[(2 * a) * 18 + 2 * a];
      d1 += et * L.Kh[(2 * a + 1) * 18 + 2 * a + 1];
    }
  diag[2 * n] = d0;
  diag[2 * n + 1] = d1;
  invdiag[2 * n] = (T)1 / d0;
  invdiag[2 * n + 1] = (T)1 / d1;
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

// coarse viscosity: mean of the four children
template <class T>
__global__ void k_coarsen_eta(int nelx_c, int nelz_c, const T* __restrict__ ef,
                              T* __restrict__ ec) {
  int E = blockIdx.x * blockDim.x + threadIdx.x;
  if (E >= nelx_c * nelz_c) return;
  int X = E % nelx_c, Z = E / nelx_c, nfx = 2 * nelx_c;
  long f = (long)(2 * Z) * nfx + 2 * X;
  ec[E] = (T)0.25 * (ef[f] + ef[f + 1] + ef[f + nfx] + ef[f + nfx + 1]);
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
__device__ __forceinline__ double pw(int I, int k) {
  switch (I & 3) {
    case 0: return k == 0 ? 1.0 : 0.0;                                    // r = -1
    case 1: return k == 0 ? 0.375 : (k == 1 ? 0.75 : -0.125);            // r = -1/2
    case 2: return k == 1 ? 1.0 : 0.0;                                    // r =  0
    default: return k == 0 ? -0.125 : (k == 1 ? 0.75 : 0.375);           // r = +1/2
  }
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

// bc = P^T rf (fixed fine residuals excluded, fixed coarse dofs zero).  Only
// the non-zero taps of the 1-D weights are visited (5 for an even coarse line,
// 3 for an odd one), in increasing offset order, and every load is issued
// unconditionally (a fixed dof contributes w * 0), so the loads do not wait on
// the mask; the sum is the same as over all 7x7 offsets.
__device__ __forceinline__ int rtaps(int i, int* off, double* w) {
  if (i & 1) {
    off[0] = -1; off[1] = 0; off[2] = 1;
    w[0] = 0.75; w[1] = 1.0; w[2] = 0.75;
    return 3;
  }
  off[0] = -3; off[1] = -1; off[2] = 0; off[3] = 1; off[4] = 3;
  w[0] = -0.125; w[1] = 0.375; w[2] = 1.0; w[3] = 0.375; w[4] = -0.125;
  return 5;
}

template <class T>
__global__ void k_restrict(int nnx_c, int nnz_c, int nnx_f, int nnz_f,
                           const uint8_t* __restrict__ fixf, const uint8_t* __restrict__ fixc,
                           const T* __restrict__ rf, T* __restrict__ bc) {
  long n = xcd_index();
  if (n >= (long)nnx_c * nnz_c) return;
  int i = (int)(n % nnx_c), j = (int)(n / nnx_c);
  int oa[5], ob[5];
  double wa[5], wb[5];
  const int na = rtaps(i, oa, wa), nb = rtaps(j, ob, wb);
  double s0 = 0, s1 = 0;
  for (int b = 0; b < nb; b++) {
    const int fj = 2 * j + ob[b];
    if (fj < 0 || fj >= nnz_f) continue;
    for (int a = 0; a < na; a++) {
      const int fi = 2 * i + oa[a];
      if (fi < 0 || fi >= nnx_f) continue;
      const long m = (long)fj * nnx_f + fi;
      const double w = wa[a] * wb[b];
      const double v0 = (double)rf[2 * m], v1 = (double)rf[2 * m + 1];
      const uint8_t f0 = fixf[2 * m], f1 = fixf[2 * m + 1];
      s0 += w * (f0 ? 0.0 : v0);
      s1 += w * (f1 ? 0.0 : v1);
    }
  }
  bc[2 * n] = fixc[2 * n] ? (T)0 : (T)s0;
  bc[2 * n + 1] = fixc[2 * n + 1] ? (T)0 : (T)s1;
}

// dense coarsest operator (FP64), row i = dof
template <class T>
__global__ void k_coarse_dense(LV<T> L, double* __restrict__ A) {
  long ndof = 2L * L.nnx * L.nnz;
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= ndof) return;
  double* row = A + i * ndof;
  for (long j = 0; j < ndof; j++) row[j] = 0.0;
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

// in-place Gauss-Jordan inverse of an SPD matrix (one workgroup); W: n x 2n work
__global__ void k_gauss_jordan(int n, const double* __restrict__ A, double* __restrict__ W) {
  const int n2 = 2 * n;
  for (long k = threadIdx.x; k < (long)n * n2; k += blockDim.x) {
    int i = (int)(k / n2), j = (int)(k % n2);
    W[k] = j < n ? A[(long)i * n + j] : (j - n == i ? 1.0 : 0.0);
  }
  __syncthreads();
  extern __shared__ double fac[];
  for (int p = 0; p < n; p++) {
    double piv = W[(long)p * n2 + p];
    __syncthreads();
    for (int j = threadIdx.x; j < n2; j += blockDim.x) W[(long)p * n2 + j] /= piv;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) fac[i] = W[(long)i * n2 + p];
    __syncthreads();
    for (long k = threadIdx.x; k < (long)n * n2; k += blockDim.x) {
      int i = (int)(k / n2), j = (int)(k % n2);
      if (i != p) W[k] -= fac[i] * W[(long)p * n2 + j];
    }
    __syncthreads();
  }
}

template <class T>
__global__ void k_extract_inverse(int n, const double* __restrict__ W, T* __restrict__ Ainv) {
  long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (k >= (long)n * n) return;
  int i = (int)(k / n), j = (int)(k % n);
  Ainv[k] = (T)W[(long)i * 2 * n + n + j];
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

// Chebyshev coefficients (c1, c2) of each smoothing step from ||D^-1 A v||^2
// (power iteration, ||v|| = 1), on the interval [0.1, 1.1] * lambda_max
template <class T>
__global__ void k_cheb_coef(const double* __restrict__ nrm2, int steps, T* __restrict__ coef,
                            double* __restrict__ lmax_out) {
  if (threadIdx.x) return;
  double lm = sqrt(nrm2[0]);
  if (lmax_out) lmax_out[0] = lm;
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
    coef[2 * k] = (T)c1;
    coef[2 * k + 1] = (T)c2;
  }
}

template <class T>
__global__ void k_copy(long n, const T* __restrict__ a, T* __restrict__ b) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// power-iteration helpers
template <class T>
__global__ void k_hash_fill(long n, T* __restrict__ v) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned h = (unsigned)(i * 2654435761u) ^ 0x9e3779b9u;
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  v[i] = (T)(0.5 + (h & 0xffff) / 65536.0);
}

template <class T>
__global__ void k_scale_invdiag_norm(long n, const T* __restrict__ invdiag, const T* __restrict__ y,
                                     const double* __restrict__ nrm2, T* __restrict__ v) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 1.0 / sqrt(nrm2[0] > 0 ? nrm2[0] : 1.0);
  v[i] = (T)((double)invdiag[i] * (double)y[i] * s);
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
          s0 += r0[2 * bb] * xu + r0[2 * bb + 1] * xw;
          s1 += r1[2 * bb] * xu + r1[2 * bb + 1] * xw;
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

template <class T>
void level_apply(const Level<T>& L, const T* x, T* y, hipStream_t st) {
  k_level_op<T, OP_APPLY><<<tiles(L), 256, 0, st>>>(view(L), x, nullptr, y, nullptr, nullptr,
                                                    nullptr, 0, 0);
}

template <class T>
Gmg<T>::~Gmg() {
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (auto* l : lev) delete l;
}

template <class T>
int Gmg<T>::setup(int nelx, int nelz, double Lx, double Lz, const uint8_t* fine_fix,
                  hipStream_t st) {
  for (auto* l : lev) delete l;
  lev.clear();
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
        (rc = L->coef.alloc(2 * (size_t)(nsmooth > 0 ? nsmooth : 1)))) {
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
    if (!can || L->ndof <= 60) break;
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
  if (ncoarse > 2048) return fail(MW_EUNSUPPORTED, "gmg: coarsest level too large for a dense solve");
  int rc;
  if ((rc = coarse_inv.alloc((size_t)ncoarse * ncoarse)) ||
      (rc = coarse_a.alloc((size_t)ncoarse * ncoarse * 3)))
    return rc;
  if (nsmooth < 1) return fail(MW_EINVAL, "gmg: Chebyshev degree must be >= 1");
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
  if (sizeof(T) == sizeof(double))
    MW_HIP(hipMemcpyAsync(F->eta.p, eta_fine, nel0 * sizeof(double), hipMemcpyDeviceToDevice, st));
  else
    k_eta_from_double<<<nblocks(nel0, 256), 256, 0, st>>>(nel0, eta_fine, (float*)F->eta.p);
  for (size_t l = 1; l < lev.size(); l++) {
    Level<T>* L = lev[l];
    k_coarsen_eta<T><<<nblocks((long)L->nelx * L->nelz, 256), 256, 0, st>>>(
        L->nelx, L->nelz, lev[l - 1]->eta.p, L->eta.p);
  }
  for (auto* L : lev)
    k_level_diag<T><<<nblocks(L->nn, 256), 256, 0, st>>>(view(*L), L->diag.p, L->invdiag.p);
  MW_CHECK_LAUNCH();
  // largest eigenvalue of D^-1 A per smoothed level (10 power iterations)
  for (size_t l = 0; l + 1 < lev.size(); l++) {
    Level<T>* L = lev[l];
    k_hash_fill<T><<<nblocks(L->ndof, 256), 256, 0, st>>>(L->ndof, L->x.p);
    for (int it = 0; it < 12; it++) {
      level_apply(*L, L->x.p, L->r.p, st);
      // v = D^-1 A v / ||v|| ; ||v|| measured before scaling
      dot<T>(L->x.p, L->x.p, L->ndof, R.result.p + 0, R, st);
      k_scale_invdiag_norm<T><<<nblocks(L->ndof, 256), 256, 0, st>>>(L->ndof, L->invdiag.p,
                                                                     L->r.p, R.result.p, L->x.p);
    }
    dot<T>(L->x.p, L->x.p, L->ndof, R.result.p + 1, R, st);
    k_cheb_coef<T><<<1, 64, 0, st>>>(R.result.p + 1, nsmooth, L->coef.p, nullptr);
  }
  // coarsest: dense inverse
  Level<T>* C = lev.back();
  int n = ncoarse;
  double* A = coarse_a.p;
  double* W = coarse_a.p + (size_t)n * n;
  k_coarse_dense<T><<<nblocks(n, 128), 128, 0, st>>>(view(*C), A);
  k_gauss_jordan<<<1, 1024, n * sizeof(double), st>>>(n, A, W);
  k_extract_inverse<T><<<nblocks((long)n * n, 256), 256, 0, st>>>(n, W, coarse_inv.p);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// Chebyshev smoothing of L.b; from zero (pre) or from L.x (post).  With an
// even degree the result always ends in L.x (ping-pong through L.x2).
template <class T>
static void chebyshev(Level<T>& L, bool zero_x, int steps, hipStream_t st) {
  T* a = L.x.p;
  T* b = L.x2.p;
  if (zero_x && (steps % 2 == 0)) {  // first output goes to x2 so the last lands in x
    a = L.x2.p;
    b = L.x.p;
  }
  for (int k = 0; k < steps; k++) {
    T* out = (k == 0 && zero_x) ? (steps % 2 == 0 ? L.x2.p : L.x.p) : b;
    const T* in = (k == 0 && zero_x) ? nullptr : a;
    if (L.fine) prof().begin(PROF_CHEB_FINE, st);
    if (L.fine)
      k_level_op<T, OP_CHEB, true><<<tiles(L), 256, 0, st>>>(
          view(L), in, L.b.p, nullptr, L.d.p, out, L.coef.p, k, (k == 0 && zero_x) ? 1 : 0);
    else
      k_level_op<T, OP_CHEB><<<tiles(L), 256, 0, st>>>(view(L), in, L.b.p, nullptr, L.d.p, out,
                                                       L.coef.p, k, (k == 0 && zero_x) ? 1 : 0);
    if (L.fine) prof().end(PROF_CHEB_FINE, st);
    a = out;
    b = (out == L.x.p) ? L.x2.p : L.x.p;
  }
  if (a != L.x.p) k_copy<T><<<nblocks(L.ndof, 256), 256, 0, st>>>(L.ndof, a, L.x.p);
}

template <class T>
int Gmg<T>::vcycle_level(int l, hipStream_t st) {
  Level<T>& L = *lev[l];
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
  chebyshev(L, true, nsmooth, st);
  k_level_op<T, OP_RESID><<<tiles(L), 256, 0, st>>>(view(L), L.x.p, L.b.p, L.r.p, nullptr,
                                                    nullptr, nullptr, 0, 0);
  Level<T>& C = *lev[l + 1];
  k_restrict<T><<<nblocks(C.nn, 256), 256, 0, st>>>(C.nnx, C.nnz, L.nnx, L.nnz, L.fix.p, C.fix.p,
                                                   L.r.p, C.b.p);
  if (l == 0 && graph_exec) {
    MW_HIP(hipGraphLaunch(graph_exec, st));
  } else {
    int rc = vcycle_level(l + 1, st);
    if (rc) return rc;
  }
  k_prolong_add<T><<<nblocks(L.nn, 256), 256, 0, st>>>(L.nnx, L.nnz, C.nnx, C.nnz, L.fix.p, C.x.p,
                                                      L.x.p);
  chebyshev(L, false, nsmooth, st);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

template <class T>
void Gmg<T>::restrict_to(int l, const T* rf, T* bc, hipStream_t st) {
  Level<T>& L = *lev[l];
  Level<T>& C = *lev[l + 1];
  k_restrict<T><<<nblocks(C.nn, 256), 256, 0, st>>>(C.nnx, C.nnz, L.nnx, L.nnz, L.fix.p, C.fix.p,
                                                   rf, bc);
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

}  // namespace mw
