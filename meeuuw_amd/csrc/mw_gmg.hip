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

// (A x) at node n = J*nnx + I, fixed dofs of x read as zero; returns the
// unmasked row sums, the caller handles fixed rows
template <class T>
__device__ __forceinline__ void apply_node(const LV<T>& L, const T* __restrict__ sK,
                                           const T* __restrict__ x, int I, int J, T& y0, T& y1) {
  int ex[2], lx[2], ez[2], lz[2];
  int cx = q2_adjacent(I, L.nelx, ex, lx);
  int cz = q2_adjacent(J, L.nelz, ez, lz);
  T a0 = 0, a1 = 0;
  for (int iz = 0; iz < cz; iz++)
    for (int ix = 0; ix < cx; ix++) {
      int e = ez[iz] * L.nelx + ex[ix];
      int a = q2_local(lx[ix], lz[iz]);
      const T* r0 = sK + (2 * a) * 18;
      const T* r1 = r0 + 18;
      T s0 = 0, s1 = 0;
      int base = (2 * ez[iz]) * L.nnx + 2 * ex[ix];
#pragma unroll
      for (int mz = 0; mz < 3; mz++)
#pragma unroll
        for (int mx = 0; mx < 3; mx++) {
          int b = q2_local(mx, mz);
          long m = base + (long)mz * L.nnx + mx;
          T xu = L.fix[2 * m] ? (T)0 : x[2 * m];
          T xw = L.fix[2 * m + 1] ? (T)0 : x[2 * m + 1];
          s0 += r0[2 * b] * xu + r0[2 * b + 1] * xw;
          s1 += r1[2 * b] * xu + r1[2 * b + 1] * xw;
        }
      T et = L.eta[e];
      a0 += et * s0;
      a1 += et * s1;
    }
  y0 = a0;
  y1 = a1;
}

template <class T>
__device__ __forceinline__ void load_Khat(const T* __restrict__ Kh, T* sK) {
  for (int i = threadIdx.x; i < 324; i += blockDim.x) sK[i] = Kh[i];
  __syncthreads();
}

// y = A x (masked: fixed rows -> diag * x)
template <class T>
__global__ void k_level_apply(LV<T> L, const T* __restrict__ x, T* __restrict__ y) {
  __shared__ T sK[324];
  load_Khat(L.Kh, sK);
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= (long)L.nnx * L.nnz) return;
  int I = (int)(n % L.nnx), J = (int)(n / L.nnx);
  T y0, y1;
  apply_node(L, sK, x, I, J, y0, y1);
  y[2 * n] = L.fix[2 * n] ? L.diag[2 * n] * x[2 * n] : y0;
  y[2 * n + 1] = L.fix[2 * n + 1] ? L.diag[2 * n + 1] * x[2 * n + 1] : y1;
}

// r = b - A x
template <class T>
__global__ void k_level_residual(LV<T> L, const T* __restrict__ x, const T* __restrict__ b,
                                 T* __restrict__ r) {
  __shared__ T sK[324];
  load_Khat(L.Kh, sK);
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= (long)L.nnx * L.nnz) return;
  int I = (int)(n % L.nnx), J = (int)(n / L.nnx);
  T y0, y1;
  apply_node(L, sK, x, I, J, y0, y1);
  if (L.fix[2 * n]) y0 = L.diag[2 * n] * x[2 * n];
  if (L.fix[2 * n + 1]) y1 = L.diag[2 * n + 1] * x[2 * n + 1];
  r[2 * n] = b[2 * n] - y0;
  r[2 * n + 1] = b[2 * n + 1] - y1;
}

// one Chebyshev step: d = c1*d + c2*D^-1 (b - A xin); xout = xin + d.
// zero_x: xin is identically zero (first pre-smoothing step)
template <class T>
__global__ void k_level_cheb(LV<T> L, const T* __restrict__ xin, const T* __restrict__ b,
                             T* __restrict__ d, T* __restrict__ xout, T c1, T c2, int zero_x) {
  __shared__ T sK[324];
  load_Khat(L.Kh, sK);
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= (long)L.nnx * L.nnz) return;
  int I = (int)(n % L.nnx), J = (int)(n / L.nnx);
  T y0 = 0, y1 = 0, x0 = 0, x1 = 0;
  if (!zero_x) {
    apply_node(L, sK, xin, I, J, y0, y1);
    x0 = xin[2 * n];
    x1 = xin[2 * n + 1];
    if (L.fix[2 * n]) y0 = L.diag[2 * n] * x0;
    if (L.fix[2 * n + 1]) y1 = L.diag[2 * n + 1] * x1;
  }
  T d0 = c2 * L.invdiag[2 * n] * (b[2 * n] - y0);
  T d1 = c2 * L.invdiag[2 * n + 1] * (b[2 * n + 1] - y1);
  if (c1 != (T)0) {
    d0 += c1 * d[2 * n];
    d1 += c1 * d[2 * n + 1];
  }
  d[2 * n] = d0;
  d[2 * n + 1] = d1;
  xout[2 * n] = x0 + d0;
  xout[2 * n + 1] = x1 + d1;
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

// 1-D quadratic interpolation from coarse line index to fine index I:
// coarse nodes and weights (up to 3)
__device__ __forceinline__ int prolong_1d(int I, int* c, double* w) {
  if (!(I & 1)) {
    c[0] = I >> 1;
    w[0] = 1.0;
    return 1;
  }
  int E = I >> 2;  // coarse element
  if ((I & 3) == 1) {  // r = -1/2
    c[0] = 2 * E; w[0] = 0.375;
    c[1] = 2 * E + 1; w[1] = 0.75;
    c[2] = 2 * E + 2; w[2] = -0.125;
  } else {  // r = +1/2
    c[0] = 2 * E; w[0] = -0.125;
    c[1] = 2 * E + 1; w[1] = 0.75;
    c[2] = 2 * E + 2; w[2] = 0.375;
  }
  return 3;
}

// transpose: fine line indices receiving from coarse index i, with weights
__device__ __forceinline__ int restrict_1d(int i, int nf_max, int* f, double* w) {
  int c = 0;
  if (i & 1) {  // coarse mid node
    f[c] = 2 * i - 1; w[c++] = 0.75;
    f[c] = 2 * i; w[c++] = 1.0;
    f[c] = 2 * i + 1; w[c++] = 0.75;
  } else {  // coarse vertex
    const int off[5] = {-3, -1, 0, 1, 3};
    const double ww[5] = {-0.125, 0.375, 1.0, 0.375, -0.125};
    for (int k = 0; k < 5; k++) {
      int I = 2 * i + off[k];
      if (I >= 0 && I <= nf_max) { f[c] = I; w[c++] = ww[k]; }
    }
  }
  return c;
}

// xf += P xc on free fine dofs
template <class T>
__global__ void k_prolong_add(int nnx_f, int nnz_f, int nnx_c, const uint8_t* __restrict__ fixf,
                              const T* __restrict__ xc, T* __restrict__ xf) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= (long)nnx_f * nnz_f) return;
  int I = (int)(n % nnx_f), J = (int)(n / nnx_f);
  int ci[3], cj[3];
  double wi[3], wj[3];
  int ni = prolong_1d(I, ci, wi), nj = prolong_1d(J, cj, wj);
  double s0 = 0, s1 = 0;
  for (int b = 0; b < nj; b++)
    for (int a = 0; a < ni; a++) {
      long m = (long)cj[b] * nnx_c + ci[a];
      double w = wi[a] * wj[b];
      s0 += w * (double)xc[2 * m];
      s1 += w * (double)xc[2 * m + 1];
    }
  if (!fixf[2 * n]) xf[2 * n] += (T)s0;
  if (!fixf[2 * n + 1]) xf[2 * n + 1] += (T)s1;
}

// bc = P^T rf (fixed fine residuals excluded, fixed coarse dofs zero)
template <class T>
__global__ void k_restrict(int nnx_c, int nnz_c, int nnx_f, int nnz_f,
                           const uint8_t* __restrict__ fixf, const uint8_t* __restrict__ fixc,
                           const T* __restrict__ rf, T* __restrict__ bc) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= (long)nnx_c * nnz_c) return;
  int i = (int)(n % nnx_c), j = (int)(n / nnx_c);
  int fi[5], fj[5];
  double wi[5], wj[5];
  int ni = restrict_1d(i, nnx_f - 1, fi, wi), nj = restrict_1d(j, nnz_f - 1, fj, wj);
  double s0 = 0, s1 = 0;
  for (int b = 0; b < nj; b++)
    for (int a = 0; a < ni; a++) {
      long m = (long)fj[b] * nnx_f + fi[a];
      double w = wi[a] * wj[b];
      if (!fixf[2 * m]) s0 += w * (double)rf[2 * m];
      if (!fixf[2 * m + 1]) s1 += w * (double)rf[2 * m + 1];
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

template <class T>
__global__ void k_dense_mv(int n, const T* __restrict__ Ainv, const T* __restrict__ b,
                           T* __restrict__ x) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < n; j++) s += (double)Ainv[(long)i * n + j] * (double)b[j];
  x[i] = (T)s;
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

template <class T>
void level_apply(const Level<T>& L, const T* x, T* y, hipStream_t st) {
  k_level_apply<T><<<nblocks((long)L.nnx * L.nnz, 256), 256, 0, st>>>(view(L), x, y);
}

template <class T>
Gmg<T>::~Gmg() {
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
        (rc = L->x2.alloc(L->ndof)) || (rc = L->d.alloc(L->ndof)) || (rc = L->b.alloc(L->ndof))) {
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
    if (!can || L->ndof <= 200) break;
    nx /= 2;
    nz /= 2;
  }
  Level<T>* C = lev.back();
  ncoarse = (int)C->ndof;
  if (ncoarse > 2048) return fail(MW_EUNSUPPORTED, "gmg: coarsest level too large for a dense solve");
  int rc;
  if ((rc = coarse_inv.alloc((size_t)ncoarse * ncoarse)) ||
      (rc = coarse_a.alloc((size_t)ncoarse * ncoarse * 3)))
    return rc;
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
    MW_HIP(hipMemcpyAsync(R.hbuf, R.result.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    MW_HIP(hipStreamSynchronize(st));
    L->lmax = sqrt(R.hbuf[1]);  // ||D^-1 A v|| with ||v|| = 1
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

// Chebyshev smoothing of L.b starting from xin (zero_x: xin == 0); result in
// the returned buffer (one of L.x / L.x2)
template <class T>
static T* chebyshev(Level<T>& L, T* xin, bool zero_x, int steps, hipStream_t st) {
  const double lmax = 1.1 * L.lmax, lmin = 0.1 * L.lmax;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
  double rho = 1.0 / sigma;
  T* a = xin;
  T* b = (xin == L.x.p) ? L.x2.p : L.x.p;
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
    const bool fine = L.fine;
    if (fine) prof().begin(PROF_CHEB_FINE, st);
    k_level_cheb<T><<<nblocks(L.nn, 256), 256, 0, st>>>(view(L), a, L.b.p, L.d.p, b, (T)c1,
                                                        (T)c2, (k == 0 && zero_x) ? 1 : 0);
    if (fine) prof().end(PROF_CHEB_FINE, st);
    T* t = a;
    a = b;
    b = t;
  }
  return a;
}

template <class T>
int Gmg<T>::vcycle_level(int l, hipStream_t st) {
  Level<T>& L = *lev[l];
  if (l == (int)lev.size() - 1) {
    k_dense_mv<T><<<nblocks(ncoarse, 64), 64, 0, st>>>(ncoarse, coarse_inv.p, L.b.p, L.x.p);
    MW_CHECK_LAUNCH();
    return MW_OK;
  }
  T* x = chebyshev(L, L.x.p, true, nsmooth, st);
  k_level_residual<T><<<nblocks(L.nn, 256), 256, 0, st>>>(view(L), x, L.b.p, L.r.p);
  Level<T>& C = *lev[l + 1];
  k_restrict<T><<<nblocks(C.nn, 256), 256, 0, st>>>(C.nnx, C.nnz, L.nnx, L.nnz, L.fix.p, C.fix.p,
                                                   L.r.p, C.b.p);
  int rc = vcycle_level(l + 1, st);
  if (rc) return rc;
  k_prolong_add<T><<<nblocks(L.nn, 256), 256, 0, st>>>(L.nnx, L.nnz, C.nnx, L.fix.p, C.x.p, x);
  x = chebyshev(L, x, false, nsmooth, st);
  if (x != L.x.p)
    MW_HIP(hipMemcpyAsync(L.x.p, x, L.ndof * sizeof(T), hipMemcpyDeviceToDevice, st));
  MW_CHECK_LAUNCH();
  return MW_OK;
}

template <class T>
int Gmg<T>::vcycle(const T* r, T* z, hipStream_t st) {
  Level<T>& L = *lev[0];
  MW_HIP(hipMemcpyAsync(L.b.p, r, L.ndof * sizeof(T), hipMemcpyDeviceToDevice, st));
  int rc = vcycle_level(0, st);
  if (rc) return rc;
  MW_HIP(hipMemcpyAsync(z, L.x.p, L.ndof * sizeof(T), hipMemcpyDeviceToDevice, st));
  return MW_OK;
}

template struct Gmg<double>;
template struct Gmg<float>;
template void level_apply<double>(const Level<double>&, const double*, double*, hipStream_t);
template void level_apply<float>(const Level<float>&, const float*, float*, hipStream_t);

}  // namespace mw
