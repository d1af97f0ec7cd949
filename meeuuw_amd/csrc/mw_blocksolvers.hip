// Block Stokes solvers of the reference on general sparse matrices:
//   uzawa3_solver_L2   uzawa3_solver_L2.py:21-99
//   cd_u_eta           cd_u_eta.py:13-99
// The matrices arrive as CSR arrays in device memory.  Where the reference
// factorises K, M_p and M_eta once with SuperLU and back-substitutes, the GPU
// solves each system with Jacobi-preconditioned conjugate gradients to a tight
// relative residual (inner_rtol), so the outer iteration follows the
// reference's sequence of iterates to that accuracy.  The outer loops, their
// formulas, the stopping test (increments below 1e-6) and the returned
// per-iteration arrays are the reference's.
#include <cmath>
#include <vector>

#include "mw_common.h"
#include "mw_solver.h"

namespace mw {

struct Csr {
  int nrows;
  const int* ptr;
  const int* col;
  const double* val;
};

// y = A x, eight lanes per row
__global__ void k_spmv(Csr A, const double* __restrict__ x, double* __restrict__ y) {
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long row = t >> 3;
  int lane = (int)(t & 7);
  double s = 0.0;
  if (row < A.nrows)
    for (int k = A.ptr[row] + lane; k < A.ptr[row + 1]; k += 8) s += A.val[k] * x[A.col[k]];
  s += __shfl_down(s, 4, 8);
  s += __shfl_down(s, 2, 8);
  s += __shfl_down(s, 1, 8);
  if (row < A.nrows && lane == 0) y[row] = s;
}

__global__ void k_csr_invdiag(Csr A, double* __restrict__ d) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= A.nrows) return;
  double v = 0.0;
  for (int k = A.ptr[row]; k < A.ptr[row + 1]; k++)
    if (A.col[k] == row) v += A.val[k];
  d[row] = v != 0.0 ? 1.0 / v : 1.0;
}

__global__ void k_mul(long n, const double* __restrict__ a, const double* __restrict__ b,
                      double* __restrict__ c) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) c[i] = a[i] * b[i];
}

// c = a / b, 0 when b is 0 (a converged CG keeps its iterate: r = 0 gives
// p = 0, p.q = 0 and no update)
__global__ void k_div1(const double* a, const double* b, double* c) {
  c[0] = b[0] != 0.0 ? a[0] / b[0] : 0.0;
}

// the inner CG reads its residual norm on the host every CG_CHECK iterations
// (one synchronisation per check instead of per iteration); the iterations
// between two checks only reduce the residual further
constexpr int CG_CHECK = 8;

static void spmv(const Csr& A, const double* x, double* y, hipStream_t st) {
  k_spmv<<<nblocks((long)A.nrows * 8, 256), 256, 0, st>>>(A, x, y);
}

struct BlockSolverWork {
  Reducer R;
  std::vector<DBuf<double>*> bufs;
  DBuf<double> sc;
  ~BlockSolverWork() {
    for (auto* b : bufs) delete b;
  }
  double* vec(long n) {
    auto* b = new DBuf<double>();
    if (b->alloc(n)) {
      delete b;
      return nullptr;
    }
    bufs.push_back(b);
    return b->p;
  }
  double hdot(const double* a, const double* b, long n, hipStream_t st) {
    dot<double>(a, b, n, R.result.p + 60, R, st);
    (void)hipMemcpyAsync(R.hbuf + 60, R.result.p + 60, 8, hipMemcpyDeviceToHost, st);
    (void)hipStreamSynchronize(st);
    return R.hbuf[60];
  }
};

// Jacobi-PCG for A x = b (x: initial guess, overwritten); relative residual rtol
static int csr_pcg(const Csr& A, const double* invd, const double* b, double* x, double rtol,
                   int maxit, BlockSolverWork& W, double* r, double* z, double* p, double* q,
                   hipStream_t st) {
  long n = A.nrows;
  double* S = W.sc.p;  // 0 rz, 1 pq, 2 alpha, 3 rz_new, 4 beta
  spmv(A, x, q, st);
  (void)hipMemcpyAsync(r, b, n * 8, hipMemcpyDeviceToDevice, st);
  axpby<double>(n, -1.0, nullptr, q, 1.0, nullptr, r, st);
  double bn = sqrt(W.hdot(b, b, n, st));
  if (bn == 0.0) {
    fill<double>(n, 0.0, x, st);
    return MW_OK;
  }
  k_mul<<<nblocks(n, 256), 256, 0, st>>>(n, invd, r, z);
  (void)hipMemcpyAsync(p, z, n * 8, hipMemcpyDeviceToDevice, st);
  dot<double>(r, z, n, S + 0, W.R, st);
  for (int k = 0; k < maxit; k++) {
    spmv(A, p, q, st);
    dot<double>(p, q, n, S + 1, W.R, st);
    k_div1<<<1, 1, 0, st>>>(S + 0, S + 1, S + 2);
    axpby<double>(n, 1.0, S + 2, p, 1.0, nullptr, x, st);
    axpby<double>(n, -1.0, S + 2, q, 1.0, nullptr, r, st);
    if ((k + 1) % CG_CHECK == 0 || k + 1 == maxit) {
      const double rn = sqrt(W.hdot(r, r, n, st));
      if (!(rn == rn)) break;  // NaN
      if (rn <= rtol * bn) return MW_OK;
    }
    k_mul<<<nblocks(n, 256), 256, 0, st>>>(n, invd, r, z);
    dot<double>(r, z, n, S + 3, W.R, st);
    k_div1<<<1, 1, 0, st>>>(S + 3, S + 0, S + 4);
    axpby<double>(n, 1.0, nullptr, z, 1.0, S + 4, p, st);
    (void)hipMemcpyAsync(S + 0, S + 3, 8, hipMemcpyDeviceToDevice, st);
  }
  return fail(MW_ENOCONV, "block solver: inner CG did not converge");
}

struct Inner {  // a CSR system with its Jacobi preconditioner and CG work vectors
  Csr A;
  double* invd;
  double *r, *z, *p, *q;
  int init(const Csr& a, BlockSolverWork& W, hipStream_t st) {
    A = a;
    long n = a.nrows;
    invd = W.vec(n); r = W.vec(n); z = W.vec(n); p = W.vec(n); q = W.vec(n);
    if (!invd || !r || !z || !p || !q) return fail(MW_EHIP, "block solver: allocation failed");
    k_csr_invdiag<<<nblocks(n, 256), 256, 0, st>>>(A, invd);
    return MW_OK;
  }
  // x = A^-1 b from a zero initial guess
  int solve(const double* b, double* x, double rtol, BlockSolverWork& W, hipStream_t st) {
    fill<double>(A.nrows, 0.0, x, st);
    return csr_pcg(A, invd, b, x, rtol, 20 * A.nrows + 1000, W, r, z, p, q, st);
  }
};

int uzawa3(int nV, int nP, Csr K, Csr G, Csr GT, Csr MP, Csr H, const double* f, const double* h,
           double* solV, double* solP, double rtol, int* niter, double* xiV, double* xiP,
           double* alph, hipStream_t st) {
  const double tol = 1e-6;
  const int maxit = 100;
  BlockSolverWork W;
  int rc;
  if ((rc = W.R.init()) || (rc = W.sc.alloc(16))) return rc;
  Inner iK, iM;
  if ((rc = iK.init(K, W, st)) || (rc = iM.init(MP, W, st))) return rc;
  double *r = W.vec(nP), *r1 = W.vec(nP), *pv = W.vec(nP), *d = W.vec(nV), *tV = W.vec(nV),
         *tP = W.vec(nP), *tP2 = W.vec(nP);
  if (!r || !r1 || !pv || !d || !tV || !tP || !tP2) return fail(MW_EHIP, "uzawa3: allocation failed");
  fill<double>(nP, 0.0, solP, st);
  if ((rc = iK.solve(f, solV, rtol, W, st))) return rc;  // V_0 = K^-1 f
  spmv(GT, solV, tP, st);                               // r_0 = M_p^-1 (G^T V - h)
  axpby<double>(nP, -1.0, nullptr, h, 1.0, nullptr, tP, st);
  if ((rc = iM.solve(tP, r, rtol, W, st))) return rc;
  (void)hipMemcpyAsync(pv, r, nP * 8, hipMemcpyDeviceToDevice, st);
  int k = 0;
  for (; k < maxit; k++) {
    spmv(G, pv, tV, st);  // d = K^-1 G p
    if ((rc = iK.solve(tV, d, rtol, W, st))) return rc;
    spmv(MP, r, tP, st);
    double num = W.hdot(r, tP, nP, st);
    spmv(H, d, tP, st);
    double den = W.hdot(pv, tP, nP, st);
    double alpha = num / den;
    alph[k] = alpha;
    axpby<double>(nP, alpha, nullptr, pv, 1.0, nullptr, solP, st);
    axpby<double>(nV, -alpha, nullptr, d, 1.0, nullptr, solV, st);
    spmv(GT, d, tP, st);  // dr = M_p^-1 (-alpha G^T d)
    axpby<double>(nP, -alpha, nullptr, tP, 0.0, nullptr, tP, st);
    if ((rc = iM.solve(tP, tP2, rtol, W, st))) return rc;
    (void)hipMemcpyAsync(r1, r, nP * 8, hipMemcpyDeviceToDevice, st);
    axpby<double>(nP, 1.0, nullptr, tP2, 1.0, nullptr, r1, st);
    spmv(MP, r1, tP, st);
    double n1 = W.hdot(r1, tP, nP, st);
    double beta = n1 / num;  // (r1, M r1) / (r, M r)
    double xp = fabs(alpha) * sqrt(W.hdot(pv, pv, nP, st));
    double xv = fabs(alpha) * sqrt(W.hdot(d, d, nV, st));
    xiP[k] = xp;
    xiV[k] = xv;
    // p_{k+1} = r_{k+1} + beta p_k (computed before the test, as in the reference)
    axpby<double>(nP, 1.0, nullptr, r1, beta, nullptr, pv, st);
    if (xp < tol && xv < tol) break;
    (void)hipMemcpyAsync(r, r1, nP * 8, hipMemcpyDeviceToDevice, st);
  }
  *niter = (k < maxit ? k : maxit - 1) + 1;
  MW_HIP(hipStreamSynchronize(st));
  return MW_OK;
}

int cd_u_eta(int nV, int nP, Csr K, Csr G, Csr GT, Csr M, Csr Meta, Csr H, const double* f,
             const double* h, double* solV, double* solP, double rtol, int* niter, double* xiV,
             double* xiP, double* alph, double* bet, hipStream_t st) {
  const double tol = 1e-6;
  const int maxit = 100;
  BlockSolverWork W;
  int rc;
  if ((rc = W.R.init()) || (rc = W.sc.alloc(16))) return rc;
  Inner iK, iM, iMe;
  if ((rc = iK.init(K, W, st)) || (rc = iM.init(M, W, st)) || (rc = iMe.init(Meta, W, st)))
    return rc;
  double *q = W.vec(nP), *q1 = W.vec(nP), *w = W.vec(nP), *w1 = W.vec(nP), *dd = W.vec(nP),
         *z = W.vec(nV), *tV = W.vec(nV), *tP = W.vec(nP), *tP2 = W.vec(nP);
  if (!q || !q1 || !w || !w1 || !dd || !z || !tV || !tP || !tP2)
    return fail(MW_EHIP, "cd_u_eta: allocation failed");
  auto rhs_q = [&](double* out) {  // out = -G^T V + M h
    spmv(GT, solV, out, st);
    spmv(M, h, tP2, st);
    axpby<double>(nP, 1.0, nullptr, tP2, -1.0, nullptr, out, st);
  };
  fill<double>(nP, 0.0, solP, st);
  if ((rc = iK.solve(f, solV, rtol, W, st))) return rc;  // (B)
  rhs_q(tP);                                             // (C) q = M^-1 (-G^T V + M h)
  if ((rc = iM.solve(tP, q, rtol, W, st))) return rc;
  if ((rc = iMe.solve(tP, w, rtol, W, st))) return rc;   // (D) w = M_eta^-1 (same rhs)
  (void)hipMemcpyAsync(dd, w, nP * 8, hipMemcpyDeviceToDevice, st);
  axpby<double>(nP, 0.0, nullptr, w, -1.0, nullptr, dd, st);  // d = -w
  int k = 0;
  for (; k < maxit; k++) {
    spmv(G, dd, tV, st);  // (E) z = K^-1 G d
    if ((rc = iK.solve(tV, z, rtol, W, st))) return rc;
    spmv(M, w, tP, st);  // (F)
    double num = W.hdot(q, tP, nP, st);
    spmv(H, z, tP, st);
    double den = W.hdot(dd, tP, nP, st);
    double alpha = num / den;
    alph[k] = alpha;
    axpby<double>(nP, alpha, nullptr, dd, 1.0, nullptr, solP, st);   // (G)
    axpby<double>(nV, -alpha, nullptr, z, 1.0, nullptr, solV, st);   // (H)
    rhs_q(tP);                                                       // (I)
    if ((rc = iM.solve(tP, q1, rtol, W, st))) return rc;
    spmv(M, q1, tP, st);                                             // (J) w1 = M_eta^-1 M q1
    if ((rc = iMe.solve(tP, w1, rtol, W, st))) return rc;
    spmv(M, w1, tP, st);                                             // (K)
    double n1 = W.hdot(q1, tP, nP, st);
    double beta = n1 / num;
    bet[k] = beta;
    // increments |alpha d_k|, |alpha z_k| with the direction before its update
    double xp = fabs(alpha) * sqrt(W.hdot(dd, dd, nP, st));
    double xv = fabs(alpha) * sqrt(W.hdot(z, z, nV, st));
    xiP[k] = xp;
    xiV[k] = xv;
    axpby<double>(nP, -1.0, nullptr, w1, beta, nullptr, dd, st);     // d = -w1 + beta d
    if (xp < tol && xv < tol) break;
    (void)hipMemcpyAsync(w, w1, nP * 8, hipMemcpyDeviceToDevice, st);
    (void)hipMemcpyAsync(q, q1, nP * 8, hipMemcpyDeviceToDevice, st);
  }
  *niter = (k < maxit ? k : maxit - 1) + 1;
  MW_HIP(hipStreamSynchronize(st));
  return MW_OK;
}

}  // namespace mw

using namespace mw;

#define CSR(name, n) Csr{(n), name##_ptr, name##_col, name##_val}

extern "C" {

int mw_uzawa3_solver_L2(int nV, int nP, const int32_t* K_ptr, const int32_t* K_col,
                        const double* K_val, const int32_t* G_ptr, const int32_t* G_col,
                        const double* G_val, const int32_t* GT_ptr, const int32_t* GT_col,
                        const double* GT_val, const int32_t* MP_ptr, const int32_t* MP_col,
                        const double* MP_val, const int32_t* H_ptr, const int32_t* H_col,
                        const double* H_val, const double* f, const double* h, double* solV,
                        double* solP, double inner_rtol, int32_t* niter_host, double* xiV_host,
                        double* xiP_host, double* alpha_host, void* stream) {
  if (nV <= 0 || nP <= 0 || !K_ptr || !G_ptr || !GT_ptr || !MP_ptr || !H_ptr || !f || !h ||
      !solV || !solP || !niter_host || !xiV_host || !xiP_host || !alpha_host)
    return fail(MW_EINVAL, "uzawa3_solver_L2: bad arguments");
  for (int k = 0; k < 100; k++) xiV_host[k] = xiP_host[k] = alpha_host[k] = 0.0;
  return uzawa3(nV, nP, CSR(K, nV), CSR(G, nV), CSR(GT, nP), CSR(MP, nP), CSR(H, nP), f, h, solV,
                solP, inner_rtol, niter_host, xiV_host, xiP_host, alpha_host, (hipStream_t)stream);
}

int mw_cd_u_eta(int nV, int nP, const int32_t* K_ptr, const int32_t* K_col, const double* K_val,
                const int32_t* G_ptr, const int32_t* G_col, const double* G_val,
                const int32_t* GT_ptr, const int32_t* GT_col, const double* GT_val,
                const int32_t* M_ptr, const int32_t* M_col, const double* M_val,
                const int32_t* Me_ptr, const int32_t* Me_col, const double* Me_val,
                const int32_t* H_ptr, const int32_t* H_col, const double* H_val, const double* f,
                const double* h, double* solV, double* solP, double inner_rtol,
                int32_t* niter_host, double* xiV_host, double* xiP_host, double* alpha_host,
                double* beta_host, void* stream) {
  if (nV <= 0 || nP <= 0 || !K_ptr || !G_ptr || !GT_ptr || !M_ptr || !Me_ptr || !H_ptr || !f ||
      !h || !solV || !solP || !niter_host || !xiV_host || !xiP_host || !alpha_host || !beta_host)
    return fail(MW_EINVAL, "cd_u_eta: bad arguments");
  for (int k = 0; k < 100; k++) xiV_host[k] = xiP_host[k] = alpha_host[k] = beta_host[k] = 0.0;
  return cd_u_eta(nV, nP, CSR(K, nV), CSR(G, nV), CSR(GT, nP), CSR(M, nP), CSR(Me, nP),
                  CSR(H, nP), f, h, solV, solP, inner_rtol, niter_host, xiV_host, xiP_host,
                  alpha_host, beta_host, (hipStream_t)stream);
}

}  // extern "C"
