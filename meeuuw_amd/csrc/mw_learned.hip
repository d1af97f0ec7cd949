// Learned Schur-complement correction of the mixed-precision Stokes
// preconditioner (solver 2).
//
// The block upper-triangular preconditioner approximates the pressure Schur
// complement S = G^T K^-1 G by S~, the 1/eta-weighted pressure mass matrix:
// Chebyshev steps on the consistent one, scaled by schur_scale (or its lumped
// inverse).  The learned stage replaces
//     z_p = -S~^-1 v_p
// by  z_p = -proj(y0 + s0 * CNN(q/s, y0/s0, log10 eta / 5, 1)),
// with q = proj(v_p), y0 = proj(S~^-1 q) (the classical block itself, so the
// network learns a correction of the block the solver uses), s and s0 their
// RMS values and proj
// the removal of the constant pressure (the null space of the free-slip
// Stokes operator).  The network is trained, on the GPU, on the pressure parts
// of the Arnoldi vectors of the method's own FGMRES solves, with the loss
//     L(q) = |q - S y|^2_W / |q|^2_W,   W = S~^-1,
// the Schur-complement residual left by y in the norm of the classical
// preconditioner (its minimiser is y = S^-1 q: sign and scale are pinned,
// which the block-triangular coupling z_u = K^-1 (v_u - G z_p) needs), S
// applied with learn_vcycles V-cycles for K^-1.  The classical preconditioner stays the
// judge: the network is used only if, on two probe systems, FGMRES with it
// needs fewer than 0.9 times the iterations and 0.95 times the device time it
// needs without it, and a
// learned solve that does not converge switches the stage off for good.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <random>
#include <vector>

#include "mw_cnn.h"
#include "mw_common.h"
#include "mw_ops.h"
#include "mw_sim.h"

namespace mw {

constexpr int LB = 1024;  // threads of the per-sample reduction kernels
constexpr int NSC = 8;    // scalars per sample: 0 mean q, 1 s, 2 s0
constexpr int MS_BLOCKS = 256;  // block partials of the inference statistics

__device__ __forceinline__ double wsum64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// sums over the block of LB threads, broadcast to every thread (fixed order)
template <int K>
__device__ __forceinline__ void block_sums(double (&v)[K], double* sh) {
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = wsum64(v[k]);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < K; k++) sh[w * K + k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) {
    double s = 0.0;
    for (int i = 0; i < LB / 64; i++) s += sh[i * K + k];
    v[k] = s;
  }
  __syncthreads();
}

// sample b: Qm = q - mean q, sc[0] = mean q, sc[1] = s = RMS(Qm)
__global__ void __launch_bounds__(LB) k_center(long n, const float* __restrict__ Q,
                                               float* __restrict__ Qm, double* __restrict__ sc) {
  __shared__ double sh[(LB / 64) * 2];
  const long o = (long)blockIdx.x * n;
  double a[1] = {0.0};
  for (long i = threadIdx.x; i < n; i += LB) a[0] += (double)Q[o + i];
  block_sums<1>(a, sh);
  const double mq = a[0] / (double)n;
  double b[1] = {0.0};
  for (long i = threadIdx.x; i < n; i += LB) {
    const double qm = (double)Q[o + i] - mq;
    Qm[o + i] = (float)qm;
    b[0] += qm * qm;
  }
  block_sums<1>(b, sh);
  if (threadIdx.x == 0) {
    sc[blockIdx.x * NSC + 0] = mq;
    sc[blockIdx.x * NSC + 1] = sqrt(b[0] / (double)n);
  }
}

// y = scale * (minv Qm) for the lumped Schur block, or y = scale * x (x: the
// Chebyshev steps' result) for the consistent one
__global__ void k_schur_y(long n, double scale, const double* __restrict__ minv,
                          const float* __restrict__ qm, const float* __restrict__ x,
                          float* __restrict__ y) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  y[i] = x ? (float)(scale * (double)x[i]) : (float)(scale * minv[i] * (double)qm[i]);
}

// network input of sample b from Qm and the classical Schur block's output Yc:
// y0 = proj(Yc), X0 = [qm/s, y0/s0, log10(eta)/5, 1], sc[2] = s0 = RMS(y0)
__global__ void __launch_bounds__(LB) k_features(long n, const float* __restrict__ Qm,
                                                 const float* __restrict__ Yc,
                                                 const float* __restrict__ leta,
                                                 float* __restrict__ Y0, float* __restrict__ X0,
                                                 double* __restrict__ sc) {
  __shared__ double sh[(LB / 64) * 2];
  const long o = (long)blockIdx.x * n;
  double b2[2] = {0.0, 0.0};
  for (long i = threadIdx.x; i < n; i += LB) {
    const double y = (double)Yc[o + i];
    b2[0] += y;
    b2[1] += y * y;
  }
  block_sums<2>(b2, sh);
  const double my = b2[0] / (double)n;
  const double s0 = sqrt(fmax(b2[1] / (double)n - my * my, 0.0));
  const double s = sc[blockIdx.x * NSC + 1];
  const double is = s > 0.0 ? 1.0 / s : 0.0, is0 = s0 > 0.0 ? 1.0 / s0 : 0.0;
  for (long i = threadIdx.x; i < n; i += LB) {
    const double qm = (double)Qm[o + i];
    const double y0 = (double)Yc[o + i] - my;
    Y0[o + i] = (float)y0;
    reinterpret_cast<float4*>(X0)[o + i] =
        make_float4((float)(qm * is), (float)(y0 * is0), 0.2f * leta[i], 1.0f);
  }
  if (threadIdx.x == 0) sc[blockIdx.x * NSC + 2] = s0;
}

// out = sign * proj(y0 + s0 * c)   (c ignored when use_c == 0)
__global__ void __launch_bounds__(LB) k_combine(long n, const float* __restrict__ Y0,
                                                const float* __restrict__ Cn,
                                                const double* __restrict__ sc, int use_c,
                                                float sign, float* __restrict__ out) {
  __shared__ double sh[LB / 64];
  const long o = (long)blockIdx.x * n;
  const double s0 = sc[blockIdx.x * NSC + 2];
  double a[1] = {0.0};
  for (long i = threadIdx.x; i < n; i += LB)
    a[0] += (double)Y0[o + i] + (use_c ? s0 * (double)Cn[o + i] : 0.0);
  block_sums<1>(a, sh);
  const double m = a[0] / (double)n;
  for (long i = threadIdx.x; i < n; i += LB)
    out[o + i] = (float)(sign * ((double)Y0[o + i] + (use_c ? s0 * (double)Cn[o + i] : 0.0) - m));
}

// L = |q - proj(Sy)|^2_W / |q|^2_W with the weight W = S~^-1 (the lumped
// viscosity-scaled mass inverse); Gd = proj(dL/d(Sy))
__global__ void __launch_bounds__(LB) k_loss(long n, const float* __restrict__ Qm,
                                             const float* __restrict__ SY,
                                             const double* __restrict__ w,
                                             float* __restrict__ Gd, double* __restrict__ loss) {
  __shared__ double sh[(LB / 64) * 2];
  const long o = (long)blockIdx.x * n;
  double a[1] = {0.0};
  for (long i = threadIdx.x; i < n; i += LB) a[0] += (double)SY[o + i];
  block_sums<1>(a, sh);
  const double m = a[0] / (double)n;
  double t[2] = {0.0, 0.0};
  for (long i = threadIdx.x; i < n; i += LB) {
    double q = Qm[o + i], e = q - ((double)SY[o + i] - m);
    t[0] += w[i] * e * e;
    t[1] += w[i] * q * q;
  }
  block_sums<2>(t, sh);
  const double L = t[1] > 0.0 ? t[0] / t[1] : 0.0;
  const double f = t[1] > 0.0 ? -2.0 / t[1] : 0.0;
  if (Gd) {
    double b[1] = {0.0};
    for (long i = threadIdx.x; i < n; i += LB)
      b[0] += f * w[i] * ((double)Qm[o + i] - ((double)SY[o + i] - m));
    block_sums<1>(b, sh);
    const double mg = b[0] / (double)n;
    for (long i = threadIdx.x; i < n; i += LB)
      Gd[o + i] = (float)(f * w[i] * ((double)Qm[o + i] - ((double)SY[o + i] - m)) - mg);
  }
  if (threadIdx.x == 0) loss[blockIdx.x] = L;
}

// dC = s0 * proj(dY) * scale
__global__ void __launch_bounds__(LB) k_dc(long n, const float* __restrict__ DY,
                                           const double* __restrict__ sc, double scale,
                                           float* __restrict__ DC) {
  __shared__ double sh[LB / 64];
  const long o = (long)blockIdx.x * n;
  double a[1] = {0.0};
  for (long i = threadIdx.x; i < n; i += LB) a[0] += (double)DY[o + i];
  block_sums<1>(a, sh);
  const double m = a[0] / (double)n;
  const double f = sc[blockIdx.x * NSC + 2] * scale;
  for (long i = threadIdx.x; i < n; i += LB) DC[o + i] = (float)(f * ((double)DY[o + i] - m));
}

// log10 of the element viscosity averaged onto the pressure nodes
__global__ void k_leta(int nelx, int nelz, const double* __restrict__ eta, float* __restrict__ le) {
  long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int nx = nelx + 1;
  if (k >= (long)nx * (nelz + 1)) return;
  const int i = (int)(k % nx), j = (int)(k / nx);
  double s = 0.0;
  int c = 0;
  for (int ej = j - 1; ej <= j; ej++)
    for (int ei = i - 1; ei <= i; ei++)
      if (ei >= 0 && ei < nelx && ej >= 0 && ej < nelz) {
        s += log10(eta[ej * nelx + ei]);
        c++;
      }
  le[k] = (float)(s / c);
}

// counter-based N(0,1) numbers (splitmix64 + Box-Muller)
__device__ __forceinline__ uint64_t splitmix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ void k_randn(long n, uint64_t seed, float* __restrict__ out) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t a = splitmix(seed ^ splitmix(2 * (uint64_t)i)), b = splitmix(seed ^ splitmix(2 * (uint64_t)i + 1));
  double u1 = ((a >> 11) + 1.0) * (1.0 / 9007199254740993.0);  // (0, 1]
  double u2 = (b >> 11) * (1.0 / 9007199254740992.0);
  out[i] = (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

// Gaussian probe with covariance S~ (q = sqrt(mdiag) * N(0,1)): the loss on
// such probes is a Hutchinson estimate of |I - S y(.)|_F^2 in the S~^-1 norm
__global__ void k_probe(long n, uint64_t seed, const double* __restrict__ mdiag,
                        float* __restrict__ q) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t a = splitmix(seed ^ splitmix(2 * (uint64_t)i)), b = splitmix(seed ^ splitmix(2 * (uint64_t)i + 1));
  double u1 = ((a >> 11) + 1.0) * (1.0 / 9007199254740993.0);
  double u2 = (b >> 11) * (1.0 / 9007199254740992.0);
  q[i] = (float)(sqrt(mdiag[i]) * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

__global__ void k_perturb(long n, const float* __restrict__ r, const float* __restrict__ z,
                          float* __restrict__ b) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) b[i] = r[i] * (1.0f + 0.3f * z[i]);
}

static inline unsigned nbl(long n) { return nblocks(n, 256); }

int Sim::learn_setup() {
  const mw_sim_config& c = cfg;
  if (c.learn_pairs < 16 || c.learn_batch < 1 || c.learn_batch > 16 || c.learn_steps < 0 ||
      c.learn_vcycles < 1 || c.learn_vcycles > 8 || !(c.learn_lr > 0))
    return fail(MW_EINVAL, "sim: bad learned-solver options");
  const long n = nnP;
  const int B = c.learn_batch;
  RC(net.init(c.nelz + 1, c.nelx + 1, B));
  RC(Qs.alloc((size_t)c.learn_pairs * n));
  for (DBuf<float>* b : {&lQ, &lQm, &lYc, &lY0, &lY, &lC, &lSY, &lGd, &lDY, &lDC})
    RC(b->alloc((size_t)B * n));
  for (DBuf<float>* b : {&lt, &lu, &ls, &lq}) RC(b->alloc(NV));
  RC(lb.alloc(NV + nnP));
  RC(leta.alloc(n));
  RC(lsc.alloc((size_t)B * NSC));
  RC(lpart.alloc(2 * MS_BLOCKS + 8));
  RC(lloss.alloc(B));
  // He-uniform initial weights (fan-in 9 c_in), zero biases, output layer x0.1
  std::mt19937 gen((unsigned)c.learn_seed);
  std::vector<float> p(CNN_NPARAM, 0.0f);
  auto fill_w = [&](int off, int cout, int cin, double scale) {
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    float lim = (float)(std::sqrt(6.0 / (9.0 * cin)) * scale);
    for (int i = 0; i < 9 * cout * cin; i++) p[off + i] = lim * u(gen);
  };
  fill_w(CNN_OFF_W1, 32, 4, 1.0);
  fill_w(CNN_OFF_W2, 32, 32, 1.0);
  fill_w(CNN_OFF_W3, 32, 32, 1.0);
  fill_w(CNN_OFF_W4, 1, 32, 0.1);
  RC(net.set_params(p.data(), st));
  return MW_OK;
}

// The network's inputs for B pressure vectors Q: the classical Schur block the
// solver uses (Chebyshev steps on the consistent 1/eta-weighted mass matrix
// times schur_scale, or the lumped inverse) applied to each centred vector
// gives y0, which the network corrects.
void Sim::learn_features(int B, const float* Q) {
  const long n = nnP;
  k_center<<<B, LB, 0, st>>>(n, Q, lQm, lsc);
  const double scale = cfg.schur_scale > 0 ? cfg.schur_scale : 1.0;
  for (int b = 0; b < B; b++) {
    const float* x = nullptr;
    if (cfg.schur_cheb >= 2) {
      float* xs = nullptr;
      (void)pmass_cheb_solve<float>(cfg.nelx, cfg.nelz, pm_w, pm_dg, pm_dinv, lQm.p + b * n,
                                    pm_xa, pm_xb, pm_d, cfg.schur_cheb, &xs, st);
      x = xs;
    }
    k_schur_y<<<nbl(n), 256, 0, st>>>(n, scale, minv, lQm.p + b * n, x, lYc.p + b * n);
  }
  k_features<<<B, LB, 0, st>>>(n, lQm, lYc, leta, lY0, net.X0, lsc);
}

// out = G^T Kt^-1 G y, Kt^-1 = learn_vcycles V-cycle iterations (not mean-free)
void Sim::schur_apply32(const float* y, float* out) {
  gtmpl_apply_u<float>(g, Gh, fixV, y, lt, false, st);
  (void)gmg32.vcycle(lt, lu, st);
  for (int k = 1; k < cfg.learn_vcycles; k++) {
    level_apply<float>(*gmg32.lev[0], lu, ls, st);
    axpby<float>(NV, 1.0, nullptr, lt, -1.0, nullptr, ls, st);
    (void)gmg32.vcycle(ls, lq, st);
    axpby<float>(NV, 1.0, nullptr, lq, 1.0, nullptr, lu, st);
  }
  gtmpl_apply_p<float>(g, Gh, fixV, lu, out, 1.0, st);
}

// ---- inference on one vector, spread over the GPU: the statistics the
// network's input and output need (means and RMS values) as block partials and
// a one-block finish, the elementwise parts one thread per pressure node (the
// per-sample single-block kernels above are for the training batches)
// partial sums of v and v^2 with v = a + c b (b may be NULL), c = cs[0] or 0
__global__ void __launch_bounds__(256) k_ms_partial(long n, const float* __restrict__ a,
                                                    const float* __restrict__ b,
                                                    const double* __restrict__ cs,
                                                    double* __restrict__ part) {
  __shared__ double sh[2][4];
  const double c = (b && cs) ? cs[0] : 0.0;
  double s1 = 0.0, s2 = 0.0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += 256L * gridDim.x) {
    const double v = (double)a[i] + (b ? c * (double)b[i] : 0.0);
    s1 += v;
    s2 += v * v;
  }
  s1 = wsum64(s1);
  s2 = wsum64(s2);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sh[0][w] = s1; sh[1][w] = s2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    part[2 * blockIdx.x + 1] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
  }
}

// out[0] = mean, out[1] = RMS about the mean, from nb partials: one wavefront,
// lane l adding partials l, l + 64, ..., then a butterfly (a fixed order)
__global__ void __launch_bounds__(64) k_ms_final(int nb, long n, const double* __restrict__ part,
                                                 double* __restrict__ out) {
  double s1 = 0.0, s2 = 0.0;
  for (int b = threadIdx.x; b < nb; b += 64) {
    s1 += part[2 * b];
    s2 += part[2 * b + 1];
  }
  s1 = wsum64(s1);
  s2 = wsum64(s2);
  if (threadIdx.x != 0) return;
  const double m = s1 / (double)n;
  out[0] = m;
  out[1] = sqrt(fmax(s2 / (double)n - m * m, 0.0));
}

// Qm = q - mean q
__global__ void k_shift(long n, const float* __restrict__ q, const double* __restrict__ mq,
                        float* __restrict__ qm) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) qm[i] = (float)((double)q[i] - mq[0]);
}

// X0 = [qm/s, y0/s0, log10(eta)/5, 1] with y0 = Yc - my; stq = {mean q, s},
// sty = {my, s0}
__global__ void k_feat1(long n, const float* __restrict__ Qm, const float* __restrict__ Yc,
                        const float* __restrict__ leta, const double* __restrict__ stq,
                        const double* __restrict__ sty, float* __restrict__ Y0,
                        float* __restrict__ X0) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double s = stq[1], s0 = sty[1];
  const double is = s > 0.0 ? 1.0 / s : 0.0, is0 = s0 > 0.0 ? 1.0 / s0 : 0.0;
  const double y0 = (double)Yc[i] - sty[0];
  Y0[i] = (float)y0;
  reinterpret_cast<float4*>(X0)[i] =
      make_float4((float)((double)Qm[i] * is), (float)(y0 * is0), 0.2f * leta[i], 1.0f);
}

// zp = -(y0 + s0 C - m), m = mean(y0 + s0 C) (stv[0])
__global__ void k_out1(long n, const float* __restrict__ Y0, const float* __restrict__ C,
                       const double* __restrict__ s0, const double* __restrict__ stv,
                       float* __restrict__ zp) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) zp[i] = (float)(-((double)Y0[i] + s0[0] * (double)C[i] - stv[0]));
}

// z_p = -proj(y0 + s0 CNN(...)) for one pressure vector (the preconditioner)
void Sim::schur_learned(const float* vp, float* zp) {
  const long n = nnP;
  const int nbk = (int)std::min<long>(MS_BLOCKS, (n + 255) / 256);
  double* part = lpart.p;
  double* stq = lpart.p + 2 * MS_BLOCKS;      // mean q, s
  double* sty = stq + 2;                      // my, s0
  double* stv = stq + 4;                      // mean of the output
  k_ms_partial<<<nbk, 256, 0, st>>>(n, vp, nullptr, nullptr, part);
  k_ms_final<<<1, 64, 0, st>>>(nbk, n, part, stq);
  k_shift<<<nbl(n), 256, 0, st>>>(n, vp, stq, lQm);
  const double scale = cfg.schur_scale > 0 ? cfg.schur_scale : 1.0;
  const float* x = nullptr;
  if (cfg.schur_cheb >= 2) {
    float* xs = nullptr;
    (void)pmass_cheb_solve<float>(cfg.nelx, cfg.nelz, pm_w, pm_dg, pm_dinv, lQm.p, pm_xa, pm_xb,
                                  pm_d, cfg.schur_cheb, &xs, st);
    x = xs;
  }
  k_schur_y<<<nbl(n), 256, 0, st>>>(n, scale, minv, lQm, x, lYc);
  k_ms_partial<<<nbk, 256, 0, st>>>(n, lYc, nullptr, nullptr, part);
  k_ms_final<<<1, 64, 0, st>>>(nbk, n, part, sty);
  k_feat1<<<nbl(n), 256, 0, st>>>(n, lQm, lYc, leta, stq, sty, lY0, net.X0);
  (void)net.forward(1, net.X0, lC, st);
  k_ms_partial<<<nbk, 256, 0, st>>>(n, lY0, lC, sty + 1, part);
  k_ms_final<<<1, 64, 0, st>>>(nbk, n, part, stv);
  k_out1<<<nbl(n), 256, 0, st>>>(n, lY0, lC, sty + 1, stv, zp);
}

// FGMRES solves with random consistent right-hand sides (A times N(0,1)) and
// randomly perturbed copies of the current residual, until `count` Arnoldi
// vectors have been collected
int Sim::collect_pairs(long count) {
  const long N = NV + nnP;
  uint64_t seed = 0x5EEDull * (uint64_t)(cfg.learn_seed + 1);
  int k = 0;
  collecting = true;
  while (ncollect < count) {
    float* b = lb.p;
    k_randn<<<nbl(N), 256, 0, st>>>(N, seed + 7919ull * (uint64_t)k, x32.p);
    if (k % 2 == 0) apply32(x32.p, b);
    else k_perturb<<<nbl(N), 256, 0, st>>>(N, r32.p, x32.p, b);
    int its = 0;
    long before = ncollect;
    RC(fgmres32(b, x32.p, cfg.fgmres_restart, 1e-6, 60, &its, -1.0));
    if (ncollect == before) break;  // nothing to collect (zero right-hand side)
    k++;
  }
  collecting = false;
  return MW_OK;
}

int Sim::learn_train() {
  auto t0 = std::chrono::steady_clock::now();
  const long n = nnP;
  const int B = cfg.learn_batch;
  const long npairs = cfg.learn_pairs;
  RC(collect_pairs(npairs));
  const long have = std::min(ncollect, npairs);
  if (have < 2 * B) return fail(MW_EINVAL, "learned solver: too few training vectors");
  const long nval = std::max<long>(B, have / 10);
  const long ntrain = have - nval;
  std::mt19937 gen((unsigned)cfg.learn_seed + 1u);
  std::uniform_int_distribution<long> pick(0, ntrain - 1);
  const int steps = cfg.learn_steps;
  for (int step = 0; step < steps; step++) {
    double lr = cfg.learn_lr * 0.5 * (1.0 + std::cos(3.141592653589793 * step / std::max(steps, 1)));
    // half of each batch: collected Arnoldi vectors; half: random probes
    for (int b = 0; b < B; b++) {
      if (2 * b < B)
        MW_HIP(hipMemcpyAsync(lQ.p + b * n, Qs.p + pick(gen) * n, n * 4, hipMemcpyDeviceToDevice, st));
      else
        k_probe<<<nbl(n), 256, 0, st>>>(n, 0xC0FFEEull + 131ull * (uint64_t)(step * B + b), mdiag,
                                        lQ.p + b * n);
    }
    learn_features(B, lQ);
    RC(net.forward(B, net.X0, lC, st));
    k_combine<<<B, LB, 0, st>>>(n, lY0, lC, lsc, 1, 1.0f, lY);
    for (int b = 0; b < B; b++) schur_apply32(lY.p + b * n, lSY.p + b * n);
    k_loss<<<B, LB, 0, st>>>(n, lQm, lSY, minv, lGd, lloss);
    for (int b = 0; b < B; b++) schur_apply32(lGd.p + b * n, lDY.p + b * n);
    k_dc<<<B, LB, 0, st>>>(n, lDY, lsc, 1.0 / B, lDC);
    RC(net.backward(B, lDC, st));
    net.adam(lr, 0.9, 0.999, 1e-8, st);
    MW_CHECK_LAUNCH();
  }
  // validation: the network against y0 alone on vectors it was not trained on
  double sum_net = 0.0, sum_base = 0.0;
  long cnt = 0;
  std::vector<double> hl(B);
  for (long v0 = ntrain; v0 + B <= have; v0 += B) {
    MW_HIP(hipMemcpyAsync(lQ.p, Qs.p + v0 * n, (size_t)B * n * 4, hipMemcpyDeviceToDevice, st));
    learn_features(B, lQ);
    RC(net.forward(B, net.X0, lC, st));
    for (int use_c = 0; use_c < 2; use_c++) {
      k_combine<<<B, LB, 0, st>>>(n, lY0, lC, lsc, use_c, 1.0f, lY);
      for (int b = 0; b < B; b++) schur_apply32(lY.p + b * n, lSY.p + b * n);
      k_loss<<<B, LB, 0, st>>>(n, lQm, lSY, minv, nullptr, lloss);
      MW_HIP(hipMemcpyAsync(hl.data(), lloss.p, B * 8, hipMemcpyDeviceToHost, st));
      MW_HIP(hipStreamSynchronize(st));
      for (int b = 0; b < B; b++) (use_c ? sum_net : sum_base) += hl[b];
    }
    cnt += B;
  }
  double ratio = sum_base > 0 ? sum_net / sum_base : 1.0;
  // the classical preconditioner judges the network: FGMRES on two probe
  // systems (the current residual, randomly perturbed, and A times N(0,1))
  // with and without the learned stage
  // with the probe solves' device time: a network that saves iterations but
  // costs more per application than it saves is not used
  int its_plain = 0, its_learn = 0;
  double ms_plain = 0.0, ms_learn = 0.0;
  {
    const long N = NV + nnP;
    hipEvent_t e0, e1;
    MW_HIP(hipEventCreate(&e0));
    MW_HIP(hipEventCreate(&e1));
    int rc = MW_OK;
    for (int probe = 0; probe < 2 && rc == MW_OK; probe++) {
      k_randn<<<nbl(N), 256, 0, st>>>(N, 0xBEEFull + (uint64_t)probe, x32.p);
      if (probe == 0) k_perturb<<<nbl(N), 256, 0, st>>>(N, r32.p, x32.p, lb.p);
      else apply32(x32.p, lb.p);
      for (int use = 0; use < 2 && rc == MW_OK; use++) {
        learned = use == 1;
        int its = 0;
        float ms = 0.0f;
        if (hipEventRecord(e0, st) != hipSuccess) rc = fail(MW_EHIP, "learn: event record");
        if (rc == MW_OK) rc = fgmres32(lb.p, x32.p, cfg.fgmres_restart, 1e-4, 200, &its, -1.0);
        if (rc == MW_OK && (hipEventRecord(e1, st) != hipSuccess ||
                            hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
          rc = fail(MW_EHIP, "learn: probe timing");
        (use ? its_learn : its_plain) += its;
        (use ? ms_learn : ms_plain) += ms;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != MW_OK) return rc;
  }
  learned = its_plain > 0 && its_learn < 0.9 * its_plain && ms_learn < 0.95 * ms_plain;
  stats[MW_STAT_LEARN_ITS_PLAIN] = its_plain;
  stats[MW_STAT_LEARN_ITS] = its_learn;
  stats[MW_STAT_LEARN_PROBE_MS_PLAIN] = ms_plain;
  stats[MW_STAT_LEARN_PROBE_MS] = ms_learn;
  learn_done = true;
  auto t1 = std::chrono::steady_clock::now();
  stats[MW_STAT_LEARN_MS] = std::chrono::duration<double, std::milli>(t1 - t0).count();
  stats[MW_STAT_LEARN_RATIO] = ratio;
  return MW_OK;
}

void Sim::update_leta() {
  k_leta<<<nbl(nnP), 256, 0, st>>>(cfg.nelx, cfg.nelz, eta_mg, leta);
}

}  // namespace mw
