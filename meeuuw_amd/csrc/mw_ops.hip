// Element-by-element operators on the structured box:
//   * the assembled Stokes and energy matrices applied straight from their COO
//     value arrays (VV_Stokes: 468 values per element in the reference order,
//     VV_T: 81 per element), gathered per node in increasing element order;
//   * matrix-free templates of the divergence operator G and of the lumped,
//     viscosity-scaled pressure mass matrix used by the preconditioners.
// No atomics: every output entry is owned by one thread.
#include "mw_common.h"
#include "mw_solver.h"
#include "mw_ops.h"

namespace mw {

// reference local node b -> (lx, lz) on the element's 3x3 node grid
__device__ __forceinline__ void q2_pos(int b, int& lx, int& lz) {
  const int X[9] = {0, 2, 2, 0, 1, 2, 1, 0, 1};
  const int Z[9] = {0, 0, 2, 2, 0, 1, 2, 1, 1};
  lx = X[b];
  lz = Z[b];
}

// ---------------------------------------------------------------------------
// assembled Stokes operator from VV
// ---------------------------------------------------------------------------
__global__ void k_stokes_vv_u(BoxGeom g, const double* __restrict__ VV,
                              const double* __restrict__ u, const double* __restrict__ p,
                              double* __restrict__ yu) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= g.nn) return;
  int I = (int)(n % g.nnx), J = (int)(n / g.nnx);
  int ex[2], lx[2], ez[2], lz[2];
  int cx = q2_adjacent(I, g.nelx, ex, lx), cz = q2_adjacent(J, g.nelz, ez, lz);
  double a0 = 0.0, a1 = 0.0;
  for (int iz = 0; iz < cz; iz++)
    for (int ix = 0; ix < cx; ix++) {
      long e = (long)ez[iz] * g.nelx + ex[ix];
      int a = q2_local(lx[ix], lz[iz]);
      const double* r0 = VV + e * 468 + (2 * a) * 26;
      const double* r1 = r0 + 26;
      long base = (long)(2 * ez[iz]) * g.nnx + 2 * ex[ix];
      double s0 = 0.0, s1 = 0.0;
      if (u) {
#pragma unroll
        for (int b = 0; b < 9; b++) {
          int bx, bz;
          q2_pos(b, bx, bz);
          long m = base + (long)bz * g.nnx + bx;
          double xu = u[2 * m], xw = u[2 * m + 1];
          s0 += r0[2 * b] * xu + r0[2 * b + 1] * xw;
          s1 += r1[2 * b] * xu + r1[2 * b + 1] * xw;
        }
      }
      if (p) {
        long pb = (long)ez[iz] * (g.nelx + 1) + ex[ix];
        const long pk[4] = {pb, pb + 1, pb + g.nelx + 2, pb + g.nelx + 1};
#pragma unroll
        for (int k = 0; k < 4; k++) {
          double pv = p[pk[k]];
          s0 += r0[18 + 2 * k] * pv;
          s1 += r1[18 + 2 * k] * pv;
        }
      }
      a0 += s0;
      a1 += s1;
    }
  yu[2 * n] = a0;
  yu[2 * n + 1] = a1;
}

__global__ void k_stokes_vv_p(BoxGeom g, const double* __restrict__ VV,
                              const double* __restrict__ u, double* __restrict__ yp) {
  long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m >= g.nnP) return;
  int i = (int)(m % (g.nelx + 1)), j = (int)(m / (g.nelx + 1));
  double s = 0.0;
  for (int ez = j - 1; ez <= j; ez++) {
    if (ez < 0 || ez >= g.nelz) continue;
    for (int ex = i - 1; ex <= i; ex++) {
      if (ex < 0 || ex >= g.nelx) continue;
      int k = (ex == i) ? (ez == j ? 0 : 3) : (ez == j ? 1 : 2);
      long e = (long)ez * g.nelx + ex;
      long base = (long)(2 * ez) * g.nnx + 2 * ex;
      const double* col = VV + e * 468 + 18 + 2 * k + 1;  // GT slot of row ikk
#pragma unroll
      for (int ikk = 0; ikk < 18; ikk++) {
        int bx, bz;
        q2_pos(ikk >> 1, bx, bz);
        long nd = base + (long)bz * g.nnx + bx;
        s += col[ikk * 26] * u[2 * nd + (ikk & 1)];
      }
    }
  }
  yp[m] = s;
}

void stokes_apply_vv(const BoxGeom& g, const double* VV, const double* u, const double* p,
                     double* yu, double* yp, hipStream_t st) {
  if (yu) {
    prof().begin(PROF_STOKES_VV, st);
    k_stokes_vv_u<<<nblocks(g.nn, 256), 256, 0, st>>>(g, VV, u, p, yu);
    prof().end(PROF_STOKES_VV, st);
  }
  if (yp) k_stokes_vv_p<<<nblocks(g.nnP, 256), 256, 0, st>>>(g, VV, u, yp);
}

__global__ void k_stokes_invdiag(BoxGeom g, const double* __restrict__ VV, double* __restrict__ d) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= g.nn) return;
  int I = (int)(n % g.nnx), J = (int)(n / g.nnx);
  int ex[2], lx[2], ez[2], lz[2];
  int cx = q2_adjacent(I, g.nelx, ex, lx), cz = q2_adjacent(J, g.nelz, ez, lz);
  double a0 = 0.0, a1 = 0.0;
  for (int iz = 0; iz < cz; iz++)
    for (int ix = 0; ix < cx; ix++) {
      long e = (long)ez[iz] * g.nelx + ex[ix];
      int a = q2_local(lx[ix], lz[iz]);
      a0 += VV[e * 468 + (2 * a) * 26 + 2 * a];
      a1 += VV[e * 468 + (2 * a + 1) * 26 + 2 * a + 1];
    }
  d[2 * n] = 1.0 / a0;
  d[2 * n + 1] = 1.0 / a1;
}

void stokes_invdiag_vv(const BoxGeom& g, const double* VV, double* invdiagK, hipStream_t st) {
  k_stokes_invdiag<<<nblocks(g.nn, 256), 256, 0, st>>>(g, VV, invdiagK);
}

// ---------------------------------------------------------------------------
// assembled energy operator from VV_T
// ---------------------------------------------------------------------------
__global__ void k_energy_vv(BoxGeom g, const double* __restrict__ VV, const double* __restrict__ x,
                            double* __restrict__ y) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= g.nn) return;
  int I = (int)(n % g.nnx), J = (int)(n / g.nnx);
  int ex[2], lx[2], ez[2], lz[2];
  int cx = q2_adjacent(I, g.nelx, ex, lx), cz = q2_adjacent(J, g.nelz, ez, lz);
  double acc = 0.0;
  for (int iz = 0; iz < cz; iz++)
    for (int ix = 0; ix < cx; ix++) {
      long e = (long)ez[iz] * g.nelx + ex[ix];
      int a = q2_local(lx[ix], lz[iz]);
      const double* row = VV + e * 81 + a * 9;
      long base = (long)(2 * ez[iz]) * g.nnx + 2 * ex[ix];
      double s = 0.0;
#pragma unroll
      for (int b = 0; b < 9; b++) {
        int bx, bz;
        q2_pos(b, bx, bz);
        s += row[b] * x[base + (long)bz * g.nnx + bx];
      }
      acc += s;
    }
  y[n] = acc;
}

void energy_apply_vv(const BoxGeom& g, const double* VV, const double* x, double* y,
                     hipStream_t st) {
  prof().begin(PROF_ENERGY_VV, st);
  k_energy_vv<<<nblocks(g.nn, 256), 256, 0, st>>>(g, VV, x, y);
  prof().end(PROF_ENERGY_VV, st);
}

__global__ void k_energy_invdiag(BoxGeom g, const double* __restrict__ VV, double* __restrict__ d) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= g.nn) return;
  int I = (int)(n % g.nnx), J = (int)(n / g.nnx);
  int ex[2], lx[2], ez[2], lz[2];
  int cx = q2_adjacent(I, g.nelx, ex, lx), cz = q2_adjacent(J, g.nelz, ez, lz);
  double acc = 0.0;
  for (int iz = 0; iz < cz; iz++)
    for (int ix = 0; ix < cx; ix++) {
      long e = (long)ez[iz] * g.nelx + ex[ix];
      int a = q2_local(lx[ix], lz[iz]);
      acc += VV[e * 81 + a * 9 + a];
    }
  d[n] = 1.0 / acc;
}

void energy_invdiag_vv(const BoxGeom& g, const double* VV, double* invdiag, hipStream_t st) {
  k_energy_invdiag<<<nblocks(g.nn, 256), 256, 0, st>>>(g, VV, invdiag);
}

// ---------------------------------------------------------------------------
// pressure: lumped Q1 mass matrix divided by the element viscosity
// ---------------------------------------------------------------------------
__global__ void k_pmass_eta(BoxGeom g, double quarter_area, const double* __restrict__ eta_e,
                            double* __restrict__ md) {
  long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m >= g.nnP) return;
  int i = (int)(m % (g.nelx + 1)), j = (int)(m / (g.nelx + 1));
  double s = 0.0;
  for (int ez = j - 1; ez <= j; ez++) {
    if (ez < 0 || ez >= g.nelz) continue;
    for (int ex = i - 1; ex <= i; ex++) {
      if (ex < 0 || ex >= g.nelx) continue;
      s += quarter_area / eta_e[(long)ez * g.nelx + ex];
    }
  }
  md[m] = s;
}

void pressure_mass_eta_lumped(const BoxGeom& g, double area, const double* eta_e, double* mdiag,
                              hipStream_t st) {
  k_pmass_eta<<<nblocks(g.nnP, 256), 256, 0, st>>>(g, 0.25 * area, eta_e, mdiag);
}

// ---------------------------------------------------------------------------
// matrix-free divergence template: Ghat (18x4) of an hx*hz rectangle,
// G_el = gscale * Ghat with rows of fixed velocity dofs zeroed
// ---------------------------------------------------------------------------
__global__ void k_ghat(double hx, double hz, double gscale, double* __restrict__ Gh) {
  __shared__ double dx[81], dz[81], np_[36], jw[9];
  int t = threadIdx.x;
  if (t < 9) {
    const double q[3] = {-sqrt(3.0 / 5.0), 0.0, sqrt(3.0 / 5.0)};
    const double w[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    int iq = t / 3, jq = t % 3;
    double dr[9], ds[9], npv[4];
    basis_V_dr(q[iq], q[jq], dr);
    basis_V_dt(q[iq], q[jq], ds);
    basis_P(q[iq], q[jq], npv);
    for (int k = 0; k < 9; k++) {
      dx[t * 9 + k] = (2.0 / hx) * dr[k];
      dz[t * 9 + k] = (2.0 / hz) * ds[k];
    }
    for (int k = 0; k < 4; k++) np_[t * 4 + k] = npv[k];
    jw[t] = (hx / 2) * (hz / 2) * (w[iq] * w[jq]);
  }
  __syncthreads();
  for (int s = t; s < 72; s += blockDim.x) {
    int i = s / 4, k = s % 4, a = i >> 1, c = i & 1;
    double sum = 0.0;
    for (int qq = 0; qq < 9; qq++)
      sum -= (c == 0 ? dx[qq * 9 + a] : dz[qq * 9 + a]) * np_[qq * 4 + k] * jw[qq];
    Gh[s] = sum * gscale;
  }
}

void ghat_compute(double hx, double hz, double gscale, double* Gh, hipStream_t st) {
  k_ghat<<<1, 128, 0, st>>>(hx, hz, gscale, Gh);
}

// yu (+)= G p  (masked rows)
template <class T>
__global__ void k_gtmpl_u(BoxGeom g, const double* __restrict__ Gh, const uint8_t* __restrict__ fix,
                          const T* __restrict__ p, T* __restrict__ yu, int accumulate) {
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= g.nn) return;
  int I = (int)(n % g.nnx), J = (int)(n / g.nnx);
  int ex[2], lx[2], ez[2], lz[2];
  int cx = q2_adjacent(I, g.nelx, ex, lx), cz = q2_adjacent(J, g.nelz, ez, lz);
  double a0 = 0.0, a1 = 0.0;
  for (int iz = 0; iz < cz; iz++)
    for (int ix = 0; ix < cx; ix++) {
      int a = q2_local(lx[ix], lz[iz]);
      long pb = (long)ez[iz] * (g.nelx + 1) + ex[ix];
      const long pk[4] = {pb, pb + 1, pb + g.nelx + 2, pb + g.nelx + 1};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        double pv = (double)p[pk[k]];
        a0 += Gh[(2 * a) * 4 + k] * pv;
        a1 += Gh[(2 * a + 1) * 4 + k] * pv;
      }
    }
  if (fix[2 * n]) a0 = 0.0;
  if (fix[2 * n + 1]) a1 = 0.0;
  if (accumulate) {
    yu[2 * n] += (T)a0;
    yu[2 * n + 1] += (T)a1;
  } else {
    yu[2 * n] = (T)a0;
    yu[2 * n + 1] = (T)a1;
  }
}

// yp = alpha * G^T u (fixed velocity dofs excluded)
template <class T>
__global__ void k_gtmpl_p(BoxGeom g, const double* __restrict__ Gh, const uint8_t* __restrict__ fix,
                          const T* __restrict__ u, T* __restrict__ yp, double alpha) {
  long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m >= g.nnP) return;
  int i = (int)(m % (g.nelx + 1)), j = (int)(m / (g.nelx + 1));
  double s = 0.0;
  for (int ez = j - 1; ez <= j; ez++) {
    if (ez < 0 || ez >= g.nelz) continue;
    for (int ex = i - 1; ex <= i; ex++) {
      if (ex < 0 || ex >= g.nelx) continue;
      int k = (ex == i) ? (ez == j ? 0 : 3) : (ez == j ? 1 : 2);
      long base = (long)(2 * ez) * g.nnx + 2 * ex;
      for (int b = 0; b < 9; b++) {
        int bx, bz;
        q2_pos(b, bx, bz);
        long nd = base + (long)bz * g.nnx + bx;
        for (int c = 0; c < 2; c++)
          if (!fix[2 * nd + c]) s += Gh[(2 * b + c) * 4 + k] * (double)u[2 * nd + c];
      }
    }
  }
  yp[m] = (T)(alpha * s);
}

template <class T>
void gtmpl_apply_u(const BoxGeom& g, const double* Gh, const uint8_t* fix, const T* p, T* yu,
                   bool accumulate, hipStream_t st) {
  k_gtmpl_u<T><<<nblocks(g.nn, 256), 256, 0, st>>>(g, Gh, fix, p, yu, accumulate ? 1 : 0);
}

template <class T>
void gtmpl_apply_p(const BoxGeom& g, const double* Gh, const uint8_t* fix, const T* u, T* yp,
                   double alpha, hipStream_t st) {
  k_gtmpl_p<T><<<nblocks(g.nnP, 256), 256, 0, st>>>(g, Gh, fix, u, yp, alpha);
}

template void gtmpl_apply_u<float>(const BoxGeom&, const double*, const uint8_t*, const float*,
                                   float*, bool, hipStream_t);
template void gtmpl_apply_u<double>(const BoxGeom&, const double*, const uint8_t*, const double*,
                                    double*, bool, hipStream_t);
template void gtmpl_apply_p<float>(const BoxGeom&, const double*, const uint8_t*, const float*,
                                   float*, double, hipStream_t);
template void gtmpl_apply_p<double>(const BoxGeom&, const double*, const uint8_t*, const double*,
                                    double*, double, hipStream_t);

// elementwise helpers used by the solvers
template <class T>
__global__ void k_diag_scale(long n, double a, const double* __restrict__ d, const T* __restrict__ x,
                             T* __restrict__ y) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) y[i] = (T)(a * d[i] * (double)x[i]);
}

template <class T>
void diag_scale(long n, double a, const double* d, const T* x, T* y, hipStream_t st) {
  k_diag_scale<T><<<nblocks(n, 256), 256, 0, st>>>(n, a, d, x, y);
}
template void diag_scale<float>(long, double, const double*, const float*, float*, hipStream_t);
template void diag_scale<double>(long, double, const double*, const double*, double*,
                                 hipStream_t);

}  // namespace mw
