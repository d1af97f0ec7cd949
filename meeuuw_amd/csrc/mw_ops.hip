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

// ---------------------------------------------------------------------------
// Matrix-free FP64 Stokes residual on the box: ru = f - (K u + G p),
// rp = h - G^T u, with K = sum_q B^T C B eta_q JxW_q and G the reference's
// -(B_0 + B_1)^T N_P JxW_q * gscale per element, the Dirichlet rows of K
// replaced by their assembled diagonal and those of G zeroed (the assembled
// operator's masking), evaluated from per-quadrature-point viscosities
// without forming the 468 values per element.  One 16x16 velocity node tile
// (with its 8x8 pressure nodes) per block over the 9x9 covering elements, in
// three phases through LDS: per (element, point) the strain rates, pressure
// and the stress terms; per (element, node) the row pair and its diagonal
// (sums over the points in order), per (element, pressure node) the
// divergence row; then every node sums its elements' contributions in
// increasing element order.  Each expression is the per-element formula's,
// so the result does not depend on how the work is split.
// tab: dx[81], dz[81] (q*9 + node), N_P[36] (q*4 + pnode), JxW[9]
constexpr int TT64 = 16;             // velocity node tile edge
constexpr int R64_E = TT64 / 2 + 1;  // covering elements per direction
constexpr int R64_N = 2 * R64_E + 1; // covering node lines per direction
constexpr int R64_P = R64_E + 1;     // covering pressure lines per direction
constexpr int R64_I = R64_E * R64_E * 9;  // (element, point) and (element, node) items
constexpr int R64_PI = R64_E * R64_E * 4; // (element, pressure node) items
constexpr int R64_BS = 256;

using ctab = const __attribute__((address_space(4))) double*;  // read-only table, scalar loads

__global__ void __launch_bounds__(R64_BS)
k_stokes_residual64(BoxGeom g, const uint8_t* __restrict__ fix, const double* __restrict__ tab,
                    const double* __restrict__ etaq, const double* __restrict__ u,
                    const double* __restrict__ p, const double* __restrict__ f,
                    const double* __restrict__ h, double gscale, double* __restrict__ ru,
                    double* __restrict__ rp) {
  __shared__ double stab[207];
  __shared__ double su[R64_N * R64_N][2];  // masked (u, w) of the covering nodes
  __shared__ double sp[R64_P * R64_P];
  // per (element, point): s0, s1, s2, gp, c, gd; then per (element, node): y0, y1, d0, d1
  __shared__ double q0[R64_I], q1[R64_I], q2[R64_I], q3[R64_I], qc[R64_I], qd[R64_I];
  const int tid = threadIdx.x;
  const int ntx = (g.nnx + TT64 - 1) / TT64;
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int I0 = (lin % ntx) * TT64, J0 = (lin / ntx) * TT64;
  const int ex0 = I0 / 2 - 1, ez0 = J0 / 2 - 1;
  const int npx = g.nelx + 1;
  for (int i = tid; i < 207; i += R64_BS) stab[i] = tab[i];
  for (int i = tid; i < R64_N * R64_N; i += R64_BS) {
    const int I = 2 * ex0 + i % R64_N, J = 2 * ez0 + i / R64_N;
    const bool ok = I >= 0 && I < g.nnx && J >= 0 && J < g.nnz;
    const long m = ok ? (long)J * g.nnx + I : 0;
    const double a = u[2 * m], b = u[2 * m + 1];
    const bool f0 = fix[2 * m] != 0, f1 = fix[2 * m + 1] != 0;
    su[i][0] = (ok && !f0) ? a : 0.0;
    su[i][1] = (ok && !f1) ? b : 0.0;
  }
  for (int i = tid; i < R64_P * R64_P; i += R64_BS) {
    const int pi = ex0 + i % R64_P, pj = ez0 + i / R64_P;
    const bool ok = pi >= 0 && pi <= g.nelx && pj >= 0 && pj <= g.nelz;
    const double v = p[ok ? (long)pj * npx + pi : 0];
    sp[i] = ok ? v : 0.0;
  }
  __syncthreads();
  const double *dx = stab, *dz = stab + 81, *np = stab + 162, *jxw = stab + 198;
  // ---- per (element, point)
  for (int i = tid; i < R64_I; i += R64_BS) {
    const int le = i / 9, q = i % 9, lx = le % R64_E, lz = le / R64_E;
    const int ex = ex0 + lx, ez = ez0 + lz;
    const bool ok = ex >= 0 && ex < g.nelx && ez >= 0 && ez < g.nelz;
    const long e = ok ? (long)ez * g.nelx + ex : 0;
    const double eq = etaq[e * 9 + q];
    double exx = 0.0, ezz = 0.0, exz = 0.0, pq = 0.0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
      int bx, bz;
      q2_pos(k, bx, bz);
      const int n = (2 * lz + bz) * R64_N + 2 * lx + bx;
      const double ux = su[n][0], uz = su[n][1];
      exx += dx[q * 9 + k] * ux;
      ezz += dz[q * 9 + k] * uz;
      exz += dz[q * 9 + k] * ux + dx[q * 9 + k] * uz;
    }
    const int pb = lz * R64_P + lx;
    const double pv[4] = {sp[pb], sp[pb + 1], sp[pb + R64_P + 1], sp[pb + R64_P]};
#pragma unroll
    for (int k = 0; k < 4; k++) pq += np[q * 4 + k] * pv[k];
    const double jw = jxw[q], c = eq * jw;
    q0[i] = 2.0 * c * exx;
    q1[i] = 2.0 * c * ezz;
    q2[i] = c * exz;
    q3[i] = -gscale * jw * pq;
    qc[i] = c;
    qd[i] = -gscale * jw * (exx + ezz);
  }
  __syncthreads();
  // ---- per (element, node) and (element, pressure node), into registers first
  constexpr int NI = (R64_I + R64_PI + R64_BS - 1) / R64_BS;
  double o0[NI], o1[NI], o2[NI], o3[NI];
#pragma unroll
  for (int k = 0; k < NI; k++) {
    const int i = tid + k * R64_BS;
    o0[k] = o1[k] = o2[k] = o3[k] = 0.0;
    if (i < R64_I) {
      const int le = i / 9, a = i % 9;
      double y0 = 0.0, y1 = 0.0, d0 = 0.0, d1 = 0.0;
#pragma unroll
      for (int q = 0; q < 9; q++) {
        const int j = le * 9 + q;
        const double ax = dx[q * 9 + a], az = dz[q * 9 + a];
        const double s0 = q0[j], s1 = q1[j], s2 = q2[j], gp = q3[j], c = qc[j];
        y0 += ax * s0 + az * s2 + ax * gp;
        y1 += az * s1 + ax * s2 + az * gp;
        d0 += (2.0 * ax * ax + az * az) * c;
        d1 += (2.0 * az * az + ax * ax) * c;
      }
      o0[k] = y0;
      o1[k] = y1;
      o2[k] = d0;
      o3[k] = d1;
    } else if (i < R64_I + R64_PI) {
      const int m = i - R64_I, le = m / 4, pk = m % 4;
      double yq = 0.0;
#pragma unroll
      for (int q = 0; q < 9; q++) yq += np[q * 4 + pk] * qd[le * 9 + q];
      o0[k] = yq;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NI; k++) {
    const int i = tid + k * R64_BS;
    if (i < R64_I) {
      q0[i] = o0[k];
      q1[i] = o1[k];
      q2[i] = o2[k];
      q3[i] = o3[k];
    } else if (i < R64_I + R64_PI) {
      qc[i - R64_I] = o0[k];
    }
  }
  __syncthreads();
  {  // velocity node of this thread
    const int I = I0 + tid % TT64, J = J0 + tid / TT64;
    if (I < g.nnx && J < g.nnz) {
      const long n = (long)J * g.nnx + I;
      int ex[2], lx[2], ez[2], lz[2];
      const int cx = q2_adjacent(I, g.nelx, ex, lx), cz = q2_adjacent(J, g.nelz, ez, lz);
      double a0 = 0.0, a1 = 0.0, d0 = 0.0, d1 = 0.0;
      for (int iz = 0; iz < cz; iz++)
        for (int ix = 0; ix < cx; ix++) {
          const int j = ((ez[iz] - ez0) * R64_E + (ex[ix] - ex0)) * 9 + q2_local(lx[ix], lz[iz]);
          a0 += q0[j];
          a1 += q1[j];
          d0 += q2[j];
          d1 += q3[j];
        }
      if (fix[2 * n]) a0 = d0 * u[2 * n];
      if (fix[2 * n + 1]) a1 = d1 * u[2 * n + 1];
      ru[2 * n] = f[2 * n] - a0;
      ru[2 * n + 1] = f[2 * n + 1] - a1;
    }
  }
  if (tid < 64) {  // pressure node of this thread
    const int i = I0 / 2 + (tid & 7), j = J0 / 2 + (tid >> 3);
    if (i <= g.nelx && j <= g.nelz) {
      double s = 0.0;
      for (int dz = 0; dz < 2; dz++)
        for (int dxx = 0; dxx < 2; dxx++) {
          const int ez = j - 1 + dz, ex = i - 1 + dxx;
          if (ez < 0 || ez >= g.nelz || ex < 0 || ex >= g.nelx) continue;
          const int k = dxx ? (dz ? 0 : 3) : (dz ? 1 : 2);
          s += qc[((ez - ez0) * R64_E + (ex - ex0)) * 4 + k];
        }
      const long m = (long)j * npx + i;
      rp[m] = h[m] - s;
    }
  }
}

__global__ void k_residual64_tables(const double* __restrict__ dNdr, const double* __restrict__ dNdt,
                                    const double* __restrict__ NP, const double* __restrict__ j00,
                                    const double* __restrict__ j01, const double* __restrict__ j10,
                                    const double* __restrict__ j11, const double* __restrict__ jxw,
                                    double* __restrict__ tab) {
  const int t = threadIdx.x;
  if (t < 81) {
    const int q = t / 9;
    tab[t] = j00[q] * dNdr[t] + j01[q] * dNdt[t];
    tab[81 + t] = j10[q] * dNdr[t] + j11[q] * dNdt[t];
  }
  if (t < 36) tab[162 + t] = NP[t];
  if (t < 9) tab[198 + t] = jxw[t];
}

int residual64_tables(const double* dNdr_V, const double* dNdt_V, const double* N_P,
                      const double* j00q, const double* j01q, const double* j10q,
                      const double* j11q, const double* JxWq, double* tab, hipStream_t st) {
  if (!dNdr_V || !dNdt_V || !N_P || !j00q || !j01q || !j10q || !j11q || !JxWq || !tab)
    return fail(MW_EINVAL, "residual64_tables: null pointer");
  k_residual64_tables<<<1, 128, 0, st>>>(dNdr_V, dNdt_V, N_P, j00q, j01q, j10q, j11q, JxWq, tab);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

int stokes_residual64(const BoxGeom& g, const uint8_t* fix, const double* tab, const double* etaq,
                      const double* u, const double* p, const double* f, const double* h,
                      double gscale, double* ru, double* rp, hipStream_t st) {
  if (!fix || !tab || !etaq || !u || !p || !f || !h || !ru || !rp)
    return fail(MW_EINVAL, "stokes_residual64: null pointer");
  const int ntx = (g.nnx + TT64 - 1) / TT64, ntz = (g.nnz + TT64 - 1) / TT64;
  prof().begin(PROF_STOKES_VV, st);
  k_stokes_residual64<<<ntx * ntz, R64_BS, 0, st>>>(g, fix, tab, etaq, u, p, f, h, gscale, ru, rp);
  prof().end(PROF_STOKES_VV, st);
  MW_CHECK_LAUNCH();
  return MW_OK;
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
