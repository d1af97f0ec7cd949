// Energy equation on the box, matrix-free in FP64.
//
// The Crank-Nicolson system of build_matrix_energy.py (A = MM + (Ka + Kd) dt/2
// on the left, MM - (Ka + Kd) dt/2 times the old temperature plus the heat
// source on the right, Dirichlet rows replaced by their diagonal and the
// fixed columns lifted to the right-hand side; build_matrix_energy.py:60-150)
// is evaluated from its quadrature-point coefficients instead of the 81
// assembled values per element, and solved with Jacobi-preconditioned
// BiCGStab in three fused kernels per iteration.
//
// Every element of the box is the same hx x hz rectangle, so the shape
// function values and Cartesian derivatives at the 9 quadrature points are
// one table for all elements.  Per element and quadrature point four
// coefficients remain (h = dt/2):
//   cm = rho0 hcapa JxW,  cu = h cm u,  cw = h cm w,  ck = h hcond JxW
// and A_e x = sum_q N (cm xq + cu gx + cw gz) + dN/dx ck gx + dN/dz ck gz with
// xq, gx, gz the value and gradient of x at the point (sum factorisation:
// 54 multiply-adds per point instead of an 81-value row per node).
//
// Kernels work on 16x16 node tiles: the 9x9 elements touching the tile are
// evaluated once each (one thread per element and quadrature point, then one
// per element and node), and every node sums its elements' contributions in
// increasing element order.  No atomics; the dot products are fixed-order
// block partials finished by one single-block kernel, so results do not
// depend on scheduling.
#include "mw_common.h"
#include "mw_solver.h"

namespace mw {

constexpr int ET = 16;            // owned node tile edge
constexpr int EE = ET / 2 + 1;    // covering elements per direction
constexpr int EN = 2 * EE + 1;    // covering node lines per direction
constexpr int EH = EN * EN;       // covering nodes
constexpr int EI = EE * EE * 9;   // (element, point) items of a tile
constexpr int E_BS = 256;
constexpr int E_ITEMS = (EI + E_BS - 1) / E_BS;  // items per thread

// device scalars of the solve (EnergyMF::S)
enum { ES_RHO = 0, ES_ALPHA, ES_OMEGA, ES_RHONEW, ES_TOL, ES_BN, ES_DONE, ES_N };
// host mirror (EnergyMF::host_dev): residual norm of the last finished
// iteration, its done flag, |b|, the true residual norm of the final check
enum { EH_RN = 0, EH_DONE, EH_BN, EH_TRUE, EH_N };

__device__ __forceinline__ double e_block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  __syncthreads();  // red may still be read by a previous call
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < E_BS / 64; w++) s += red[w];
  return s;  // thread 0
}

struct ETile {
  int I0, J0, ex0, ez0, lin;
};

__device__ __forceinline__ ETile e_tile(const BoxGeom& g) {
  const int ntx = (g.nnx + ET - 1) / ET;
  ETile t;
  t.lin = xcd_remap(blockIdx.x, gridDim.x);
  t.I0 = (t.lin % ntx) * ET;
  t.J0 = (t.lin / ntx) * ET;
  t.ex0 = t.I0 / 2 - 1;
  t.ez0 = t.J0 / 2 - 1;
  return t;
}

// value and gradient of the staged nodal field sx at point q of tile element le
__device__ __forceinline__ void e_point(const double* sx, const double* sT, int le, int q,
                                        double& xq, double& gx, double& gz) {
  const int base = (2 * (le / EE)) * EN + 2 * (le % EE);
  xq = gx = gz = 0.0;
#pragma unroll
  for (int c = 0; c < 9; c++) {
    int lx, lz;
    q2_pos(c, lx, lz);
    const double v = sx[base + lz * EN + lx];
    xq += sT[q * 9 + c] * v;
    gx += sT[81 + q * 9 + c] * v;
    gz += sT[162 + q * 9 + c] * v;
  }
}

// element row r of tile element le from the per-point (s, kx, kz)
__device__ __forceinline__ double e_row(const double* sT, const double* ss, const double* skx,
                                        const double* skz, int le, int r) {
  double y = 0.0;
#pragma unroll
  for (int q = 0; q < 9; q++)
    y += sT[q * 9 + r] * ss[le * 9 + q] + sT[81 + q * 9 + r] * skx[le * 9 + q] +
         sT[162 + q * 9 + r] * skz[le * 9 + q];
  return y;
}

// sum of the per-element values sy over the elements of owned node (I, J)
__device__ __forceinline__ double e_gather(const BoxGeom& g, const ETile& t, const double* sy,
                                          int I, int J) {
  int ex[2], lx[2], ez[2], lz[2];
  const int cx = q2_adjacent(I, g.nelx, ex, lx), cz = q2_adjacent(J, g.nelz, ez, lz);
  double a = 0.0;
  for (int iz = 0; iz < cz; iz++)
    for (int ix = 0; ix < cx; ix++)
      a += sy[((ez[iz] - t.ez0) * EE + (ex[ix] - t.ex0)) * 9 + q2_local(lx[ix], lz[iz])];
  return a;
}

__device__ __forceinline__ bool e_valid(const BoxGeom& g, const ETile& t, int le, long* e) {
  const int ex = t.ex0 + le % EE, ez = t.ez0 + le / EE;
  const bool ok = ex >= 0 && ex < g.nelx && ez >= 0 && ez < g.nelz;
  *e = ok ? (long)ez * g.nelx + ex : 0;
  return ok;
}

// ---------------------------------------------------------------------------
// right-hand side, diagonal and coefficients (once per solve)
// ---------------------------------------------------------------------------
struct ERhsIn {
  BoxGeom g;
  const uint8_t* fix;
  const double *tab, *Told, *val, *uq, *wq, *hcapaq, *hcondq, *hprodq;
  double rho0, dt;
  double *coef, *b, *d, *invd;
};

__global__ void __launch_bounds__(E_BS) k_energy_rhs(ERhsIn a) {
  __shared__ double sT[243];
  __shared__ double sx[EH], sv[EH];
  __shared__ double sc[4][EI];
  __shared__ double ss[EI], skx[EI], skz[EI], sy[EI], sd[EI];
  const BoxGeom& g = a.g;
  const ETile t = e_tile(g);
  const int tid = threadIdx.x;
  const double h = 0.5 * a.dt;
  for (int i = tid; i < 243; i += E_BS) sT[i] = a.tab[i];
  for (int i = tid; i < EH; i += E_BS) {
    const int I = t.I0 - 2 + i % EN, J = t.J0 - 2 + i / EN;
    const bool ok = I >= 0 && I < g.nnx && J >= 0 && J < g.nnz;
    const long n = ok ? (long)J * g.nnx + I : 0;
    const double x = a.Told[n], v = a.val[n];
    const bool f = a.fix[n] != 0;
    sx[i] = ok ? x : 0.0;
    sv[i] = (ok && f) ? v : 0.0;
  }
  // coefficients of every (element, point) item
  for (int k = 0; k < E_ITEMS; k++) {
    const int i = tid + k * E_BS;
    if (i >= EI) break;
    const int le = i / 9, q = i % 9;
    long e;
    const bool ok = e_valid(g, t, le, &e);
    const long qi = e * 9 + q;
    const double jw = a.tab[243 + q];
    const double cm = a.rho0 * a.hcapaq[qi] * jw;
    double c[4] = {cm, h * cm * a.uq[qi], h * cm * a.wq[qi], h * a.hcondq[qi] * jw};
    const double src = a.dt * a.hprodq[qi] * jw;
#pragma unroll
    for (int m = 0; m < 4; m++) sc[m][i] = ok ? c[m] : 0.0;
    ss[i] = ok ? src : 0.0;  // the source, completed below
    const int lx = le % EE, lz = le / EE;
    if (ok && lx >= 1 && lz >= 1)  // elements owned by this tile
#pragma unroll
      for (int m = 0; m < 4; m++) a.coef[e * 36 + m * 9 + q] = c[m];
  }
  __syncthreads();
  // per point: E T_old (E = MM - (Ka + Kd) h) minus A applied to the fixed values
  for (int k = 0; k < E_ITEMS; k++) {
    const int i = tid + k * E_BS;
    if (i >= EI) break;
    const int le = i / 9, q = i % 9;
    double xq, gx, gz, vq, vx, vz;
    e_point(sx, sT, le, q, xq, gx, gz);
    e_point(sv, sT, le, q, vq, vx, vz);
    const double cm = sc[0][i], cu = sc[1][i], cw = sc[2][i], ck = sc[3][i];
    ss[i] += (cm * xq - cu * gx - cw * gz) - (cm * vq + cu * vx + cw * vz);
    skx[i] = -ck * gx - ck * vx;
    skz[i] = -ck * gz - ck * vz;
  }
  __syncthreads();
  // per element row: right-hand side and the diagonal of A
  for (int k = 0; k < E_ITEMS; k++) {
    const int i = tid + k * E_BS;
    if (i >= EI) break;
    const int le = i / 9, r = i % 9;
    sy[i] = e_row(sT, ss, skx, skz, le, r);
    double dg = 0.0;
#pragma unroll
    for (int q = 0; q < 9; q++) {
      const double N = sT[q * 9 + r], dx = sT[81 + q * 9 + r], dz = sT[162 + q * 9 + r];
      const int j = le * 9 + q;
      dg += N * N * sc[0][j] + N * (sc[1][j] * dx + sc[2][j] * dz) + sc[3][j] * (dx * dx + dz * dz);
    }
    sd[i] = dg;
  }
  __syncthreads();
  const int I = t.I0 + tid % ET, J = t.J0 + tid / ET;
  if (I < g.nnx && J < g.nnz) {
    const long n = (long)J * g.nnx + I;
    const double dn = e_gather(g, t, sd, I, J);
    const double y = e_gather(g, t, sy, I, J);
    const int hi = (tid / ET + 2) * EN + tid % ET + 2;
    a.b[n] = a.fix[n] ? dn * sv[hi] : y;
    a.d[n] = dn;
    a.invd[n] = 1.0 / dn;
  }
}

// ---------------------------------------------------------------------------
// operator kernels of the solve
// ---------------------------------------------------------------------------
enum { EM_RES = 0, EM_P = 1, EM_S = 2 };

struct EOpIn {
  BoxGeom g;
  const uint8_t* fix;
  const double *tab, *coef, *d, *invd, *S;
  // EM_RES: r = b - A x (init: rh = r, p0 = v0 = 0); partials r.r, b.b
  // EM_P:   p1 = r + beta (p0 - omega v0), ph = invd p1, v1 = A ph; partial rh.v1
  // EM_S:   s = r - alpha v1, sh = invd s, t = A sh; partials t.s, t.t
  const double *x, *b, *r, *rh, *p0, *v0, *v1;
  double *r_out, *rh_out, *p0_out, *v0_out, *p1, *ph, *v1_out, *t;
  double* part;
  int init;
};

// the operator's input at node n before (z) and after (y) the Jacobi scaling
template <int MODE>
__device__ __forceinline__ void e_input(const EOpIn& a, long n, double alpha, double beta,
                                        double omega, double& z, double& y) {
  if (MODE == EM_RES) {
    z = a.x[n];
    y = z;
  } else if (MODE == EM_P) {
    const double r = a.r[n], p = a.p0[n], v = a.v0[n], id = a.invd[n];
    z = r + beta * (p - omega * v);
    y = id * z;
  } else {
    const double r = a.r[n], v = a.v1[n], id = a.invd[n];
    z = r - alpha * v;
    y = id * z;
  }
}

template <int MODE>
__global__ void __launch_bounds__(E_BS) k_energy_op(EOpIn a) {
  // 22 KB of LDS: seven blocks per CU keep a 256x256 mesh's 1089 tiles resident at once
  __shared__ double sT[243];
  __shared__ double sx[EH];  // operator input, fixed nodes zeroed
  __shared__ double ss[EI], skx[EI], skz[EI];  // per point; ss then holds the element rows
  __shared__ double red[E_BS / 64];
  const double* S = a.S;
  double alpha = 0.0, beta = 0.0, omega = 0.0;
  if (MODE != EM_RES) {
    if (S[ES_DONE] != 0.0) return;  // converged: an iteration queued ahead of the check
    alpha = S[ES_ALPHA];
    omega = S[ES_OMEGA];
    beta = (S[ES_RHONEW] / S[ES_RHO]) * (alpha / omega);
  }
  const BoxGeom& g = a.g;
  const ETile t = e_tile(g);
  const int tid = threadIdx.x;
  // coefficients of this thread's items, loaded before anything waits
  double cf[E_ITEMS][4];
#pragma unroll
  for (int k = 0; k < E_ITEMS; k++) {
    const int i = min(tid + k * E_BS, EI - 1);
    long e;
    const bool ok = e_valid(g, t, i / 9, &e);
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const double c = a.coef[e * 36 + m * 9 + i % 9];
      cf[k][m] = ok ? c : 0.0;
    }
  }
  for (int i = tid; i < 243; i += E_BS) sT[i] = a.tab[i];
  for (int i = tid; i < EH; i += E_BS) {
    const int I = t.I0 - 2 + i % EN, J = t.J0 - 2 + i / EN;
    const bool ok = I >= 0 && I < g.nnx && J >= 0 && J < g.nnz;
    const long n = ok ? (long)J * g.nnx + I : 0;
    const bool f = a.fix[n] != 0;
    double z, y;
    e_input<MODE>(a, n, alpha, beta, omega, z, y);
    sx[i] = (ok && !f) ? y : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < E_ITEMS; k++) {
    const int i = tid + k * E_BS;
    if (i < EI) {
      double xq, gx, gz;
      e_point(sx, sT, i / 9, i % 9, xq, gx, gz);
      ss[i] = cf[k][0] * xq + cf[k][1] * gx + cf[k][2] * gz;
      skx[i] = cf[k][3] * gx;
      skz[i] = cf[k][3] * gz;
    }
  }
  __syncthreads();
  double yr[E_ITEMS];
#pragma unroll
  for (int k = 0; k < E_ITEMS; k++) {
    const int i = tid + k * E_BS;
    yr[k] = i < EI ? e_row(sT, ss, skx, skz, i / 9, i % 9) : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < E_ITEMS; k++) {
    const int i = tid + k * E_BS;
    if (i < EI) ss[i] = yr[k];
  }
  __syncthreads();
  double q0 = 0.0, q1 = 0.0;
  const int I = t.I0 + tid % ET, J = t.J0 + tid / ET;
  if (I < g.nnx && J < g.nnz) {
    const long n = (long)J * g.nnx + I;
    double z, y0;
    e_input<MODE>(a, n, alpha, beta, omega, z, y0);
    const double ay = a.fix[n] ? a.d[n] * y0 : e_gather(g, t, ss, I, J);
    if (MODE == EM_RES) {
      const double b = a.b[n], r = b - ay;
      a.r_out[n] = r;
      if (a.init) {
        a.rh_out[n] = r;
        a.p0_out[n] = 0.0;
        a.v0_out[n] = 0.0;
      }
      q0 = r * r;
      q1 = b * b;
    } else if (MODE == EM_P) {
      a.p1[n] = z;
      a.ph[n] = y0;
      a.v1_out[n] = ay;
      q0 = a.rh[n] * ay;
    } else {
      a.t[n] = ay;
      q0 = ay * z;
      q1 = ay * ay;
    }
  }
  const int nblk = gridDim.x;
  q0 = e_block_sum(q0, red);
  if (tid == 0) a.part[blockIdx.x] = q0;
  if (MODE != EM_P) {
    q1 = e_block_sum(q1, red);
    if (tid == 0) a.part[nblk + blockIdx.x] = q1;
  }
}

// pointwise end of an iteration: x += alpha ph + omega sh, r = s - omega t with
// s = r - alpha v, sh = invd s; partials rh.r and r.r
__global__ void __launch_bounds__(E_BS)
k_energy_update(long nn, const double* __restrict__ S, const double* __restrict__ invd,
                const double* __restrict__ v, const double* __restrict__ ph,
                const double* __restrict__ tt, const double* __restrict__ rh,
                double* __restrict__ x, double* __restrict__ r, double* __restrict__ part) {
  __shared__ double red[E_BS / 64];
  if (S[ES_DONE] != 0.0) return;
  const double alpha = S[ES_ALPHA], omega = S[ES_OMEGA];
  const long n = blockIdx.x * (long)E_BS + threadIdx.x;
  double q0 = 0.0, q1 = 0.0;
  if (n < nn) {
    const double s = r[n] - alpha * v[n];
    const double sh = invd[n] * s;
    x[n] = (x[n] + alpha * ph[n]) + omega * sh;
    const double rn = s - omega * tt[n];
    r[n] = rn;
    q0 = rh[n] * rn;
    q1 = rn * rn;
  }
  q0 = e_block_sum(q0, red);
  if (threadIdx.x == 0) part[blockIdx.x] = q0;
  q1 = e_block_sum(q1, red);
  if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = q1;
}

// single-block finish of the partial sums; what = which step of the solve
enum { EF_START = 0, EF_ALPHA, EF_OMEGA, EF_END, EF_TRUE };

// fixed-order sums of one or two rows of block partials (thread 0 gets them);
// every load is issued before the first add
__device__ __forceinline__ void e_sum_parts(const double* part, int nparts, bool two, double* red,
                                            double& s0, double& s1) {
  constexpr int U = 8;
  s0 = s1 = 0.0;
  for (int base = threadIdx.x; base < nparts; base += U * E_BS) {
    double v0[U], v1[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = base + u * E_BS;
      const int c = min(i, nparts - 1);
      const double a = part[c], b = two ? part[nparts + c] : 0.0;
      v0[u] = i < nparts ? a : 0.0;
      v1[u] = i < nparts ? b : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      s0 += v0[u];
      s1 += v1[u];
    }
  }
  s0 = e_block_sum(s0, red);
  if (two) s1 = e_block_sum(s1, red);
}

__global__ void __launch_bounds__(E_BS)
k_energy_final(int what, int nparts, const double* __restrict__ part, double tol_rel,
               double* __restrict__ S, double* __restrict__ host) {
  __shared__ double red[E_BS / 64];
  if (what != EF_START && what != EF_TRUE && S[ES_DONE] != 0.0) return;
  double s0, s1;
  e_sum_parts(part, nparts, what != EF_ALPHA, red, s0, s1);
  if (threadIdx.x != 0) return;
  if (what == EF_START) {  // s0 = r.r, s1 = b.b
    const double bn = sqrt(s1), rn = sqrt(s0);
    const double tol = tol_rel * (bn > 0.0 ? bn : 1.0);
    S[ES_RHO] = S[ES_ALPHA] = S[ES_OMEGA] = 1.0;
    S[ES_RHONEW] = s0;
    S[ES_TOL] = tol;
    S[ES_BN] = bn;
    const double done = rn <= tol ? 1.0 : 0.0;
    S[ES_DONE] = done;
    host[EH_RN] = rn;
    host[EH_BN] = bn;
    host[EH_DONE] = done;
  } else if (what == EF_ALPHA) {  // s0 = rh.v
    S[ES_ALPHA] = S[ES_RHONEW] / s0;
  } else if (what == EF_OMEGA) {  // s0 = t.s, s1 = t.t
    S[ES_OMEGA] = s0 / s1;
  } else if (what == EF_END) {  // s0 = rh.r, s1 = r.r
    S[ES_RHO] = S[ES_RHONEW];
    S[ES_RHONEW] = s0;
    const double rn = sqrt(s1);
    const double done = rn <= S[ES_TOL] ? 1.0 : 0.0;
    S[ES_DONE] = done;
    host[EH_RN] = rn;
    host[EH_DONE] = done;
  } else {  // EF_TRUE: s0 = |b - A x|^2
    host[EH_TRUE] = sqrt(s0);
  }
}

__global__ void k_energy_tables(const double* __restrict__ N, const double* __restrict__ dNdr,
                                const double* __restrict__ dNds, const double* __restrict__ j00,
                                const double* __restrict__ j01, const double* __restrict__ j10,
                                const double* __restrict__ j11, const double* __restrict__ jxw,
                                double* __restrict__ tab) {
  const int t = threadIdx.x;
  if (t < 81) {
    const int q = t / 9;
    tab[t] = N[t];
    tab[81 + t] = j00[q] * dNdr[t] + j01[q] * dNds[t];
    tab[162 + t] = j10[q] * dNdr[t] + j11[q] * dNds[t];
  }
  if (t < 9) tab[243 + t] = jxw[t];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static unsigned e_tiles(const BoxGeom& g) {
  return (unsigned)(((g.nnx + ET - 1) / ET) * ((g.nnz + ET - 1) / ET));
}

long energy_mf_partials(const BoxGeom& g) {
  const long a = e_tiles(g), b = nblocks(g.nn, E_BS);
  return 2 * (a > b ? a : b);
}

int energy_mf_tables(const double* N_V, const double* dNdr_V, const double* dNdt_V,
                     const double* j00q, const double* j01q, const double* j10q,
                     const double* j11q, const double* JxWq, double* tab, hipStream_t st) {
  if (!N_V || !dNdr_V || !dNdt_V || !j00q || !j01q || !j10q || !j11q || !JxWq || !tab)
    return fail(MW_EINVAL, "energy_mf_tables: null pointer");
  k_energy_tables<<<1, 128, 0, st>>>(N_V, dNdr_V, dNdt_V, j00q, j01q, j10q, j11q, JxWq, tab);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

int energy_mf_rhs(const EnergyMF& E, const double* Told, const double* val, const double* uq,
                  const double* wq, const double* hcapaq, const double* hcondq,
                  const double* hprodq, double rho0, double dt, double* b, hipStream_t st) {
  if (!E.fix || !E.tab || !E.coef || !E.d || !E.invd || !Told || !val || !uq || !wq || !hcapaq ||
      !hcondq || !hprodq || !b)
    return fail(MW_EINVAL, "energy_mf_rhs: null pointer");
  ERhsIn a;
  a.g = E.g; a.fix = E.fix; a.tab = E.tab; a.Told = Told; a.val = val; a.uq = uq; a.wq = wq;
  a.hcapaq = hcapaq; a.hcondq = hcondq; a.hprodq = hprodq; a.rho0 = rho0; a.dt = dt;
  a.coef = E.coef; a.b = b; a.d = E.d; a.invd = E.invd;
  k_energy_rhs<<<e_tiles(E.g), E_BS, 0, st>>>(a);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

static EOpIn e_op(const EnergyMF& E) {
  EOpIn a{};
  a.g = E.g; a.fix = E.fix; a.tab = E.tab; a.coef = E.coef; a.d = E.d; a.invd = E.invd;
  a.S = E.S; a.part = E.part;
  return a;
}

// r = b - A x and the norms; init: also rh = r, p0 = v0 = 0 and the solver scalars
static int e_residual(const EnergyMF& E, const double* x, const double* b, double* r, double* rh,
                      double* p0, double* v0, int init, double tol_rel, hipStream_t st) {
  EOpIn a = e_op(E);
  a.x = x; a.b = b; a.r_out = r; a.rh_out = rh; a.p0_out = p0; a.v0_out = v0; a.init = init;
  const unsigned nt = e_tiles(E.g);
  prof().begin(PROF_ENERGY_VV, st);
  k_energy_op<EM_RES><<<nt, E_BS, 0, st>>>(a);
  prof().end(PROF_ENERGY_VV, st);
  k_energy_final<<<1, E_BS, 0, st>>>(init ? EF_START : EF_TRUE, (int)nt, E.part, tol_rel, E.S,
                                    E.host_dev);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

int energy_mf_solve(const EnergyMF& E, double* x, const double* b, double* const w[8],
                    double tol_rel, int maxit, hipEvent_t ev[2], int* its, hipStream_t st) {
  if (!E.fix || !E.tab || !E.coef || !E.d || !E.invd || !E.S || !E.host || !E.host_dev ||
      !E.part || !x || !b || !its)
    return fail(MW_EINVAL, "energy_mf_solve: null pointer");
  for (int i = 0; i < 8; i++)
    if (!w[i]) return fail(MW_EINVAL, "energy_mf_solve: null work vector");
  double *r = w[0], *rh = w[1], *ph = w[6], *tt = w[7];
  double* P[2] = {w[2], w[3]};
  double* V[2] = {w[4], w[5]};
  const unsigned nt = e_tiles(E.g), nu = nblocks(E.g.nn, E_BS);
  RC(e_residual(E, x, b, r, rh, P[0], V[0], 1, tol_rel, st));
  MW_HIP(hipEventRecord(ev[0], st));
  // Iteration k is queued before the residual of iteration k-1 is read back,
  // so the host never leaves the device idle; an iteration queued after
  // convergence finds the device-side done flag set and does nothing.
  int k = 1;
  for (;; k++) {
    if (k <= maxit) {
      EOpIn a = e_op(E);
      a.r = r; a.rh = rh; a.p0 = P[(k - 1) & 1]; a.v0 = V[(k - 1) & 1];
      a.p1 = P[k & 1]; a.ph = ph; a.v1_out = V[k & 1];
      prof().begin(PROF_ENERGY_VV, st);
      k_energy_op<EM_P><<<nt, E_BS, 0, st>>>(a);
      prof().end(PROF_ENERGY_VV, st);
      k_energy_final<<<1, E_BS, 0, st>>>(EF_ALPHA, (int)nt, E.part, tol_rel, E.S, E.host_dev);
      EOpIn c = e_op(E);
      c.r = r; c.v1 = V[k & 1]; c.t = tt;
      prof().begin(PROF_ENERGY_VV, st);
      k_energy_op<EM_S><<<nt, E_BS, 0, st>>>(c);
      prof().end(PROF_ENERGY_VV, st);
      k_energy_final<<<1, E_BS, 0, st>>>(EF_OMEGA, (int)nt, E.part, tol_rel, E.S, E.host_dev);
      k_energy_update<<<nu, E_BS, 0, st>>>(E.g.nn, E.S, E.invd, V[k & 1], ph, tt, rh, x, r,
                                           E.part);
      k_energy_final<<<1, E_BS, 0, st>>>(EF_END, (int)nu, E.part, tol_rel, E.S, E.host_dev);
      MW_CHECK_LAUNCH();
      MW_HIP(hipEventRecord(ev[k & 1], st));
    }
    MW_HIP(hipEventSynchronize(ev[(k - 1) & 1]));
    const double rn = E.host[EH_RN];
    if (E.host[EH_DONE] != 0.0 || !(rn == rn)) break;  // converged (or NaN) after k-1
    if (k > maxit) break;
  }
  *its = k - 1;
  // true residual of the result, read back with the step's diagnostics
  return e_residual(E, x, b, r, nullptr, nullptr, nullptr, 0, tol_rel, st);
}

double energy_mf_true_residual(const EnergyMF& E) { return E.host[EH_TRUE]; }
double energy_mf_bnorm(const EnergyMF& E) { return E.host[EH_BN]; }

}  // namespace mw
