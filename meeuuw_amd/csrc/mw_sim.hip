// Device-resident timestep (box geometry), its solvers and the C entry points.
//
// The stages follow the reference time loop, meeuuw.py:1105-2875, and reuse
// the drop-in kernels (assembly, PIC, projections, reductions).  Two Stokes
// solvers:
//   solver 0 (plain FP64): Uzawa conjugate gradients on the pressure Schur
//     complement, preconditioned by the lumped pressure mass matrix scaled by
//     1/eta (the preconditioner of cd_u_eta.py:43-76), every K^-1 an FP64 CG
//     preconditioned by one geometric-multigrid V-cycle; the structure of
//     uzawa3_solver_L2.py:31-95.
//   solver 1 (mixed): FGMRES in FP32 on the whole saddle-point system with an
//     upper block-triangular preconditioner (FP32 multigrid V-cycle for K,
//     lumped eta-scaled pressure mass matrix for the Schur complement),
//     wrapped in FP64 defect correction against the assembled FP64 matrix.
// Both stop on the relative residual of the full assembled system.
#include <cmath>
#include <algorithm>
#include <cstring>
#include <vector>

#include "mw_capi_impl.h"
#include "mw_common.h"
#include "mw_ops.h"
#include "mw_sim.h"

namespace mw {

// ---------------------------------------------------------------- setup ----
__global__ void k_mesh_nodes(int nnx, int nnz, double hx, double hz, double* __restrict__ x,
                             double* __restrict__ z) {
  long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (c >= (long)nnx * nnz) return;
  int i = (int)(c % nnx), j = (int)(c / nnx);
  x[c] = i * hx / 2;  // build_velocity_mesh.py:116-117
  z[c] = j * hz / 2;
}

__global__ void k_mesh_elems(int nelx, int nelz, int* __restrict__ iv, int* __restrict__ ip,
                             uint8_t* __restrict__ L, uint8_t* __restrict__ R,
                             uint8_t* __restrict__ B, uint8_t* __restrict__ T) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  int nel = nelx * nelz;
  if (c >= nel) return;
  int i = c % nelx, j = c / nelx, nnx = 2 * nelx + 1;
  // build_velocity_connectivity.py:67-75, build_pressure_connectivity.py:19-22
  iv[0 * nel + c] = i * 2 + 1 + j * 2 * nnx - 1;
  iv[1 * nel + c] = i * 2 + 3 + j * 2 * nnx - 1;
  iv[2 * nel + c] = i * 2 + 3 + j * 2 * nnx + nnx * 2 - 1;
  iv[3 * nel + c] = i * 2 + 1 + j * 2 * nnx + nnx * 2 - 1;
  iv[4 * nel + c] = i * 2 + 2 + j * 2 * nnx - 1;
  iv[5 * nel + c] = i * 2 + 3 + j * 2 * nnx + nnx - 1;
  iv[6 * nel + c] = i * 2 + 2 + j * 2 * nnx + nnx * 2 - 1;
  iv[7 * nel + c] = i * 2 + 1 + j * 2 * nnx + nnx - 1;
  iv[8 * nel + c] = i * 2 + 2 + j * 2 * nnx + nnx - 1;
  ip[0 * nel + c] = i + j * (nelx + 1);
  ip[1 * nel + c] = i + 1 + j * (nelx + 1);
  ip[2 * nel + c] = i + 1 + (j + 1) * (nelx + 1);
  ip[3 * nel + c] = i + (j + 1) * (nelx + 1);
  L[c] = i == 0;
  R[c] = i == nelx - 1;
  B[c] = j == 0;
  T[c] = j == nelz - 1;
}

// reference-element tables (basis_functions_setup_q.py:48-63, _Vnodes.py)
__global__ void k_tables(double* NV, double* dr, double* dt, double* NP, double* dpr, double* dpt,
                         double* NPn, double* drn, double* dtn, double* qc, double* qw) {
  int t = threadIdx.x;
  const double q[3] = {-sqrt(3.0 / 5.0), 0.0, sqrt(3.0 / 5.0)};
  const double w[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  const double rv[9] = {-1, 1, 1, -1, 0, 1, 0, -1, 0};
  const double tv[9] = {-1, -1, 1, 1, -1, 0, 1, 0, 0};
  if (t < 9) {
    int iq = t / 3, jq = t % 3;
    double r = q[iq], s = q[jq];
    basis_V(r, s, NV + t * 9);
    basis_V_dr(r, s, dr + t * 9);
    basis_V_dt(r, s, dt + t * 9);
    basis_P(r, s, NP + t * 4);
    dpr[t * 4 + 0] = -0.25 * (1 - s);
    dpr[t * 4 + 1] = 0.25 * (1 - s);
    dpr[t * 4 + 2] = 0.25 * (1 + s);
    dpr[t * 4 + 3] = -0.25 * (1 + s);
    dpt[t * 4 + 0] = -0.25 * (1 - r);
    dpt[t * 4 + 1] = -0.25 * (1 + r);
    dpt[t * 4 + 2] = 0.25 * (1 + r);
    dpt[t * 4 + 3] = 0.25 * (1 - r);
    basis_P(rv[t], tv[t], NPn + t * 4);
    basis_V_dr(rv[t], tv[t], drn + t * 9);
    basis_V_dt(rv[t], tv[t], dtn + t * 9);
  }
  if (t < 3) {
    qc[t] = q[t];
    qw[t] = w[t];
  }
}

// Jacobian inverse and JxW at np points per element (basis_functions_setup_q.py:65-75)
__global__ void k_jac(int nel, const double* __restrict__ dr, const double* __restrict__ dt,
                      const double* __restrict__ qw, int with_w, const double* __restrict__ x,
                      const double* __restrict__ z, const int* __restrict__ icon,
                      double* __restrict__ jxw, double* __restrict__ a00,
                      double* __restrict__ a01, double* __restrict__ a10,
                      double* __restrict__ a11) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= 9L * nel) return;
  int e = (int)(idx / 9), k = (int)(idx % 9);
  double xe[9], ze[9];
  for (int m = 0; m < 9; m++) {
    xe[m] = x[icon[m * (long)nel + e]];
    ze[m] = z[icon[m * (long)nel + e]];
  }
  double a = dot9(dr + k * 9, xe), b = dot9(dr + k * 9, ze);
  double c = dot9(dt + k * 9, xe), d = dot9(dt + k * 9, ze);
  double det = a * d - b * c;
  if (with_w) jxw[idx] = det * (qw[k / 3] * qw[k % 3]);
  a00[idx] = d / det;
  a01[idx] = -b / det;
  a10[idx] = -c / det;
  a11[idx] = a / det;
}

// free slip on all sides and T boundary values (experiment0.py:128-176)
__global__ void k_bc(long nn, const double* __restrict__ x, const double* __restrict__ z,
                     double Lx, double Lz, double Tb, double Tt, uint8_t* __restrict__ fixV,
                     double* __restrict__ valV, uint8_t* __restrict__ fixT,
                     double* __restrict__ valT) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const double eps = 1e-8;
  bool fu = (x[i] / Lx < eps) || (x[i] / Lx > (1 - eps));
  bool fw = (z[i] / Lz < eps) || (z[i] / Lz > (1 - eps));
  fixV[2 * i] = fu;
  fixV[2 * i + 1] = fw;
  valV[2 * i] = 0.0;
  valV[2 * i + 1] = 0.0;
  uint8_t ft = 0;
  double vt = 0.0;
  if (z[i] < eps) { ft = 1; vt = Tb; }
  if (z[i] > (Lz - eps)) { ft = 1; vt = Tt; }
  fixT[i] = ft;
  valT[i] = vt;
}

// experiment0.py:91-95
__global__ void k_init_T(long nn, const double* __restrict__ x, const double* __restrict__ z,
                         double Lx, double Lz, double Tb, double Tt, double* __restrict__ T) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const double pi = 3.141592653589793;
  T[i] = (Tb - Tt) * (Lz - z[i]) / Lz + Tt + 0.01 * cos(pi * x[i] / Lx) * sin(pi * z[i] / Lz);
}

// element-major placement from reduced coordinates (swarm_coordinates_setup.py:71-86)
__global__ void k_place(int nel, int ppe, const double* __restrict__ r, const double* __restrict__ t,
                        const double* __restrict__ xv, const double* __restrict__ zv,
                        const int* __restrict__ icon, double* __restrict__ x,
                        double* __restrict__ z, double* __restrict__ sr, double* __restrict__ st,
                        int* __restrict__ iel, uint8_t* __restrict__ act) {
  long ip = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (ip >= (long)nel * ppe) return;
  int e = (int)(ip / ppe);
  double N[9], xe[9], ze[9];
  basis_V(r[ip], t[ip], N);
  for (int m = 0; m < 9; m++) {
    xe[m] = xv[icon[m * (long)nel + e]];
    ze[m] = zv[icon[m * (long)nel + e]];
  }
  x[ip] = dot9(N, xe);
  z[ip] = dot9(N, ze);
  sr[ip] = r[ip];
  st[ip] = t[ip];
  iel[ip] = e;
  act[ip] = 1;
}

// particles collocated with the quadrature points (swarm_coordinates_setup.py:51-65)
__global__ void k_place_qpts(int nel, const double* __restrict__ qc, const double* __restrict__ NV,
                             const double* __restrict__ xv, const double* __restrict__ zv,
                             const int* __restrict__ icon, double* __restrict__ x,
                             double* __restrict__ z, double* __restrict__ sr,
                             double* __restrict__ st, int* __restrict__ iel,
                             uint8_t* __restrict__ act) {
  long ip = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (ip >= 9L * nel) return;
  int e = (int)(ip / 9), c = (int)(ip % 9);
  double xe[9], ze[9];
  for (int m = 0; m < 9; m++) {
    xe[m] = xv[icon[m * (long)nel + e]];
    ze[m] = zv[icon[m * (long)nel + e]];
  }
  x[ip] = dot9(NV + c * 9, xe);
  z[ip] = dot9(NV + c * 9, ze);
  sr[ip] = qc[c / 3];
  st[ip] = qc[c % 3];
  iel[ip] = e;
  act[ip] = 1;
}

// ------------------------------------------------------------ per step -----
// Temperature on the particles (interpolate_field_on_particles,
// pic_functions.py:84-101; inactive particles get 0), then the Boussinesq
// material model (experiment0.py:230-264) with Frank-Kamenetskii viscosity
// (experiment3.py:46, 72) when contrast != 1.  One thread per slot j of the
// element-sorted order: particle order[j] is evaluated and its properties and
// reduced coordinates are stored at j, so the projections that follow read
// every particle field contiguously.  Slots from seg[nel] on hold the particles
// no projection reads (inactive ones).  The box grid's own node numbers
// replace icon_V (the same numbers).
// Two-material mode (mat != NULL): the particle's material m gives eta = meta[m]
// and rho = mrho[m] (1 - alpha T).
struct MatLaw {
  const uint8_t* mat;
  double meta[2], mrho[2];
  // visco-plastic law of Tosi et al. (2015) (tosi: case 1..5, 0 off)
  int tosi;
  double gT, gy, sy, estar;
  const double *exx, *ezz, *exz;  // nodal strain rate
  const double* z;                // particle z (particle order)
};

// viscosity(), experiment3.py:67-110
__device__ inline double tosi_viscosity(const MatLaw& ml, double T, double exx, double eyy,
                                        double exy, double y) {
  double val;
  if (ml.tosi == 1) {
    val = exp(-ml.gT * T);
  } else if (ml.tosi == 3) {
    val = exp(-ml.gT * T + ml.gy * (1 - y));
  } else {
    double e = sqrt(0.5 * (exx * exx + eyy * eyy) + exy * exy);
    e = e > 1e-12 ? e : 1e-12;
    const double lin = ml.tosi == 2 ? exp(-ml.gT * T) : exp(-ml.gT * T + ml.gy * (1 - y));
    const double plast = ml.estar + ml.sy / (sqrt(2.0) * e);
    val = 2.0 / (1.0 / lin + 1.0 / plast);
  }
  val = val < 2.0 ? val : 2.0;
  val = val > 1.0e-5 ? val : 1.0e-5;
  return val;
}

__global__ void k_material_sorted(int np, int nel, int nelx, const int* __restrict__ order,
                                  const int* __restrict__ seg, const int* __restrict__ siel,
                                  const double* __restrict__ r, const double* __restrict__ t,
                                  const double* __restrict__ T, double rho0, double alpha,
                                  double lncontrast, double emin, double emax, int iso,
                                  MatLaw ml, double* __restrict__ rho, double* __restrict__ eta,
                                  double* __restrict__ hcond, double* __restrict__ hcapa,
                                  double* __restrict__ hprod, double* __restrict__ al,
                                  double* __restrict__ rs, double* __restrict__ ts) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= np) return;
  double tp = 0.0, rr = 0.0, tt = 0.0, exx = 0.0, ezz = 0.0, exz = 0.0;
  const int ip = order[j];
  if (j < seg[nel]) {
    const int e = siel[ip];
    rr = r[ip];
    tt = t[ip];
    double N[9], v[9];
    basis_V(rr, tt, N);
    const int nnx = 2 * nelx + 1;
#pragma unroll
    for (int k = 0; k < 9; k++) v[k] = T[grid_node(e, nelx, nnx, k)];
    tp = dot9(N, v);
    if (ml.tosi) {  // strain rate on the particle (meeuuw.py:1116-1118)
      double a[9], b[9];
#pragma unroll
      for (int k = 0; k < 9; k++) {
        const int n = grid_node(e, nelx, nnx, k);
        v[k] = ml.exx[n];
        a[k] = ml.ezz[n];
        b[k] = ml.exz[n];
      }
      exx = dot9(N, v);
      ezz = dot9(N, a);
      exz = dot9(N, b);
    }
  }
  double ev = 1.0;
  if (ml.tosi) {
    rho[j] = rho0 * (1 - alpha * tp);
    ev = tosi_viscosity(ml, tp, exx, ezz, exz, ml.z[ip]);
  } else if (ml.mat) {
    const int m = ml.mat[ip] ? 1 : 0;
    rho[j] = ml.mrho[m] * (1 - alpha * tp);
    ev = ml.meta[m];
  } else {
    rho[j] = rho0 * (1 - alpha * tp);
    if (!iso) {
      ev = exp(-lncontrast * tp);
      ev = ev < emin ? emin : (ev > emax ? emax : ev);
    }
  }
  eta[j] = ev;
  hcond[j] = 1.0;
  hcapa[j] = 1.0;
  hprod[j] = 0.0;
  al[j] = alpha;
  rs[j] = rr;
  ts[j] = tt;
}

__global__ void k_centres(int nel, int nelx, const double* __restrict__ x_V,
                          const double* __restrict__ z_V, double* __restrict__ xe,
                          double* __restrict__ ze) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  const int n = grid_node(e, nelx, 2 * nelx + 1, 8);
  xe[e] = x_V[n];
  ze[e] = z_V[n];
}

// 'least_squares_P1' / '_Q1' values at the quadrature points (meeuuw.py:1516-1530,
// 1541-1555): a + b (xq - x_e) + c (zq - z_e) [+ d (xq - x_e)(zq - z_e)], for
// rho (which & 1) and eta (which & 2)
__global__ void k_ls_qpoints(int nel, int nb, int which, const double* __restrict__ cr,
                             const double* __restrict__ ce, const double* __restrict__ xq,
                             const double* __restrict__ zq, const double* __restrict__ xe,
                             const double* __restrict__ ze, double* __restrict__ rhoq,
                             double* __restrict__ etaq) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= 9L * nel) return;
  const long e = i / 9;
  const double dx = xq[i] - xe[e], dz = zq[i] - ze[e];
  for (int f = 0; f < 2; f++) {
    if (!(which & (1 << f))) continue;
    const double* c = (f ? ce : cr) + e * nb;
    double v = c[0] + c[1] * dx + c[2] * dz;
    if (nb == 4) v = v + c[3] * dx * dz;
    (f ? etaq : rhoq)[i] = v;
  }
}

// 'elemental' (meeuuw.py:1509-1511) or 'qpts' (meeuuw.py:1503-1508) projection
__global__ void k_qfields(int nel, int qpts, const double* __restrict__ rho_e,
                          const double* __restrict__ eta_e, const double* __restrict__ srho,
                          const double* __restrict__ seta, double* __restrict__ rhoq,
                          double* __restrict__ etaq) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= 9L * nel) return;
  if (qpts) {
    rhoq[i] = srho[i];
    etaq[i] = seta[i];
  } else {
    rhoq[i] = rho_e[i / 9];
    etaq[i] = eta_e[i / 9];
  }
}

__global__ void k_mean9(int nel, const double* __restrict__ q, double* __restrict__ e) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  double s = 0.0;
  for (int k = 0; k < 9; k++) s += q[i * 9L + k];
  e[i] = s / 9.0;
}

__global__ void k_const(long n, double v, double* __restrict__ a) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// max |v| over nodes, sums of rho, hcond, hcapa over all particles
__global__ void k_dt_partial(long nn, const double* __restrict__ u, int np,
                             const double* __restrict__ rho, const double* __restrict__ hc,
                             const double* __restrict__ hp, double* __restrict__ part) {
  __shared__ double sh[4][256];
  double vmax = 0.0, s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < nn; i += gridDim.x * 256L) {
    double v = sqrt(u[2 * i] * u[2 * i] + u[2 * i + 1] * u[2 * i + 1]);
    vmax = v > vmax ? v : vmax;
  }
  // the particle sums in the same order (i, i + stride, ...), U loads of each
  // field in flight at once (a missing term adds an exact 0)
  constexpr int U = 8;
  const long stride = gridDim.x * 256L;
  for (long i0 = blockIdx.x * 256L + threadIdx.x; i0 < np; i0 += U * stride) {
    double a[U], b[U], c[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const long i = i0 + u * stride;
      const bool ok = i < np;
      a[u] = ok ? rho[i] : 0.0;
      b[u] = ok ? hc[i] : 0.0;
      c[u] = ok ? hp[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      s0 += a[u];
      s1 += b[u];
      s2 += c[u];
    }
  }
  sh[0][threadIdx.x] = vmax;
  sh[1][threadIdx.x] = s0;
  sh[2][threadIdx.x] = s1;
  sh[3][threadIdx.x] = s2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sh[0][threadIdx.x] = fmax(sh[0][threadIdx.x], sh[0][threadIdx.x + w]);
      for (int k = 1; k < 4; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; k++) part[blockIdx.x * 4 + k] = sh[k][0];
}

// nb <= 256 block partials, staged in LDS by the whole block (one round trip
// instead of one per partial), then summed in block order by one thread
__global__ void k_dt_final(int nb, const double* __restrict__ part, double* __restrict__ out,
                           const double* __restrict__ poff, double* __restrict__ host) {
  __shared__ double sp[256 * 4];
  for (int i = threadIdx.x; i < nb * 4; i += blockDim.x) sp[i] = part[i];
  __syncthreads();
  // one thread per quantity, each summing (or taking the maximum) in block order
  const int k = threadIdx.x;
  if (k >= 4) return;
  double a = 0.0;
  if (k == 0)
    for (int b = 0; b < nb; b++) a = fmax(a, sp[b * 4]);
  else
    for (int b = 0; b < nb; b++) a += sp[b * 4 + k];
  out[k] = a;
  // pinned host mirror: [0] the pressure offset, [1..4] the above
  host[1 + k] = a;
  if (k == 0) host[0] = poff[0];
}

// the step's diagnostics into the pinned host mirror: 9 global quantities,
// 13 Nusselt/heat-flux values and the 8 failure flags
__global__ void k_step_readback(const double* __restrict__ gq, const double* __restrict__ nu,
                                const int* __restrict__ flags,
                                const unsigned long long* __restrict__ ee,
                                double* __restrict__ host) {
  const int t = threadIdx.x;
  if (t < 9) host[8 + t] = gq[t];
  if (t < 13) host[24 + t] = nu[t];
  if (t < 8) reinterpret_cast<int*>(host + 56)[t] = flags[t];
  if (t < 2) reinterpret_cast<unsigned long long*>(host + 60)[t] = ee[t];
}

__global__ void k_sub_scalar(long n, const double* __restrict__ s, double* __restrict__ a) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) a[i] -= s[0];
}

// Per-element particle counts: flags[4] += empty elements, flags[5] = min and
// flags[7] = max over elements (the reference's ptcl_stats, meeuuw.py:1198-1202),
// flags[6] += active particles (every active particle is counted in its
// element); ee[0], ee[1] = min and max of the positive eta_e as ordered bit
// patterns (meeuuw.py:1203).  256 threads per block; one atomic per block and flag.
__global__ __launch_bounds__(256) void k_count_check(int nel, const int* __restrict__ cnt,
                                                     const double* __restrict__ eta_e,
                                                     int* __restrict__ flags,
                                                     unsigned long long* __restrict__ ee) {
  __shared__ int smin[4], smax[4], ssum[4], semp[4];
  __shared__ unsigned long long slo[4], shi[4];
  // a grid-stride loop over few blocks: every block ends in atomics on the same
  // counters, which the memory system serialises
  int mn = 0x7fffffff, mx = 0, sm = 0, em = 0;
  unsigned long long lo = ~0ull, hi = 0ull;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nel; e += gridDim.x * blockDim.x) {
    const int c = cnt[e];
    mn = min(mn, c);
    mx = max(mx, c);
    sm += c;
    em += c == 0 ? 1 : 0;
    // positive doubles order as their bit patterns; a non-positive or NaN
    // value (an empty element's 0/0) is left out
    const double ev = eta_e[e];
    if (ev > 0.0) {
      const unsigned long long b = (unsigned long long)__double_as_longlong(ev);
      lo = b < lo ? b : lo;
      hi = b > hi ? b : hi;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    mn = min(mn, __shfl_xor(mn, o));
    mx = max(mx, __shfl_xor(mx, o));
    sm += __shfl_xor(sm, o);
    em += __shfl_xor(em, o);
    const unsigned long long l2 = __shfl_xor(lo, o), h2 = __shfl_xor(hi, o);
    lo = l2 < lo ? l2 : lo;
    hi = h2 > hi ? h2 : hi;
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    smin[wv] = mn; smax[wv] = mx; ssum[wv] = sm; semp[wv] = em; slo[wv] = lo; shi[wv] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; k++) {
      mn = min(mn, smin[k]); mx = max(mx, smax[k]);
      lo = slo[k] < lo ? slo[k] : lo; hi = shi[k] > hi ? shi[k] : hi;
    }
    mn = min(smin[0], mn);
    mx = max(smax[0], mx);
    lo = slo[0] < lo ? slo[0] : lo;
    hi = shi[0] > hi ? shi[0] : hi;
    sm = ssum[0] + ssum[1] + ssum[2] + ssum[3];
    em = semp[0] + semp[1] + semp[2] + semp[3];
    if (em) atomicAdd(&flags[4], em);
    atomicMin(&flags[5], mn);
    atomicAdd(&flags[6], sm);
    atomicMax(&flags[7], mx);
    atomicMin(&ee[0], lo);
    atomicMax(&ee[1], hi);
  }
}

// Step flags: all zero except the running minimum of flags[5]; the eta_e
// extrema start empty.
__global__ void k_flags_init(int* __restrict__ a, unsigned long long* __restrict__ ee) {
  const int i = threadIdx.x;
  if (i < 8) a[i] = (i == 5) ? 0x7fffffff : 0;
  if (i == 8) ee[0] = ~0ull;
  if (i == 9) ee[1] = 0ull;
}

// scalar helpers on device scalars
__global__ void k_div(const double* a, const double* b, double* c) { c[0] = a[0] / b[0]; }

__global__ void k_recip(long n, const double* __restrict__ a, double* __restrict__ b) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) b[i] = 1.0 / a[i];
}

__global__ void k_split_uv(long nn, const double* __restrict__ u, double* __restrict__ a,
                           double* __restrict__ b) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= nn) return;
  a[i] = u[2 * i];
  b[i] = u[2 * i + 1];
}

// ---------------------------------------------------------------- helpers --
static inline unsigned nb(long n) { return nblocks(n, 256); }

double Sim::host_dot(const double* a, const double* b, long n) {
  dot<double>(a, b, n, R.hbuf_dev + 40, R, st);  // the final reduction writes the pinned mirror
  (void)hipStreamSynchronize(st);
  return R.hbuf[40];
}

// squared norms of k <= 4 vectors with one synchronisation (each as host_dot)
int Sim::host_norms2(int k, const double* const* v, const long* n, double* out) {
  if (k < 1 || k > 4) return fail(MW_EINVAL, "host_norms2: 1..4 vectors");
  // the final reductions write the pinned mirror directly (no copy launch)
  for (int i = 0; i < k; i++) dot<double>(v[i], v[i], n[i], R.hbuf_dev + 50 + i, R, st);
  MW_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < k; i++) out[i] = R.hbuf[50 + i];
  return MW_OK;
}

int Sim::create(const mw_sim_config& c, hipStream_t s) {
  cfg = c;
  st = s;
  if (c.nelx < 2 || c.nelz < 2 || !(c.Lx > 0) || !(c.Lz > 0))
    return fail(MW_EINVAL, "sim: bad mesh size");
  if (!(c.rk == 1 || c.rk == 2 || c.rk == 4 || c.rk == -1))
    return fail(MW_EINVAL, "sim: rk must be 1, 2, 4 or -1");
  if (c.rk != -1 && c.ppe < 1) return fail(MW_EINVAL, "sim: ppe must be >= 1");
  if (c.material_mode < 0 || c.material_mode > 1 ||
      (c.material_mode == 1 && !(c.mat_eta[0] > 0 && c.mat_eta[1] > 0 && c.mat_rho[0] > 0 &&
                                 c.mat_rho[1] > 0)))
    return fail(MW_EINVAL, "sim: material_mode must be 0 or 1 (1: positive mat_eta, mat_rho)");
  if (!(c.gravity > 0) && !(c.alphaT != 0))
    return fail(MW_EINVAL, "sim: gravity <= 0 needs alphaT != 0 (g_z = -Ra / alphaT)");
  if (c.rheology < 0 || c.rheology > 1 ||
      (c.rheology == 1 && (c.tosi_case < 1 || c.tosi_case > 5 || c.material_mode != 0)))
    return fail(MW_EINVAL, "sim: rheology must be 0 or 1 (1: tosi_case 1..5, material_mode 0)");
  if (c.nonlinear && (c.niter_nl > 1000 || c.tol_nl < 0 || c.tol_nl != c.tol_nl))
    return fail(MW_EINVAL, "sim: niter_nl must be <= 1000 and tol_nl >= 0");
  if (c.averaging < 0 || c.averaging > 2 || c.nodal_projection_type < 1 ||
      c.nodal_projection_type > 4 || c.pressure_normalisation < 0 ||
      c.pressure_normalisation > 2 || c.solver < 0 || c.solver > 2)
    return fail(MW_EINVAL, "sim: bad option");
  nel = c.nelx * c.nelz;
  np = c.rk == -1 ? 9 * nel : nel * c.ppe;
  g.nelx = c.nelx; g.nelz = c.nelz; g.nnx = 2 * c.nelx + 1; g.nnz = 2 * c.nelz + 1;
  g.nn = (long)g.nnx * g.nnz;
  g.nnP = (long)(c.nelx + 1) * (c.nelz + 1);
  nn = g.nn; nnP = g.nnP; NV = 2 * nn;
  hx = c.Lx / c.nelx; hz = c.Lz / c.nelz;
  volume = c.Lx * c.Lz;
  L_ref = (c.Lx + c.Lz) / 2;  // meeuuw.py:343
  RC(R.init());
  RC(x_V.alloc(nn)); RC(z_V.alloc(nn));
  RC(icon_V.alloc(9L * nel)); RC(icon_P.alloc(4L * nel)); RC(l2g.alloc(18L * nel));
  RC(left.alloc(nel)); RC(right.alloc(nel)); RC(top.alloc(nel)); RC(bottom.alloc(nel));
  RC(N_V.alloc(81)); RC(dNdr_V.alloc(81)); RC(dNdt_V.alloc(81)); RC(N_P.alloc(36));
  RC(dNdr_P.alloc(36)); RC(dNdt_P.alloc(36)); RC(N_P_n.alloc(36)); RC(dNdr_n.alloc(81));
  RC(dNdt_n.alloc(81)); RC(qc.alloc(3)); RC(qw.alloc(3));
  for (DBuf<double>* b : {&JxWq, &j00q, &j01q, &j10q, &j11q, &j00n, &j01n, &j10n, &j11n, &gxq,
                          &gzq, &rhoq, &etaq, &Tq, &uq, &wq, &hcondq, &hcapaq, &hprodq, &alphaq,
                          &zq})
    RC(b->alloc(9L * nel));
  RC(fixV.alloc(NV)); RC(valV.alloc(NV)); RC(fixT.alloc(nn)); RC(valT.alloc(nn));
  k_mesh_nodes<<<nb(nn), 256, 0, st>>>(g.nnx, g.nnz, hx, hz, x_V, z_V);
  k_mesh_elems<<<nb(nel), 256, 0, st>>>(c.nelx, c.nelz, icon_V, icon_P, left, right, bottom, top);
  k_tables<<<1, 64, 0, st>>>(N_V, dNdr_V, dNdt_V, N_P, dNdr_P, dNdt_P, N_P_n, dNdr_n, dNdt_n, qc,
                             qw);
  k_jac<<<nb(9L * nel), 256, 0, st>>>(nel, dNdr_V, dNdt_V, qw, 1, x_V, z_V, icon_V, JxWq, j00q,
                                      j01q, j10q, j11q);
  k_jac<<<nb(9L * nel), 256, 0, st>>>(nel, dNdr_n, dNdt_n, qw, 0, x_V, z_V, icon_V, nullptr, j00n,
                                      j01n, j10n, j11n);
  k_bc<<<nb(nn), 256, 0, st>>>(nn, x_V, z_V, c.Lx, c.Lz, c.Tbottom, c.Ttop, fixV, valV, fixT, valT);
  k_const<<<nb(9L * nel), 256, 0, st>>>(9L * nel, 0.0, gxq);
  // g_z = -Ra / alphaT (experiment0.py:274), or -gravity where it is given
  k_const<<<nb(9L * nel), 256, 0, st>>>(9L * nel, c.gravity > 0 ? -c.gravity : -c.Ra / c.alphaT,
                                        gzq);
  k_const<<<nb(9L * nel), 256, 0, st>>>(9L * nel, 0.0, zq);
  MW_CHECK_LAUNCH();
  RC(local_to_global(nel, icon_V, l2g, st));
  // incidence lists: dofs of l2g, pressure nodes, velocity nodes
  {
    size_t wb = std::max(std::max(incidence_temp_bytes((int)NV), incidence_temp_bytes((int)nnP)),
                         incidence_temp_bytes((int)nn)) +
                sizeof(int) * 2 * (NV + 1) + 512;
    DBuf<double> w;
    RC(w.alloc(wb / 8 + 1));
    RC(incv_ptr.alloc(NV + 1)); RC(incv_list.alloc(18L * nel));
    RC(incp_ptr.alloc(nnP + 1)); RC(incp_list.alloc(4L * nel));
    RC(incn_ptr.alloc(nn + 1)); RC(incn_list.alloc(9L * nel));
    RC(incidence_build(nel, 18, l2g, (int)NV, incv_ptr, incv_list, w.p, st));
    RC(incidence_build(nel, 4, icon_P, (int)nnP, incp_ptr, incp_list, w.p, st));
    RC(incidence_build(nel, 9, icon_V, (int)nn, incn_ptr, incn_list, w.p, st));
  }
  // swarm and fields
  for (DBuf<double>* b : {&sx, &sz, &sr, &st_, &sr_s, &st_s, &srho, &seta, &shcond, &shcapa,
                          &shprod, &salpha, &su, &sw})
    RC(b->alloc(np));
  RC(siel.alloc(np)); RC(order.alloc(np)); RC(seg.alloc(nel + 1)); RC(sact.alloc(np));
  RC(smat.alloc(np));
  MW_HIP(hipMemsetAsync(smat, 0, np, st));
  RC(sortwork.alloc(sort_work_bytes(np, nel) / 8 + 64));
  for (DBuf<double>* b : {&T, &q, &rho_n, &eta_n, &hcond_n, &hcapa_n, &hprod_n, &alpha_n,
                          &dTdx, &dTdz, &qxn, &qzn})
    RC(b->alloc(nn));
  RC(u.alloc(NV)); RC(p.alloc(nnP)); RC(rho_e.alloc(nel)); RC(eta_e.alloc(nel));
  for (DBuf<double>* b : {&exxn, &ezzn, &exzn}) {  // zero before the first Stokes solve
    RC(b->alloc(nn));
    MW_HIP(hipMemsetAsync(b->p, 0, nn * 8, st));
  }
  RC(exxq.alloc(9L * nel)); RC(ezzq.alloc(9L * nel)); RC(exzq.alloc(9L * nel));
  if (c.rho_projection < 0 || c.rho_projection > 3 || c.eta_projection < 0 ||
      c.eta_projection > 3 || (c.rk == -1 && (c.rho_projection || c.eta_projection)) ||
      (c.rho_projection >= 2 && c.eta_projection >= 2 && c.rho_projection != c.eta_projection))
    return fail(MW_EINVAL, "sim: rho/eta_projection must be 0..3 (rk > 0), least squares of one kind");
  if (c.population_control) {
    if (c.rk <= 0) return fail(MW_EINVAL, "sim: population control needs moving particles (rk > 0)");
    RC(pc_need.alloc(nel)); RC(pc_offs.alloc(nel + 1)); RC(pc_flag.alloc(np));
    RC(pc_pos.alloc(np + 1)); RC(pc_avail.alloc(np)); RC(pc_e.alloc(np)); RC(pc_tmpl.alloc(np));
    RC(pc_tsum.alloc(scan_tile_sums_ints(std::max(np, nel)))); RC(pc_bad.alloc(1));
    RC(pc_x.alloc(np)); RC(pc_z.alloc(np)); RC(pc_r.alloc(np)); RC(pc_t.alloc(np));
  }
  if (c.nonlinear) {  // u_mem, w_mem, p_mem start at zero (meeuuw.py:1095-1097)
    RC(Tm.alloc(nn)); RC(u_mem.alloc(NV)); RC(p_mem.alloc(nnP)); RC(nlsum.alloc(8 + 8 * 256));
    MW_HIP(hipMemsetAsync(u_mem.p, 0, NV * 8, st));
    MW_HIP(hipMemsetAsync(p_mem.p, 0, nnP * 8, st));
  }
  if (c.extrapolate < 0 || c.extrapolate > 2) return fail(MW_EINVAL, "sim: extrapolate must be 0, 1 or 2");
  if (c.warm_start && c.extrapolate) { RC(u_prev.alloc(NV)); RC(p_prev.alloc(nnP)); }
  if (c.warm_start && c.extrapolate == 2) { RC(u_prev2.alloc(NV)); RC(p_prev2.alloc(nnP)); }
  RC(eta_mg.alloc(nel));
  RC(cnt.alloc(nel));
  if (c.rho_projection >= 2 || c.eta_projection >= 2) {
    // x_e, z_e: the centre node (compute_center_coordinates.py:17-19); xq, zq
    // as meeuuw.py:796-797 forms them (the Q2 interpolation of the node coordinates)
    RC(xe.alloc(nel)); RC(ze.alloc(nel)); RC(xqc.alloc(9L * nel)); RC(zqc.alloc(9L * nel));
    RC(lscr.alloc(4L * nel)); RC(lsce.alloc(4L * nel)); RC(lsmm.alloc(4L * nel));
    k_centres<<<nb(nel), 256, 0, st>>>(nel, c.nelx, x_V, z_V, xe, ze);
    const double* phi[2] = {x_V, z_V};
    double* out[2] = {xqc, zqc};
    RC(nodal_to_qpoints(2, phi, nel, 9, N_V, icon_V, out, st, c.nelx));
  }
  // the assembled Stokes values are formed only for the FP64 Uzawa solver; the
  // Krylov solvers evaluate the FP64 residual matrix-free
  if (c.solver == 0) RC(VV.alloc(468L * nel));
  RC(rhs_f.alloc(NV)); RC(rhs_h.alloc(nnP)); RC(r64tab.alloc(207));
  RC(residual64_tables(dNdr_V, dNdt_V, N_P, j00q, j01q, j10q, j11q, JxWq, r64tab, st));
  if (c.energy_mf) {
    // matrix-free energy operator: 36 coefficients per element instead of 81 values
    RC(etab.alloc(252)); RC(ecoef.alloc(36L * nel)); RC(edT.alloc(nn));
    RC(epart.alloc(energy_mf_partials(g))); RC(escal.alloc(8));
    RC(energy_mf_tables(N_V, dNdr_V, dNdt_V, j00q, j01q, j10q, j11q, JxWq, etab, st));
    MW_HIP(hipHostMalloc((void**)&ehost, 4 * sizeof(double), hipHostMallocDefault));
  } else {
    RC(VVT.alloc(81L * nel));
  }
  RC(rhsT.alloc(nn)); RC(awork.alloc(22L * nel));
  if (c.energy_mf) RC(Tprev.alloc(nn));
  RC(nodework.alloc(std::max(48L * nel, 1024L * 24)));
  RC(un.alloc(nn)); RC(wn.alloc(nn));
  RC(Gh.alloc(72)); RC(mdiag.alloc(nnP)); RC(minv.alloc(nnP)); RC(invdT.alloc(nn));
  for (auto& b : wu) RC(b.alloc(NV));
  for (auto& b : wp) RC(b.alloc(nnP));
  for (auto& b : wT) RC(b.alloc(nn));
  if (c.energy_mf) {  // every array it points to is allocated above
    emf.g = g; emf.fix = fixT; emf.tab = etab; emf.coef = ecoef; emf.d = edT; emf.invd = invdT;
    emf.part = epart; emf.S = escal; emf.host = ehost;
    MW_HIP(hipHostGetDevicePointer((void**)&emf.host_dev, ehost, 0));
  }
  RC(flags.alloc(8)); RC(eebits.alloc(2)); RC(dscal.alloc(64)); RC(hcol.alloc(2 * FGMRES_MAXM));
  RC(hzero.alloc(FGMRES_MAXM + 1));
  fill<double>(FGMRES_MAXM + 1, 0.0, hzero, st);
  if (c.gs_passes < 0 || c.gs_passes > 2) return fail(MW_EINVAL, "sim: gs_passes must be 0, 1 or 2");
  gs_once = c.gs_passes == 1 || (c.gs_passes == 0 && nel >= 192 * 192);
  ghat_compute(hx, hz, 1.0 / L_ref, Gh, st);  // eta_ref = 1
  k_init_T<<<nb(nn), 256, 0, st>>>(nn, x_V, z_V, c.Lx, c.Lz, c.Tbottom, c.Ttop, T);
  if (c.rk == -1)
    k_place_qpts<<<nb(np), 256, 0, st>>>(nel, qc, N_V, x_V, z_V, icon_V, sx, sz, sr, st_, siel,
                                          sact);
  swarm_set = (c.rk == -1);
  fill<double>(NV, 0.0, u, st);
  fill<double>(nnP, 0.0, p, st);
  MW_CHECK_LAUNCH();
  gmg64.fuse_tail = gmg32.fuse_tail = c.gmg_tail != 0;
  if (c.gmg_bottom < 0 || c.gmg_bottom > 3) return fail(MW_EINVAL, "sim: gmg_bottom must be in [0, 3]");
  gmg64.allow_bottom = gmg32.allow_bottom = c.gmg_bottom;
  if (c.gmg_power_its < 0 || c.gmg_power_its > 64)
    return fail(MW_EINVAL, "sim: gmg_power_its must be in [0, 64]");
  gmg64.power_its_warm = gmg32.power_its_warm = c.gmg_power_its;
  if (c.gmg_eta_coarsen < 0 || c.gmg_eta_coarsen > 2)
    return fail(MW_EINVAL, "sim: gmg_eta_coarsen must be 0, 1 or 2");
  gmg64.eta_coarsen = gmg32.eta_coarsen = c.gmg_eta_coarsen;
  if (c.gmg_coarse_dofs > 2304) return fail(MW_EINVAL, "sim: gmg_coarse_dofs must be <= 2304");
  gmg64.coarse_max_dof = gmg32.coarse_max_dof = c.gmg_coarse_dofs > 0 ? c.gmg_coarse_dofs : 60;
  if (c.solver == 0) {
    gmg64.nsmooth = c.cheb_degree;
    gmg64.nsmooth_coarse = c.cheb_degree_coarse;
    RC(gmg64.setup(c.nelx, c.nelz, c.Lx, c.Lz, fixV, st));
  } else {
    gmg32.nsmooth = c.cheb_degree;
    gmg32.nsmooth_coarse = c.cheb_degree_coarse;
    gmg32.solver_pumps = true;  // FGMRES queues the coarsest inverse's sweeps
    RC(gmg32.setup(c.nelx, c.nelz, c.Lx, c.Lz, fixV, st));
    int m = c.fgmres_restart;
    if (m < 2 || m > FGMRES_MAXM) return fail(MW_EINVAL, "sim: fgmres_restart must be in [2, 64]");
    if (c.pc_vcycles < 1 || c.pc_vcycles > 8) return fail(MW_EINVAL, "sim: pc_vcycles must be in [1, 8]");
    long N = NV + nnP;
    RC(V32.alloc((size_t)(m + 1) * N)); RC(Z32.alloc((size_t)m * N));
    RC(f32a.alloc(N)); RC(f32b.alloc(N)); RC(f32c.alloc(N)); RC(f32d.alloc(N));
    RC(r32.alloc(N)); RC(x32.alloc(N));
    if (c.schur_cheb == 1 || c.schur_cheb < 0 || c.schur_cheb > 32)
      return fail(MW_EINVAL, "sim: schur_cheb must be 0 or in [2, 32]");
    if (c.schur_cheb >= 2) {
      RC(pm_xa.alloc(nnP)); RC(pm_xb.alloc(nnP)); RC(pm_d.alloc(nnP));
      RC(pm_w.alloc(nel)); RC(pm_dg.alloc(nnP)); RC(pm_dinv.alloc(nnP));
    }
    RC(gm.alloc((size_t)(m + 1) * m + 5 * m + 16));
    MW_HIP(hipHostMalloc((void**)&hest, 2 * FGMRES_MAXM * sizeof(double), hipHostMallocDefault));
    MW_HIP(hipHostGetDevicePointer((void**)&hest_dev, hest, 0));
    for (auto& e : ev) MW_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (c.solver == 2) RC(learn_setup());
    if (c.fgmres_graphs) RC(build_iteration_graphs(m));  // after hest: Arnoldi steps write it
  }
  if (!ev[0])
    for (auto& e : ev) MW_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  MW_HIP(hipStreamSynchronize(st));
  return MW_OK;
}

int Sim::set_temperature(const double* Th) {
  tm_set = false;  // T_mem = T.copy() after the initial temperature (meeuuw.py:708, 719)
  have_tprev = false;
  MW_HIP(hipMemcpyAsync(T, Th, nn * sizeof(double), hipMemcpyHostToDevice, st));
  MW_HIP(hipStreamSynchronize(st));
  return MW_OK;
}

int Sim::set_swarm_reduced(const double* rh, const double* th) {
  if (cfg.rk == -1) return fail(MW_EINVAL, "sim: particles are fixed on the qpoints (rk = -1)");
  DBuf<double> r, t;
  RC(r.alloc(np)); RC(t.alloc(np));
  MW_HIP(hipMemcpyAsync(r, rh, np * sizeof(double), hipMemcpyHostToDevice, st));
  MW_HIP(hipMemcpyAsync(t, th, np * sizeof(double), hipMemcpyHostToDevice, st));
  k_place<<<nb(np), 256, 0, st>>>(nel, cfg.ppe, r, t, x_V, z_V, icon_V, sx, sz, sr, st_, siel, sact);
  MW_CHECK_LAUNCH();
  swarm_set = true;
  MW_HIP(hipStreamSynchronize(st));
  return MW_OK;
}

int Sim::set_swarm_material(const uint8_t* mh) {
  MW_HIP(hipMemcpyAsync(smat, mh, np, hipMemcpyHostToDevice, st));
  MW_HIP(hipStreamSynchronize(st));
  return MW_OK;
}

int Sim::set_swarm(long n, const double* xh, const double* zh, const double* rh, const double* th,
                   const int* ielh, const uint8_t* acth) {
  if (n != np) return fail(MW_EINVAL, "sim: particle count does not match nel*ppe");
  for (long i = 0; i < n; i++)  // the element of every active particle indexes the mesh
    if (acth[i] && (ielh[i] < 0 || ielh[i] >= nel))
      return fail(MW_EINVAL, "sim_set_swarm: active particle " + std::to_string(i) +
                                 " has an element index outside [0, nel)");
  MW_HIP(hipMemcpyAsync(sx, xh, n * 8, hipMemcpyHostToDevice, st));
  MW_HIP(hipMemcpyAsync(sz, zh, n * 8, hipMemcpyHostToDevice, st));
  MW_HIP(hipMemcpyAsync(sr, rh, n * 8, hipMemcpyHostToDevice, st));
  MW_HIP(hipMemcpyAsync(st_, th, n * 8, hipMemcpyHostToDevice, st));
  MW_HIP(hipMemcpyAsync(siel, ielh, n * 4, hipMemcpyHostToDevice, st));
  MW_HIP(hipMemcpyAsync(sact, acth, n, hipMemcpyHostToDevice, st));
  MW_HIP(hipStreamSynchronize(st));
  swarm_set = true;
  return MW_OK;
}

// ------------------------------------------------------------ Stokes -------
// full residual of the assembled system: ru = f - K u - G p, rp = h - G^T u
// ru, rp = f - (K u + G p), h - G^T u; with bnrm: also |f, h| (one synchronisation)
int Sim::stokes_residual(double* ru, double* rp, double* nrm2, double* bnrm) {
  if (cfg.solver == 0) {
    stokes_apply_vv(g, VV, u, p, ru, rp, st);
    axpby<double>(NV, 1.0, nullptr, rhs_f, -1.0, nullptr, ru, st);
    axpby<double>(nnP, 1.0, nullptr, rhs_h, -1.0, nullptr, rp, st);
  } else {
    RC(stokes_residual64(g, fixV, r64tab, etaq, u, p, rhs_f, rhs_h, 1.0 / L_ref, ru, rp, st));
  }
  const double* v[4] = {ru, rp, rhs_f, rhs_h};
  const long n[4] = {NV, nnP, NV, nnP};
  double q[4];
  RC(host_norms2(bnrm ? 4 : 2, v, n, q));
  *nrm2 = q[0] + q[1];
  if (bnrm) *bnrm = sqrt(q[2] + q[3]);
  return MW_OK;
}

// CG for K x = b (assembled K from VV), one FP64 V-cycle as preconditioner
int Sim::pcg_K(const double* b, double* x, double rtol, int maxit, int* its) {
  double *r = wu[2], *z = wu[3], *pv = wu[4], *qv = wu[5];
  double* S = dscal.p;  // 0 rz, 1 pq, 2 alpha, 3 rz_new, 4 beta
  stokes_apply_vv(g, VV, x, nullptr, qv, nullptr, st);
  axpby<double>(NV, 1.0, nullptr, b, -1.0, nullptr, qv, st);  // qv = b - Kx
  MW_HIP(hipMemcpyAsync(r, qv, NV * 8, hipMemcpyDeviceToDevice, st));
  double bn = sqrt(host_dot(b, b, NV));
  if (bn == 0.0) {
    fill<double>(NV, 0.0, x, st);
    *its = 0;
    return MW_OK;
  }
  RC(gmg64.vcycle(r, z, st));
  int nvc = 1;
  MW_HIP(hipMemcpyAsync(pv, z, NV * 8, hipMemcpyDeviceToDevice, st));
  dot<double>(r, z, NV, S + 0, R, st);
  int k = 0;
  for (; k < maxit; k++) {
    stokes_apply_vv(g, VV, pv, nullptr, qv, nullptr, st);
    dot<double>(pv, qv, NV, S + 1, R, st);
    k_div<<<1, 1, 0, st>>>(S + 0, S + 1, S + 2);
    axpby<double>(NV, 1.0, S + 2, pv, 1.0, nullptr, x, st);
    axpby<double>(NV, -1.0, S + 2, qv, 1.0, nullptr, r, st);
    double rn = sqrt(host_dot(r, r, NV));
    if (rn <= rtol * bn) break;
    RC(gmg64.vcycle(r, z, st));
    nvc++;
    dot<double>(r, z, NV, S + 3, R, st);
    k_div<<<1, 1, 0, st>>>(S + 3, S + 0, S + 4);
    axpby<double>(NV, 1.0, nullptr, z, 1.0, S + 4, pv, st);
    MW_HIP(hipMemcpyAsync(S + 0, S + 3, 8, hipMemcpyDeviceToDevice, st));
  }
  *its = nvc;
  if (k == maxit) return fail(MW_ENOCONV, "stokes: inner K solve did not converge");
  return MW_OK;
}

// x <- x + a (x - prev), prev <- x (old value)
__global__ void k_extrapolate(long n, double a, double* __restrict__ x, double* __restrict__ prev) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double t = x[i];
  x[i] = t + a * (t - prev[i]);
  prev[i] = t;
}

// x <- l2 x + l1 prev + l0 prev2, prev2 <- prev, prev <- x (old values)
__global__ void k_extrapolate2(long n, double l0, double l1, double l2, double* __restrict__ x,
                               double* __restrict__ prev, double* __restrict__ prev2) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double t = x[i], a = prev[i];
  x[i] = l2 * t + l1 * a + l0 * prev2[i];
  prev2[i] = a;
  prev[i] = t;
}

// Initial guess of a Stokes solve: zero (cold), the last solution (warm), or
// its extrapolation in time: linear from the last two solutions, u_n +
// (dt_n / dt_{n-1}) (u_n - u_{n-1}), or quadratic (Lagrange) from the last
// three.  The solve still stops on the same residual; the guess only changes
// where it starts.
int Sim::initial_guess() {
  if (nl_keep_guess && cfg.warm_start && warm) return MW_OK;  // the last Picard iterate
  if (!(cfg.warm_start && warm)) {
    fill<double>(NV, 0.0, u, st);
    fill<double>(nnP, 0.0, p, st);
    have_prev = 0;
    return MW_OK;
  }
  if (!cfg.extrapolate) return MW_OK;
  const double h1 = dt_hist[0], h2 = dt_hist[1], h3 = dt_hist[2];  // t_{n-1}-t_{n-2}, .., t_{n+1}-t_n
  if (cfg.extrapolate == 2 && have_prev >= 2 && h1 > 0.0 && h2 > 0.0) {
    const double l0 = (h2 + h3) * h3 / (h1 * (h1 + h2));
    const double l1 = -(h1 + h2 + h3) * h3 / (h1 * h2);
    const double l2 = (h1 + h2 + h3) * (h2 + h3) / ((h1 + h2) * h2);
    k_extrapolate2<<<nb(NV), 256, 0, st>>>(NV, l0, l1, l2, u, u_prev, u_prev2);
    k_extrapolate2<<<nb(nnP), 256, 0, st>>>(nnP, l0, l1, l2, p, p_prev, p_prev2);
  } else if (have_prev >= 1 && h2 > 0.0) {
    if (cfg.extrapolate == 2) {
      MW_HIP(hipMemcpyAsync(u_prev2, u_prev, NV * 8, hipMemcpyDeviceToDevice, st));
      MW_HIP(hipMemcpyAsync(p_prev2, p_prev, nnP * 8, hipMemcpyDeviceToDevice, st));
    }
    const double a = h3 / h2;
    k_extrapolate<<<nb(NV), 256, 0, st>>>(NV, a, u, u_prev);
    k_extrapolate<<<nb(nnP), 256, 0, st>>>(nnP, a, p, p_prev);
    have_prev = 2;
  } else {
    MW_HIP(hipMemcpyAsync(u_prev, u, NV * 8, hipMemcpyDeviceToDevice, st));
    MW_HIP(hipMemcpyAsync(p_prev, p, nnP * 8, hipMemcpyDeviceToDevice, st));
    have_prev = 1;
  }
  MW_CHECK_LAUNCH();
  return MW_OK;
}

int Sim::stokes_solve_plain() {
  const double tol = cfg.stokes_tol;
  RC(initial_guess());
  RC(gmg64.update(eta_mg, R, st));
  pressure_mass_eta_lumped(g, hx * hz, eta_mg, mdiag, st);
  k_recip<<<nb(nnP), 256, 0, st>>>(nnP, mdiag, minv);
  double bnrm = sqrt(host_dot(rhs_f, rhs_f, NV) + host_dot(rhs_h, rhs_h, nnP));
  const double rtol_in = 1e-11;
  const int pcg_maxit = cfg.pcg_maxit > 0 ? cfg.pcg_maxit : 500;
  int its = 0, vc = 0, outer = 0;
  history.clear();
  double *t = wu[0], *w = wu[1];
  double *rho = wp[0], *z = wp[1], *d = wp[2], *Sd = wp[3];
  double* S = dscal.p + 8;  // 0 rz, 1 dSd, 2 alpha, 3 rz_new, 4 beta
  // u = K^-1 (f - G p)
  stokes_apply_vv(g, VV, nullptr, p, t, nullptr, st);
  axpby<double>(NV, 1.0, nullptr, rhs_f, -1.0, nullptr, t, st);
  RC(pcg_K(t, u, rtol_in, pcg_maxit, &its));
  vc += its;
  // rho = G^T u - h
  stokes_apply_vv(g, VV, u, nullptr, nullptr, rho, st);
  axpby<double>(nnP, -1.0, nullptr, rhs_h, 1.0, nullptr, rho, st);
  diag_scale<double>(nnP, 1.0, minv, rho, z, st);
  MW_HIP(hipMemcpyAsync(d, z, nnP * 8, hipMemcpyDeviceToDevice, st));
  dot<double>(rho, z, nnP, S + 0, R, st);
  double res2 = 0.0;
  for (outer = 0; outer < 500; outer++) {
    double rn = sqrt(host_dot(rho, rho, nnP));
    if (!(rn == rn)) break;  // NaN: reported as no convergence below
    if (rn <= 0.3 * tol * bnrm) {
      RC(stokes_residual(wu[6], wp[4], &res2));
      if (sqrt(res2) <= tol * bnrm) break;
      // velocity block not converged enough: correct u with one more K solve
      fill<double>(NV, 0.0, w, st);
      RC(pcg_K(wu[6], w, rtol_in, pcg_maxit, &its));
      vc += its;
      axpby<double>(NV, 1.0, nullptr, w, 1.0, nullptr, u, st);
      stokes_apply_vv(g, VV, u, nullptr, nullptr, rho, st);
      axpby<double>(nnP, -1.0, nullptr, rhs_h, 1.0, nullptr, rho, st);
      RC(stokes_residual(wu[6], wp[4], &res2));
      if (sqrt(res2) <= tol * bnrm) break;
    }
    stokes_apply_vv(g, VV, nullptr, d, t, nullptr, st);  // t = G d
    fill<double>(NV, 0.0, w, st);
    RC(pcg_K(t, w, rtol_in, pcg_maxit, &its));
    vc += its;
    stokes_apply_vv(g, VV, w, nullptr, nullptr, Sd, st);  // Sd = G^T w
    dot<double>(d, Sd, nnP, S + 1, R, st);
    k_div<<<1, 1, 0, st>>>(S + 0, S + 1, S + 2);
    axpby<double>(nnP, 1.0, S + 2, d, 1.0, nullptr, p, st);
    axpby<double>(NV, -1.0, S + 2, w, 1.0, nullptr, u, st);
    axpby<double>(nnP, -1.0, S + 2, Sd, 1.0, nullptr, rho, st);
    diag_scale<double>(nnP, 1.0, minv, rho, z, st);
    dot<double>(rho, z, nnP, S + 3, R, st);
    k_div<<<1, 1, 0, st>>>(S + 3, S + 0, S + 4);
    axpby<double>(nnP, 1.0, nullptr, z, 1.0, S + 4, d, st);
    MW_HIP(hipMemcpyAsync(S + 0, S + 3, 8, hipMemcpyDeviceToDevice, st));
  }
  if (res2 == 0.0) RC(stokes_residual(wu[6], wp[4], &res2));
  stats[MW_STAT_STOKES_OUTER] = outer;
  stats[MW_STAT_STOKES_VCYCLES] = vc;
  stats[MW_STAT_STOKES_RELRES] = sqrt(res2) / bnrm;
  warm = true;
  if (!(sqrt(res2) <= tol * bnrm)) return fail(MW_ENOCONV, "stokes (plain): no convergence");
  return MW_OK;
}

// ----------------------------------------------------- mixed precision -----
// y = A32 z: [K z_u + G z_p ; G^T z_u] (matrix-free templates, FP32)
void Sim::apply32(const float* z, float* y) {
  stokes_apply_fine<float>(*gmg32.lev[0], Gh, z, y, NV, st);
}

// z = P^-1 v, P = [K G; 0 -S], S ~ lumped M/eta; K^-1 ~ pc_vcycles V-cycles
void Sim::prec32(const float* v, float* z) {
  float* t = f32c.p;
  if (learned) {
    schur_learned(v + NV, z + NV);  // z_p = -(learned S^-1) v_p
    gtmpl_apply_u<float>(g, Gh, fixV, z + NV, t, false, st);  // t = G z_p
    axpby<float>(NV, 1.0, nullptr, v, -1.0, nullptr, t, st);  // t = v_u - G z_p
  } else if (cfg.schur_cheb >= 2 && cfg.schur_cheb <= SCHUR_FUSED_MAX && !split_schur) {
    (void)schur_front_cheb<float>(*gmg32.lev[0], Gh, pm_w, pm_dg, pm_dinv, v, z + NV, t, NV,
                                  cfg.schur_cheb, cfg.schur_scale, st);  // one launch
  } else if (cfg.schur_cheb >= 2) {  // S^-1 ~ Chebyshev on the consistent mass matrix
    float* xs = nullptr;
    (void)pmass_cheb_solve<float>(cfg.nelx, cfg.nelz, pm_w, pm_dg, pm_dinv, v + NV, pm_xa, pm_xb,
                                  pm_d, cfg.schur_cheb, &xs, st);
    schur_front<float>(*gmg32.lev[0], Gh, minv, v, z + NV, t, NV, st, xs, cfg.schur_scale);
  } else {
    schur_front<float>(*gmg32.lev[0], Gh, minv, v, z + NV, t, NV, st, nullptr,
                       cfg.schur_scale);  // same, one launch
  }
  (void)gmg32.vcycle(t, z, st);
  for (int k = 1; k < cfg.pc_vcycles; k++) {  // z_u += V(t - K z_u)
    float *s = f32a.p, *q = f32d.p;
    level_apply<float>(*gmg32.lev[0], z, s, st);
    axpby<float>(NV, 1.0, nullptr, t, -1.0, nullptr, s, st);
    (void)gmg32.vcycle(s, q, st);
    axpby<float>(NV, 1.0, nullptr, q, 1.0, nullptr, z, st);
  }
}

__global__ void k_gs_update(long n, int k, const float* __restrict__ V, long ld,
                            const double* __restrict__ h, float* __restrict__ w) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = (double)w[i];
  for (int j = 0; j < k; j++) s -= h[j] * (double)V[j * ld + i];
  w[i] = (float)s;
}

// Device state of one FGMRES(m) cycle, laid out in Sim::gm:
// H (m+1) x m row-major | cs m | sn m | g m+1 | y m | est m | scalars 16
struct GmView {
  double *H, *cs, *sn, *g, *y, *est, *sc;  // sc[0] ||w||^2, sc[1] beta^2, sc[2] 1/||w||, sc[3] 1/beta
  GmView(double* b, int m) {
    H = b; cs = H + (m + 1) * m; sn = cs + m; g = sn + m; y = g + m + 1; est = y + m; sc = est + m;
  }
};

__global__ void k_fgmres_start(int m, GmView v) {
  if (threadIdx.x != 0) return;
  double beta = sqrt(v.sc[1]);
  v.g[0] = beta;
  for (int i = 1; i <= m; i++) v.g[i] = 0.0;
  v.sc[3] = beta > 0.0 ? 1.0 / beta : 0.0;
}

// Column j of the Hessenberg matrix from the two Gram-Schmidt passes and
// ||w||^2, the previous Givens rotations applied to it, a new rotation that
// zeroes H[j+1][j], and the updated least-squares right-hand side.  With
// pyth, ||w||^2 is that of the once-projected w (h2[j+1]) less |h2|^2 (the
// second projection removes exactly that much from an orthonormal basis), so
// the last Gram-Schmidt sweep needs no reduction and can write the scaled
// basis vector directly.  The residual estimate is also stored straight into
// the host's pinned mirror (est_host, a device view of it), which the host
// reads after the step's event.  With partial (pyth only), h2 and the norm
// are first summed from the block partials of the projection sweep (nblk
// blocks, j + 2 interleaved columns), as its final reduction would, and
// stored to h2.
__global__ void k_arnoldi(int j, int m, const double* __restrict__ h1, double* __restrict__ h2,
                          GmView v, double* est_host, int pyth, const double* __restrict__ partial,
                          int nblk) {
  // the inputs are staged in LDS by the whole block first, so the serial
  // part below waits on LDS, not on one memory round trip per rotation
  __shared__ double sh1[FGMRES_MAXM + 1], sh2[FGMRES_MAXM + 1], scs[FGMRES_MAXM],
      ssn[FGMRES_MAXM], col[FGMRES_MAXM + 1];
  if (partial) {
    const int kc = j + 2, lane = threadIdx.x & 63;
    for (int c = threadIdx.x >> 6; c < kc; c += blockDim.x >> 6) {
      const double s = reduce_column(partial, nblk, kc, c, lane);
      if (lane == 0) {
        sh2[c] = s;
        h2[c] = s;
      }
    }
  }
  for (int i = threadIdx.x; i <= j + 1; i += blockDim.x) {
    sh1[i] = i <= j ? h1[i] : 0.0;
    if (!partial) sh2[i] = (i <= j || pyth) ? h2[i] : 0.0;  // h2[j + 1]: the projected norm (pyth only)
  }
  for (int i = threadIdx.x; i < j; i += blockDim.x) {
    scs[i] = v.cs[i];
    ssn[i] = v.sn[i];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int i = 0; i <= j; i++) col[i] = sh1[i] + sh2[i];
  double hn2 = v.sc[0];
  if (pyth) {
    hn2 = sh2[j + 1];
    for (int i = 0; i <= j; i++) hn2 -= sh2[i] * sh2[i];
  }
  double hn = sqrt(hn2 > 0.0 ? hn2 : 0.0);
  for (int i = 0; i < j; i++) {
    double a = col[i], b = col[i + 1];
    col[i] = scs[i] * a + ssn[i] * b;
    col[i + 1] = -ssn[i] * a + scs[i] * b;
  }
  double a = col[j], b = hn;
  double rr = sqrt(a * a + b * b);
  const double c = rr > 0 ? a / rr : 1.0, sn = rr > 0 ? b / rr : 0.0;
  v.cs[j] = c;
  v.sn[j] = sn;
  for (int i = 0; i < j; i++) v.H[i * m + j] = col[i];
  v.H[j * m + j] = rr;
  v.H[(j + 1) * m + j] = 0.0;
  const double gj = v.g[j];
  v.g[j + 1] = -sn * gj;
  v.g[j] = c * gj;
  v.est[j] = fabs(-sn * gj);
  est_host[j] = v.est[j];
  v.sc[2] = hn > 0.0 ? 1.0 / hn : 0.0;
}

// x = Z y with y = H(0:k,0:k)^-1 g(0:k) (H upper triangular after the
// rotations; back substitution from the last row, each y[i] from the rows
// above it in increasing column order).  Every block solves the small system
// itself from H and g staged in LDS (one launch instead of a one-thread solve
// followed by the product); all blocks repeat the same sequential arithmetic,
// so they agree on y, which block 0 also stores.
__global__ void __launch_bounds__(256) k_backsub_lincomb(long n, int k, int m, GmView v,
                                                         const float* __restrict__ Z, long ld,
                                                         float* __restrict__ x) {
  __shared__ double sH[FGMRES_MAXM * (FGMRES_MAXM + 1) / 2 + FGMRES_MAXM];
  __shared__ double sy[FGMRES_MAXM];
  double* sg = sH + FGMRES_MAXM * (FGMRES_MAXM + 1) / 2;
  // upper triangle of H(0:k,0:k), row i from column i, packed
  for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
    const int i = e / k, l = e % k;
    if (l >= i) sH[i * k - i * (i - 1) / 2 + (l - i)] = v.H[i * m + l];
  }
  for (int i = threadIdx.x; i < k; i += blockDim.x) sg[i] = v.g[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = k - 1; i >= 0; i--) {
      const double* row = sH + i * k - i * (i - 1) / 2;  // row[l - i] = H[i][l]
      double s = sg[i];
      for (int l = i + 1; l < k; l++) s -= row[l - i] * sy[l];
      sy[i] = s / row[0];
      if (blockIdx.x == 0) v.y[i] = sy[i];
    }
  }
  __syncthreads();
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int j = 0; j < k; j++) acc += sy[j] * (double)Z[j * ld + i];
  x[i] = (float)acc;
}

// threads of k_arnoldi summing the j + 2 partial columns: a wavefront per
// column, up to 16
static int arnoldi_threads(int j) { return 64 * (j + 2 < 16 ? j + 2 : 16); }

// FGMRES waits for an iteration's residual estimate before queuing the next
// one when the estimate is predicted within this factor of the limit
constexpr double FGMRES_PREDICT_MARGIN = 2.0;

// One FGMRES iteration on the device: z_j = P^-1 v_j, w = A z_j, classical
// Gram-Schmidt twice, column j of the Hessenberg matrix and its rotation,
// v_{j+1} = w / |w|, and the residual estimate copied to pinned host memory.
int Sim::arnoldi_step(int j, int m) {
  if (!hest_dev) return fail(MW_EINVAL, "fgmres: residual-estimate mirror not allocated");
  const long N = NV + nnP;
  float* Vb = V32.p;
  float* Zb = Z32.p;
  float* w = f32a.p;
  GmView gv(gm.p, m);
  double* h1 = hcol.p;
  double* h2 = hcol.p + FGMRES_MAXM;
  float* Vj = Vb + (long)j * N;
  prec32(Vj, Zb + (long)j * N);
  apply32(Zb + (long)j * N, w);
  if (gs_once && j + 1 <= 32) {
    // classical Gram-Schmidt once: h2 = V^T w and |w|^2 in one sweep, then
    // v_{j+1} = (w - V h2) / h_{j+1,j} with the norm by Pythagoras
    prof().begin(PROF_GS, st);
    RC(gs_project<float>(Vb, N, j + 1, nullptr, w, w, N, h2, R, st, true, false));
    prof().end(PROF_GS, st);
    k_arnoldi<<<1, arnoldi_threads(j), 0, st>>>(j, m, hzero, h2, gv, hest_dev, 1, R.partial,
                                                dot_blocks(N));
    prof().begin(PROF_GS, st);
    RC(gs_update_scale<float>(Vb, N, j + 1, h2, w, N, gv.sc + 2, Vb + (long)(j + 1) * N, st));
    prof().end(PROF_GS, st);
  } else if (j + 1 <= 32) {
    // classical Gram-Schmidt, twice
    // fused passes: h1 = V^T w; w -= V h1, h2 = V^T w, |w|^2; then (Hessenberg
    // column, norm by Pythagoras) v_{j+1} = (w - V h2) / h_{j+1,j}
    RC(multi_dot<float>(Vb, N, j + 1, w, N, h1, R, st));
    prof().begin(PROF_GS, st);
    RC(gs_project<float>(Vb, N, j + 1, h1, w, w, N, h2, R, st, true, false));
    prof().end(PROF_GS, st);
    k_arnoldi<<<1, arnoldi_threads(j), 0, st>>>(j, m, h1, h2, gv, hest_dev, 1, R.partial,
                                                dot_blocks(N));
    prof().begin(PROF_GS, st);
    RC(gs_update_scale<float>(Vb, N, j + 1, h2, w, N, gv.sc + 2, Vb + (long)(j + 1) * N, st));
    prof().end(PROF_GS, st);
  } else {
    for (int pass = 0; pass < 2; pass++) {
      double* h = pass ? h2 : h1;
      RC(multi_dot<float>(Vb, N, 32, w, N, h, R, st));
      RC(multi_dot<float>(Vb + 32 * N, N, j + 1 - 32, w, N, h + 32, R, st));
      prof().begin(PROF_GS, st);
      k_gs_update<<<nb(N), 256, 0, st>>>(N, j + 1, Vb, N, h, w);
      prof().end(PROF_GS, st);
    }
    dot<float>(w, w, N, gv.sc + 0, R, st);
    k_arnoldi<<<1, 64, 0, st>>>(j, m, h1, h2, gv, hest_dev, 0, nullptr, 0);
    axpby<float>(N, 1.0, gv.sc + 2, w, 0.0, nullptr, Vb + (long)(j + 1) * N, st);
  }
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// Each Arnoldi step j of FGMRES(m) touches fixed buffers only (basis column j,
// the work vectors, the multigrid hierarchy, whose operators change in place),
// so it is captured once as a graph: one launch per iteration instead of ~60,
// which keeps the GPU fed when the kernels are shorter than their launches.
// The multigrid coarse levels are captured inline (no nested graph launch).
int Sim::build_iteration_graphs(int m) {
  hipStream_t cs;
  MW_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipStream_t keep = st;
  Prof::muted++;
  gmg32.inline_coarse = true;
  it_exec.assign(m, nullptr);
  int rc = MW_OK;
  MW_HIP(hipStreamSynchronize(keep));
  for (int j = 0; j < m && rc == MW_OK; j++) {
    st = cs;
    hipGraph_t gph = nullptr;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) rc = arnoldi_step(j, m);
    hipError_t e2 = e == hipSuccess ? hipStreamEndCapture(cs, &gph) : e;
    st = keep;
    if (rc == MW_OK && e2 != hipSuccess)
      rc = fail(MW_EHIP, std::string("fgmres capture: ") + hipGetErrorString(e2));
    if (rc == MW_OK) {
      hipError_t e3 = hipGraphInstantiate(&it_exec[j], gph, nullptr, nullptr, 0);
      if (e3 != hipSuccess) rc = fail(MW_EHIP, std::string("fgmres instantiate: ") + hipGetErrorString(e3));
    }
    if (gph) (void)hipGraphDestroy(gph);
  }
  st = keep;
  gmg32.inline_coarse = false;
  Prof::muted--;
  (void)hipStreamDestroy(cs);
  use_graphs = rc == MW_OK;
  return rc;
}

// Restarted FGMRES(m) on A32 x = r0, x0 = 0; stops at ||r|| <= tol ||r0||.
// The Hessenberg least-squares problem lives on the device.  The host polls
// the residual estimate of iteration j-1 while iteration j is queued, so the
// device never waits for the host; convergence costs one speculative
// iteration, whose column is not used.
int Sim::fgmres32(const float* r0, float* x, int m, double tol, int maxit, int* its,
                  double hist_scale) {
  long N = NV + nnP;
  float* Vb = V32.p;
  float* Zb = Z32.p;
  float* w = f32a.p;
  float* res = f32b.p;
  GmView gv(gm.p, m);
  double* h1 = hcol.p;
  double* h2 = hcol.p + FGMRES_MAXM;
  double r0n = 0;
  int total = 0;
  bool have_x = false;  // x holds a correction (the first one is written, not added)
  for (int cycle = 0; total < maxit; cycle++) {
    const float* rc = cycle == 0 ? r0 : res;  // the first cycle starts from r0 itself
    dot<float>(rc, rc, N, gv.sc + 1, R, st);
    MW_HIP(hipMemcpyAsync(hest + FGMRES_MAXM, gv.sc + 1, 8, hipMemcpyDeviceToHost, st));
    // |r0| of the first cycle is read back with the first iteration's polling
    // (a zero residual is caught there, before any column is used); a restart
    // waits for its residual norm
    double lim = -1.0;
    double cyc0 = 0.0;  // residual norm at the start of this cycle
    if (cycle > 0) {
      MW_HIP(hipStreamSynchronize(st));
      const double beta = sqrt(hest[FGMRES_MAXM]);
      if (beta <= tol * r0n || beta == 0.0) break;
      lim = tol * r0n;
      cyc0 = beta;
    }
    bool zero = false;
    auto first_norm = [&]() {
      if (lim >= 0.0) return;
      r0n = sqrt(hest[FGMRES_MAXM]);
      zero = r0n == 0.0;
      lim = tol * r0n;
      cyc0 = r0n;
    };
    k_fgmres_start<<<1, 64, 0, st>>>(m, gv);
    axpby<float>(N, 1.0, gv.sc + 3, rc, 0.0, nullptr, Vb, st);
    int j = 0, k = -1;
    for (; j < m && total < maxit; j++, total++) {
      float* Vj = Vb + (long)j * N;
      if (j >= 2) {
        // The estimate of iteration j-1 predicted from the last two known ones
        // (geometric contraction): when it is expected to meet the limit, wait
        // for it before queuing iteration j, which would otherwise run (a whole
        // V-cycle and operator) only to be discarded.  A wrong guess costs the
        // device one host round trip of idle time; the columns used, and so the
        // result, are the same either way.
        const double prev = j >= 3 ? hest[j - 3] : cyc0;
        const double last = hest[j - 2];
        const double pred = prev > 0.0 ? last * (last / prev) : 0.0;
        if (pred <= FGMRES_PREDICT_MARGIN * lim) {
          MW_HIP(hipEventSynchronize(ev[(j - 1) & 1]));
          if (hest[j - 1] <= lim) {
            k = j;  // columns 0..j-1; no speculative iteration queued
            break;
          }
        }
      }
      if (collecting && ncollect < cfg.learn_pairs) {
        MW_HIP(hipMemcpyAsync(Qs.p + ncollect * nnP, Vj + NV, nnP * 4, hipMemcpyDeviceToDevice, st));
        ncollect++;
      }
      if (use_graphs && !learned && !collecting && !prof().on) {
        MW_HIP(hipGraphLaunch(it_exec[j], st));
      } else {
        RC(arnoldi_step(j, m));
      }
      MW_HIP(hipEventRecord(ev[j & 1], st));
      // a few sweeps of the next coarsest inverse on the multigrid's side
      // stream, queued while the device works on this iteration
      RC(gmg32.pump_side(3));
      if (j >= 1) {
        MW_HIP(hipEventSynchronize(ev[(j - 1) & 1]));
        first_norm();
        if (zero) {
          k = 0;
          break;
        }
        if (hest[j - 1] <= lim) {
          k = j;  // columns 0..j-1; iteration j was speculative
          break;
        }
      }
    }
    bool done = k >= 0;
    if (!done) {
      k = j;
      MW_HIP(hipStreamSynchronize(st));
      first_norm();
      if (zero) k = 0;
      done = zero || (k > 0 && hest[k - 1] <= lim);
    }
    if (hist_scale >= 0 && !zero)
      for (int i = 0; i < k; i++) history.push_back(hist_scale * hest[i] / r0n);
    if (k > 0) {
      k_backsub_lincomb<<<nb(N), 256, 0, st>>>(N, k, m, gv, Zb, N, have_x ? w : x);
      if (have_x) axpby<float>(N, 1.0, nullptr, w, 1.0, nullptr, x, st);
      have_x = true;
    }
    if (done) break;
    // true residual for the restart: res = r0 - A x
    if (!have_x) fill<float>(N, 0.0f, x, st);
    have_x = true;
    apply32(x, w);
    MW_HIP(hipMemcpyAsync(res, r0, N * 4, hipMemcpyDeviceToDevice, st));
    axpby<float>(N, -1.0, nullptr, w, 1.0, nullptr, res, st);
  }
  if (!have_x) fill<float>(N, 0.0f, x, st);
  *its = total;
  return MW_OK;
}

__global__ void k_gs_sub64(long n, int k, const double* __restrict__ V, long ld,
                           const double* __restrict__ h, double* __restrict__ w) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = w[i];
  for (int j = 0; j < k; j++) s -= h[j] * V[j * ld + i];
  w[i] = s;
}

__global__ void k_lincomb_add64(long nv, long np_, int k, const float* __restrict__ Z, long ld,
                                const double* __restrict__ y, double* __restrict__ u,
                                double* __restrict__ p) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= nv + np_) return;
  double s = 0.0;
  for (int j = 0; j < k; j++) s += y[j] * (double)Z[j * ld + i];
  if (i < nv) u[i] += s;
  else p[i - nv] += s;
}

// FGMRES(m) in FP64 on the matrix-free FP64 Stokes operator, right-
// preconditioned by the FP32 block preconditioner (prec32 on the FP32 copy of
// each basis vector).  The Krylov recurrences, the operator and the solution
// update are FP64, so the residual it minimises is the true one whatever the
// FP32 operator's accuracy; the preconditioner only needs to be a fair
// approximation.  Classical Gram-Schmidt twice, Hessenberg and Givens
// rotations on the host (one synchronisation per iteration).
int Sim::fgmres64(double abs_tol, int maxit, double bnrm, double* res2, int* its) {
  const long N = NV + nnP;
  const int m = std::min(std::max(cfg.fgmres_restart, 2), 32);
  if (!V64.p) {
    RC(V64.alloc((size_t)(m + 1) * N)); RC(zb64.alloc(N)); RC(w64.alloc(N));
    RC(zero64.alloc(N)); RC(h64.alloc(72));
    fill<double>(N, 0.0, zero64, st);
  }
  std::vector<double> H((m + 1) * m), cs(m), sn(m), gv(m + 1), y(m);
  double hh[72];
  int total = 0;
  double rn = sqrt(*res2);
  while (total < maxit && rn > abs_tol) {
    // residual of the current iterate (wu[0], wp[0] hold it) as v_0
    MW_HIP(hipMemcpyAsync(V64.p, wu[0], NV * 8, hipMemcpyDeviceToDevice, st));
    MW_HIP(hipMemcpyAsync(V64.p + NV, wp[0], nnP * 8, hipMemcpyDeviceToDevice, st));
    axpby<double>(N, 1.0 / rn, nullptr, V64.p, 0.0, nullptr, V64.p, st);
    std::fill(gv.begin(), gv.end(), 0.0);
    gv[0] = rn;
    int k = 0;
    for (int j = 0; j < m && total < maxit; j++, total++) {
      double* Vj = V64.p + (long)j * N;
      float* Zj = Z32.p + (long)j * N;
      convert<double, float>(N, Vj, f32c.p, st);
      prec32(f32c.p, Zj);
      convert<float, double>(N, Zj, zb64.p, st);
      // w = A z = -(0 - A z)
      RC(stokes_residual64(g, fixV, r64tab, etaq, zb64.p, zb64.p + NV, zero64.p, zero64.p + NV,
                           1.0 / L_ref, w64.p, w64.p + NV, st));
      axpby<double>(N, -1.0, nullptr, w64.p, 0.0, nullptr, w64.p, st);
      RC(multi_dot<double>(V64.p, N, j + 1, w64.p, N, h64.p, R, st));
      k_gs_sub64<<<nb(N), 256, 0, st>>>(N, j + 1, V64.p, N, h64.p, w64.p);
      RC(multi_dot<double>(V64.p, N, j + 1, w64.p, N, h64.p + 33, R, st));
      k_gs_sub64<<<nb(N), 256, 0, st>>>(N, j + 1, V64.p, N, h64.p + 33, w64.p);
      RC(dot<double>(w64.p, w64.p, N, h64.p + 66, R, st));
      MW_HIP(hipMemcpyAsync(hh, h64.p, sizeof(hh), hipMemcpyDeviceToHost, st));
      MW_HIP(hipStreamSynchronize(st));
      const double hn = sqrt(hh[66]);
      for (int i = 0; i <= j; i++) H[i * m + j] = hh[i] + hh[33 + i];
      H[(j + 1) * m + j] = hn;
      if (hn > 0) axpby<double>(N, 1.0 / hn, nullptr, w64.p, 0.0, nullptr, V64.p + (long)(j + 1) * N, st);
      for (int i = 0; i < j; i++) {  // earlier rotations
        const double a = H[i * m + j], b = H[(i + 1) * m + j];
        H[i * m + j] = cs[i] * a + sn[i] * b;
        H[(i + 1) * m + j] = -sn[i] * a + cs[i] * b;
      }
      const double a = H[j * m + j], b = H[(j + 1) * m + j];
      const double r = sqrt(a * a + b * b);
      cs[j] = r > 0 ? a / r : 1.0;
      sn[j] = r > 0 ? b / r : 0.0;
      H[j * m + j] = r;
      H[(j + 1) * m + j] = 0.0;
      gv[j + 1] = -sn[j] * gv[j];
      gv[j] = cs[j] * gv[j];
      k = j + 1;
      history.push_back(fabs(gv[j + 1]) / bnrm);
      if (fabs(gv[j + 1]) <= abs_tol || hn == 0.0) break;
    }
    for (int i = k - 1; i >= 0; i--) {  // H y = g
      double sacc = gv[i];
      for (int l = i + 1; l < k; l++) sacc -= H[i * m + l] * y[l];
      y[i] = sacc / H[i * m + i];
    }
    MW_HIP(hipMemcpyAsync(h64.p, y.data(), k * 8, hipMemcpyHostToDevice, st));
    k_lincomb_add64<<<nb(N), 256, 0, st>>>(NV, nnP, k, Z32.p, N, h64.p, u, p);
    MW_CHECK_LAUNCH();
    RC(stokes_residual(wu[0], wp[0], res2));
    rn = sqrt(*res2);
    history.push_back(rn / bnrm);
    if (!(rn == rn)) break;
  }
  *its = total;
  return MW_OK;
}

__global__ void k_pack_residual(long nv, long np_, const double* __restrict__ ru,
                                const double* __restrict__ rp, double inv, float* __restrict__ r) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < nv) r[i] = (float)(ru[i] * inv);
  else if (i < nv + np_) r[i] = (float)(rp[i - nv] * inv);
}

__global__ void k_unpack_add(long nv, long np_, const float* __restrict__ d, double s,
                             double* __restrict__ u, double* __restrict__ p) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < nv) u[i] += s * (double)d[i];
  else if (i < nv + np_) p[i - nv] += s * (double)d[i];
}

int Sim::stokes_solve_mixed() {
  const double tol = cfg.stokes_tol;
  RC(initial_guess());
  RC(gmg32.update(eta_mg, R, st));
  pressure_mass_eta_lumped(g, hx * hz, eta_mg, mdiag, st);
  if (cfg.schur_cheb >= 2)
    RC(pmass_prepare(cfg.nelx, cfg.nelz, hx * hz, eta_mg, pm_w, pm_dg, pm_dinv, st));
  k_recip<<<nb(nnP), 256, 0, st>>>(nnP, mdiag, minv);
  if (cfg.solver == 2) {
    update_leta();
    collecting = !learn_done;  // the first solve provides training vectors
  }
  stats[MW_STAT_LEARNED] = learned ? 1.0 : 0.0;
  stats[MW_STAT_LEARN_MS] = 0.0;
  double bnrm = 0;
  long N = NV + nnP;
  double res2 = 0;
  int outer = 0, inner = 0;
  RC(stokes_residual(wu[0], wp[0], &res2, &bnrm));
  history.assign(1, sqrt(res2) / bnrm);
  // Defect correction contracts by the accuracy of the FP32 solve: a few
  // orders per step on smooth viscosity, but not at all once eps32 times the
  // operator's condition number nears 1 (viscosity contrasts of 1e5-1e6 with
  // sharp, plastic weak zones).  Two steps in a row that gain < 10 % hand over
  // to FP64 FGMRES with the same FP32 preconditioner.
  int slow = 0;
  while (sqrt(res2) > tol * bnrm && outer < 20 && slow < 2) {
    double rn = sqrt(res2);
    k_pack_residual<<<nb(N), 256, 0, st>>>(NV, nnP, wu[0], wp[0], 1.0 / rn, r32.p);
    int its = 0;
    // Ask FGMRES for what this outer step needs (a fraction of the remaining
    // reduction, leaving room for the FP32 estimate's error), but never beyond
    // what FP32 Krylov can deliver (1e-5).
    const double frac = cfg.fgmres_inner > 0.0 ? cfg.fgmres_inner : 0.8;
    double itol = std::min(0.5, std::max(1e-5, frac * tol * bnrm / rn));
    const int maxit = cfg.fgmres_maxit > 0 ? cfg.fgmres_maxit : 300;
    RC(fgmres32(r32.p, x32.p, cfg.fgmres_restart, itol, maxit, &its, rn / bnrm));
    inner += its;
    if (learned && its >= maxit) learned = false;  // the classical preconditioner takes over
    k_unpack_add<<<nb(N), 256, 0, st>>>(NV, nnP, x32.p, rn, u, p);
    outer++;
    RC(stokes_residual(wu[0], wp[0], &res2));
    history.push_back(sqrt(res2) / bnrm);
    if (!(res2 == res2)) break;  // NaN: reported as no convergence below
    slow = sqrt(res2) > 0.9 * rn ? slow + 1 : 0;
  }
  RC(gmg32.pump_side(1 << 30));  // the rest of the next coarsest inverse
  if (sqrt(res2) > tol * bnrm && res2 == res2) {
    int its64 = 0;
    const int maxit = 5 * (cfg.fgmres_maxit > 0 ? cfg.fgmres_maxit : 300);
    RC(fgmres64(tol * bnrm, maxit, bnrm, &res2, &its64));
    inner += its64;
    stats[MW_STAT_STOKES_FGMRES64] = its64;
  } else {
    stats[MW_STAT_STOKES_FGMRES64] = 0;
  }
  stats[MW_STAT_STOKES_OUTER] = outer;
  stats[MW_STAT_STOKES_VCYCLES] = inner;
  stats[MW_STAT_STOKES_RELRES] = sqrt(res2) / bnrm;
  warm = true;
  collecting = false;
  if (!(sqrt(res2) <= tol * bnrm)) return fail(MW_ENOCONV, "stokes (mixed): no convergence");
  if (cfg.solver == 2 && !learn_done) RC(learn_train());
  return MW_OK;
}

// ------------------------------------------------------------- energy ------
__global__ void k_bicg_p(long n, const double* __restrict__ r, const double* __restrict__ v,
                         const double* __restrict__ S, double* __restrict__ pv) {
  // beta = (rho_new/rho)*(alpha/omega); p = r + beta*(p - omega*v)
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double beta = (S[3] / S[0]) * (S[1] / S[2]);
  pv[i] = r[i] + beta * (pv[i] - S[2] * v[i]);
}

__global__ void k_bicg_scalars_rho(double* S) { S[0] = S[3]; }

// T <- T + a (T - prev) on the free nodes (the fixed ones keep their values),
// prev <- T (old values)
__global__ void k_extrapolate_free(long n, double a, const uint8_t* __restrict__ fix,
                                   double* __restrict__ x, double* __restrict__ prev) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double t = x[i];
  x[i] = fix[i] ? t : t + a * (t - prev[i]);
  prev[i] = t;
}

int Sim::energy_solve(double dt) {
  if (cfg.energy_mf) {
    double* w[8];
    for (int i = 0; i < 8; i++) w[i] = wT[i];
    int its = 0;
    // BiCGStab starts from the old temperature extrapolated in time (with
    // warm_start and extrapolate, outside Picard iterations): the
    // right-hand side already holds the old temperature, and the solve stops
    // on the same residual, so only where it starts changes
    if (cfg.warm_start && cfg.extrapolate && !cfg.nonlinear) {
      const double h = dt_hist[2];  // the previous step's time step
      if (have_tprev && h > 0.0) {
        k_extrapolate_free<<<nb(nn), 256, 0, st>>>(nn, dt / h, fixT, T, Tprev);
      } else {
        MW_HIP(hipMemcpyAsync(Tprev, T, nn * 8, hipMemcpyDeviceToDevice, st));
        have_tprev = true;
      }
      MW_CHECK_LAUNCH();
    }
    RC(energy_mf_solve(emf, T, rhsT, w, cfg.energy_tol, 1000, ev, &its, st));
    stats[MW_STAT_ENERGY_ITS] = its;
    energy_pending = true;
    return MW_OK;
  }
  double *r = wT[0], *rh = wT[1], *pv = wT[2], *v = wT[3], *s = wT[4], *tt = wT[5], *ph = wT[6],
         *sh = wT[7];
  double* S = dscal.p + 20;  // 0 rho, 1 alpha, 2 omega, 3 rho_new, 4 (rh,v), 5 (t,s), 6 (t,t)
  energy_invdiag_vv(g, VVT, invdT, st);
  // x0 = T (old temperature), r = b - A x0
  energy_apply_vv(g, VVT, T, r, st);
  axpby<double>(nn, 1.0, nullptr, rhsT, -1.0, nullptr, r, st);
  MW_HIP(hipMemcpyAsync(rh, r, nn * 8, hipMemcpyDeviceToDevice, st));
  fill<double>(nn, 0.0, pv, st);
  fill<double>(nn, 0.0, v, st);
  k_const<<<1, 64, 0, st>>>(3, 1.0, S);
  double bn = sqrt(host_dot(rhsT, rhsT, nn));
  double tol = cfg.energy_tol * (bn > 0 ? bn : 1.0);
  int k = 0;
  double rn = sqrt(host_dot(r, r, nn));
  for (; k < 1000 && rn > tol; k++) {
    dot<double>(rh, r, nn, S + 3, R, st);
    k_bicg_p<<<nb(nn), 256, 0, st>>>(nn, r, v, S, pv);
    diag_scale<double>(nn, 1.0, invdT, pv, ph, st);
    energy_apply_vv(g, VVT, ph, v, st);
    dot<double>(rh, v, nn, S + 4, R, st);
    k_div<<<1, 1, 0, st>>>(S + 3, S + 4, S + 1);  // alpha
    MW_HIP(hipMemcpyAsync(s, r, nn * 8, hipMemcpyDeviceToDevice, st));
    axpby<double>(nn, -1.0, S + 1, v, 1.0, nullptr, s, st);
    diag_scale<double>(nn, 1.0, invdT, s, sh, st);
    energy_apply_vv(g, VVT, sh, tt, st);
    dot<double>(tt, s, nn, S + 5, R, st);
    dot<double>(tt, tt, nn, S + 6, R, st);
    k_div<<<1, 1, 0, st>>>(S + 5, S + 6, S + 2);  // omega
    axpby<double>(nn, 1.0, S + 1, ph, 1.0, nullptr, T, st);
    axpby<double>(nn, 1.0, S + 2, sh, 1.0, nullptr, T, st);
    MW_HIP(hipMemcpyAsync(r, s, nn * 8, hipMemcpyDeviceToDevice, st));
    axpby<double>(nn, -1.0, S + 2, tt, 1.0, nullptr, r, st);
    k_bicg_scalars_rho<<<1, 1, 0, st>>>(S);
    rn = sqrt(host_dot(r, r, nn));
  }
  // true residual
  energy_apply_vv(g, VVT, T, r, st);
  axpby<double>(nn, 1.0, nullptr, rhsT, -1.0, nullptr, r, st);
  double tr = sqrt(host_dot(r, r, nn));
  stats[MW_STAT_ENERGY_ITS] = k;
  stats[MW_STAT_ENERGY_RELRES] = tr / (bn > 0 ? bn : 1.0);
  if (!(tr <= 10 * tol)) return fail(MW_ENOCONV, "energy: no convergence");
  return MW_OK;
}

// -------------------------------------------------------------- step -------
// After a failed solve: MW_EEMPTY if this step's projection found an element
// without particles (flags[4], set before the solves), else MW_OK.
int Sim::check_empty_elements() {
  int h[8];
  MW_HIP(hipStreamSynchronize(st));
  MW_HIP(hipMemcpy(h, flags.p, sizeof(h), hipMemcpyDeviceToHost));
  if (h[4]) return fail(MW_EEMPTY, "ABORT: an element contains no particle!");
  return MW_OK;
}

// ---------------------------------------------------- population control ---
// population_control, pic_functions.py:1623-1825.
__global__ void k_pc_need(int nel, int nmin, const int* __restrict__ cnt, int* __restrict__ need) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nel) need[e] = cnt[e] < nmin ? nmin - cnt[e] : 0;
}

__global__ void k_pc_flag(int np, const uint8_t* __restrict__ act, int* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) flag[i] = act[i] ? 0 : 1;
}

// the inactive particles in increasing index order (list_available_particles)
__global__ void k_pc_avail(int np, const uint8_t* __restrict__ act, const int* __restrict__ pos,
                           int* __restrict__ avail) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np && !act[i]) avail[pos[i]] = i;
}

// splitmix64 of a (seed, step, draw) counter, top 53 bits as a double in [0, 1)
__device__ inline double pc_uniform01(unsigned long long seed, unsigned long long istep,
                                      unsigned long long draw) {
  unsigned long long z = seed * 0x9E3779B97F4A7C15ull + istep * 0xD1B54A32D192ED03ull + draw;
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-53;
}

// New particle k: its element (offs[e] <= k < offs[e+1]), reduced coordinates
// (random.uniform(-0.95, 0.95), r then t), position from the Q2 map, and the
// nearest particle among those of the element that are active (the element's
// segment of this step's sort, in increasing index; first minimum as np.argmin)
__global__ void k_pc_place(int needed, int nel, int nelx, const int* __restrict__ offs,
                           const int* __restrict__ order, const int* __restrict__ seg,
                           const uint8_t* __restrict__ act, const double* __restrict__ sx,
                           const double* __restrict__ sz, const double* __restrict__ x_V,
                           const double* __restrict__ z_V, unsigned long long seed, int istep,
                           double* __restrict__ nx, double* __restrict__ nz,
                           double* __restrict__ nr, double* __restrict__ nt, int* __restrict__ ne,
                           int* __restrict__ tmpl, int* __restrict__ bad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= needed) return;
  int lo = 0, hi = nel;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (offs[mid] <= k) lo = mid;
    else hi = mid;
  }
  const int e = lo;
  const double r = -0.95 + 1.9 * pc_uniform01(seed, (unsigned)istep, 2ull * k);
  const double t = -0.95 + 1.9 * pc_uniform01(seed, (unsigned)istep, 2ull * k + 1);
  double N[9], xv[9], zv[9];
  basis_V(r, t, N);
  const int nnx = 2 * nelx + 1;
  for (int q = 0; q < 9; q++) {
    const int n = grid_node(e, nelx, nnx, q);
    xv[q] = x_V[n];
    zv[q] = z_V[n];
  }
  const double x = dot9(N, xv), z = dot9(N, zv);
  int best = -1;
  double bd = 0.0;
  for (int j = seg[e]; j < seg[e + 1]; j++) {
    const int q = order[j];
    if (!act[q]) continue;
    const double dx = sx[q] - x, dz = sz[q] - z;
    const double d = dx * dx + dz * dz;
    if (best < 0 || d < bd) {
      bd = d;
      best = q;
    }
  }
  if (best < 0) atomicOr(bad, 1);
  nx[k] = x;
  nz[k] = z;
  nr[k] = r;
  nt[k] = t;
  ne[k] = e;
  tmpl[k] = best;
}

__global__ void k_pc_apply(int needed, const int* __restrict__ avail, const int* __restrict__ ne,
                           const double* __restrict__ nx, const double* __restrict__ nz,
                           const double* __restrict__ nr, const double* __restrict__ nt,
                           const int* __restrict__ tmpl, uint8_t* __restrict__ act,
                           int* __restrict__ iel, double* __restrict__ sx, double* __restrict__ sz,
                           double* __restrict__ sr, double* __restrict__ st,
                           uint8_t* __restrict__ mat) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= needed) return;
  const int p = avail[k];
  act[p] = 1;
  iel[p] = ne[k];
  sx[p] = nx[k];
  sz[p] = nz[k];
  sr[p] = nr[k];
  st[p] = nt[k];
  mat[p] = tmpl[k] >= 0 ? mat[tmpl[k]] : 0;
}

int Sim::population_control() {
  const mw_sim_config& c = cfg;
  const int nmin = c.nparticle_min > 0 ? c.nparticle_min : (int)(0.85 * c.ppe);
  k_pc_need<<<nb(nel), 256, 0, st>>>(nel, nmin, cnt, pc_need);
  RC(exclusive_scan_i32(nel, pc_need, pc_offs, pc_tsum, st));
  k_pc_flag<<<nb(np), 256, 0, st>>>(np, sact, pc_flag);
  RC(exclusive_scan_i32(np, pc_flag, pc_pos, pc_tsum, st));
  k_pc_avail<<<nb(np), 256, 0, st>>>(np, sact, pc_pos, pc_avail);
  MW_CHECK_LAUNCH();
  int needed = 0, available = 0;
  MW_HIP(hipMemcpyAsync(&needed, pc_offs.p + nel, 4, hipMemcpyDeviceToHost, st));
  MW_HIP(hipMemcpyAsync(&available, pc_pos.p + np, 4, hipMemcpyDeviceToHost, st));
  MW_HIP(hipStreamSynchronize(st));
  stats[MW_STAT_PC_NEEDED] = needed;
  stats[MW_STAT_PC_AVAILABLE] = available;
  stats[MW_STAT_PC_ACTIVE] = np - available;
  if (needed == 0) return MW_OK;
  if (needed > available) return fail(MW_ERANGE, "population_control: not available yet");
  MW_HIP(hipMemsetAsync(pc_bad.p, 0, 4, st));
  k_pc_place<<<nb(needed), 256, 0, st>>>(needed, nel, c.nelx, pc_offs, order, seg, sact, sx, sz,
                                          x_V, z_V, (unsigned long long)(unsigned)c.pc_seed,
                                          istep, pc_x, pc_z, pc_r, pc_t, pc_e, pc_tmpl, pc_bad);
  k_pc_apply<<<nb(needed), 256, 0, st>>>(needed, pc_avail, pc_e, pc_x, pc_z, pc_r, pc_t, pc_tmpl,
                                          sact, siel, sx, sz, sr, st_, smat);
  MW_CHECK_LAUNCH();
  int bad = 0;
  MW_HIP(hipMemcpyAsync(&bad, pc_bad.p, 4, hipMemcpyDeviceToHost, st));
  MW_HIP(hipStreamSynchronize(st));
  if (bad) return fail(MW_ERANGE, "population_control: mask empty");
  return MW_OK;
}

// Sums of the Picard assessment (assess_steady_state.py:34-47): per block,
// (u_mem - u)^2 and u^2 over the x and z velocity entries, the same for p and
// for T against T_mem; 8 partial sums per block.
__global__ void __launch_bounds__(256) k_nl_partial(long NV, long nnP, long nn,
                                                    const double* __restrict__ u,
                                                    const double* __restrict__ um,
                                                    const double* __restrict__ p,
                                                    const double* __restrict__ pm,
                                                    const double* __restrict__ T,
                                                    const double* __restrict__ Tm,
                                                    double* __restrict__ part) {
  __shared__ double sh[8][256];
  double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const long n = NV > nnP ? (NV > nn ? NV : nn) : (nnP > nn ? nnP : nn);
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += 256L * gridDim.x) {
    if (i < NV) {
      const double d = um[i] - u[i];
      const int c = (int)(i & 1) * 2;
      a[c] += d * d;
      a[c + 1] += u[i] * u[i];
    }
    if (i < nnP) {
      const double d = pm[i] - p[i];
      a[4] += d * d;
      a[5] += p[i] * p[i];
    }
    if (i < nn) {
      const double d = Tm[i] - T[i];
      a[6] += d * d;
      a[7] += T[i] * T[i];
    }
  }
  for (int k = 0; k < 8; k++) sh[k][threadIdx.x] = a[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < 8; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 8) part[blockIdx.x * 8 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_nl_final(int nblk, const double* __restrict__ part, double* __restrict__ out) {
  if (threadIdx.x < 8) {
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += part[b * 8 + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// assess_nlconvergence (assess_steady_state.py:10-74): relative changes of u,
// w, p against the previous iterate and of T against T_mem; converged when all
// four are below tol_nl.  The iterate becomes the new u_mem, p_mem; T_mem = T
// once the iterations end (before this iteration's energy solve, as the
// reference orders it).
int Sim::assess_nl(int iter, bool* inside) {
  (void)iter;
  const int NB = 256;
  k_nl_partial<<<NB, 256, 0, st>>>(NV, nnP, nn, u, u_mem, p, p_mem, T, Tm, nlsum.p + 8);
  k_nl_final<<<1, 64, 0, st>>>(NB, nlsum.p + 8, nlsum.p);
  MW_CHECK_LAUNCH();
  double s8[8];
  MW_HIP(hipMemcpyAsync(s8, nlsum.p, sizeof(s8), hipMemcpyDeviceToHost, st));
  MW_HIP(hipStreamSynchronize(st));
  double xi[4];
  for (int k = 0; k < 4; k++) xi[k] = sqrt(s8[2 * k]) / sqrt(s8[2 * k + 1]);
  nl_hist.insert(nl_hist.end(), xi, xi + 4);
  const double tol = cfg.tol_nl > 0 ? cfg.tol_nl : 1e-3;
  if (xi[0] < tol && xi[1] < tol && xi[2] < tol && xi[3] < tol) *inside = false;
  MW_HIP(hipMemcpyAsync(u_mem.p, u.p, NV * 8, hipMemcpyDeviceToDevice, st));
  MW_HIP(hipMemcpyAsync(p_mem.p, p.p, nnP * 8, hipMemcpyDeviceToDevice, st));
  if (!*inside) MW_HIP(hipMemcpyAsync(Tm.p, T.p, nn * 8, hipMemcpyDeviceToDevice, st));
  return MW_OK;
}

// Stages 1-12 of the time loop, one Picard iteration: material, projections,
// Stokes, pressure, dt, energy (meeuuw.py:1112-2250).  *inside: on entry,
// whether further iterations are allowed; cleared when the iteration has
// converged (assess_steady_state.py:60-66).
int Sim::iteration(int iter_nl, bool nonlinear, bool* inside, double* dt_out, double* poff_out) {
  const mw_sim_config& c = cfg;
  const int avg = c.averaging;
  // 1. particles grouped by element (stable counting sort): every particle
  // field below is formed and stored in this order
  RC(sort_by_element(np, nel, sact, siel, order, seg, sortwork.p, st, sort_clean));
  sort_clean = true;  // the sort leaves its counters cleared
  // 2. temperature on the particles and the material model (meeuuw.py:1118-1125,
  // 1150-1180); the interpolated pressure of the reference feeds no material law
  // here and is not formed
  {
    MatLaw ml{};
    if (c.material_mode == 1) {
      ml.mat = smat;
      for (int m = 0; m < 2; m++) { ml.meta[m] = c.mat_eta[m]; ml.mrho[m] = c.mat_rho[m]; }
    }
    if (c.rheology == 1) {  // experiment3.py:67-110 with the last nodal strain rate
      ml.tosi = c.tosi_case;
      ml.gT = c.tosi_gamma_T; ml.gy = c.tosi_gamma_y; ml.sy = c.tosi_sigma_y;
      ml.estar = c.tosi_eta_star;
      ml.exx = exxn; ml.ezz = ezzn; ml.exz = exzn; ml.z = sz;
    }
    k_material_sorted<<<nb(np), 256, 0, st>>>(np, nel, c.nelx, order, seg, siel, sr, st_, T,
                                              c.rho0, c.alphaT, log(c.eta_contrast), c.eta_min,
                                              c.eta_max, c.eta_contrast == 1.0 ? 1 : 0, ml, srho,
                                              seta, shcond, shcapa, shprod, salpha, sr_s, st_s);
  }
  mark(15);  // material model on the particles (with the sort and the T interpolation)
  // 3. particles -> elements (meeuuw.py:1194-1199)
  RC(project_on_elements(nel, seg, nullptr, srho, seta, avg, rho_e, eta_e, cnt, st));
  k_count_check<<<std::min(nb(nel), 64u), 256, 0, st>>>(nel, cnt, eta_e, flags, eebits);
  mark(17);
  // 4. particles -> nodes (meeuuw.py:1344-1405)
  {
    const int t = c.nodal_projection_type;
    if (t == 2) {
      const double* phi[6] = {srho, seta, shcond, shcapa, shprod, salpha};
      double* out[6] = {rho_n, eta_n, hcond_n, hcapa_n, hprod_n, alpha_n};
      const int av[6] = {0, avg, 0, 0, 0, 0};
      RC(project_on_nodes(2, nel, (int)nn, 6, seg, nullptr, phi, av, icon_V, sr_s, st_s,
                          incn_ptr, incn_list, nodework, out, st));
    } else {
      const double* phi[2] = {srho, seta};
      double* out[2] = {rho_n, eta_n};
      const int av[2] = {0, avg};
      RC(project_on_nodes(t, nel, (int)nn, 2, seg, nullptr, phi, av, icon_V, sr_s, st_s,
                          incn_ptr, incn_list, nodework, out, st));
      const double* phi2[4] = {shcond, shcapa, shprod, salpha};
      double* out2[4] = {hcond_n, hcapa_n, hprod_n, alpha_n};
      const int av2[4] = {0, 0, 0, 0};
      RC(project_on_nodes(2, nel, (int)nn, 4, seg, nullptr, phi2, av2, icon_V, sr_s, st_s,
                          incn_ptr, incn_list, nodework, out2, st));
    }
  }
  mark(18);
  // 5. qpoint fields (meeuuw.py:1503-1590)
  k_qfields<<<nb(9L * nel), 256, 0, st>>>(nel, c.rk == -1 ? 1 : 0, rho_e, eta_e, srho, seta, rhoq,
                                          etaq);
  // 'nodal' (meeuuw.py:1514-1515, 1539-1540): the nodal fields through the Q1 functions
  if (c.rho_projection == 1) {
    const double* phi[1] = {rho_n};
    double* out[1] = {rhoq};
    RC(nodal_to_qpoints(1, phi, nel, 4, N_P, icon_V, out, st, c.nelx));
  }
  if (c.eta_projection == 1) {
    const double* phi[1] = {eta_n};
    double* out[1] = {etaq};
    RC(nodal_to_qpoints(1, phi, nel, 4, N_P, icon_V, out, st, c.nelx));
  }
  // least squares (meeuuw.py:1224-1290): fits over the element's active
  // particles, the slope limiter for P1 (meeuuw.py:1243-1249, h = hx)
  if (c.rho_projection >= 2 || c.eta_projection >= 2) {
    const int nbf = (c.rho_projection == 2 || c.eta_projection == 2) ? 3 : 4;
    RC(ls_coefficients_sorted(nbf, nel, xe, ze, seg, order, sx, sz, srho, seta, lscr, lsce, lsmm,
                              st));
    if (nbf == 3) {
      RC(limiter_strided(nel, lscr, lscr.p + 1, lscr.p + 2, lsmm, lsmm.p + 1, hx, 3, 4, st));
      RC(limiter_strided(nel, lsce, lsce.p + 1, lsce.p + 2, lsmm.p + 2, lsmm.p + 3, hx, 3, 4, st));
    }
    const int which = (c.rho_projection >= 2 ? 1 : 0) | (c.eta_projection >= 2 ? 2 : 0);
    k_ls_qpoints<<<nb(9L * nel), 256, 0, st>>>(nel, nbf, which, lscr, lsce, xqc, zqc, xe, ze,
                                               rhoq, etaq);
  }
  // element viscosity seen by the multigrid levels and the Schur preconditioner
  if (c.rk == -1 || c.eta_projection != 0)
    k_mean9<<<nb(nel), 256, 0, st>>>(nel, etaq, eta_mg);
  else
    MW_HIP(hipMemcpyAsync(eta_mg, eta_e, nel * 8, hipMemcpyDeviceToDevice, st));
  {
    const double* phi[1] = {T};
    double* out[1] = {Tq};
    RC(nodal_to_qpoints(1, phi, nel, 9, N_V, icon_V, out, st, c.nelx));
    const double* phi4[4] = {hcapa_n, hcond_n, hprod_n, alpha_n};
    double* out4[4] = {hcapaq, hcondq, hprodq, alphaq};
    RC(nodal_to_qpoints(4, phi4, nel, 4, N_P, icon_V, out4, st, c.nelx));
  }
  mark(21);
  // 6. Stokes assembly (build_matrix_stokes)
  {
    StokesArgs a{};
    a.nel = nel; a.Nfem_V = (int)NV; a.Nfem_P = (int)nnP; a.blocks = 0;
    a.rhs_only = c.solver == 0 ? 0 : 1;
    a.icon_P = icon_P; a.l2g = l2g; a.rhoq = rhoq; a.etaq = etaq; a.JxWq = JxWq;
    a.gxq = gxq; a.gyq = gzq; a.N_V = N_V; a.N_P = N_P; a.dNdr_V = dNdr_V; a.dNdt_V = dNdt_V;
    a.dNdr_P = dNdr_P; a.dNdt_P = dNdt_P; a.j00 = j00q; a.j01 = j01q; a.j10 = j10q;
    a.j11 = j11q; a.eta_e = eta_e; a.eta_ref = 1.0; a.L_ref = L_ref; a.bc_fix = fixV;
    a.bc_val = valV; a.inc_v_ptr = incv_ptr; a.inc_v_list = incv_list; a.inc_p_ptr = incp_ptr;
    a.inc_p_list = incp_list; a.VV = c.solver == 0 ? VV.p : nullptr; a.rhs_f = rhs_f;
    a.rhs_h = rhs_h;
    a.VV_K = a.VV_G = a.VV_GT = a.VV_M = a.VV_Me = a.VV_H = nullptr; a.work = awork;
    // the box's Dirichlet values are all zero (free slip, k_bc): the right-hand
    // sides need no lifting
    RC(a.rhs_only ? stokes_rhs_zero_lift(a, st) : build_matrix_stokes(a, st));
  }
  mark(1);
  // 7. Stokes solve; an element without particles (0/0 viscosity) is reported as
  // such, as the reference aborts right after the projection (meeuuw.py:1198)
  {
    const int rc = c.solver == 0 ? stokes_solve_plain() : stokes_solve_mixed();
    if (rc != MW_OK) {
      RC(check_empty_elements());
      return rc;
    }
  }
  mark(2);
  // 8. pressure normalisation (meeuuw.py:1958-1977), p scaled by eta_ref/L_ref
  if (L_ref != 1.0) axpby<double>(nnP, 0.0, nullptr, p, 1.0 / L_ref, nullptr, p, st);
  RC(pressure_offset(c.pressure_normalisation, top, nel, c.nelx, N_P, JxWq, p, icon_P, volume,
                     dscal.p + 30, nodework, st));
  k_sub_scalar<<<nb(nnP), 256, 0, st>>>(nnP, dscal.p + 30, p);
  mark(12);
  // 9. nodal pressure
  RC(nodal_pressure(nel, (int)nn, icon_V, icon_P, p, N_P_n, incn_ptr, incn_list, nodework, q, st));
  mark(3);
  // assess the Picard iteration (meeuuw.py:2096-2101)
  if (nonlinear) {
    RC(assess_nl(iter_nl, inside));
    mark(38);
  }
  // 10. time step (meeuuw.py:2112-2139)
  {
    int nbk = 256;
    k_dt_partial<<<nbk, 256, 0, st>>>(nn, u, np, srho, shcond, shcapa, nodework);
    k_dt_final<<<1, 256, 0, st>>>(nbk, nodework, dscal.p + 44, dscal.p + 30, R.hbuf_dev + 43);
    mark(19);
    MW_HIP(hipStreamSynchronize(st));
  }
  const double poff = R.hbuf[43];
  double vmax = R.hbuf[44], srho_ = R.hbuf[45], shc = R.hbuf[46], shp = R.hbuf[47];
  double h = hx < hz ? hx : hz;
  double dt1 = c.CFL * h / vmax;
  double dt2 = c.CFL * h * h / ((shc / np) / (shp / np) / (srho_ / np));
  dt1 = std::min(dt1, 1.25 * dt1_mem);
  dt2 = std::min(dt2, 1.25 * dt2_mem);
  double dt = std::min(std::min(dt1, dt2), 1e30);
  if (!*inside) {  // the limiter's memory moves on with the time step only (meeuuw.py:2141-2151)
    dt1_mem = dt1;
    dt2_mem = dt2;
  }
  stats[MW_STAT_DT1] = dt1;
  stats[MW_STAT_DT2] = dt2;
  // 11. uq, wq (meeuuw.py:2162-2164) from the nodal velocity components
  {
    k_split_uv<<<nb(nn), 256, 0, st>>>(nn, u, un, wn);
    const double* phi[2] = {un, wn};
    double* out[2] = {uq, wq};
    RC(nodal_to_qpoints(2, phi, nel, 9, N_V, icon_V, out, st, c.nelx));
  }
  mark(21);
  // 12. energy assembly (build_matrix_energy with T_mem, meeuuw.py:2187-2193) and
  // solve.  Without Picard iterations T_mem is the previous step's temperature,
  // which T still holds until the solve; with them it is the T_mem buffer.
  const double* Told = nonlinear ? Tm.p : T.p;
  if (c.energy_mf) {
    RC(energy_mf_rhs(emf, Told, valT, uq, wq, hcapaq, hcondq, hprodq, c.rho0, dt, rhsT, st));
  } else {
    EnergyArgs a;
    a.nel = nel; a.Nfem_T = (int)nn; a.eba = 0; a.T = Told; a.icon_T = icon_V;
    a.etaq = etaq; a.Tq = Tq; a.uq = uq; a.wq = wq; a.hcondq = hcondq; a.hcapaq = hcapaq;
    a.alphaq = alphaq; a.hprodq = hprodq; a.exxq = zq; a.ezzq = zq; a.exzq = zq; a.JxWq = JxWq;
    a.N_T = N_V; a.dNdr_T = dNdr_V; a.dNds_T = dNdt_V; a.j00 = j00q; a.j01 = j01q; a.j10 = j10q;
    a.j11 = j11q; a.bc_fix = fixT; a.bc_val = valT; a.dt = dt; a.rho0 = c.rho0;
    a.inc_ptr = incn_ptr; a.inc_list = incn_list; a.VV = VVT; a.rhs = rhsT; a.work = awork;
    RC(build_matrix_energy(a, st));
  }
  mark(4);
  {
    const int rc = energy_solve(dt);
    if (rc != MW_OK) {
      RC(check_empty_elements());
      return rc;
    }
  }
  mark(5);
  // nodal strain rate (meeuuw.py:2386-2389, compute_strain_rate.py:50-73): the
  // global quantities of this step and the strain-rate rheology of the next
  // material evaluation use it
  RC(mw_nodal_gradients(0, nel, (int)nn, icon_V, un, wn, dNdr_n, dNdt_n, j00n, j01n, j10n, j11n,
                        incn_ptr, incn_list, exxn, ezzn, exzn, nodework, st));
  mark(11);
  *dt_out = dt;
  *poff_out = poff;
  return MW_OK;
}

int Sim::step() {
  const mw_sim_config& c = cfg;
  if (!swarm_set) return fail(MW_EINVAL, "sim_step: particles not set (mw_sim_set_swarm*)");
  nsev = 0;
  mark(0);
  k_flags_init<<<1, 64, 0, st>>>(flags, eebits);
  const bool nonlinear = c.nonlinear != 0;
  const int niter = nonlinear ? (c.niter_nl > 0 ? c.niter_nl : 25) : 1;
  if (nonlinear && !tm_set) {  // T_mem = T.copy() after the initial temperature (meeuuw.py:708)
    MW_HIP(hipMemcpyAsync(Tm.p, T.p, nn * 8, hipMemcpyDeviceToDevice, st));
    tm_set = true;
  }
  nl_hist.clear();
  double dt = 0.0, poff = 0.0;
  int iter_nl = 0;
  for (;; iter_nl++) {
    // meeuuw.py:2084-2089: inside the iterations unless this is the last allowed one
    bool inside = nonlinear && iter_nl < niter - 1;
    nl_keep_guess = iter_nl > 0;
    const int rc = iteration(iter_nl, nonlinear, &inside, &dt, &poff);
    nl_keep_guess = false;
    if (rc != MW_OK) return rc;
    if (!inside) break;
  }
  stats[MW_STAT_NL_ITS] = iter_nl + 1;
  stats[MW_STAT_PC_NEEDED] = stats[MW_STAT_PC_AVAILABLE] = stats[MW_STAT_PC_ACTIVE] = 0.0;
  for (int k = 0; k < 4; k++)
    stats[MW_STAT_NL_XI_U + k] = nl_hist.empty() ? 0.0 : nl_hist[nl_hist.size() - 4 + k];
  // 13. heat flux, Nu (meeuuw.py:2260-2300)
  RC(nodal_heat_flux(icon_V, T, hcond_n, (int)nn, nel, dNdr_n, dNdt_n, j00n, j01n, j10n, j11n,
                     incn_ptr, incn_list, nodework, dTdx, dTdz, qxn, qzn, st));
  RC(compute_Nu(c.Lx, c.Lz, x_V, z_V, nel, left, right, top, bottom, icon_V, T, dTdx, dTdz, qc, qw,
                hcond_n, dscal.p + 48, nodework, st));
  mark(7);
  // 14. strain rate on the quadrature points (meeuuw.py:2496-2498), then vrms
  // and the global quantities (meeuuw.py:2512-2514) with Tq of this step
  {
    const double* phi[3] = {exxn, ezzn, exzn};
    double* out[3] = {exxq, ezzq, exzq};
    RC(nodal_to_qpoints(3, phi, nel, 9, N_V, icon_V, out, st, c.nelx));
  }
  RC(global_quantities(nel, uq, wq, Tq, rhoq, hcapaq, etaq, exxq, ezzq, exzq, volume, JxWq, gxq,
                       gzq, dscal.p + 32, nodework + 1024 * 13, st));
  mark(6);
  // 15. advection and locate (meeuuw.py:2719-2864)
  if (c.rk > 0) {
    const double* uu = un;
    const double* ww = wn;
    RC(advect_box(c.rk, dt, np, sx, sz, sact, uu, ww, c.Lx, c.Lz, hx, hz, c.nelx, nel, icon_V,
                  x_V, z_V, su, sw, flags.p + 3, st, u));  // u: (u, w) per node
    mark(13);
    if (c.population_control) {  // meeuuw.py:2847-2851, before the locate
      RC(population_control());
      mark(39);
    }
    RC(locate_box(np, sact, sx, sz, hx, hz, x_V, z_V, icon_V, nel, c.nelx, sr, st_, siel, flags,
                  st));
    mark(16);
  }
  // diagnostics and failure flags (one synchronisation)
  k_step_readback<<<1, 64, 0, st>>>(dscal.p + 32, dscal.p + 48, flags, eebits, R.hbuf_dev);
  MW_HIP(hipStreamSynchronize(st));
  int hf[8];
  for (int i = 0; i < 8; i++) hf[i] = reinterpret_cast<const int*>(R.hbuf + 56)[i];
  prof().collect();
  collect_stage_times();
  stats[MW_STAT_DT] = dt;
  dt_hist[0] = dt_hist[1];
  dt_hist[1] = dt_hist[2];
  dt_hist[2] = dt;
  stats[MW_STAT_VRMS] = R.hbuf[8];
  stats[MW_STAT_TAVRG] = R.hbuf[8 + 7];
  stats[MW_STAT_NU] = R.hbuf[24 + 12];
  stats[MW_STAT_TIME] = time;
  stats[MW_STAT_POFFSET] = poff;
  time += dt;
  istep++;
  stats[MW_STAT_MINCOUNT] = hf[5];
  stats[MW_STAT_NACTIVE] = hf[6];
  stats[MW_STAT_MAXCOUNT] = hf[7];
  {
    const unsigned long long* eb = reinterpret_cast<const unsigned long long*>(R.hbuf + 60);
    double lo, hi;
    memcpy(&lo, &eb[0], 8);
    memcpy(&hi, &eb[1], 8);
    stats[MW_STAT_ETAE_MIN] = eb[0] == ~0ull ? 0.0 : lo;
    stats[MW_STAT_ETAE_MAX] = hi;
  }
  stats[MW_STAT_EK] = R.hbuf[8 + 1];
  stats[MW_STAT_WAG] = R.hbuf[8 + 2];
  stats[MW_STAT_TVD] = R.hbuf[8 + 3];
  stats[MW_STAT_TM] = R.hbuf[8 + 6];
  stats[MW_STAT_ETA_AVRG] = R.hbuf[8 + 8];
  if (hf[4]) return fail(MW_EEMPTY, "ABORT: an element contains no particle!");
  if (energy_pending) {  // the matrix-free solve's true residual, read with the diagnostics
    energy_pending = false;
    const double bn = energy_mf_bnorm(emf), tr = energy_mf_true_residual(emf);
    stats[MW_STAT_ENERGY_RELRES] = tr / (bn > 0 ? bn : 1.0);
    if (!(tr <= 10 * cfg.energy_tol * (bn > 0 ? bn : 1.0)))
      return fail(MW_ENOCONV, "energy: no convergence");
  }
  if (hf[0] || hf[3]) return fail(MW_EOUTSIDE, "particle position outside the mesh");
  if (hf[1]) return fail(MW_ERANGE, "r value out of bounds");
  if (hf[2]) return fail(MW_ERANGE, "t value out of bounds");
  return MW_OK;
}

__global__ void k_mask_fix(long n, const uint8_t* __restrict__ fix, const float* __restrict__ x,
                           float* __restrict__ y) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) y[i] = fix[i] ? 0.0f : x[i];
}

int Sim::op(int which, const float* x, float* y) {
  if (cfg.solver == 0) return fail(MW_EINVAL, "sim_operator: needs the mixed solver");
  const bool transfer = which == MW_OP_RESTRICT32 || which == MW_OP_PROLONG32 ||
                        which == MW_OP_MASK32;
  if (transfer && gmg32.lev.size() < 2)
    return fail(MW_EINVAL, "sim_operator: the transfers need at least two multigrid levels");
  const Level<float>& L1 = *gmg32.lev[transfer ? 1 : 0];
  switch (which) {
    case MW_OP_STOKES32: apply32(x, y); break;
    case MW_OP_PREC32: prec32(x, y); break;
    case MW_OP_PREC32_SPLIT:
      split_schur = true;
      prec32(x, y);
      split_schur = false;
      break;
    case MW_OP_RESTRICT32: gmg32.restrict_to(0, x, y, st); break;
    case MW_OP_PROLONG32:
      fill<float>(NV, 0.0f, y, st);
      gmg32.prolong_add(0, x, y, st);
      break;
    case MW_OP_MASK32: k_mask_fix<<<nb(L1.ndof), 256, 0, st>>>(L1.ndof, L1.fix.p, x, y); break;
    default: return fail(MW_EINVAL, "sim_operator: unknown operator");
  }
  MW_CHECK_LAUNCH();
  MW_HIP(hipStreamSynchronize(st));
  return MW_OK;
}

int Sim::set_stage_timing(bool on) {
  if (on && !sev[0])
    for (int i = 0; i < NSEV; i++) MW_HIP(hipEventCreate(&sev[i]));
  stage_timing = on;
  nsev = 0;
  return MW_OK;
}

void Sim::collect_stage_times() {
  for (int i = 0; i < NTIMINGS; i++) stage_s[i] = 0.0;
  for (int i = 1; i < nsev; i++) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, sev[i - 1], sev[i]) == hipSuccess)
      stage_s[sev_slot[i]] += 1e-3 * ms;
  }
  nsev = 0;
}

Sim::~Sim() {
  for (auto& g : it_exec)
    if (g) (void)hipGraphExecDestroy(g);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : sev)
    if (e) (void)hipEventDestroy(e);
  if (hest) (void)hipHostFree(hest);
  if (ehost) (void)hipHostFree(ehost);
}

long Sim::size(int f) const {
  switch (f) {
    case MW_F_T: case MW_F_U: case MW_F_W: case MW_F_Q: case MW_F_RHO_N: case MW_F_ETA_N:
    case MW_F_DTDX: case MW_F_DTDZ:
      return nn;
    case MW_F_P: case MW_F_RHS_H: return nnP;
    case MW_F_SX: case MW_F_SZ: case MW_F_SR: case MW_F_ST: case MW_F_SIEL: case MW_F_SACTIVE:
      return np;
    case MW_F_RHO_E: case MW_F_ETA_E: return nel;
    case MW_F_VV: return 468L * nel;
    case MW_F_RHS_F: return NV;
    case MW_F_STOKES_HISTORY: return (long)history.size();
    case MW_F_EXX_N: case MW_F_EZZ_N: case MW_F_EXZ_N: return nn;
    case MW_F_NL_HISTORY: return (long)nl_hist.size();
    case MW_F_SMAT: return np;
    default: return -1;
  }
}

int Sim::get(int f, void* out, long n) {
  if (n != size(f) || n < 0) return fail(MW_EINVAL, "sim_get_field: size mismatch");
  const void* src = nullptr;
  size_t bytes = n * 8;
  DBuf<double> tmp;
  switch (f) {
    case MW_F_T: src = T.p; break;
    case MW_F_U: case MW_F_W:
      RC(tmp.alloc(2 * nn));
      k_split_uv<<<nb(nn), 256, 0, st>>>(nn, u, tmp.p, tmp.p + nn);
      src = f == MW_F_U ? tmp.p : tmp.p + nn;
      break;
    case MW_F_P: src = p.p; break;
    case MW_F_Q: src = q.p; break;
    case MW_F_SX: src = sx.p; break;
    case MW_F_SZ: src = sz.p; break;
    case MW_F_SR: src = sr.p; break;
    case MW_F_ST: src = st_.p; break;
    case MW_F_SIEL: src = siel.p; bytes = n * 4; break;
    case MW_F_SACTIVE: src = sact.p; bytes = n; break;
    case MW_F_RHO_E: src = rho_e.p; break;
    case MW_F_ETA_E: src = eta_e.p; break;
    case MW_F_RHO_N: src = rho_n.p; break;
    case MW_F_ETA_N: src = eta_n.p; break;
    case MW_F_DTDX: src = dTdx.p; break;
    case MW_F_DTDZ: src = dTdz.p; break;
    case MW_F_VV:
      if (!VV.p) return fail(MW_EINVAL, "field VV: the assembled Stokes values are formed by the "
                                        "plain FP64 solver only");
      src = VV.p;
      break;
    case MW_F_RHS_F: src = rhs_f.p; break;
    case MW_F_RHS_H: src = rhs_h.p; break;
    case MW_F_STOKES_HISTORY:
      std::copy(history.begin(), history.end(), (double*)out);
      return MW_OK;
    case MW_F_SMAT: src = smat.p; bytes = n; break;
    case MW_F_EXX_N: src = exxn.p; break;
    case MW_F_EZZ_N: src = ezzn.p; break;
    case MW_F_EXZ_N: src = exzn.p; break;
    case MW_F_NL_HISTORY:
      std::copy(nl_hist.begin(), nl_hist.end(), (double*)out);
      return MW_OK;
    default: return fail(MW_EINVAL, "sim_get_field: unknown field");
  }
  MW_HIP(hipMemcpyAsync(out, src, bytes, hipMemcpyDeviceToHost, st));
  MW_HIP(hipStreamSynchronize(st));
  return MW_OK;
}

}  // namespace mw

using namespace mw;

extern "C" {

int mw_sim_create(const mw_sim_config* cfg, void* stream, void** sim) {
  if (!cfg || !sim) return fail(MW_EINVAL, "sim_create: null argument");
  Sim* s = new Sim();
  int rc = s->create(*cfg, (hipStream_t)stream);
  if (rc) {
    delete s;
    *sim = nullptr;
    return rc;
  }
  *sim = s;
  return MW_OK;
}

int mw_sim_operator(void* sim, int op, const float* x, float* y) {
  if (!sim || !x || !y) return fail(MW_EINVAL, "sim_operator: null argument");
  return ((Sim*)sim)->op(op, x, y);
}

int mw_sim_time_kernel(void* sim, int which, int reps, double* avg_us) {
  if (!sim || !avg_us || reps < 1) return fail(MW_EINVAL, "sim_time_kernel: bad argument");
  Sim* s = (Sim*)sim;
  if (which != 0) return fail(MW_EINVAL, "sim_time_kernel: unknown kernel");
  if (s->cfg.solver == 0 || s->gmg32.lev.empty())
    return fail(MW_EINVAL, "sim_time_kernel: needs the mixed solver");
  hipEvent_t e0, e1;
  MW_HIP(hipEventCreate(&e0));
  MW_HIP(hipEventCreate(&e1));
  s->gmg32.smooth_steps(0, 2, s->st);  // warm
  MW_HIP(hipEventRecord(e0, s->st));
  s->gmg32.smooth_steps(0, reps, s->st);
  MW_HIP(hipEventRecord(e1, s->st));
  MW_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  MW_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_us = 1e3 * ms / reps;
  return MW_OK;
}

int mw_sim_destroy(void* sim) {
  if (sim) {
    Sim* s = (Sim*)sim;
    (void)hipStreamSynchronize(s->st);
    delete s;
  }
  return MW_OK;
}

int mw_sim_set_temperature(void* sim, const double* T_host) {
  if (!sim || !T_host) return fail(MW_EINVAL, "sim_set_temperature: null argument");
  return ((Sim*)sim)->set_temperature(T_host);
}

int mw_sim_set_swarm_reduced(void* sim, const double* r_host, const double* t_host) {
  if (!sim || !r_host || !t_host) return fail(MW_EINVAL, "sim_set_swarm_reduced: null argument");
  return ((Sim*)sim)->set_swarm_reduced(r_host, t_host);
}

int mw_sim_set_swarm_material(void* sim, const uint8_t* mat_host) {
  if (!sim || !mat_host) return fail(MW_EINVAL, "sim_set_swarm_material: null argument");
  return ((Sim*)sim)->set_swarm_material(mat_host);
}

int mw_sim_set_swarm(void* sim, long n, const double* x_host, const double* z_host,
                     const double* r_host, const double* t_host, const int32_t* iel_host,
                     const uint8_t* active_host) {
  if (!sim || !x_host || !z_host || !r_host || !t_host || !iel_host || !active_host)
    return fail(MW_EINVAL, "sim_set_swarm: null argument");
  return ((Sim*)sim)->set_swarm(n, x_host, z_host, r_host, t_host, iel_host, active_host);
}

int mw_sim_step(void* sim, double* stats_host) {
  if (!sim) return fail(MW_EINVAL, "sim_step: null sim");
  Sim* s = (Sim*)sim;
  int rc = s->step();
  if (stats_host) memcpy(stats_host, s->stats, sizeof(s->stats));
  return rc;
}

int mw_profile_enable(int on) {
  prof().collect();
  prof().on = on != 0;
  return MW_OK;
}

int mw_profile_read(double* out_host, int reset) {
  if (!out_host) return fail(MW_EINVAL, "profile_read: null output");
  prof().collect();
  for (int c = 0; c < PROF_NCLASSES; c++) {
    out_host[2 * c] = (double)prof().count[c];
    out_host[2 * c + 1] = prof().total_ms[c];
  }
  if (reset) prof().reset();
  return MW_OK;
}

long mw_sim_field_size(void* sim, int field) {
  if (!sim) return -1;
  return ((Sim*)sim)->size(field);
}

int mw_sim_get_field(void* sim, int field, void* host_out, long n) {
  if (!sim || !host_out) return fail(MW_EINVAL, "sim_get_field: null argument");
  return ((Sim*)sim)->get(field, host_out, n);
}

int mw_sim_stage_timing(void* sim, int on) {
  if (!sim) return fail(MW_EINVAL, "sim_stage_timing: null sim");
  return ((Sim*)sim)->set_stage_timing(on != 0);
}

int mw_sim_stage_times(void* sim, double* seconds_host) {
  if (!sim || !seconds_host) return fail(MW_EINVAL, "sim_stage_times: null argument");
  const Sim* s = (const Sim*)sim;
  memcpy(seconds_host, s->stage_s, sizeof(s->stage_s));
  return MW_OK;
}

}  // extern "C"
