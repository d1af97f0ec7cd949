// Nodal <-> qpoint projections and the FP64 diagnostics of a timestep.
//
// Reference behaviour restated here:
//   project_nodal_{P,V,T}field_onto_qpoints  project_nodal_field_onto_qpoints.py:11-62
//   compute_pressure_offset (plane)          compute_pressure_offset.py:42-61
//   compute_nodal_pressure                   compute_nodal_pressure.py:7-21
//   compute_nodal_heat_flux                  compute_nodal_heat_flux.py:22-38
//   compute_global_quantities                postprocessors.py:53-83
//   compute_Nu (box)                         postprocessors.py:152-248
//
// Reductions are two-pass and deterministic: fixed-size blocks write partial
// sums, one block adds the partials in block order.  Nodal averages gather the
// per-element contributions of each node in increasing element order, which
// reproduces the reference's element-loop accumulation exactly.
#include "mw_common.h"
#include "mw_capi_impl.h"

namespace mw {

constexpr int RBS = 256;       // reduction block size
constexpr int RMAX = 13;       // max number of simultaneous sums
constexpr int RBLOCKS = 1024;  // max partial blocks

struct FPtrs {
  const double* p[8];
};
struct OPtrs {
  double* p[8];
};

// ---------------------------------------------------------------------------
// nodal -> qpoints
// ---------------------------------------------------------------------------
// One element per thread: its m nodes are read once per field and its 9
// points formed in turn (sum over the nodes in reference order per point).
// GRID: the nodes follow from the element index (the box mesh's own node grid).
constexpr int NQ_BS = 128;
template <int M, bool GRID>
__global__ void __launch_bounds__(NQ_BS)
k_nodal_to_q_elem(int nf, FPtrs phi, int nel, int nelx, const double* __restrict__ N,
                  const int* __restrict__ icon, OPtrs out) {
  const int e = blockIdx.x * NQ_BS + threadIdx.x;
  if (e >= nel) return;
  int nd[M];
#pragma unroll
  for (int k = 0; k < M; k++)
    nd[k] = GRID ? grid_node(e, nelx, 2 * nelx + 1, k) : icon[k * (long)nel + e];
  for (int f = 0; f < nf; f++) {
    double v[M];
#pragma unroll
    for (int k = 0; k < M; k++) v[k] = phi.p[f][nd[k]];
    double* o = out.p[f] + 9L * e;
#pragma unroll
    for (int q = 0; q < 9; q++) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < M; k++) s += N[q * M + k] * v[k];
      o[q] = s;
    }
  }
}

int nodal_to_qpoints(int nf, const double* const* phi, int nel, int m, const double* N,
                     const int* icon, double* const* phiq, hipStream_t st, int grid_nelx) {
  if (nf < 1 || nf > 8 || nel <= 0 || (m != 4 && m != 9) || !N || (!icon && grid_nelx <= 0) ||
      (grid_nelx > 0 && nel % grid_nelx))
    return fail(MW_EINVAL, "nodal_to_qpoints: bad arguments");
  FPtrs fp;
  OPtrs op;
  for (int f = 0; f < nf; f++) {
    if (!phi[f] || !phiq[f]) return fail(MW_EINVAL, "nodal_to_qpoints: null field");
    fp.p[f] = phi[f];
    op.p[f] = phiq[f];
  }
  const unsigned nbq = (unsigned)((nel + NQ_BS - 1) / NQ_BS);
  if (grid_nelx > 0) {
    if (m == 9)
      k_nodal_to_q_elem<9, true><<<nbq, NQ_BS, 0, st>>>(nf, fp, nel, grid_nelx, N, icon, op);
    else
      k_nodal_to_q_elem<4, true><<<nbq, NQ_BS, 0, st>>>(nf, fp, nel, grid_nelx, N, icon, op);
  } else if (m == 9) {
    k_nodal_to_q_elem<9, false><<<nbq, NQ_BS, 0, st>>>(nf, fp, nel, 0, N, icon, op);
  } else {
    k_nodal_to_q_elem<4, false><<<nbq, NQ_BS, 0, st>>>(nf, fp, nel, 0, N, icon, op);
  }
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// ---------------------------------------------------------------------------
// generic deterministic reduction
// ---------------------------------------------------------------------------
template <int NS>
__device__ void block_reduce_store(double (&v)[NS], double* __restrict__ partial) {
  __shared__ double sh[NS][RBS];
  for (int s = 0; s < NS; s++) sh[s][threadIdx.x] = v[s];
  __syncthreads();
  for (int w = RBS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int s = 0; s < NS; s++) sh[s][threadIdx.x] += sh[s][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int s = 0; s < NS; s++) partial[blockIdx.x * NS + s] = sh[s][0];
}

template <int NS>
__global__ void k_reduce_final(int nblk, const double* __restrict__ partial,
                               double* __restrict__ out) {
  __shared__ double sh[NS][RBS];
  // every partial of this thread (at most RBLOCKS / RBS per quantity) loaded
  // before the first add, then summed in block order as before
  constexpr int U = RBLOCKS / RBS;
  double v[NS][U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int b = threadIdx.x + u * RBS;
#pragma unroll
    for (int s = 0; s < NS; s++) v[s][u] = b < nblk ? partial[b * NS + s] : 0.0;
  }
#pragma unroll
  for (int s = 0; s < NS; s++) {
    double a = 0.0;
#pragma unroll
    for (int u = 0; u < U; u++)
      if ((int)threadIdx.x + u * RBS < nblk) a += v[s][u];
    sh[s][threadIdx.x] = a;
  }
  __syncthreads();
  for (int w = RBS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int s = 0; s < NS; s++) sh[s][threadIdx.x] += sh[s][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int s = 0; s < NS; s++) out[s] = sh[s][0];
}

static int red_blocks(long n) {
  long b = (n + RBS - 1) / RBS;
  if (b > RBLOCKS) b = RBLOCKS;
  if (b < 1) b = 1;
  return (int)b;
}

// ---------------------------------------------------------------------------
// pressure offset
// ---------------------------------------------------------------------------
__global__ void k_poff_surface(int nel, const uint8_t* __restrict__ top,
                               const double* __restrict__ p, const int* __restrict__ icon_P,
                               double* __restrict__ partial) {
  double v[1] = {0.0};
  for (long e = blockIdx.x * (long)RBS + threadIdx.x; e < nel; e += (long)gridDim.x * RBS)
    if (top[e]) v[0] += 0.5 * (p[icon_P[2L * nel + e]] + p[icon_P[3L * nel + e]]);
  block_reduce_store<1>(v, partial);
}

__global__ void k_poff_volume(int nel, const double* __restrict__ N_P,
                              const double* __restrict__ JxWq, const double* __restrict__ p,
                              const int* __restrict__ icon_P, double* __restrict__ partial) {
  double v[1] = {0.0};
  for (long e = blockIdx.x * (long)RBS + threadIdx.x; e < nel; e += (long)gridDim.x * RBS) {
    double pv[4];
    for (int k = 0; k < 4; k++) pv[k] = p[icon_P[k * (long)nel + e]];
    for (int q = 0; q < 9; q++) {
      double s = 0.0;
      for (int k = 0; k < 4; k++) s += N_P[q * 4 + k] * pv[k];
      v[0] += s * JxWq[e * 9 + q];
    }
  }
  block_reduce_store<1>(v, partial);
}

// axisymmetric branches (compute_pressure_offset.py:33-39, 49-55): the top
// edges' mean pressure times sin of the colatitude at the edge midpoint times
// its angular width, halved; the volume integral weighted by 2 pi x.  Neither
// is divided afterwards.
__global__ void k_poff_surface_axi(int nel, const uint8_t* __restrict__ top,
                                   const double* __restrict__ p, const int* __restrict__ icon_P,
                                   const double* __restrict__ theta_P,
                                   double* __restrict__ partial) {
  double v[1] = {0.0};
  for (long e = blockIdx.x * (long)RBS + threadIdx.x; e < nel; e += (long)gridDim.x * RBS)
    if (top[e]) {
      const int n2 = icon_P[2L * nel + e], n3 = icon_P[3L * nel + e];
      const double pmean = 0.5 * (p[n2] + p[n3]);
      const double theta_mean = M_PI / 2 - 0.5 * (theta_P[n2] + theta_P[n3]);
      const double dtheta = theta_P[n3] - theta_P[n2];
      v[0] += 0.5 * pmean * sin(theta_mean) * dtheta;
    }
  block_reduce_store<1>(v, partial);
}

__global__ void k_poff_volume_axi(int nel, const double* __restrict__ N_P,
                                  const double* __restrict__ JxWq, const double* __restrict__ p,
                                  const int* __restrict__ icon_P, const double* __restrict__ xq,
                                  double* __restrict__ partial) {
  double v[1] = {0.0};
  for (long e = blockIdx.x * (long)RBS + threadIdx.x; e < nel; e += (long)gridDim.x * RBS) {
    double pv[4];
    for (int k = 0; k < 4; k++) pv[k] = p[icon_P[k * (long)nel + e]];
    for (int q = 0; q < 9; q++) {
      double s = 0.0;
      for (int k = 0; k < 4; k++) s += N_P[q * 4 + k] * pv[k];
      v[0] += s * JxWq[e * 9 + q] * 2 * M_PI * xq[e * 9 + q];
    }
  }
  block_reduce_store<1>(v, partial);
}

int pressure_offset_axi(int mode, const uint8_t* top_element, int nel, const double* N_P,
                        const double* JxWq, const double* p, const int* icon_P,
                        const double* theta_P, const double* xq, double* out, double* work,
                        hipStream_t st) {
  if (mode < 0 || mode > 2) return fail(MW_EINVAL, "compute_pressure_offset: unknown mode");
  if (nel <= 0 || !p || !icon_P || !out || !work)
    return fail(MW_EINVAL, "pressure_offset: bad arguments");
  const int nb = red_blocks(nel);
  if (mode == 0) {
    MW_HIP(hipMemsetAsync(out, 0, sizeof(double), st));
    return MW_OK;
  }
  if (mode == 1) {
    if (!top_element || !theta_P)
      return fail(MW_EINVAL, "pressure_offset: axisymmetric surface needs top_element, theta_P");
    k_poff_surface_axi<<<nb, RBS, 0, st>>>(nel, top_element, p, icon_P, theta_P, work);
  } else {
    if (!N_P || !JxWq || !xq)
      return fail(MW_EINVAL, "pressure_offset: axisymmetric volume needs N_P, JxWq, xq");
    k_poff_volume_axi<<<nb, RBS, 0, st>>>(nel, N_P, JxWq, p, icon_P, xq, work);
  }
  MW_CHECK_LAUNCH();
  k_reduce_final<1><<<1, RBS, 0, st>>>(nb, work, out);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

__global__ void k_scale(double* x, double a) { x[0] = x[0] * a; }
__global__ void k_div(double* x, double a) { x[0] = x[0] / a; }

// work: RBLOCKS doubles; result written to out[0] (device)
int pressure_offset(int mode, const uint8_t* top_element, int nel, int nelx, const double* N_P,
                    const double* JxWq, const double* p, const int* icon_P, double volume,
                    double* out, double* work, hipStream_t st) {
  if (mode < 0 || mode > 2) return fail(MW_EINVAL, "compute_pressure_offset: unknown mode");
  if (nel <= 0 || !p || !icon_P || !out || !work)
    return fail(MW_EINVAL, "pressure_offset: bad arguments");
  int nb = red_blocks(nel);
  if (mode == 0) {
    MW_HIP(hipMemsetAsync(out, 0, sizeof(double), st));
    return MW_OK;
  }
  if (mode == 1) {
    if (!top_element || nelx <= 0) return fail(MW_EINVAL, "pressure_offset: surface needs top");
    k_poff_surface<<<nb, RBS, 0, st>>>(nel, top_element, p, icon_P, work);
    MW_CHECK_LAUNCH();
    k_reduce_final<1><<<1, RBS, 0, st>>>(nb, work, out);
    MW_CHECK_LAUNCH();
    k_div<<<1, 1, 0, st>>>(out, (double)nelx);
  } else {
    if (!N_P || !JxWq) return fail(MW_EINVAL, "pressure_offset: volume needs N_P, JxWq");
    k_poff_volume<<<nb, RBS, 0, st>>>(nel, N_P, JxWq, p, icon_P, work);
    MW_CHECK_LAUNCH();
    k_reduce_final<1><<<1, RBS, 0, st>>>(nb, work, out);
    MW_CHECK_LAUNCH();
    k_div<<<1, 1, 0, st>>>(out, volume);
  }
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// ---------------------------------------------------------------------------
// nodal pressure
// ---------------------------------------------------------------------------
__global__ void k_np_elem(int nel, const int* __restrict__ icon_P, const double* __restrict__ p,
                          const double* __restrict__ N_P_n, double* __restrict__ work) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= 9L * nel) return;
  int e = (int)(idx / 9), k = (int)(idx % 9);
  double s = 0.0;
  for (int j = 0; j < 4; j++) s += N_P_n[k * 4 + j] * p[icon_P[j * (long)nel + e]];
  work[idx] = s;
}

__global__ void k_np_node(int nn, const int* __restrict__ ptr, const int* __restrict__ list,
                          const double* __restrict__ work, double* __restrict__ q) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  int ent[INC4];
  const int cn = inc_entries4(ptr, list, n, ent);
  double s = 0.0;
  int c = 0;
  if (cn <= INC4) {  // every load first, then the sum in list order
    double w[INC4];
#pragma unroll
    for (int k = 0; k < INC4; k++)
      w[k] = k < cn ? work[(long)(ent[k] >> 5) * 9 + (ent[k] & 31)] : 0.0;
#pragma unroll
    for (int k = 0; k < INC4; k++)
      if (k < cn) {
        s += w[k];
        c++;
      }
  } else {
    for (int i = ptr[n]; i < ptr[n + 1]; i++) {
      int v = list[i];
      s += work[(long)(v >> 5) * 9 + (v & 31)];
      c++;
    }
  }
  q[n] = s / (double)c;
}

int nodal_pressure(int nel, int nn, const int* icon_V, const int* icon_P, const double* p,
                   const double* N_P_n, const int* inc_ptr, const int* inc_list, double* work,
                   double* q, hipStream_t st) {
  (void)icon_V;
  if (nel <= 0 || nn <= 0 || !icon_P || !p || !N_P_n || !inc_ptr || !inc_list || !work || !q)
    return fail(MW_EINVAL, "nodal_pressure: bad arguments");
  k_np_elem<<<nblocks(9L * nel, 256), 256, 0, st>>>(nel, icon_P, p, N_P_n, work);
  MW_CHECK_LAUNCH();
  k_np_node<<<nblocks(nn, 256), 256, 0, st>>>(nn, inc_ptr, inc_list, work, q);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// ---------------------------------------------------------------------------
// nodal heat flux
// ---------------------------------------------------------------------------
__global__ void k_hf_elem(int nel, const int* __restrict__ icon, const double* __restrict__ T,
                          const double* __restrict__ dNdr, const double* __restrict__ dNds,
                          const double* __restrict__ j00, const double* __restrict__ j01,
                          const double* __restrict__ j10, const double* __restrict__ j11,
                          double* __restrict__ work) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= 9L * nel) return;
  int e = (int)(idx / 9), i = (int)(idx % 9);
  double Tv[9], dx[9], dy[9];
  for (int k = 0; k < 9; k++) Tv[k] = T[icon[k * (long)nel + e]];
  double a = j00[idx], b = j01[idx], c = j10[idx], d = j11[idx];
  for (int k = 0; k < 9; k++) {
    dx[k] = a * dNdr[i * 9 + k] + b * dNds[i * 9 + k];
    dy[k] = c * dNdr[i * 9 + k] + d * dNds[i * 9 + k];
  }
  work[idx * 2] = dot9(dx, Tv);
  work[idx * 2 + 1] = dot9(dy, Tv);
}

__global__ void k_hf_node(int nn, const int* __restrict__ ptr, const int* __restrict__ list,
                          const double* __restrict__ work, const double* __restrict__ hc,
                          double* __restrict__ dTdx, double* __restrict__ dTdy,
                          double* __restrict__ qx, double* __restrict__ qy) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  double sx = 0.0, sy = 0.0, fx = 0.0, fy = 0.0, c = 0.0;
  double h = hc[n];
  for (int i = ptr[n]; i < ptr[n + 1]; i++) {
    int v = list[i];
    long o = ((long)(v >> 5) * 9 + (v & 31)) * 2;
    double gx = work[o], gy = work[o + 1];
    sx -= gx;
    sy -= gy;
    fx -= h * gx;
    fy -= h * gy;
    c += 1;
  }
  qx[n] = fx / c;
  qy[n] = fy / c;
  dTdx[n] = sx / c;
  dTdy[n] = sy / c;
}

int nodal_heat_flux(const int* icon_T, const double* T, const double* hcond_n, int nn, int nel,
                    const double* dNdr_n, const double* dNds_n, const double* j00,
                    const double* j01, const double* j10, const double* j11, const int* inc_ptr,
                    const int* inc_list, double* work, double* dTdx, double* dTdy, double* qx,
                    double* qy, hipStream_t st) {
  if (nel <= 0 || nn <= 0 || !icon_T || !T || !hcond_n || !dNdr_n || !dNds_n || !j00 || !j01 ||
      !j10 || !j11 || !inc_ptr || !inc_list || !work || !dTdx || !dTdy || !qx || !qy)
    return fail(MW_EINVAL, "nodal_heat_flux: bad arguments");
  k_hf_elem<<<nblocks(9L * nel, 256), 256, 0, st>>>(nel, icon_T, T, dNdr_n, dNds_n, j00, j01, j10,
                                                     j11, work);
  MW_CHECK_LAUNCH();
  k_hf_node<<<nblocks(nn, 256), 256, 0, st>>>(nn, inc_ptr, inc_list, work, hcond_n, dTdx, dTdy,
                                               qx, qy);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// ---------------------------------------------------------------------------
// global quantities: [TM, EK, WAG, TVD, ITE, vrms_sum, Tavrg_sum, num, den]
// ---------------------------------------------------------------------------
__global__ void k_gq(long nq, const double* __restrict__ uq, const double* __restrict__ wq,
                     const double* __restrict__ Tq, const double* __restrict__ rhoq,
                     const double* __restrict__ hcapaq, const double* __restrict__ etaq,
                     const double* __restrict__ exxq, const double* __restrict__ ezzq,
                     const double* __restrict__ exzq, const double* __restrict__ JxWq,
                     const double* __restrict__ gxq, const double* __restrict__ gzq,
                     double* __restrict__ partial) {
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (long e = blockIdx.x * (long)RBS + threadIdx.x; e < nq; e += (long)gridDim.x * RBS) {
    double exx = exxq[e], ezz = ezzq[e], exz = exzq[e], jw = JxWq[e];
    double eII = exx * exx + ezz * ezz + 2 * exz * exz;
    double u2 = uq[e] * uq[e] + wq[e] * wq[e];
    v[0] += rhoq[e] * jw;
    v[1] += 0.5 * rhoq[e] * u2 * jw;
    v[2] -= rhoq[e] * (uq[e] * gxq[e] + wq[e] * gzq[e]) * jw;
    v[3] += 2 * etaq[e] * (2 * exx * exx / 3 + 2 * ezz * ezz / 3 - 2 * exx * ezz / 3 + 2 * exz * exz) * jw;
    v[4] += rhoq[e] * hcapaq[e] * Tq[e] * jw;
    v[5] += u2 * jw;
    v[6] += Tq[e] * jw;
    v[7] += etaq[e] * eII * jw;
    v[8] += eII * jw;
  }
  block_reduce_store<9>(v, partial);
}

__global__ void k_gq_finish(const double* __restrict__ s, double volume, double* __restrict__ out) {
  // out = [vrms, EK, WAG, TVD, GPE, ITE, TM, Tavrg, eta_avrg]
  out[0] = sqrt(s[5] / volume);
  out[1] = s[1];
  out[2] = s[2];
  out[3] = s[3];
  out[4] = 0.0;
  out[5] = s[4];
  out[6] = s[0];
  out[7] = s[6] / volume;
  out[8] = s[8] > 0 ? s[7] / s[8] : 0.0;
}

// work: RBLOCKS*9 + 9 doubles; out: 9 doubles (device)
int global_quantities(int nel, const double* uq, const double* wq, const double* Tq,
                      const double* rhoq, const double* hcapaq, const double* etaq,
                      const double* exxq, const double* ezzq, const double* exzq, double volume,
                      const double* JxWq, const double* gxq, const double* gzq, double* out,
                      double* work, hipStream_t st) {
  if (nel <= 0 || !uq || !wq || !Tq || !rhoq || !hcapaq || !etaq || !exxq || !ezzq || !exzq ||
      !JxWq || !gxq || !gzq || !out || !work)
    return fail(MW_EINVAL, "global_quantities: bad arguments");
  long nq = 9L * nel;
  int nb = red_blocks(nq);
  k_gq<<<nb, RBS, 0, st>>>(nq, uq, wq, Tq, rhoq, hcapaq, etaq, exxq, ezzq, exzq, JxWq, gxq, gzq,
                           work);
  MW_CHECK_LAUNCH();
  double* sums = work + RBLOCKS * 9;
  k_reduce_final<9><<<1, RBS, 0, st>>>(nb, work, sums);
  MW_CHECK_LAUNCH();
  k_gq_finish<<<1, 1, 0, st>>>(sums, volume, out);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// ---------------------------------------------------------------------------
// Nusselt number and boundary averages (box)
// sums: [T_l, T_r, T_b, T_t, dTdx_l, dTdx_r, dTdz_b, dTdz_t, qx_l, qx_r, qz_b, qz_t]
// ---------------------------------------------------------------------------
__global__ void k_nu(int nel, const double* __restrict__ x_V, const double* __restrict__ z_V,
                     const uint8_t* __restrict__ left, const uint8_t* __restrict__ right,
                     const uint8_t* __restrict__ top, const uint8_t* __restrict__ bottom,
                     const int* __restrict__ icon, const double* __restrict__ T,
                     const double* __restrict__ dTdx, const double* __restrict__ dTdz,
                     const double* __restrict__ qc, const double* __restrict__ qw,
                     const double* __restrict__ hc, double* __restrict__ partial) {
  double v[12];
  for (int s = 0; s < 12; s++) v[s] = 0.0;
  for (long e = blockIdx.x * (long)RBS + threadIdx.x; e < nel; e += (long)gridDim.x * RBS) {
    bool L = left[e], R = right[e], Tp = top[e], B = bottom[e];
    if (!(L || R || Tp || B)) continue;
    double Tv[9], hv[9], gx[9], gz[9], N[9];
    int nd[9];
    for (int k = 0; k < 9; k++) {
      nd[k] = icon[k * (long)nel + e];
      Tv[k] = T[nd[k]];
      hv[k] = hc[nd[k]];
      gx[k] = dTdx[nd[k]];
      gz[k] = dTdz[nd[k]];
    }
    if (L) {
      double jc = (z_V[nd[3]] - z_V[nd[0]]) / 2;
      for (int iq = 0; iq < 3; iq++) {
        basis_V(-1.0, qc[iq], N);
        double Tq = dot9(N, Tv), hq = dot9(N, hv), dq = dot9(N, gx);
        v[0] += Tq * jc * qw[iq];
        v[4] += dq * jc * qw[iq] * -1;
        v[8] -= hq * dq * jc * qw[iq] * -1;
      }
    }
    if (R) {
      double jc = (z_V[nd[2]] - z_V[nd[1]]) / 2;
      for (int iq = 0; iq < 3; iq++) {
        basis_V(1.0, qc[iq], N);
        double Tq = dot9(N, Tv), hq = dot9(N, hv), dq = dot9(N, gx);
        v[1] += Tq * jc * qw[iq];
        v[5] += dq * jc * qw[iq] * +1;
        v[9] -= hq * dq * jc * qw[iq] * +1;
      }
    }
    if (Tp) {
      double jc = (x_V[nd[2]] - x_V[nd[3]]) / 2;
      for (int iq = 0; iq < 3; iq++) {
        basis_V(qc[iq], 1.0, N);
        double Tq = dot9(N, Tv), hq = dot9(N, hv), dq = dot9(N, gz);
        v[3] += Tq * jc * qw[iq];
        v[7] += dq * jc * qw[iq] * +1;
        v[11] -= hq * dq * jc * qw[iq] * +1;
      }
    }
    if (B) {
      double jc = (x_V[nd[1]] - x_V[nd[0]]) / 2;
      for (int iq = 0; iq < 3; iq++) {
        basis_V(qc[iq], -1.0, N);
        double Tq = dot9(N, Tv), hq = dot9(N, hv), dq = dot9(N, gz);
        v[2] += Tq * jc * qw[iq];
        v[6] += dq * jc * qw[iq] * -1;
        v[10] -= hq * dq * jc * qw[iq] * -1;
      }
    }
  }
  block_reduce_store<12>(v, partial);
}

__global__ void k_nu_finish(const double* __restrict__ s, double Lx, double Lz,
                            double* __restrict__ out) {
  double o[12];
  for (int k = 0; k < 12; k++) o[k] = s[k];
  o[0] /= Lz; o[1] /= Lz; o[2] /= Lx; o[3] /= Lx;
  o[4] /= Lz; o[5] /= Lz; o[6] /= Lx; o[7] /= Lx;
  for (int k = 0; k < 12; k++) out[k] = o[k];
  out[12] = fabs(o[7]) / (o[2] - o[3]) * Lz;
}

// work: RBLOCKS*12 + 12 doubles; out: 13 doubles (device)
int compute_Nu(double Lx, double Lz, const double* x_V, const double* z_V, int nel,
               const uint8_t* left, const uint8_t* right, const uint8_t* top,
               const uint8_t* bottom, const int* icon, const double* T, const double* dTdx,
               const double* dTdz, const double* qc, const double* qw, const double* hcond_n,
               double* out, double* work, hipStream_t st) {
  if (nel <= 0 || !x_V || !z_V || !left || !right || !top || !bottom || !icon || !T || !dTdx ||
      !dTdz || !qc || !qw || !hcond_n || !out || !work)
    return fail(MW_EINVAL, "compute_Nu: bad arguments");
  int nb = red_blocks(nel);
  k_nu<<<nb, RBS, 0, st>>>(nel, x_V, z_V, left, right, top, bottom, icon, T, dTdx, dTdz, qc, qw,
                           hcond_n, work);
  MW_CHECK_LAUNCH();
  double* sums = work + RBLOCKS * 12;
  k_reduce_final<12><<<1, RBS, 0, st>>>(nb, work, sums);
  MW_CHECK_LAUNCH();
  k_nu_finish<<<1, 1, 0, st>>>(sums, Lx, Lz, out);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

}  // namespace mw

// ---------------------------------------------------------------------------
// derived fields: nodal strain rate, nodal pressure gradient
// (compute_strain_rate.py:50-73, compute_nodal_pressure_gradient.py:11-31) and
// elemental strain rate (compute_strain_rate.py:17-42)
// ---------------------------------------------------------------------------
namespace mw {

// per (element, local node): mode 0 -> (exx, eyy, exy) of (u, v); mode 1 ->
// (dq/dx, dq/dy) of q = u
__global__ void k_ngrad_elem(int mode, int nel, const int* __restrict__ icon,
                             const double* __restrict__ u, const double* __restrict__ v,
                             const double* __restrict__ dNdr, const double* __restrict__ dNds,
                             const double* __restrict__ j00, const double* __restrict__ j01,
                             const double* __restrict__ j10, const double* __restrict__ j11,
                             double* __restrict__ work) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= 9L * nel) return;
  int e = (int)(idx / 9), i = (int)(idx % 9);
  double uv[9], vv[9], dx[9], dy[9];
  for (int k = 0; k < 9; k++) {
    long n = icon[k * (long)nel + e];
    uv[k] = u[n];
    vv[k] = mode == 0 ? v[n] : 0.0;
  }
  double a = j00[idx], b = j01[idx], c = j10[idx], d = j11[idx];
  for (int k = 0; k < 9; k++) {
    dx[k] = a * dNdr[i * 9 + k] + b * dNds[i * 9 + k];
    dy[k] = c * dNdr[i * 9 + k] + d * dNds[i * 9 + k];
  }
  if (mode == 0) {
    work[idx * 3] = dot9(dx, uv);
    work[idx * 3 + 1] = dot9(dy, vv);
    work[idx * 3 + 2] = 0.5 * dot9(dx, vv) + 0.5 * dot9(dy, uv);
  } else {
    work[idx * 3] = dot9(dx, uv);
    work[idx * 3 + 1] = dot9(dy, uv);
  }
}

__global__ void k_ngrad_node(int nout, int nn, const int* __restrict__ ptr,
                             const int* __restrict__ list, const double* __restrict__ work,
                             double* __restrict__ o0, double* __restrict__ o1,
                             double* __restrict__ o2) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, c = 0.0;
  for (int i = ptr[n]; i < ptr[n + 1]; i++) {
    int vv = list[i];
    long o = ((long)(vv >> 5) * 9 + (vv & 31)) * 3;
    s0 += work[o];
    s1 += work[o + 1];
    s2 += work[o + 2];
    c += 1;
  }
  o0[n] = s0 / c;
  o1[n] = s1 / c;
  if (nout == 3) o2[n] = s2 / c;
}

__global__ void k_elem_strain(int nel, const int* __restrict__ icon, const double* __restrict__ u,
                              const double* __restrict__ v, const double* __restrict__ x,
                              const double* __restrict__ y, double* __restrict__ exx,
                              double* __restrict__ eyy, double* __restrict__ exy) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  double dNdr[9], dNdt[9], xe[9], ye[9], ue[9], ve[9], dx[9], dy[9];
  basis_V_dr(0.0, 0.0, dNdr);
  basis_V_dt(0.0, 0.0, dNdt);
  for (int k = 0; k < 9; k++) {
    long n = icon[k * (long)nel + e];
    xe[k] = x[n];
    ye[k] = y[n];
    ue[k] = u[n];
    ve[k] = v[n];
  }
  double a = dot9(dNdr, xe), b = dot9(dNdr, ye), c = dot9(dNdt, xe), d = dot9(dNdt, ye);
  double det = a * d - b * c;
  double i00 = d / det, i01 = -b / det, i10 = -c / det, i11 = a / det;
  for (int k = 0; k < 9; k++) {
    dx[k] = i00 * dNdr[k] + i01 * dNdt[k];
    dy[k] = i10 * dNdr[k] + i11 * dNdt[k];
  }
  exx[e] = dot9(dx, ue);
  eyy[e] = dot9(dy, ve);
  exy[e] = dot9(dx, ve) * 0.5 + dot9(dy, ue) * 0.5;
}

}  // namespace mw

extern "C" {

int mw_nodal_gradients(int mode, int nel, int nn, const int32_t* icon_V, const double* u,
                       const double* v, const double* dNdr_n, const double* dNds_n,
                       const double* jcbi00, const double* jcbi01, const double* jcbi10,
                       const double* jcbi11, const int32_t* inc_ptr, const int32_t* inc_list,
                       double* o0, double* o1, double* o2, double* work, void* stream) {
  using namespace mw;
  hipStream_t st = (hipStream_t)stream;
  if ((mode != 0 && mode != 1) || nel <= 0 || nn <= 0 || !icon_V || !u || (mode == 0 && !v) ||
      !dNdr_n || !dNds_n || !jcbi00 || !jcbi01 || !jcbi10 || !jcbi11 || !inc_ptr || !inc_list ||
      !o0 || !o1 || (mode == 0 && !o2) || !work)
    return fail(MW_EINVAL, "nodal_gradients: bad arguments");
  k_ngrad_elem<<<nblocks(9L * nel, 256), 256, 0, st>>>(mode, nel, icon_V, u, v, dNdr_n, dNds_n,
                                                        jcbi00, jcbi01, jcbi10, jcbi11, work);
  MW_CHECK_LAUNCH();
  k_ngrad_node<<<nblocks(nn, 256), 256, 0, st>>>(mode == 0 ? 3 : 2, nn, inc_ptr, inc_list, work,
                                                  o0, o1, o2);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

int mw_elemental_strain_rate(int nel, const int32_t* icon_V, const double* u, const double* v,
                             const double* x_V, const double* y_V, double* exx, double* eyy,
                             double* exy, void* stream) {
  using namespace mw;
  if (nel <= 0 || !icon_V || !u || !v || !x_V || !y_V || !exx || !eyy || !exy)
    return fail(MW_EINVAL, "elemental_strain_rate: bad arguments");
  k_elem_strain<<<nblocks(nel, 256), 256, 0, (hipStream_t)stream>>>(nel, icon_V, u, v, x_V, y_V,
                                                                     exx, eyy, exy);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

}  // extern "C"
