// COO index generation, node/dof incidence lists and element assembly of the
// Q2xQ1 Stokes and Q2 Crank-Nicolson energy matrices.
//
// Reference behaviour restated here:
//   local_to_globalV and the II/JJ fills       meeuuw.py:874-1015
//   build_matrix_stokes (plane strain)         build_matrix_stokes.py:99-276
//   build_matrix_energy                        build_matrix_energy.py:55-139
//
// Element kernels stage the per-qpoint shape-function derivatives of EPB
// elements in LDS and let every thread produce whole output slots of the COO
// value array, so VV is written once, in order, fully coalesced.  Right-hand
// sides are summed without atomics: each element writes its local vector to a
// workspace and a dof-parallel pass adds the contributions of the elements
// that touch a dof in increasing element order -- the same order in which the
// reference's element loop accumulates them.

#include "mw_common.h"
#include "mw_capi_impl.h"
#include "mw_solver.h"

namespace mw {

// ---------------------------------------------------------------------------
// COO index arrays
// ---------------------------------------------------------------------------
__global__ void k_local_to_global(int nel, const int* __restrict__ icon_V, int* __restrict__ l2g) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= 18L * nel) return;
  int ikk = (int)(idx / nel), iel = (int)(idx % nel);
  l2g[idx] = 2 * icon_V[(ikk >> 1) * (long)nel + iel] + (ikk & 1);
}

// Stokes: per element, for ikk in 0..17: 18 K entries, then for each of the 4
// pressure dofs the pair (m1, P+NV), (P+NV, m1)  -> 26 entries per ikk.
__global__ void k_coo_stokes(int nel, const int* __restrict__ l2g, const int* __restrict__ icon_P,
                             int Nfem_V, int* __restrict__ II, int* __restrict__ JJ) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= 468L * nel) return;
  int iel = (int)(idx / 468), s = (int)(idx % 468);
  int row = s / 26, col = s % 26;
  int m1 = l2g[row * (long)nel + iel];
  int a, b;
  if (col < 18) {
    a = m1;
    b = l2g[col * (long)nel + iel];
  } else {
    int g = col - 18;
    int m2 = icon_P[(g >> 1) * (long)nel + iel] + Nfem_V;
    if (g & 1) { a = m2; b = m1; } else { a = m1; b = m2; }
  }
  II[idx] = a;
  JJ[idx] = b;
}

// generic (r x c) outer-product fill: II = rowmap[i], JJ = colmap[j], i outer
__global__ void k_coo_outer(int nel, int nr, int nc, const int* __restrict__ rmap, int roff,
                            const int* __restrict__ cmap, int coff, int* __restrict__ II,
                            int* __restrict__ JJ) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long per = (long)nr * nc;
  if (idx >= per * nel) return;
  int iel = (int)(idx / per), s = (int)(idx % per);
  int i = s / nc, j = s % nc;
  II[idx] = rmap[i * (long)nel + iel] + roff;
  JJ[idx] = cmap[j * (long)nel + iel] + coff;
}

int local_to_global(int nel, const int* icon_V, int* l2g, hipStream_t st) {
  if (nel <= 0 || !icon_V || !l2g) return fail(MW_EINVAL, "local_to_global: bad arguments");
  k_local_to_global<<<nblocks(18L * nel, 256), 256, 0, st>>>(nel, icon_V, l2g);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

int coo_stokes(int nel, const int* l2g, const int* icon_P, int Nfem_V, int* II, int* JJ,
               hipStream_t st) {
  if (nel <= 0 || !l2g || !icon_P || !II || !JJ) return fail(MW_EINVAL, "coo_stokes: bad arguments");
  k_coo_stokes<<<nblocks(468L * nel, 256), 256, 0, st>>>(nel, l2g, icon_P, Nfem_V, II, JJ);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

int coo_outer(int nel, int nr, int nc, const int* rmap, const int* cmap, int* II, int* JJ,
              hipStream_t st) {
  if (nel <= 0 || nr <= 0 || nc <= 0 || !rmap || !cmap || !II || !JJ)
    return fail(MW_EINVAL, "coo_outer: bad arguments");
  k_coo_outer<<<nblocks((long)nr * nc * nel, 256), 256, 0, st>>>(nel, nr, nc, rmap, 0, cmap, 0, II,
                                                                  JJ);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// ---------------------------------------------------------------------------
// incidence lists: for every node n of a connectivity (m, nel), the sorted
// list of (iel << 5 | k) with icon[k, iel] == n.
// ---------------------------------------------------------------------------
__global__ void k_inc_count(int nel, int m, const int* __restrict__ icon, int nn,
                            int* __restrict__ cnt, int* __restrict__ bad) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= (long)m * nel) return;
  int n = icon[idx];
  if (n < 0 || n >= nn) { atomicAdd(bad, 1); return; }
  atomicAdd(&cnt[n], 1);
}

__global__ void k_inc_fill(int nel, int m, const int* __restrict__ icon, int nn,
                           const int* __restrict__ ptr, int* __restrict__ cur,
                           int* __restrict__ list) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= (long)m * nel) return;
  int k = (int)(idx / nel), iel = (int)(idx % nel);
  int n = icon[idx];
  if (n < 0 || n >= nn) return;
  int pos = atomicAdd(&cur[n], 1);
  list[ptr[n] + pos] = (iel << 5) | k;
}

__global__ void k_inc_sort(int nn, const int* __restrict__ ptr, int* __restrict__ list) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  int b = ptr[n], e = ptr[n + 1];
  for (int i = b + 1; i < e; i++) {
    int v = list[i], j = i - 1;
    while (j >= b && list[j] > v) { list[j + 1] = list[j]; j--; }
    list[j + 1] = v;
  }
}

// Exclusive prefix sum of n ints over tiles of 4096 (a 1024-thread block per
// tile, four consecutive values per thread): the first kernel sums each tile,
// the second adds the sums of the tiles before its own (read by the whole
// block) to the scan of its tile.  Two launches, every load coalesced.
constexpr int SCAN_TILE = 4096;

__device__ __forceinline__ int block_scan_1024(int v, int* wsum, int& total) {
  // exclusive scan of v over a 1024-thread block; total: the block's sum
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  if (w == 0) {
    int u = lane < 16 ? wsum[lane] : 0;
    for (int o = 1; o < 16; o <<= 1) {
      const int y = __shfl_up(u, o);
      if (lane >= o) u += y;
    }
    if (lane < 16) wsum[lane] = u;
  }
  __syncthreads();
  total = wsum[15];
  return x - v + (w > 0 ? wsum[w - 1] : 0);
}

__device__ __forceinline__ void tile_load(int n, const int* __restrict__ in, long i0, int* v) {
#pragma unroll
  for (int q = 0; q < 4; q++) v[q] = (i0 + q < n) ? in[i0 + q] : 0;
}

__global__ __launch_bounds__(1024) void k_scan_tile_sums(int n, const int* __restrict__ in,
                                                         int* __restrict__ tsum) {
  __shared__ int wsum[16];
  int v[4];
  tile_load(n, in, (long)blockIdx.x * SCAN_TILE + 4L * threadIdx.x, v);
  int total;
  (void)block_scan_1024(v[0] + v[1] + v[2] + v[3], wsum, total);
  if (threadIdx.x == 0) tsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_tiles(int n, const int* __restrict__ in,
                                                     const int* __restrict__ tsum,
                                                     int* __restrict__ out) {
  __shared__ int wsum[16];
  __shared__ int sbase;
  const int b = blockIdx.x;
  // sum of the tiles before this one
  int pre = 0;
  for (int i = threadIdx.x; i < b; i += 1024) pre += tsum[i];
  int total;
  (void)block_scan_1024(pre, wsum, total);
  if (threadIdx.x == 0) sbase = total;
  __syncthreads();  // wsum is reused below
  const int base = sbase;
  int v[4];
  const long i0 = (long)b * SCAN_TILE + 4L * threadIdx.x;
  tile_load(n, in, i0, v);
  int run = base + block_scan_1024(v[0] + v[1] + v[2] + v[3], wsum, total);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    if (i0 + q < n) out[i0 + q] = run;
    run += v[q];
  }
  if (b == (int)gridDim.x - 1 && threadIdx.x == 0) out[n] = base + total;
}

int exclusive_scan_i32(int n, const int* in, int* out, int* tsum, hipStream_t st) {
  if (n < 0 || !in || !out || !tsum) return fail(MW_EINVAL, "exclusive_scan: bad arguments");
  const int tiles = n > 0 ? (n + SCAN_TILE - 1) / SCAN_TILE : 1;
  k_scan_tile_sums<<<tiles, 1024, 0, st>>>(n, in, tsum);
  MW_CHECK_LAUNCH();
  k_scan_tiles<<<tiles, 1024, 0, st>>>(n, in, tsum, out);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

size_t scan_tile_sums_ints(int n) { return (size_t)(n + SCAN_TILE - 1) / SCAN_TILE + 1; }

size_t incidence_temp_bytes(int nn) { return 256; }

// ptr: nn+1 ints, list: m*nel ints, work: 2*(nn+1) ints + incidence_temp_bytes
int incidence_build(int nel, int m, const int* icon, int nn, int* ptr, int* list, void* work,
                    hipStream_t st) {
  if (nel <= 0 || m <= 0 || m > 32 || nn <= 0 || !icon || !ptr || !list || !work)
    return fail(MW_EINVAL, "incidence_build: bad arguments");
  if ((long)nel >= (1L << 26)) return fail(MW_EINVAL, "incidence_build: nel too large");
  int* cnt = (int*)(((uintptr_t)work + 15) & ~(uintptr_t)15);  // 16-byte aligned for the scan
  int* cur = cnt + (nn + 1);
  int* bad = cur + (nn + 1) - 1;  // reuse last slot of cur: never a node index
  MW_HIP(hipMemsetAsync(cnt, 0, sizeof(int) * 2 * (nn + 1), st));
  k_inc_count<<<nblocks((long)m * nel, 256), 256, 0, st>>>(nel, m, icon, nn, cnt, bad);
  MW_CHECK_LAUNCH();
  RC(exclusive_scan_i32(nn, cnt, ptr, cur, st));  // cur: the tile sums until it is cleared
  int hbad = 0;
  MW_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st));
  MW_HIP(hipStreamSynchronize(st));
  if (hbad) return fail(MW_EINVAL, "incidence_build: connectivity entry outside [0, nn)");
  MW_HIP(hipMemsetAsync(cur, 0, sizeof(int) * (nn + 1), st));
  k_inc_fill<<<nblocks((long)m * nel, 256), 256, 0, st>>>(nel, m, icon, nn, ptr, cur, list);
  MW_CHECK_LAUNCH();
  k_inc_sort<<<nblocks(nn, 256), 256, 0, st>>>(nn, ptr, list);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// out[n] = sum over (iel,k) in inc(n), increasing iel, of vel[iel*m + k]
__global__ void k_gather_sum(int nn, int m, const int* __restrict__ ptr,
                             const int* __restrict__ list, const double* __restrict__ vel,
                             double* __restrict__ out) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  int ent[INC4];
  const int c = inc_entries4(ptr, list, n, ent);
  double s = 0.0;
  if (c <= INC4) {  // every load first, then the sum in list order
    double w[INC4];
#pragma unroll
    for (int q = 0; q < INC4; q++)
      w[q] = q < c ? vel[(long)(ent[q] >> 5) * m + (ent[q] & 31)] : 0.0;
#pragma unroll
    for (int q = 0; q < INC4; q++)
      if (q < c) s += w[q];
  } else {
    for (int i = ptr[n]; i < ptr[n + 1]; i++) {
      int v = list[i];
      s += vel[(long)(v >> 5) * m + (v & 31)];
    }
  }
  out[n] = s;
}

int gather_sum(int nn, int m, const int* ptr, const int* list, const double* vel, double* out,
               hipStream_t st) {
  k_gather_sum<<<nblocks(nn, 256), 256, 0, st>>>(nn, m, ptr, list, vel, out);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// ---------------------------------------------------------------------------
// Stokes element kernel
// ---------------------------------------------------------------------------
constexpr int S_EPB = 4;    // elements per block (matrices and right-hand sides)
constexpr int S_EPB_RHS = 11;  // elements per block for the right-hand sides alone (22 threads each)
constexpr int S_BS = 256;   // threads per block

struct StokesIn {
  int nel, Nfem_V, Nfem_P, blocks, rhs_only;
  const int *icon_P, *l2g;
  const double *rhoq, *etaq, *JxWq, *gxq, *gyq;
  const double *N_V, *N_P, *dNdr_V, *dNdt_V, *dNdr_P, *dNdt_P;
  const double *j00, *j01, *j10, *j11, *eta_e;
  double gscale;  // eta_ref / L_ref
  const uint8_t* bc_fix;
  const double* bc_val;
  double *VV, *VV_K, *VV_G, *VV_GT, *VV_M, *VV_Me, *VV_H;
  double *fel, *hel;  // workspace: nel*18, nel*4
  const double* theta;
  const uint8_t *bot_el, *top_el;
  int bot_fs, top_fs;
  int axi;            // axisymmetric branch
  const double* xq;   // radius at the quadrature points (axi)
};

// AXI: the axisymmetric branch (build_matrix_stokes.py:108-125): B rows
// (du/dx, u/x, dw/dz, du/dz + dw/dx), C = diag(2, 2, 2, 1), every term weighted
// by 2 pi x at the quadrature point, no M or H blocks (written as zeros).
template <int EPB, bool AXI>
__global__ void __launch_bounds__(S_BS) k_stokes_elements(StokesIn a) {
  __shared__ double sNV[81], sNP[36];
  __shared__ double sdx[EPB][81], sdz[EPB][81];  // [q*9 + node]
  __shared__ double spx[EPB][36], spz[EPB][36];  // [q*4 + pnode]
  __shared__ double seta[EPB][9], sjw[EPB][9], srx[EPB][9], sry[EPB][9];
  __shared__ double sxq[AXI ? EPB : 1][9];
  __shared__ uint8_t sfix[EPB][18];
  __shared__ double sval[EPB][18];

  const int tid = threadIdx.x;
  const int e0 = blockIdx.x * EPB;
  const int nel = a.nel;
  for (int i = tid; i < 81; i += S_BS) sNV[i] = a.N_V[i];
  for (int i = tid; i < 36; i += S_BS) sNP[i] = a.N_P[i];
  for (int i = tid; i < EPB * 81; i += S_BS) {
    int le = i / 81, s = i % 81, q = s / 9, k = s % 9;
    int iel = e0 + le;
    if (iel >= nel) continue;
    long qi = (long)iel * 9 + q;
    double j00 = a.j00[qi], j01 = a.j01[qi], j10 = a.j10[qi], j11 = a.j11[qi];
    sdx[le][s] = j00 * a.dNdr_V[q * 9 + k] + j01 * a.dNdt_V[q * 9 + k];
    sdz[le][s] = j10 * a.dNdr_V[q * 9 + k] + j11 * a.dNdt_V[q * 9 + k];
    if (k < 4) {
      spx[le][q * 4 + k] = j00 * a.dNdr_P[q * 4 + k] + j01 * a.dNdt_P[q * 4 + k];
      spz[le][q * 4 + k] = j10 * a.dNdr_P[q * 4 + k] + j11 * a.dNdt_P[q * 4 + k];
    }
    if (k == 0) {
      seta[le][q] = a.etaq[qi];
      sjw[le][q] = a.JxWq[qi];
      srx[le][q] = a.gxq[qi];
      sry[le][q] = a.gyq[qi];
      if constexpr (AXI) sxq[le][q] = a.xq[qi];
    }
  }
  for (int i = tid; i < EPB * 18; i += S_BS) {
    int le = i / 18, d = i % 18, iel = e0 + le;
    if (iel >= nel) continue;
    int m1 = a.l2g[d * (long)nel + iel];
    sfix[le][d] = a.bc_fix[m1];
    sval[le][d] = a.bc_val[m1];
  }
  __syncthreads();
  // B entry (row kk, column i) at qpoint q of local element le
  auto Bent = [&](int le, int q, int kk, int i) -> double {
    int nd = i >> 1, c = i & 1;
    if constexpr (AXI) {
      if (kk == 0) return c == 0 ? sdx[le][q * 9 + nd] : 0.0;
      if (kk == 1) return c == 0 ? sNV[q * 9 + nd] / sxq[le][q] : 0.0;
      if (kk == 2) return c == 1 ? sdz[le][q * 9 + nd] : 0.0;
      return c == 0 ? sdz[le][q * 9 + nd] : sdx[le][q * 9 + nd];
    }
    if (kk == 0) return c == 0 ? sdx[le][q * 9 + nd] : 0.0;
    if (kk == 1) return c == 1 ? sdz[le][q * 9 + nd] : 0.0;
    return c == 0 ? sdz[le][q * 9 + nd] : sdx[le][q * 9 + nd];
  };
  auto coeffq = [&](int le, int q) -> double {
    if constexpr (AXI) return 2.0 * M_PI * sxq[le][q];
    return 1.0;
  };
  auto Kval = [&](int le, int i, int j) -> double {
    double s = 0.0;
    for (int q = 0; q < 9; q++) {
      if constexpr (AXI) {
        double btcb = Bent(le, q, 0, i) * (2.0 * Bent(le, q, 0, j)) +
                      Bent(le, q, 1, i) * (2.0 * Bent(le, q, 1, j)) +
                      Bent(le, q, 2, i) * (2.0 * Bent(le, q, 2, j)) +
                      Bent(le, q, 3, i) * (1.0 * Bent(le, q, 3, j));
        s += btcb * seta[le][q] * sjw[le][q] * coeffq(le, q);
      } else {
        double btcb = Bent(le, q, 0, i) * (2.0 * Bent(le, q, 0, j)) +
                      Bent(le, q, 1, i) * (2.0 * Bent(le, q, 1, j)) +
                      Bent(le, q, 2, i) * (1.0 * Bent(le, q, 2, j));
        s += btcb * seta[le][q] * sjw[le][q];
      }
    }
    return s;
  };
  auto Gval = [&](int le, int i, int p) -> double {
    double s = 0.0;
    for (int q = 0; q < 9; q++) {
      double np_ = sNP[q * 4 + p];
      if constexpr (AXI) {
        double bn = Bent(le, q, 0, i) * np_ + Bent(le, q, 1, i) * np_ +
                    Bent(le, q, 2, i) * np_ + Bent(le, q, 3, i) * 0.0;
        s -= bn * sjw[le][q] * coeffq(le, q);
      } else {
        double bn = Bent(le, q, 0, i) * np_ + Bent(le, q, 1, i) * np_ + Bent(le, q, 2, i) * 0.0;
        s -= bn * sjw[le][q];
      }
    }
    return s * a.gscale;
  };
  // final (Dirichlet-masked) K and G
  auto Kfin = [&](int le, int i, int j) -> double {
    if (sfix[le][i] || sfix[le][j]) return (i == j) ? Kval(le, i, i) : 0.0;
    return Kval(le, i, j);
  };
  auto Gfin = [&](int le, int i, int p) -> double { return sfix[le][i] ? 0.0 : Gval(le, i, p); };

  if (a.rhs_only) {
    // right-hand sides only
  } else if (!a.blocks) {
    for (int i = tid; i < EPB * 468; i += S_BS) {
      int le = i / 468, s = i % 468, iel = e0 + le;
      if (iel >= nel) continue;
      int row = s / 26, col = s % 26;
      double v = col < 18 ? Kfin(le, row, col) : Gfin(le, row, (col - 18) >> 1);
      a.VV[(long)iel * 468 + s] = v;
    }
  } else {
    for (int i = tid; i < EPB * 324; i += S_BS) {
      int le = i / 324, s = i % 324, iel = e0 + le;
      if (iel >= nel) continue;
      a.VV_K[(long)iel * 324 + s] = Kfin(le, s / 18, s % 18);
    }
    for (int i = tid; i < EPB * 72; i += S_BS) {
      int le = i / 72, s = i % 72, iel = e0 + le;
      if (iel >= nel) continue;
      a.VV_G[(long)iel * 72 + s] = Gfin(le, s / 4, s % 4);
      int p = s / 18, d = s % 18;
      a.VV_GT[(long)iel * 72 + s] = Gfin(le, d, p);
      // H[p][d] = sum_q (dPdx*N + dPdz*0) * JxW for d even, (dPdx*0 + dPdz*N)*JxW for d odd
      double h = 0.0;
      for (int q = 0; !AXI && q < 9; q++) {
        double nv = sNV[q * 9 + (d >> 1)];
        double bb0 = (d & 1) ? 0.0 : nv, bb1 = (d & 1) ? nv : 0.0;
        h += (spx[le][q * 4 + p] * bb0 + spz[le][q * 4 + p] * bb1) * sjw[le][q];
      }
      a.VV_H[(long)iel * 72 + s] = h;
    }
    for (int i = tid; i < EPB * 16; i += S_BS) {
      int le = i / 16, s = i % 16, iel = e0 + le;
      if (iel >= nel) continue;
      int p = s / 4, r = s % 4;
      double mm = 0.0;
      for (int q = 0; !AXI && q < 9; q++) mm += sNP[q * 4 + p] * sNP[q * 4 + r] * sjw[le][q];
      a.VV_M[(long)iel * 16 + s] = mm;
      a.VV_Me[(long)iel * 16 + s] = mm / a.eta_e[iel];
    }
  }
  // element right-hand sides with Dirichlet lifting (progressive, ascending ikk)
  for (int i = tid; i < EPB * 22; i += S_BS) {
    int le = i / 22, s = i % 22, iel = e0 + le;
    if (iel >= nel) continue;
    if (s < 18) {
      int nd = s >> 1, c = s & 1;
      double f = 0.0;
      for (int q = 0; q < 9; q++) {
        long qi = (long)iel * 9 + q;
        double g = c ? sry[le][q] : srx[le][q];
        if constexpr (AXI)
          f += sNV[q * 9 + nd] * sjw[le][q] * coeffq(le, q) * a.rhoq[qi] * g;
        else
          f += sNV[q * 9 + nd] * sjw[le][q] * a.rhoq[qi] * g;
      }
      if (sfix[le][s]) {
        f = Kval(le, s, s) * sval[le][s];
      } else {
        for (int j = 0; j < 18; j++)
          if (sfix[le][j]) f -= Kval(le, s, j) * sval[le][j];
      }
      a.fel[(long)iel * 18 + s] = f;
    } else {
      int p = s - 18;
      double h = 0.0;
      for (int j = 0; j < 18; j++)
        if (sfix[le][j]) h -= Gval(le, j, p) * sval[le][j];
      a.hel[(long)iel * 4 + p] = h;
    }
  }
}

// Element matrices of build_matrix_stokes with one thread per (element, row
// of K): 7 elements per 128-thread block share the staged derivatives, each
// thread forms its row's 18 K entries, 4 G entries and right-hand side entry
// and writes them contiguously.  The strain-operator rows are specialised to
// the components of the row and column (B's structural zeros are not
// multiplied); every remaining product and sum is the one the full
// B^T C B expression evaluates, in the same order, so the results are those of
// k_stokes_elements bit for bit for finite inputs (a dropped term is an exact
// zero, which can change only the sign of a zero partial sum, and every
// accumulator starts at +0).
constexpr int SR_EPB = 7;  // 7 * 18 = 126 of 128 threads
constexpr int SR_BS = 128;

template <bool AXI>
__global__ void __launch_bounds__(SR_BS, 4) k_stokes_rows(StokesIn a) {
  __shared__ double sNV[81], sNP[36];
  __shared__ double sdx[SR_EPB][81], sdz[SR_EPB][81];  // [q*9 + node]
  __shared__ double spx[SR_EPB][36], spz[SR_EPB][36];  // [q*4 + pnode]
  __shared__ double seta[SR_EPB][9], sjw[SR_EPB][9], srx[SR_EPB][9], sry[SR_EPB][9];
  __shared__ double sxq[AXI ? SR_EPB : 1][9];
  __shared__ uint8_t sfix[SR_EPB][18];
  __shared__ double sval[SR_EPB][18];
  __shared__ double sG[SR_EPB][18][4];  // unmasked G rows, for h and G^T

  const int tid = threadIdx.x;
  const int e0 = blockIdx.x * SR_EPB;
  const int nel = a.nel;
  for (int i = tid; i < 81; i += SR_BS) sNV[i] = a.N_V[i];
  for (int i = tid; i < 36; i += SR_BS) sNP[i] = a.N_P[i];
  for (int i = tid; i < SR_EPB * 81; i += SR_BS) {
    const int le = i / 81, s = i % 81, q = s / 9, k = s % 9;
    const int iel = e0 + le;
    if (iel >= nel) continue;
    const long qi = (long)iel * 9 + q;
    const double j00 = a.j00[qi], j01 = a.j01[qi], j10 = a.j10[qi], j11 = a.j11[qi];
    sdx[le][s] = j00 * a.dNdr_V[q * 9 + k] + j01 * a.dNdt_V[q * 9 + k];
    sdz[le][s] = j10 * a.dNdr_V[q * 9 + k] + j11 * a.dNdt_V[q * 9 + k];
    if (k < 4) {
      spx[le][q * 4 + k] = j00 * a.dNdr_P[q * 4 + k] + j01 * a.dNdt_P[q * 4 + k];
      spz[le][q * 4 + k] = j10 * a.dNdr_P[q * 4 + k] + j11 * a.dNdt_P[q * 4 + k];
    }
    if (k == 0) {
      seta[le][q] = a.etaq[qi];
      sjw[le][q] = a.JxWq[qi];
      srx[le][q] = a.gxq[qi];
      sry[le][q] = a.gyq[qi];
      if constexpr (AXI) sxq[le][q] = a.xq[qi];
    }
  }
  for (int i = tid; i < SR_EPB * 18; i += SR_BS) {
    const int le = i / 18, d = i % 18, iel = e0 + le;
    if (iel >= nel) continue;
    const int m1 = a.l2g[d * (long)nel + iel];
    sfix[le][d] = a.bc_fix[m1];
    sval[le][d] = a.bc_val[m1];
  }
  __syncthreads();

  // the first wavefront takes the even (x) rows, the second the odd (z) rows,
  // so the component branches below are uniform within each wavefront
  const int half = tid >> 6, t = tid & 63;
  const int le = t / 9, i = 2 * (t % 9) + half;
  const int iel = e0 + le;
  const bool active = t < SR_EPB * 9 && iel < nel;
  if (active) {
    const int ni = i >> 1, ci = i & 1;
    // K(i, j) and G(i, p) before the Dirichlet masking, all columns at once:
    // the quadrature points run outermost so the 18 + 4 sums are independent
    // chains (each still adds its points in ascending order)
    double kr[18], g[4];
#pragma unroll
    for (int j = 0; j < 18; j++) kr[j] = 0.0;
#pragma unroll
    for (int p = 0; p < 4; p++) g[p] = 0.0;
#pragma unroll 1
    for (int q = 0; q < 9; q++) {
      const double e = seta[le][q], w = sjw[le][q];
      const double dxq = sdx[le][q * 9 + ni], dzq = sdz[le][q * 9 + ni];
      double cqq = 1.0, hiq = 0.0;
      if constexpr (AXI) {
        cqq = 2.0 * M_PI * sxq[le][q];
        hiq = sNV[q * 9 + ni] / sxq[le][q];
      }
#pragma unroll
      for (int nj = 0; nj < 9; nj++) {
        const double dxj = sdx[le][q * 9 + nj], dzj = sdz[le][q * 9 + nj];
        double b0, b1;  // B^T C B for the column pair (2 nj, 2 nj + 1)
        if constexpr (AXI) {
          const double hj = sNV[q * 9 + nj] / sxq[le][q];
          if (ci == 0) {
            b0 = (dxq * (2.0 * dxj) + hiq * (2.0 * hj)) + dzq * (1.0 * dzj);
            b1 = dzq * (1.0 * dxj);
          } else {
            b0 = dxq * (1.0 * dzj);
            b1 = dzq * (2.0 * dzj) + dxq * (1.0 * dxj);
          }
          kr[2 * nj] += b0 * e * w * cqq;
          kr[2 * nj + 1] += b1 * e * w * cqq;
        } else {
          if (ci == 0) {
            b0 = dxq * (2.0 * dxj) + dzq * (1.0 * dzj);
            b1 = dzq * (1.0 * dxj);
          } else {
            b0 = dxq * (1.0 * dzj);
            b1 = dzq * (2.0 * dzj) + dxq * (1.0 * dxj);
          }
          kr[2 * nj] += b0 * e * w;
          kr[2 * nj + 1] += b1 * e * w;
        }
      }
#pragma unroll
      for (int p = 0; p < 4; p++) {
        const double np_ = sNP[q * 4 + p];
        if constexpr (AXI) {
          const double bn = ci == 0 ? dxq * np_ + hiq * np_ : dzq * np_;
          g[p] -= bn * w * cqq;
        } else {
          const double bn = ci == 0 ? dxq * np_ : dzq * np_;
          g[p] -= bn * w;
        }
      }
    }
#pragma unroll
    for (int p = 0; p < 4; p++) {
      g[p] = g[p] * a.gscale;
      sG[le][i][p] = g[p];
    }
    // gravity term of f
    double f = 0.0;
#pragma unroll
    for (int q = 0; q < 9; q++) {
      const long qi = (long)iel * 9 + q;
      const double gq = ci ? sry[le][q] : srx[le][q];
      if constexpr (AXI)
        f += sNV[q * 9 + ni] * sjw[le][q] * (2.0 * M_PI * sxq[le][q]) * a.rhoq[qi] * gq;
      else
        f += sNV[q * 9 + ni] * sjw[le][q] * a.rhoq[qi] * gq;
    }
    const bool fixi = sfix[le][i];
    double fl = f;  // Dirichlet lifting, ascending columns
#pragma unroll
    for (int j = 0; j < 18; j++) {
      const double k = kr[j];
      const bool fixj = sfix[le][j];
      kr[j] = (fixi || fixj) ? (i == j ? k : 0.0) : k;
      if (!fixi && fixj) fl -= k * sval[le][j];
      if (fixi && j == i) f = k * sval[le][i];
    }
    double* krow = a.blocks ? a.VV_K + (long)iel * 324 + i * 18 : a.VV + (long)iel * 468 + i * 26;
#pragma unroll
    for (int j = 0; j < 18; j++) krow[j] = kr[j];
    a.fel[(long)iel * 18 + i] = fixi ? f : fl;
    if (a.blocks) {
#pragma unroll
      for (int p = 0; p < 4; p++) {
        const double gm = fixi ? 0.0 : g[p];
        a.VV_G[(long)iel * 72 + i * 4 + p] = gm;
        a.VV_GT[(long)iel * 72 + p * 18 + i] = gm;
        // H[p][i] = sum_q (dPdx N for an x dof, dPdz N for a z dof) JxW
        double h = 0.0;
        for (int q = 0; !AXI && q < 9; q++) {
          const double nv = sNV[q * 9 + ni];
          const double bb0 = ci ? 0.0 : nv, bb1 = ci ? nv : 0.0;
          h += (spx[le][q * 4 + p] * bb0 + spz[le][q * 4 + p] * bb1) * sjw[le][q];
        }
        a.VV_H[(long)iel * 72 + p * 18 + i] = h;
      }
      if (i < 16) {
        const int p = i / 4, r = i % 4;
        double mm = 0.0;
        for (int q = 0; !AXI && q < 9; q++) mm += sNP[q * 4 + p] * sNP[q * 4 + r] * sjw[le][q];
        a.VV_M[(long)iel * 16 + i] = mm;
        a.VV_Me[(long)iel * 16 + i] = mm / a.eta_e[iel];
      }
    } else {
#pragma unroll
      for (int p = 0; p < 4; p++) {
        const double gm = fixi ? 0.0 : g[p];
        krow[18 + 2 * p] = gm;
        krow[19 + 2 * p] = gm;
      }
    }
  }
  __syncthreads();
  if (active && i < 4) {  // pressure right-hand side: -sum over fixed rows, ascending
    double h = 0.0;
    for (int j = 0; j < 18; j++)
      if (sfix[le][j]) h -= sG[le][j][i] * sval[le][j];
    a.hel[(long)iel * 4 + i] = h;
  }
}

// Free slip on a curved boundary (build_matrix_stokes.py:163-203), one
// 64-thread block per element; blocks of elements without a free-slip side
// return at once.  The element's K, G and f are formed again in LDS with the
// expressions of k_stokes_elements, then at each boundary node whose two dofs
// are free (top nodes 2, 3, 6, then bottom nodes 0, 1, 4) the system is rotated
// to (normal, tangential) components, K <- RR K RR^T, G <- RR G, f <- RR f
// (RR: identity but RR[n][n] = RR[t][t] = cos(theta), RR[n][t] = sin(theta),
// RR[t][n] = -sin(theta)), and the normal dof is eliminated; the Dirichlet
// rows follow in ascending local order as in the element loop.  The element's
// outputs overwrite what k_stokes_elements wrote for it.  The node of local
// velocity dof 2i is l2g[2i] / 2 (local_to_globalV = ndof_V * icon_V + k,
// meeuuw.py:875-880).
constexpr int FS_BS = 64;

__global__ void __launch_bounds__(FS_BS) k_stokes_free_slip(StokesIn a) {
  const int iel = blockIdx.x, nel = a.nel, tid = threadIdx.x;
  const bool top = a.top_fs && a.top_el[iel], bot = a.bot_fs && a.bot_el[iel];
  if (!top && !bot) return;
  __shared__ double sNV[81], sNP[36], sdx[81], sdz[81], seta[9], sjw[9], srx[9], sry[9], sxq[9];
  __shared__ double K[18][18], G[18][4], f[18], h[4], sval[18];
  __shared__ uint8_t sfix[18];
  __shared__ int snode[18];
  for (int i = tid; i < 81; i += FS_BS) {
    sNV[i] = a.N_V[i];
    const int q = i / 9, k = i % 9;
    const long qi = (long)iel * 9 + q;
    const double j00 = a.j00[qi], j01 = a.j01[qi], j10 = a.j10[qi], j11 = a.j11[qi];
    sdx[i] = j00 * a.dNdr_V[q * 9 + k] + j01 * a.dNdt_V[q * 9 + k];
    sdz[i] = j10 * a.dNdr_V[q * 9 + k] + j11 * a.dNdt_V[q * 9 + k];
  }
  if (tid < 36) sNP[tid] = a.N_P[tid];
  if (tid < 9) {
    const long qi = (long)iel * 9 + tid;
    seta[tid] = a.etaq[qi];
    sjw[tid] = a.JxWq[qi];
    srx[tid] = a.gxq[qi];
    sry[tid] = a.gyq[qi];
    sxq[tid] = a.axi ? a.xq[qi] : 1.0;
  }
  if (tid < 18) {
    const int m1 = a.l2g[tid * (long)nel + iel];
    sfix[tid] = a.bc_fix[m1];
    sval[tid] = a.bc_val[m1];
    snode[tid] = m1 >> 1;
  }
  if (tid < 4) h[tid] = 0.0;
  __syncthreads();
  // plane-strain rows (du/dx, dw/dz, shear) or, axisymmetric, (du/dx, u/x,
  // dw/dz, shear), as in k_stokes_elements
  const bool axi = a.axi != 0;
  auto Bent = [&](int q, int kk, int i) -> double {
    const int nd = i >> 1, c = i & 1;
    if (axi) {
      if (kk == 0) return c == 0 ? sdx[q * 9 + nd] : 0.0;
      if (kk == 1) return c == 0 ? sNV[q * 9 + nd] / sxq[q] : 0.0;
      if (kk == 2) return c == 1 ? sdz[q * 9 + nd] : 0.0;
      return c == 0 ? sdz[q * 9 + nd] : sdx[q * 9 + nd];
    }
    if (kk == 0) return c == 0 ? sdx[q * 9 + nd] : 0.0;
    if (kk == 1) return c == 1 ? sdz[q * 9 + nd] : 0.0;
    return c == 0 ? sdz[q * 9 + nd] : sdx[q * 9 + nd];
  };
  for (int e = tid; e < 324; e += FS_BS) {
    const int i = e / 18, j = e % 18;
    double s = 0.0;
    for (int q = 0; q < 9; q++) {
      if (axi) {
        const double btcb = Bent(q, 0, i) * (2.0 * Bent(q, 0, j)) +
                            Bent(q, 1, i) * (2.0 * Bent(q, 1, j)) +
                            Bent(q, 2, i) * (2.0 * Bent(q, 2, j)) +
                            Bent(q, 3, i) * (1.0 * Bent(q, 3, j));
        s += btcb * seta[q] * sjw[q] * (2.0 * M_PI * sxq[q]);
      } else {
        const double btcb = Bent(q, 0, i) * (2.0 * Bent(q, 0, j)) +
                            Bent(q, 1, i) * (2.0 * Bent(q, 1, j)) +
                            Bent(q, 2, i) * (1.0 * Bent(q, 2, j));
        s += btcb * seta[q] * sjw[q];
      }
    }
    K[i][j] = s;
  }
  for (int e = tid; e < 72; e += FS_BS) {
    const int i = e / 4, p = e % 4;
    double s = 0.0;
    for (int q = 0; q < 9; q++) {
      const double np_ = sNP[q * 4 + p];
      if (axi) {
        const double bn = Bent(q, 0, i) * np_ + Bent(q, 1, i) * np_ + Bent(q, 2, i) * np_ +
                          Bent(q, 3, i) * 0.0;
        s -= bn * sjw[q] * (2.0 * M_PI * sxq[q]);
      } else {
        const double bn = Bent(q, 0, i) * np_ + Bent(q, 1, i) * np_ + Bent(q, 2, i) * 0.0;
        s -= bn * sjw[q];
      }
    }
    G[i][p] = s * a.gscale;
  }
  if (tid < 18) {
    const int nd = tid >> 1, c = tid & 1;
    double fv = 0.0;
    for (int q = 0; q < 9; q++) {
      const long qi = (long)iel * 9 + q;
      if (axi)
        fv += sNV[q * 9 + nd] * sjw[q] * (2.0 * M_PI * sxq[q]) * a.rhoq[qi] * (c ? sry[q] : srx[q]);
      else
        fv += sNV[q * 9 + nd] * sjw[q] * a.rhoq[qi] * (c ? sry[q] : srx[q]);
    }
    f[tid] = fv;
  }
  __syncthreads();
  for (int side = 0; side < 2; side++) {
    if (side == 0 ? !top : !bot) continue;
    for (int k = 0; k < 3; k++) {
      const int i = side == 0 ? (k == 2 ? 6 : 2 + k) : (k == 2 ? 4 : k);
      const int n = 2 * i, t = n + 1;
      if (sfix[n] || sfix[t]) continue;
      const double th = a.theta[snode[n]];
      const double c = cos(th), s = sin(th);
      if (tid < 18) {  // K RR^T: columns n and t of row tid
        const double kn = K[tid][n], kt = K[tid][t];
        K[tid][n] = kn * c + kt * s;
        K[tid][t] = kn * (-s) + kt * c;
      }
      __syncthreads();
      if (tid < 18) {  // RR (K RR^T): rows n and t of column tid
        const double rn = K[n][tid], rt = K[t][tid];
        K[n][tid] = c * rn + s * rt;
        K[t][tid] = (-s) * rn + c * rt;
      } else if (tid < 22) {
        const int j = tid - 18;
        const double gn = G[n][j], gt = G[t][j];
        G[n][j] = c * gn + s * gt;
        G[t][j] = (-s) * gn + c * gt;
      } else if (tid == 22) {
        const double fn = f[n], ft = f[t];
        f[n] = c * fn + s * ft;
        f[t] = (-s) * fn + c * ft;
      }
      __syncthreads();
      const double kref = K[n][n];
      __syncthreads();
      if (tid < 18) {
        K[n][tid] = 0.0;
        K[tid][n] = 0.0;
      } else if (tid < 22) {
        G[n][tid - 18] = 0.0;
      }
      __syncthreads();
      if (tid == 0) {
        K[n][n] = kref;
        f[n] = 0.0;
      }
      __syncthreads();
    }
  }
  for (int ikk = 0; ikk < 18; ikk++) {
    if (!sfix[ikk]) continue;
    const double kref = K[ikk][ikk], v = sval[ikk];
    if (tid < 18) f[tid] -= K[tid][ikk] * v;
    __syncthreads();
    if (tid < 18) {
      K[ikk][tid] = 0.0;
      K[tid][ikk] = 0.0;
    } else if (tid < 22) {
      h[tid - 18] -= G[ikk][tid - 18] * v;
      G[ikk][tid - 18] = 0.0;
    }
    __syncthreads();
    if (tid == 0) {
      K[ikk][ikk] = kref;
      f[ikk] = kref * v;
    }
    __syncthreads();
  }
  if (!a.rhs_only) {
    if (!a.blocks) {
      for (int e = tid; e < 468; e += FS_BS) {
        const int row = e / 26, col = e % 26;
        a.VV[(long)iel * 468 + e] = col < 18 ? K[row][col] : G[row][(col - 18) >> 1];
      }
    } else {
      for (int e = tid; e < 324; e += FS_BS) a.VV_K[(long)iel * 324 + e] = K[e / 18][e % 18];
      for (int e = tid; e < 72; e += FS_BS) {
        a.VV_G[(long)iel * 72 + e] = G[e / 4][e % 4];
        a.VV_GT[(long)iel * 72 + e] = G[e % 18][e / 18];
      }
    }
  }
  if (tid < 18) a.fel[(long)iel * 18 + tid] = f[tid];
  else if (tid < 22) a.hel[(long)iel * 4 + tid - 18] = h[tid - 18];
}

// ---------------------------------------------------------------------------
// energy element kernel
// ---------------------------------------------------------------------------
constexpr int T_EPB = 8;
constexpr int T_BS = 256;

struct EnergyIn {
  int nel, eba;
  const double* T;
  const int* icon_T;
  const double *etaq, *Tq, *uq, *wq, *hcondq, *hcapaq, *alphaq, *hprodq, *exxq, *ezzq, *exzq,
      *JxWq;
  const double *N_T, *dNdr_T, *dNds_T, *j00, *j01, *j10, *j11;
  const uint8_t* bc_fix;
  const double* bc_val;
  double dt, rho0;
  double* VV;
  double* bel;  // workspace nel*9
};

__global__ void __launch_bounds__(T_BS) k_energy_elements(EnergyIn a) {
  __shared__ double sN[81];
  __shared__ double sdx[T_EPB][81], sdy[T_EPB][81];
  __shared__ double su[T_EPB][9], sw[T_EPB][9], scap[T_EPB][9], scond[T_EPB][9], sjw[T_EPB][9];
  __shared__ double sT[T_EPB][9];
  __shared__ double sA[T_EPB][81], sE[T_EPB][81];
  __shared__ uint8_t sfix[T_EPB][9];
  __shared__ double sval[T_EPB][9];
  const int tid = threadIdx.x;
  const int e0 = blockIdx.x * T_EPB, nel = a.nel;
  for (int i = tid; i < 81; i += T_BS) sN[i] = a.N_T[i];
  for (int i = tid; i < T_EPB * 81; i += T_BS) {
    int le = i / 81, s = i % 81, q = s / 9, k = s % 9, iel = e0 + le;
    if (iel >= nel) continue;
    long qi = (long)iel * 9 + q;
    sdx[le][s] = a.j00[qi] * a.dNdr_T[q * 9 + k] + a.j01[qi] * a.dNds_T[q * 9 + k];
    sdy[le][s] = a.j10[qi] * a.dNdr_T[q * 9 + k] + a.j11[qi] * a.dNds_T[q * 9 + k];
    if (k == 0) {
      su[le][q] = a.uq[qi];
      sw[le][q] = a.wq[qi];
      scap[le][q] = a.hcapaq[qi];
      scond[le][q] = a.hcondq[qi];
      sjw[le][q] = a.JxWq[qi];
    }
    if (q == 0) {
      int n = a.icon_T[k * (long)nel + iel];
      sT[le][k] = a.T[n];
      sfix[le][k] = a.bc_fix[n];
      sval[le][k] = a.bc_val[n];
    }
  }
  __syncthreads();
  const double rho0 = a.rho0, dt = a.dt;
  // unmasked A and the explicit part (MM - (Ka+Kd)*dt*0.5)
  for (int i = tid; i < T_EPB * 81; i += T_BS) {
    int le = i / 81, s = i % 81, r = s / 9, c = s % 9, iel = e0 + le;
    if (iel >= nel) continue;
    double MM = 0.0, Kd = 0.0, Ka = 0.0;
    for (int q = 0; q < 9; q++) {
      double Nr = sN[q * 9 + r], Nc = sN[q * 9 + c];
      MM += Nr * Nc * rho0 * scap[le][q] * sjw[le][q];
      Kd += (sdx[le][q * 9 + r] * sdx[le][q * 9 + c] + sdy[le][q * 9 + r] * sdy[le][q * 9 + c]) *
            scond[le][q] * sjw[le][q];
      double vb = su[le][q] * sdx[le][q * 9 + c] + sw[le][q] * sdy[le][q * 9 + c];
      Ka += Nr * vb * rho0 * scap[le][q] * sjw[le][q];
    }
    double A = 0.0;
    A += MM + (Ka + Kd) * dt * 0.5;
    sA[le][s] = A;
    sE[le][s] = MM - (Ka + Kd) * dt * 0.5;
  }
  __syncthreads();
  // element rhs: b = dt*(source) + Bexp . T, then Dirichlet lifting
  for (int i = tid; i < T_EPB * 9; i += T_BS) {
    int le = i / 9, r = i % 9, iel = e0 + le;
    if (iel >= nel) continue;
    double b = 0.0;
    for (int q = 0; q < 9; q++) {
      long qi = (long)iel * 9 + q;
      double Nr = sN[q * 9 + r], jw = sjw[le][q];
      b += Nr * a.hprodq[qi] * jw;
      if (a.eba) {
        double ex = a.exxq[qi], ez = a.ezzq[qi], exz = a.exzq[qi];
        double sr = 2.0 / 3.0 * (ex * ex) + 2.0 / 3.0 * (ez * ez) - 2.0 / 3.0 * ex * ez +
                    2 * (exz * exz);
        b += Nr * 2 * a.etaq[qi] * sr * jw;
        b += Nr * a.alphaq[qi] * a.Tq[qi] * (-sw[le][q] * rho0 * 10) * jw;
      }
    }
    b *= dt;
    double s = 0.0;
    for (int c = 0; c < 9; c++) s += sE[le][r * 9 + c] * sT[le][c];
    b += s;
    if (sfix[le][r]) {
      b = sA[le][r * 9 + r] * sval[le][r];
    } else {
      for (int j = 0; j < 9; j++)
        if (sfix[le][j]) b -= sA[le][r * 9 + j] * sval[le][j];
    }
    a.bel[(long)iel * 9 + r] = b;
  }
  for (int i = tid; i < T_EPB * 81; i += T_BS) {
    int le = i / 81, s = i % 81, r = s / 9, c = s % 9, iel = e0 + le;
    if (iel >= nel) continue;
    double v = sA[le][s];
    if (sfix[le][r] || sfix[le][c]) v = (r == c) ? v : 0.0;
    a.VV[(long)iel * 81 + s] = v;
  }
}

// ---------------------------------------------------------------------------
// host entry points
// ---------------------------------------------------------------------------
int build_matrix_stokes(const StokesArgs& s, hipStream_t st) {
  if (s.nel <= 0 || !s.icon_P || !s.l2g || !s.rhoq || !s.etaq || !s.JxWq || !s.gxq || !s.gyq ||
      !s.N_V || !s.N_P || !s.dNdr_V || !s.dNdt_V || !s.dNdr_P || !s.dNdt_P || !s.j00 || !s.j01 ||
      !s.j10 || !s.j11 || !s.bc_fix || !s.bc_val || !s.rhs_f || !s.rhs_h || !s.work ||
      !s.inc_v_ptr || !s.inc_v_list || !s.inc_p_ptr || !s.inc_p_list)
    return fail(MW_EINVAL, "build_matrix_stokes: null pointer argument");
  if (s.rhs_only && s.blocks)
    return fail(MW_EINVAL, "build_matrix_stokes: rhs_only with blocks");
  if (!s.rhs_only &&
      (s.blocks ? (!s.VV_K || !s.VV_G || !s.VV_GT || !s.VV_M || !s.VV_Me || !s.VV_H || !s.eta_e)
                : !s.VV))
    return fail(MW_EINVAL, "build_matrix_stokes: null output array");
  StokesIn a;
  a.nel = s.nel; a.Nfem_V = s.Nfem_V; a.Nfem_P = s.Nfem_P; a.blocks = s.blocks;
  a.rhs_only = s.rhs_only;
  a.icon_P = s.icon_P; a.l2g = s.l2g;
  a.rhoq = s.rhoq; a.etaq = s.etaq; a.JxWq = s.JxWq; a.gxq = s.gxq; a.gyq = s.gyq;
  a.N_V = s.N_V; a.N_P = s.N_P; a.dNdr_V = s.dNdr_V; a.dNdt_V = s.dNdt_V;
  a.dNdr_P = s.dNdr_P; a.dNdt_P = s.dNdt_P;
  a.j00 = s.j00; a.j01 = s.j01; a.j10 = s.j10; a.j11 = s.j11; a.eta_e = s.eta_e;
  a.gscale = s.eta_ref / s.L_ref;
  a.bc_fix = s.bc_fix; a.bc_val = s.bc_val;
  a.VV = s.VV; a.VV_K = s.VV_K; a.VV_G = s.VV_G; a.VV_GT = s.VV_GT;
  a.VV_M = s.VV_M; a.VV_Me = s.VV_Me; a.VV_H = s.VV_H;
  a.fel = s.work;
  a.hel = s.work + 18L * s.nel;
  a.theta = s.theta; a.bot_el = s.bot_el; a.top_el = s.top_el;
  a.bot_fs = s.bot_fs; a.top_fs = s.top_fs;
  a.axi = s.axi; a.xq = s.xq;
  const bool free_slip = s.theta && (s.bot_fs || s.top_fs);
  if (free_slip && ((s.bot_fs && !s.bot_el) || (s.top_fs && !s.top_el)))
    return fail(MW_EINVAL, "build_matrix_stokes: free slip needs bot_element/top_element");
  if (s.axi && !s.xq) return fail(MW_EINVAL, "build_matrix_stokes: axisymmetric needs xq");
  prof().begin(PROF_ASSEMBLY, st);
  if (s.axi && s.rhs_only)
    k_stokes_elements<S_EPB_RHS, true><<<nblocks(s.nel, S_EPB_RHS), S_BS, 0, st>>>(a);
  else if (s.axi)
    k_stokes_rows<true><<<nblocks(s.nel, SR_EPB), SR_BS, 0, st>>>(a);
  else if (s.rhs_only)
    k_stokes_elements<S_EPB_RHS, false><<<nblocks(s.nel, S_EPB_RHS), S_BS, 0, st>>>(a);
  else
    k_stokes_rows<false><<<nblocks(s.nel, SR_EPB), SR_BS, 0, st>>>(a);
  if (free_slip) k_stokes_free_slip<<<s.nel, FS_BS, 0, st>>>(a);
  prof().end(PROF_ASSEMBLY, st);
  MW_CHECK_LAUNCH();
  int rc = gather_sum(s.Nfem_V, 18, s.inc_v_ptr, s.inc_v_list, a.fel, s.rhs_f, st);
  if (rc) return rc;
  return gather_sum(s.Nfem_P, 4, s.inc_p_ptr, s.inc_p_list, a.hel, s.rhs_h, st);
}

// Velocity right-hand side of build_matrix_stokes when every Dirichlet value
// is zero (the free-slip box of the timestep): the lifting terms are then
// products with 0 and the fixed rows K_ii * 0, so f_e is the gravity term
// alone, formed with the same expression and order as k_stokes_elements
// (and the pressure right-hand side is zero); one thread per element dof.
__global__ void k_stokes_rhs_zero_lift(int nel, const int* __restrict__ l2g,
                                       const uint8_t* __restrict__ bc_fix,
                                       const double* __restrict__ N_V,
                                       const double* __restrict__ JxWq,
                                       const double* __restrict__ rhoq,
                                       const double* __restrict__ gxq,
                                       const double* __restrict__ gyq, double* __restrict__ fel) {
  // one thread per (element, component): the element's point values are read
  // once for its nine rows
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= 2L * nel) return;
  const int iel = (int)(t >> 1), c = (int)(t & 1);
  const double* g = c ? gyq : gxq;
  double jw[9], rh[9], gg[9];
#pragma unroll
  for (int q = 0; q < 9; q++) {
    const long qi = (long)iel * 9 + q;
    jw[q] = JxWq[qi];
    rh[q] = rhoq[qi];
    gg[q] = g[qi];
  }
  uint8_t fx[9];
#pragma unroll
  for (int nd = 0; nd < 9; nd++) fx[nd] = bc_fix[l2g[(2 * nd + c) * (long)nel + iel]];
#pragma unroll
  for (int nd = 0; nd < 9; nd++) {
    double f = 0.0;
#pragma unroll
    for (int q = 0; q < 9; q++) f += N_V[q * 9 + nd] * jw[q] * rh[q] * gg[q];
    fel[(long)iel * 18 + 2 * nd + c] = fx[nd] ? 0.0 : f;
  }
}

int stokes_rhs_zero_lift(const StokesArgs& s, hipStream_t st) {
  if (s.nel <= 0 || !s.l2g || !s.bc_fix || !s.N_V || !s.JxWq || !s.rhoq || !s.gxq || !s.gyq ||
      !s.rhs_f || !s.rhs_h || !s.work || !s.inc_v_ptr || !s.inc_v_list)
    return fail(MW_EINVAL, "stokes_rhs_zero_lift: null pointer argument");
  prof().begin(PROF_ASSEMBLY, st);
  k_stokes_rhs_zero_lift<<<nblocks(2L * s.nel, 256), 256, 0, st>>>(
      s.nel, s.l2g, s.bc_fix, s.N_V, s.JxWq, s.rhoq, s.gxq, s.gyq, s.work);
  prof().end(PROF_ASSEMBLY, st);
  MW_CHECK_LAUNCH();
  RC(gather_sum(s.Nfem_V, 18, s.inc_v_ptr, s.inc_v_list, s.work, s.rhs_f, st));
  MW_HIP(hipMemsetAsync(s.rhs_h, 0, sizeof(double) * (size_t)s.Nfem_P, st));
  return MW_OK;
}

int build_matrix_energy(const EnergyArgs& s, hipStream_t st) {
  if (s.nel <= 0 || !s.T || !s.icon_T || !s.uq || !s.wq || !s.hcondq || !s.hcapaq ||
      !s.hprodq || !s.JxWq || !s.N_T || !s.dNdr_T || !s.dNds_T || !s.j00 || !s.j01 || !s.j10 ||
      !s.j11 || !s.bc_fix || !s.bc_val || !s.VV || !s.rhs || !s.work || !s.inc_ptr ||
      !s.inc_list)
    return fail(MW_EINVAL, "build_matrix_energy: null pointer argument");
  if (s.eba && (!s.etaq || !s.Tq || !s.alphaq || !s.exxq || !s.ezzq || !s.exzq))
    return fail(MW_EINVAL, "build_matrix_energy: EBA needs etaq, Tq, alphaq, exxq, ezzq, exzq");
  EnergyIn a;
  a.nel = s.nel; a.eba = s.eba; a.T = s.T; a.icon_T = s.icon_T;
  a.etaq = s.etaq; a.Tq = s.Tq; a.uq = s.uq; a.wq = s.wq; a.hcondq = s.hcondq;
  a.hcapaq = s.hcapaq; a.alphaq = s.alphaq; a.hprodq = s.hprodq;
  a.exxq = s.exxq; a.ezzq = s.ezzq; a.exzq = s.exzq; a.JxWq = s.JxWq;
  a.N_T = s.N_T; a.dNdr_T = s.dNdr_T; a.dNds_T = s.dNds_T;
  a.j00 = s.j00; a.j01 = s.j01; a.j10 = s.j10; a.j11 = s.j11;
  a.bc_fix = s.bc_fix; a.bc_val = s.bc_val; a.dt = s.dt; a.rho0 = s.rho0;
  a.VV = s.VV; a.bel = s.work;
  k_energy_elements<<<nblocks(s.nel, T_EPB), T_BS, 0, st>>>(a);
  MW_CHECK_LAUNCH();
  return gather_sum(s.Nfem_T, 9, s.inc_ptr, s.inc_list, a.bel, s.rhs, st);
}

}  // namespace mw
