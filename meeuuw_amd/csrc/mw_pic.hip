// Particle-in-cell kernels: locate, interpolate, RK advection, projections.
//
// Reference behaviour restated here:
//   locate_particles___box               pic_functions.py:107-137
//   interpolate_field_on_particles       pic_functions.py:84-101
//   interpolate_vel_on_pt___box          pic_functions.py:17-37
//   advect_particles___box (RK1/2/4)     pic_functions.py:179-292
//   project_particles_on_elements        pic_functions.py:602-650
//   project_particle_field_on_nodes_1-4  pic_functions.py:658-1053
//
// One thread per particle for the streaming kernels.  Projections never use
// floating-point atomics: particles are stably sorted by element (a counting
// sort on the element index, inactive particles last), every element reduces its
// own particles in increasing particle order, and nodes then add the partial
// sums of their elements in increasing element order.  Results are therefore
// deterministic; the elemental projection is bit-identical to a sequential
// loop over particles.

#include "mw_common.h"
#include "mw_capi_impl.h"
#include "mw_solver.h"

namespace mw {

constexpr int MAXF = 8;  // max fields fused in one interpolation/projection pass

struct FieldPtrs {
  const double* p[MAXF];
};
struct OutPtrs {
  double* p[MAXF];
};

// ---------------------------------------------------------------------------
// locate
// ---------------------------------------------------------------------------
__global__ void k_locate_box(int np, const uint8_t* __restrict__ active,
                             const double* __restrict__ x, const double* __restrict__ z, double hx,
                             double hz, const double* __restrict__ x_V,
                             const double* __restrict__ z_V, const int* __restrict__ icon,
                             int nel, int nelx, double* __restrict__ r, double* __restrict__ t,
                             int* __restrict__ iel_out, int* __restrict__ flags) {
  int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  double rr = 0.0, tt = 0.0;
  int iel = 0;
  if (active[ip]) {
    double xp = x[ip], zp = z[ip];
    const double fx = xp / hx, fz = zp / hz;
    // a NaN or huge position has no element (its conversion to int would be
    // undefined): flagged as outside the mesh
    const bool conv = fabs(fx) < 2147483647.0 && fabs(fz) < 2147483647.0;
    const long ie = conv ? (long)nelx * (int)fz + (int)fx : -1;
    iel = conv ? (int)ie : 0;
    if (ie < 0 || ie >= nel) {
      atomicAdd(&flags[0], 1);
      rr = tt = __builtin_nan("");
    } else {
      int n0 = icon[iel];
      rr = ((xp - x_V[n0]) / hx - 0.5) * 2;
      tt = ((zp - z_V[n0]) / hz - 0.5) * 2;
      if (!(rr >= -1 && rr <= 1)) atomicAdd(&flags[1], 1);
      if (!(tt >= -1 && tt <= 1)) atomicAdd(&flags[2], 1);
    }
  }
  r[ip] = rr;
  t[ip] = tt;
  iel_out[ip] = iel;
}

int locate_box(int np, const uint8_t* active, const double* x, const double* z, double hx,
               double hz, const double* x_V, const double* z_V, const int* icon_V, int nel,
               int nelx, double* r, double* t, int* iel, int* flags, hipStream_t st) {
  if (np < 0 || nel <= 0 || nelx <= 0 || !(hx > 0) || !(hz > 0) || !flags)
    return fail(MW_EINVAL, "locate_box: bad arguments");
  if (np == 0) return MW_OK;
  if (!active || !x || !z || !x_V || !z_V || !icon_V || !r || !t || !iel)
    return fail(MW_EINVAL, "locate_box: null pointer argument");
  k_locate_box<<<nblocks(np, 256), 256, 0, st>>>(np, active, x, z, hx, hz, x_V, z_V, icon_V, nel,
                                                 nelx, r, t, iel, flags);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// ---------------------------------------------------------------------------
// interpolation (nf fields fused)
// ---------------------------------------------------------------------------
__global__ void k_interpolate(int np, int nf, const uint8_t* __restrict__ active,
                              const double* __restrict__ r, const double* __restrict__ t,
                              const int* __restrict__ iel, FieldPtrs phi,
                              const int* __restrict__ icon, int nel, OutPtrs out) {
  int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  if (!active[ip]) {
    for (int f = 0; f < nf; f++) out.p[f][ip] = 0.0;
    return;
  }
  int e = iel[ip];
  if (e < 0 || e >= nel) {  // invalid element index: no out-of-bounds read
    for (int f = 0; f < nf; f++) out.p[f][ip] = __builtin_nan("");
    return;
  }
  double N[9];
  basis_V(r[ip], t[ip], N);
  int nd[9];
#pragma unroll
  for (int k = 0; k < 9; k++) nd[k] = icon[k * (long)nel + e];
  for (int f = 0; f < nf; f++) {
    double v[9];
#pragma unroll
    for (int k = 0; k < 9; k++) v[k] = phi.p[f][nd[k]];
    out.p[f][ip] = dot9(N, v);
  }
}

int interpolate_fields(int np, int nf, const uint8_t* active, const double* r, const double* t,
                       const int* iel, const double* const* phi, const int* icon, int nel,
                       double* const* out, hipStream_t st) {
  if (np < 0 || nf < 1 || nf > MAXF || nel <= 0)
    return fail(MW_EINVAL, "interpolate_fields: bad sizes");
  if (np == 0) return MW_OK;
  if (!active || !r || !t || !iel || !icon) return fail(MW_EINVAL, "interpolate_fields: null");
  FieldPtrs fp;
  OutPtrs op;
  for (int f = 0; f < nf; f++) {
    if (!phi[f] || !out[f]) return fail(MW_EINVAL, "interpolate_fields: null field");
    fp.p[f] = phi[f];
    op.p[f] = out[f];
  }
  k_interpolate<<<nblocks(np, 256), 256, 0, st>>>(np, nf, active, r, t, iel, fp, icon, nel, op);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// ---------------------------------------------------------------------------
// RK advection
// ---------------------------------------------------------------------------
struct VelCtx {
  const double *u, *w, *x_V, *z_V;
  const int* icon;
  double hx, hz;
  int nelx, nel;
  const double* uw;  // grid numbering: (u, w) interleaved per node
  int nnx;           // grid numbering: nodes per row
};

// interpolate_vel_on_pt___box; false when the element index leaves the mesh.
// GRID: the nodes of an element follow from its index (the box mesh's own
// numbering) and the velocity is read as (u, w) pairs, the same values and
// arithmetic as through icon and separate u, w arrays.
template <bool GRID>
__device__ __forceinline__ bool vel_on_pt(const VelCtx& c, double xp, double zp, double& up,
                                          double& wp) {
  int ielx = (int)(xp / c.hx);
  int ielz = (int)(zp / c.hz);
  int iel = c.nelx * ielz + ielx;
  if (iel < 0 || iel >= c.nel) {
    up = wp = __builtin_nan("");
    return false;
  }
  int n0 = GRID ? grid_node(iel, c.nelx, c.nnx, 0) : c.icon[iel];
  double rp = ((xp - c.x_V[n0]) / c.hx - 0.5) * 2;
  double tp = ((zp - c.z_V[n0]) / c.hz - 0.5) * 2;
  double N[9], uv[9], wv[9];
  basis_V(rp, tp, N);
#pragma unroll
  for (int k = 0; k < 9; k++) {
    if (GRID) {
      const double2 v = reinterpret_cast<const double2*>(c.uw)[grid_node(iel, c.nelx, c.nnx, k)];
      uv[k] = v.x;
      wv[k] = v.y;
    } else {
      int n = c.icon[k * (long)c.nel + iel];
      uv[k] = c.u[n];
      wv[k] = c.w[n];
    }
  }
  up = dot9(N, uv);
  wp = dot9(N, wv);
  return true;
}

template <bool GRID>
__global__ void k_advect_box(int rk, double dt, int np, double* __restrict__ x,
                             double* __restrict__ z, uint8_t* __restrict__ active, VelCtx c,
                             double Lx, double Lz, double* __restrict__ su,
                             double* __restrict__ sw, int* __restrict__ flags) {
  int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  double ou = 0.0, ow = 0.0;
  bool act = active[ip];
  double xA = x[ip], zA = z[ip];
  double xo = xA, zo = zA;
  int bad = 0;
  if (act) {
    if (rk == 1) {
      double uu, ww;
      bad += !vel_on_pt<GRID>(c, xA, zA, uu, ww);
      ou = uu;
      ow = ww;
      xo = xA + uu * dt;
      zo = zA + ww * dt;
      if (xo < 0 || xo > Lx || zo < 0 || zo > Lz) act = false;
    } else if (rk == 2) {
      double uA, wA, uB, wB;
      bad += !vel_on_pt<GRID>(c, xA, zA, uA, wA);
      double xB = xA + uA * dt / 2.0, yB = zA + wA * dt / 2.0;
      if (xB < 0 || xB > Lx || yB < 0 || yB > Lz) {
        act = false;
      } else {
        bad += !vel_on_pt<GRID>(c, xB, yB, uB, wB);
        xo = xA + uB * dt;
        zo = zA + wB * dt;
        ou = uB;
        ow = wB;
      }
    } else {
      double uA, wA, uB, wB, uC, wC, uD, wD;
      bad += !vel_on_pt<GRID>(c, xA, zA, uA, wA);
      double xB = xA + uA * dt / 2.0, yB = zA + wA * dt / 2.0;
      if (xB < 0 || xB > Lx || yB < 0 || yB > Lz) {
        act = false;
      } else {
        bad += !vel_on_pt<GRID>(c, xB, yB, uB, wB);
        double xC = xA + uB * dt / 2.0, yC = zA + wB * dt / 2.0;
        if (xC < 0 || xC > Lx || yC < 0 || yC > Lz) {
          act = false;
        } else {
          bad += !vel_on_pt<GRID>(c, xC, yC, uC, wC);
          double xD = xA + uC * dt, yD = zA + wC * dt;
          if (xD < 0 || xD > Lx || yD < 0 || yD > Lz) {
            act = false;
          } else {
            bad += !vel_on_pt<GRID>(c, xD, yD, uD, wD);
            ou = (uA + 2 * uB + 2 * uC + uD) / 6;
            ow = (wA + 2 * wB + 2 * wC + wD) / 6;
            xo = xA + ou * dt;
            zo = zA + ow * dt;
          }
        }
      }
    }
  }
  if (bad) atomicAdd(&flags[0], 1);
  if (!act) {
    xo = 0;
    zo = 0;
  }
  x[ip] = xo;
  z[ip] = zo;
  active[ip] = act;
  su[ip] = ou;
  sw[ip] = ow;
}

int advect_box(int rk, double dt, int np, double* x, double* z, uint8_t* active, const double* u,
               const double* w, double Lx, double Lz, double hx, double hz, int nelx, int nel,
               const int* icon, const double* x_V, const double* z_V, double* su, double* sw,
               int* flags, hipStream_t st, const double* uw) {
  if (rk != 1 && rk != 2 && rk != 4) return fail(MW_EINVAL, "advect_box: RKorder must be 1, 2 or 4");
  if (np < 0 || nel <= 0 || nelx <= 0 || !flags) return fail(MW_EINVAL, "advect_box: bad sizes");
  if (np == 0) return MW_OK;
  if (!x || !z || !active || (!uw && (!u || !w || !icon)) || !x_V || !z_V || !su || !sw)
    return fail(MW_EINVAL, "advect_box: null pointer argument");
  VelCtx c{u, w, x_V, z_V, icon, hx, hz, nelx, nel, uw, 2 * nelx + 1};
  prof().begin(PROF_ADVECT, st);
  if (uw)
    k_advect_box<true><<<nblocks(np, 256), 256, 0, st>>>(rk, dt, np, x, z, active, c, Lx, Lz, su,
                                                         sw, flags);
  else
    k_advect_box<false><<<nblocks(np, 256), 256, 0, st>>>(rk, dt, np, x, z, active, c, Lx, Lz,
                                                          su, sw, flags);
  prof().end(PROF_ADVECT, st);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// ---------------------------------------------------------------------------
// stable sort of particles by element (a counting sort)
// ---------------------------------------------------------------------------
// Key of a particle: its element index when it is active (active == NULL: all
// particles are) and the index lies in [0, nel); nel otherwise, the bucket of
// the particles no projection reads, which is placed last.
__device__ __forceinline__ int sort_key(int ip, int nel, const uint8_t* __restrict__ active,
                                        const int* __restrict__ iel) {
  const int e = iel[ip];
  return ((!active || active[ip]) && e >= 0 && e < nel) ? e : nel;
}

__global__ void k_zero_i32(int n, int* __restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// Runs of equal keys within a wavefront (consecutive particles mostly share an
// element, so a wavefront holds a few runs): lane `lead` starts the run of this
// lane, `len` is the run's length, counted only on the leading lane.  Lanes
// past np carry key -1 and form runs of their own that are never used.
struct KeyRun {
  int lead, len;
};
__device__ __forceinline__ KeyRun key_run(int k) {
  const int lane = threadIdx.x & 63;
  const int kprev = __shfl_up(k, 1);
  const unsigned long long starts = __ballot(lane == 0 || k != kprev);
  const unsigned long long upto = starts & (~0ull >> (63 - lane));        // bits 0..lane
  const unsigned long long after = lane == 63 ? 0ull : starts & (~0ull << (lane + 1));
  KeyRun r;
  r.lead = 63 - __clzll(upto);
  r.len = (after ? __ffsll(after) - 1 : 64) - lane;
  return r;
}

// cnt[key] = particles per key (integer atomics, one per run of equal keys in a
// wavefront: the counts are exact)
__global__ void k_sort_count(int np, int nel, const uint8_t* __restrict__ active,
                             const int* __restrict__ iel, int* __restrict__ cnt) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = ip < np ? sort_key(ip, nel, active, iel) : -1;
  const KeyRun run = key_run(k);
  if (k >= 0 && run.lead == (int)(threadIdx.x & 63)) atomicAdd(&cnt[k], run.len);
}

// Every particle takes a distinct slot of its key's range: the leading lane of
// a run takes len slots at once (counting cnt[key] back down, so the counters
// end cleared) and the run's lanes fill them in lane order.  Runs of one key
// from different wavefronts land in the order of their atomics; k_sort_rank
// makes the order stable.
__global__ void k_sort_scatter(int np, int nel, const uint8_t* __restrict__ active,
                               const int* __restrict__ iel, const int* __restrict__ seg,
                               int* __restrict__ cnt, int* __restrict__ tmp) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = ip < np ? sort_key(ip, nel, active, iel) : -1;
  const KeyRun run = key_run(k);
  const int lane = threadIdx.x & 63;
  int top = 0;
  if (k >= 0 && run.lead == lane) top = atomicSub(&cnt[k], run.len);
  top = __shfl(top, run.lead);
  if (k < 0) return;
  const int base = k < nel ? seg[k] : seg[nel];
  // the run occupies [top - len_run, top) of the key's remaining count
  const int len_run = __shfl(run.len, run.lead);
  tmp[base + top - len_run + (lane - run.lead)] = ip;
}

// Stable order within each element: a particle's rank is the number of the
// element's particles with a smaller index (an element holds tens of particles;
// the range is read through the cache by all its lanes).  The last bucket is
// copied as it is (its order is unspecified).
__global__ void k_sort_rank(int np, int nel, const uint8_t* __restrict__ active,
                            const int* __restrict__ iel, const int* __restrict__ seg,
                            const int* __restrict__ tmp, int* __restrict__ order) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= np) return;
  const int v = tmp[j];
  const int k = sort_key(v, nel, active, iel);
  if (k == nel) {
    order[j] = v;
    return;
  }
  const int b = seg[k], en = seg[k + 1];
  int rank = 0;
  for (int i = b; i < en; i++) rank += tmp[i] < v ? 1 : 0;
  order[b + rank] = v;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t sort_work_bytes(int np, int nel) {
  return align256((size_t)(nel + 1) * sizeof(int)) + align256(scan_tile_sums_ints(nel) * sizeof(int)) +
         align256((size_t)np * sizeof(int)) + 256;
}

int sort_by_element_scratch(int np, int nel, const uint8_t* active, const int* iel, int* order,
                            int* seg, int* cnt, int* tsum, int* tmp, bool cnt_zero,
                            hipStream_t st) {
  if (np <= 0 || nel <= 0 || !iel || !order || !seg || !cnt || !tsum || !tmp)
    return fail(MW_EINVAL, "sort_by_element: bad arguments");
  if (!cnt_zero) {
    k_zero_i32<<<nblocks(nel + 1, 256), 256, 0, st>>>(nel + 1, cnt);
    MW_CHECK_LAUNCH();
  }
  k_sort_count<<<nblocks(np, 256), 256, 0, st>>>(np, nel, active, iel, cnt);
  MW_CHECK_LAUNCH();
  RC(exclusive_scan_i32(nel, cnt, seg, tsum, st));
  k_sort_scatter<<<nblocks(np, 256), 256, 0, st>>>(np, nel, active, iel, seg, cnt, tmp);
  MW_CHECK_LAUNCH();
  k_sort_rank<<<nblocks(np, 256), 256, 0, st>>>(np, nel, active, iel, seg, tmp, order);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

int sort_by_element(int np, int nel, const uint8_t* active, const int* iel, int* order, int* seg,
                    void* work, hipStream_t st, bool cnt_zero) {
  if (!work) return fail(MW_EINVAL, "sort_by_element: no workspace");
  uintptr_t b = ((uintptr_t)work + 255) & ~(uintptr_t)255;
  int* cnt = (int*)b;
  b += align256((size_t)(nel + 1) * sizeof(int));
  int* tsum = (int*)b;
  b += align256(scan_tile_sums_ints(nel) * sizeof(int));
  int* tmp = (int*)b;
  return sort_by_element_scratch(np, nel, active, iel, order, seg, cnt, tsum, tmp, cnt_zero, st);
}

// ---------------------------------------------------------------------------
// particle -> element
// ---------------------------------------------------------------------------
__global__ void k_project_elements(int nel, const int* __restrict__ seg,
                                   const int* __restrict__ order, const double* __restrict__ rho,
                                   const double* __restrict__ eta, int averaging,
                                   double* __restrict__ rho_e, double* __restrict__ eta_e,
                                   int* __restrict__ cnt) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  int b = seg[e], en = seg[e + 1];
  double sr = 0.0, se = 0.0;
  for (int j = b; j < en; j++) {
    const int ip = order ? order[j] : j;
    sr += rho[ip];
    double v = eta[ip];
    if (averaging == 0) se += v;
    else if (averaging == 1) se += log10(v);
    else se += 1.0 / v;
  }
  int c = en - b;
  double dc = (double)c;
  rho_e[e] = sr / dc;
  if (averaging == 0) eta_e[e] = se / dc;
  else if (averaging == 1) eta_e[e] = pow(10.0, se / dc);
  else eta_e[e] = dc / se;
  cnt[e] = c;
}

// The same sums when the particles are stored in element order (order ==
// NULL): a 64-thread block takes PE_EPB consecutive elements, whose particles
// are one contiguous range; the block stages that range in LDS with one
// coalesced load per lane and field (more chunks only for unusually full
// elements), and lane l < PE_EPB adds element l's values from LDS in particle
// order, the arithmetic of k_project_elements (bit-identical).  Per element
// the sum is a sequential chain, so the kernel is bound by that chain's load
// latency: the particles of an element are loaded together up front instead
// of one dependent load per particle, and many small blocks keep the CUs busy.
constexpr int PE_CHUNK = 512, PE_EPB = 16;
__global__ void __launch_bounds__(64) k_project_elements_seq(int nel, const int* __restrict__ seg,
                                                             const double* __restrict__ rho,
                                                             const double* __restrict__ eta,
                                                             int averaging,
                                                             double* __restrict__ rho_e,
                                                             double* __restrict__ eta_e,
                                                             int* __restrict__ cnt) {
  __shared__ double sr_[PE_CHUNK], se_[PE_CHUNK];
  const int e0 = blockIdx.x * PE_EPB, lane = threadIdx.x;
  const int e = e0 + lane;
  const bool mine = lane < PE_EPB && e < nel;
  const int elast = min(e0 + PE_EPB, nel);
  const int p0 = seg[e0], p1 = seg[elast];
  const int b = mine ? seg[e] : p1, en = mine ? seg[e + 1] : p1;
  double sr = 0.0, se = 0.0;
  for (int c0 = p0; c0 < p1; c0 += PE_CHUNK) {
    const int cn = min(PE_CHUNK, p1 - c0);
    double vr[PE_CHUNK / 64], ve[PE_CHUNK / 64];
#pragma unroll
    for (int k = 0; k < PE_CHUNK / 64; k++) {
      const int i = lane + 64 * k;
      vr[k] = i < cn ? rho[c0 + i] : 0.0;
      ve[k] = i < cn ? eta[c0 + i] : 0.0;
    }
    __syncthreads();  // the previous chunk's reads are done
    // the viscosity's transform (the term each particle adds) in parallel here;
    // the lanes below only add, in particle order
#pragma unroll
    for (int k = 0; k < PE_CHUNK / 64; k++) {
      const double v = ve[k];
      sr_[lane + 64 * k] = vr[k];
      se_[lane + 64 * k] = averaging == 0 ? v : (averaging == 1 ? log10(v) : 1.0 / v);
    }
    __syncthreads();
    const int j0 = max(b, c0), j1 = min(en, c0 + cn);
#pragma unroll 4
    for (int j = j0; j < j1; j++) {
      sr += sr_[j - c0];
      se += se_[j - c0];
    }
  }
  if (!mine) return;
  const int c = en - b;
  const double dc = (double)c;
  rho_e[e] = sr / dc;
  if (averaging == 0) eta_e[e] = se / dc;
  else if (averaging == 1) eta_e[e] = pow(10.0, se / dc);
  else eta_e[e] = dc / se;
  cnt[e] = c;
}

int project_on_elements(int nel, const int* seg, const int* order, const double* rho,
                        const double* eta, int averaging, double* rho_e, double* eta_e, int* cnt,
                        hipStream_t st) {
  if (nel <= 0 || averaging < 0 || averaging > 2)
    return fail(MW_EINVAL, "project_on_elements: bad arguments");
  if (!order)
    k_project_elements_seq<<<(nel + PE_EPB - 1) / PE_EPB, 64, 0, st>>>(nel, seg, rho, eta,
                                                                       averaging, rho_e, eta_e,
                                                                       cnt);
  else
    k_project_elements<<<nblocks(nel, 256), 256, 0, st>>>(nel, seg, order, rho, eta, averaging,
                                                          rho_e, eta_e, cnt);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// ---------------------------------------------------------------------------
// particle -> corner nodes (types 1-4), then Q1 fill of the other Q2 nodes
// ---------------------------------------------------------------------------
struct AvgCodes {
  int a[MAXF];
};

// per element e, field f, corner k: work[((e*nf + f)*4 + k)*2 + {0: value sum, 1: weight sum}]
// Per element: the (weighted) sums of each field and of the weights at its 4
// corners over the element's particles.  A 32-lane group serves one element:
// each lane stages one particle's row in LDS (its 4 corner weights, which
// corners its quadrant uses, and its nf transformed field values: log10 or
// 1/x for geometric or harmonic averaging), then lane c forms output c (field
// c / 4 at corner c % 4, or for c >= 4 nf the weight sum of corner c - 4 nf)
// from the rows in sorted order: v * N or v (unweighted types), 0 where the
// quadrant does not use that corner, added in particle order: a sequential
// sum per element and output, as the reference's particle loop orders it.
constexpr int NP_GROUP = 32;
constexpr int NP_EPB = 8;               // elements (groups) per 256-thread block
constexpr int NP_PITCH = 5 + MAXF;      // N[4], corner-use bits, values; odd pitch

__global__ void __launch_bounds__(NP_EPB * NP_GROUP)
k_node_partials(int type, int nel, int nf, const int* __restrict__ seg,
                const int* __restrict__ order, FieldPtrs phi, AvgCodes av,
                const double* __restrict__ r, const double* __restrict__ t,
                double* __restrict__ work) {
  __shared__ double sc[NP_EPB][NP_GROUP][NP_PITCH];
  __shared__ int spass[NP_EPB];
  const int g = threadIdx.x / NP_GROUP, lane = threadIdx.x % NP_GROUP;
  const int e = blockIdx.x * NP_EPB + g;
  const bool valid = e < nel;
  const int b0 = valid ? seg[e] : 0, b1 = valid ? seg[e + 1] : 0;
  const int nout = 4 * nf + 4;
  if (lane == 0) spass[g] = (b1 - b0 + NP_GROUP - 1) / NP_GROUP;
  __syncthreads();
  int npass = 0;
  for (int k = 0; k < NP_EPB; k++) npass = max(npass, spass[k]);
  const bool plain = type == 1 || type == 3;
  // this lane's outputs: column lane (and lane + 32)
  int f0 = -1, k0 = 0, f1 = -1, k1 = 0;
  const bool has0 = lane < nout, has1 = lane + NP_GROUP < nout;
  if (has0) {
    if (lane < 4 * nf) { f0 = lane >> 2; k0 = lane & 3; } else { k0 = lane - 4 * nf; }
  }
  if (has1) {
    const int c = lane + NP_GROUP;
    if (c < 4 * nf) { f1 = c >> 2; k1 = c & 3; } else { k1 = c - 4 * nf; }
  }
  double acc0 = 0.0, acc1 = 0.0;
  for (int pass = 0; pass < npass; pass++) {
    const int j = b0 + pass * NP_GROUP + lane;
    if (j < b1) {
      const int ip = order ? order[j] : j;
      const double rr = r[ip], tt = t[ip];
      double N[4];
      bool use[4] = {true, true, true, true};
      if (type == 1) {
        N[0] = N[1] = N[2] = N[3] = 1.0;
      } else if (type == 2) {
        N[0] = 0.25 * (1 - rr) * (1 - tt);
        N[1] = 0.25 * (1 + rr) * (1 - tt);
        N[2] = 0.25 * (1 + rr) * (1 + tt);
        N[3] = 0.25 * (1 - rr) * (1 + tt);
      } else {
        use[0] = (rr <= 0 && tt <= 0);
        use[1] = (rr >= 0 && tt <= 0);
        use[2] = (rr >= 0 && tt >= 0);
        use[3] = (rr <= 0 && tt >= 0);
        if (type == 3) {
          N[0] = N[1] = N[2] = N[3] = 1.0;
        } else {
          N[0] = -rr * -tt;
          N[1] = (1 - rr) * -tt;
          N[2] = (1 - rr) * (1 - tt);
          N[3] = -rr * (1 - tt);
        }
      }
      double* row = sc[g][lane];
#pragma unroll
      for (int k = 0; k < 4; k++) row[k] = N[k];
      row[4] = (double)((use[0] ? 1 : 0) | (use[1] ? 2 : 0) | (use[2] ? 4 : 0) | (use[3] ? 8 : 0));
#pragma unroll
      for (int f = 0; f < MAXF; f++) {
        if (f >= nf) break;
        double v = phi.p[f][ip];
        if (av.a[f] == 1) v = log10(v);
        else if (av.a[f] == 2) v = 1.0 / v;
        row[5 + f] = v;
      }
    }
    __syncthreads();
    const int cnt = min(NP_GROUP, b1 - b0 - pass * NP_GROUP);
    if (has0) {
      for (int q = 0; q < cnt; q++) {
        const double* row = sc[g][q];
        const bool u = ((int)row[4] >> k0) & 1;
        const double N = row[k0];
        if (f0 >= 0) {
          const double v = row[5 + f0];
          acc0 += u ? (plain ? v : v * N) : 0.0;
        } else {
          acc0 += u ? (plain ? 1.0 : 1 * N) : 0.0;
        }
      }
    }
    if (has1) {
      for (int q = 0; q < cnt; q++) {
        const double* row = sc[g][q];
        const bool u = ((int)row[4] >> k1) & 1;
        const double N = row[k1];
        if (f1 >= 0) {
          const double v = row[5 + f1];
          acc1 += u ? (plain ? v : v * N) : 0.0;
        } else {
          acc1 += u ? (plain ? 1.0 : 1 * N) : 0.0;
        }
      }
    }
    __syncthreads();
  }
  if (!valid) return;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int c = lane + h * NP_GROUP;
    const double a = h ? acc1 : acc0;
    if (c < 4 * nf) {
      work[(((long)e * nf + c / 4) * 4 + c % 4) * 2] = a;
    } else if (c < nout) {
      const int k = c - 4 * nf;
      for (int f = 0; f < nf; f++) work[(((long)e * nf + f) * 4 + k) * 2 + 1] = a;
    }
  }
}

// corner nodes: sum partials of incident elements (increasing element), then average
__global__ void k_node_finish(int nn, int nf, const int* __restrict__ ptr,
                              const int* __restrict__ list, const double* __restrict__ work,
                              AvgCodes av, OutPtrs out) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const int i0 = ptr[n], i1 = ptr[n + 1];
  // a box node has at most four incident elements: their entries are read
  // once, up front, and every field's partials loaded together; longer lists
  // (other meshes) take the loop below
  constexpr int NI = 4;
  int ent[NI];
#pragma unroll
  for (int q = 0; q < NI; q++) ent[q] = i0 + q < i1 ? list[i0 + q] : 4;  // local 4: not a corner
  for (int f = 0; f < nf; f++) {
    double s = 0.0, c = 0.0;
    if (i1 - i0 <= NI) {
      double ws[NI], wc[NI];
#pragma unroll
      for (int q = 0; q < NI; q++) {
        const int e = ent[q] >> 5, k = ent[q] & 31;
        const long o = (((long)e * nf + f) * 4 + (k < 4 ? k : 0)) * 2;
        const bool use = k < 4;
        ws[q] = use ? work[o] : 0.0;
        wc[q] = use ? work[o + 1] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < NI; q++) {
        if ((ent[q] & 31) >= 4) continue;
        s += ws[q];
        c += wc[q];
      }
    } else {
      for (int i = i0; i < i1; i++) {
        int v = list[i], e = v >> 5, k = v & 31;
        if (k >= 4) continue;
        long o = (((long)e * nf + f) * 4 + k) * 2;
        s += work[o];
        c += work[o + 1];
      }
    }
    double r;
    if (av.a[f] == 0) r = s / c;
    else if (av.a[f] == 1) r = pow(10.0, s / c);
    else r = c / s;
    out.p[f][n] = r;
  }
}

// edge and centre nodes from the corners (pic_functions.py:715-722)
__global__ void k_node_fill(int nel, int nf, const int* __restrict__ icon, OutPtrs out) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  int n[9];
  for (int k = 0; k < 9; k++) n[k] = icon[k * (long)nel + e];
  for (int f = 0; f < nf; f++) {
    double* p = out.p[f];
    double a0 = p[n[0]], a1 = p[n[1]], a2 = p[n[2]], a3 = p[n[3]];
    p[n[4]] = 0.50 * (a0 + a1);
    p[n[5]] = 0.50 * (a1 + a2);
    p[n[6]] = 0.50 * (a2 + a3);
    p[n[7]] = 0.50 * (a0 + a3);
    p[n[8]] = 0.25 * (a0 + a1 + a2 + a3);
  }
}

int project_on_nodes(int type, int nel, int nn, int nf, const int* seg, const int* order,
                     const double* const* phi, const int* averaging, const int* icon,
                     const double* r, const double* t, const int* inc_ptr, const int* inc_list,
                     double* work, double* const* out, hipStream_t st) {
  if (type < 1 || type > 4) return fail(MW_EINVAL, "project_on_nodes: type must be 1..4");
  if (nel <= 0 || nn <= 0 || nf < 1 || nf > MAXF || !seg || !icon || !r || !t ||
      !inc_ptr || !inc_list || !work)
    return fail(MW_EINVAL, "project_on_nodes: bad arguments");
  FieldPtrs fp;
  OutPtrs op;
  AvgCodes av;
  for (int f = 0; f < nf; f++) {
    if (!phi[f] || !out[f] || averaging[f] < 0 || averaging[f] > 2)
      return fail(MW_EINVAL, "project_on_nodes: bad field");
    fp.p[f] = phi[f];
    op.p[f] = out[f];
    av.a[f] = averaging[f];
  }
  k_node_partials<<<nblocks(nel, NP_EPB), NP_EPB * NP_GROUP, 0, st>>>(type, nel, nf, seg, order,
                                                                    fp, av, r, t, work);
  MW_CHECK_LAUNCH();
  k_node_finish<<<nblocks(nn, 256), 256, 0, st>>>(nn, nf, inc_ptr, inc_list, work, av, op);
  MW_CHECK_LAUNCH();
  k_node_fill<<<nblocks(nel, 256), 256, 0, st>>>(nel, nf, icon, op);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

}  // namespace mw
