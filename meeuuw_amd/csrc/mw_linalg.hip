// Vector kernels and deterministic reductions used by the Krylov solvers.
#include <type_traits>

#include "mw_common.h"
#include "mw_solver.h"

namespace mw {

template <class T>
int DBuf<T>::alloc(size_t count) {
  release();
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(&p, count * sizeof(T));
  if (e != hipSuccess) {
    p = nullptr;
    return fail(MW_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  n = count;
  return MW_OK;
}
template struct DBuf<double>;
template struct DBuf<float>;
template struct DBuf<int>;
template struct DBuf<uint8_t>;
template struct DBuf<uint16_t>;
template struct DBuf<unsigned long long>;

int Reducer::init() {
  int rc = partial.alloc(DOT_BLOCKS * 33);  // up to 32 columns and a norm
  if (rc) return rc;
  rc = result.alloc(64);
  if (rc) return rc;
  if (!hbuf) {
    MW_HIP(hipHostMalloc((void**)&hbuf, 64 * sizeof(double), hipHostMallocDefault));
    MW_HIP(hipHostGetDevicePointer((void**)&hbuf_dev, hbuf, 0));
  }
  return MW_OK;
}

Reducer::~Reducer() {
  if (hbuf) (void)hipHostFree(hbuf);
}

// sums k interleaved partial arrays: out[j] = sum_b partial[b*k + j], a
// wavefront per column (reduce_column)
// (a separate launch: a last-block-done finish inside the partial kernel needs
// device-scope fences, which cost far more than this kernel on this GPU)
__global__ void k_dot_final(const double* __restrict__ partial, int nblk, int k,
                            double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  for (int j = threadIdx.x >> 6; j < k; j += blockDim.x >> 6) {
    const double s = reduce_column(partial, nblk, k, j, lane);
    if (lane == 0) out[j] = s;
  }
}

// threads of the final kernel for k columns: a wavefront per column, up to 16
static int final_threads(int k) { return 64 * (k < 16 ? k : 16); }

template <class T>
__global__ void __launch_bounds__(DOT_BS) k_dot_partial(const T* __restrict__ a,
                                                        const T* __restrict__ b, long n,
                                                        double* __restrict__ partial) {
  double s = 0.0;
  for (long i = blockIdx.x * (long)DOT_BS + threadIdx.x; i < n; i += (long)gridDim.x * DOT_BS)
    s += (double)a[i] * (double)b[i];
  __shared__ double sh[DOT_BS / 64];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < DOT_BS / 64; w++) t += sh[w];
    partial[blockIdx.x] = t;
  }
}

template <class T>
int dot(const T* a, const T* b, long n, double* out, Reducer& R, hipStream_t st) {
  int nb = dot_blocks(n);
  k_dot_partial<T><<<nb, DOT_BS, 0, st>>>(a, b, n, R.partial);
  k_dot_final<<<1, 64, 0, st>>>(R.partial, nb, 1, out);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// The multi-column kernels below read KMAX columns per element with every
// load unconditional (columns past k re-read column k-1, which the cache
// serves), so a wavefront issues all its loads before the first wait; only
// the arithmetic is limited to j < k.  KMAX runs in steps of 4.
template <class T, int KMAX>
__device__ __forceinline__ void load_cols(const T* __restrict__ V, long ld, int k, long i,
                                          double* v) {
#pragma unroll
  for (int j = 0; j < KMAX; j++) v[j] = (double)V[(long)(j < k ? j : k - 1) * ld + i];
}

template <class T, int KMAX>
__global__ void __launch_bounds__(DOT_BS) k_mdot_partial(const T* __restrict__ V, long ld, int k,
                                                         const T* __restrict__ w, long n,
                                                         double* __restrict__ partial) {
  double s[KMAX];
#pragma unroll
  for (int j = 0; j < KMAX; j++) s[j] = 0.0;
  for (long i = blockIdx.x * (long)DOT_BS + threadIdx.x; i < n; i += (long)gridDim.x * DOT_BS) {
    double v[KMAX];
    load_cols<T, KMAX>(V, ld, k, i, v);
    const double wi = (double)w[i];
#pragma unroll
    for (int j = 0; j < KMAX; j++) s[j] += v[j] * wi;  // s[j >= k] is never used
  }
  __shared__ double sh[DOT_BS / 64][KMAX];
#pragma unroll
  for (int j = 0; j < KMAX; j++) {
    double t = wave_sum(s[j]);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][j] = t;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < k; j += DOT_BS) {
    double t = 0.0;
    for (int wv = 0; wv < DOT_BS / 64; wv++) t += sh[wv][j];
    partial[(long)blockIdx.x * k + j] = t;
  }
}

// KMAX of a k-column pass: the multiple of 4 at or above k (k <= 32)
template <class F>
static void kmax_dispatch(int k, F f) {
  switch ((k + 3) / 4) {
    case 1: f(std::integral_constant<int, 4>()); break;
    case 2: f(std::integral_constant<int, 8>()); break;
    case 3: f(std::integral_constant<int, 12>()); break;
    case 4: f(std::integral_constant<int, 16>()); break;
    case 5: f(std::integral_constant<int, 20>()); break;
    case 6: f(std::integral_constant<int, 24>()); break;
    case 7: f(std::integral_constant<int, 28>()); break;
    default: f(std::integral_constant<int, 32>()); break;
  }
}

template <class T>
int multi_dot(const T* V, long ld, int k, const T* w, long n, double* out, Reducer& R,
              hipStream_t st) {
  if (k <= 0) return MW_OK;
  if (k > 32) return fail(MW_EINVAL, "multi_dot: at most 32 columns");
  int nb = dot_blocks(n);
  kmax_dispatch(k, [&](auto km) {
    k_mdot_partial<T, decltype(km)::value><<<nb, DOT_BS, 0, st>>>(V, ld, k, w, n, R.partial);
  });
  k_dot_final<<<1, final_threads(k), 0, st>>>(R.partial, nb, k, out);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// Classical Gram-Schmidt passes fused with the next reduction: one sweep over
// the basis per pass.  Per element the arithmetic is that of a separate
// update (w - sum_j h_j V_j, in order, in FP64, rounded to T) followed by a
// separate multi_dot / dot over the stored T values, and the partial sums use
// the same grid and reduction tree, so the results are identical to the
// unfused sequence.
template <class T, int KMAX>
__global__ void __launch_bounds__(DOT_BS) k_gs_project(const T* __restrict__ V, long ld, int k,
                                                       const double* __restrict__ h,
                                                       const T* w_in, T* w_out, long n,
                                                       double* __restrict__ partial, int norm) {
  __shared__ double sh_h[KMAX];
  if (threadIdx.x < KMAX) sh_h[threadIdx.x] = (h && (int)threadIdx.x < k) ? h[threadIdx.x] : 0.0;
  __syncthreads();
  double s[KMAX], sn = 0.0;
#pragma unroll
  for (int j = 0; j < KMAX; j++) s[j] = 0.0;
  for (long i = blockIdx.x * (long)DOT_BS + threadIdx.x; i < n; i += (long)gridDim.x * DOT_BS) {
    double v[KMAX];
    load_cols<T, KMAX>(V, ld, k, i, v);
    double wi = (double)w_in[i];
    if (h) {
#pragma unroll
      for (int j = 0; j < KMAX; j++) {
        const double t = wi - sh_h[j] * v[j];
        wi = j < k ? t : wi;
      }
    }
    const T wt = (T)wi;
    if (h) w_out[i] = wt;
    const double wd = (double)wt;
#pragma unroll
    for (int j = 0; j < KMAX; j++) s[j] += v[j] * wd;  // s[j >= k] is never used
    sn += wd * wd;
  }
  __shared__ double sh[DOT_BS / 64][KMAX + 1];
#pragma unroll
  for (int j = 0; j < KMAX; j++) {
    double t = wave_sum(s[j]);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][j] = t;
  }
  {
    double t = wave_sum(sn);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][KMAX] = t;
  }
  __syncthreads();
  const int kc = k + norm;  // columns written: the k projections, then |w|^2
  for (int j = threadIdx.x; j < kc; j += DOT_BS) {
    double t = 0.0;
    for (int wv = 0; wv < DOT_BS / 64; wv++) t += sh[wv][j < k ? j : KMAX];
    partial[(long)blockIdx.x * kc + j] = t;
  }
}

// w -= V h (k columns), partial sums of <w, w>
template <class T, int KMAX>
__global__ void __launch_bounds__(DOT_BS) k_gs_update_norm(const T* __restrict__ V, long ld, int k,
                                                           const double* __restrict__ h, T* w,
                                                           long n, double* __restrict__ partial) {
  __shared__ double sh_h[KMAX];
  if (threadIdx.x < KMAX) sh_h[threadIdx.x] = (int)threadIdx.x < k ? h[threadIdx.x] : 0.0;
  __syncthreads();
  double s = 0.0;
  for (long i = blockIdx.x * (long)DOT_BS + threadIdx.x; i < n; i += (long)gridDim.x * DOT_BS) {
    double v[KMAX];
    load_cols<T, KMAX>(V, ld, k, i, v);
    double wi = (double)w[i];
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
      const double t = wi - sh_h[j] * v[j];
      wi = j < k ? t : wi;
    }
    const T wt = (T)wi;
    w[i] = wt;
    s += (double)wt * (double)wt;
  }
  __shared__ double sh[DOT_BS / 64];
  double t = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    double u = 0.0;
    for (int wv = 0; wv < DOT_BS / 64; wv++) u += sh[wv];
    partial[blockIdx.x] = u;
  }
}

template <class T>
int gs_project(const T* V, long ld, int k, const double* h, const T* w_in, T* w_out, long n,
               double* out, Reducer& R, hipStream_t st, bool with_norm, bool finish) {
  if (k <= 0) return MW_OK;
  if (k > 32) return fail(MW_EINVAL, "gs_project: at most 32 columns");
  int nb = dot_blocks(n);
  const int norm = with_norm ? 1 : 0;
  kmax_dispatch(k, [&](auto km) {
    k_gs_project<T, decltype(km)::value><<<nb, DOT_BS, 0, st>>>(V, ld, k, h, w_in, w_out, n,
                                                                 R.partial, norm);
  });
  if (finish) k_dot_final<<<1, final_threads(k + norm), 0, st>>>(R.partial, nb, k + norm, out);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

// out = (w - V h) * scale[0] (k columns, FP64 sums rounded to T, then the
// scaled value rounded again: the stored update followed by a separate
// scaling pass, in one sweep)
template <class T, int KMAX>
__global__ void __launch_bounds__(DOT_BS) k_gs_update_scale(const T* __restrict__ V, long ld, int k,
                                                            const double* __restrict__ h,
                                                            const T* __restrict__ w, long n,
                                                            const double* __restrict__ scale,
                                                            T* __restrict__ out) {
  __shared__ double sh_h[KMAX];
  if (threadIdx.x < KMAX) sh_h[threadIdx.x] = (int)threadIdx.x < k ? h[threadIdx.x] : 0.0;
  __syncthreads();
  const double sc = scale[0];
  for (long i = blockIdx.x * (long)DOT_BS + threadIdx.x; i < n; i += (long)gridDim.x * DOT_BS) {
    double v[KMAX];
    load_cols<T, KMAX>(V, ld, k, i, v);
    double wi = (double)w[i];
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
      const double t = wi - sh_h[j] * v[j];
      wi = j < k ? t : wi;
    }
    const T wt = (T)wi;
    out[i] = (T)(sc * (double)wt);
  }
}

template <class T>
int gs_update_scale(const T* V, long ld, int k, const double* h, const T* w, long n,
                    const double* scale, T* out, hipStream_t st) {
  if (k <= 0 || k > 32) return fail(MW_EINVAL, "gs_update_scale: 1 to 32 columns");
  int nb = dot_blocks(n);
  kmax_dispatch(k, [&](auto km) {
    k_gs_update_scale<T, decltype(km)::value><<<nb, DOT_BS, 0, st>>>(V, ld, k, h, w, n, scale,
                                                                      out);
  });
  MW_CHECK_LAUNCH();
  return MW_OK;
}

template <class T>
int gs_update_norm(const T* V, long ld, int k, const double* h, T* w, long n, double* out,
                   Reducer& R, hipStream_t st) {
  if (k <= 0 || k > 32) return fail(MW_EINVAL, "gs_update_norm: 1 to 32 columns");
  int nb = dot_blocks(n);
  kmax_dispatch(k, [&](auto km) {
    k_gs_update_norm<T, decltype(km)::value><<<nb, DOT_BS, 0, st>>>(V, ld, k, h, w, n,
                                                                     R.partial);
  });
  k_dot_final<<<1, 64, 0, st>>>(R.partial, nb, 1, out);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

template int gs_project<float>(const float*, long, int, const double*, const float*, float*, long,
                               double*, Reducer&, hipStream_t, bool, bool);
template int gs_update_scale<float>(const float*, long, int, const double*, const float*, long,
                                     const double*, float*, hipStream_t);
template int gs_update_norm<float>(const float*, long, int, const double*, float*, long, double*,
                                   Reducer&, hipStream_t);

template <class T>
__global__ void k_axpby(long n, double a, const double* __restrict__ da, const T* __restrict__ x,
                        double b, const double* __restrict__ db, T* __restrict__ y) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double aa = da ? a * da[0] : a;
  double bb = db ? b * db[0] : b;
  y[i] = (T)(aa * (double)x[i] + (bb == 0.0 ? 0.0 : bb * (double)y[i]));
}

template <class T>
void axpby(long n, double a, const double* da, const T* x, double b, const double* db, T* y,
           hipStream_t st) {
  k_axpby<T><<<nblocks(n, 256), 256, 0, st>>>(n, a, da, x, b, db, y);
}

template <class T>
__global__ void k_fill(long n, T v, T* __restrict__ y) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}

template <class T>
void fill(long n, T v, T* y, hipStream_t st) {
  k_fill<T><<<nblocks(n, 256), 256, 0, st>>>(n, v, y);
}

template <class T, class U>
__global__ void k_convert(long n, const T* __restrict__ x, U* __restrict__ y) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) y[i] = (U)x[i];
}

template <class T, class U>
void convert(long n, const T* x, U* y, hipStream_t st) {
  k_convert<T, U><<<nblocks(n, 256), 256, 0, st>>>(n, x, y);
}

template int dot<double>(const double*, const double*, long, double*, Reducer&, hipStream_t);
template int dot<float>(const float*, const float*, long, double*, Reducer&, hipStream_t);
template int multi_dot<double>(const double*, long, int, const double*, long, double*, Reducer&,
                               hipStream_t);
template int multi_dot<float>(const float*, long, int, const float*, long, double*, Reducer&,
                              hipStream_t);
template void axpby<double>(long, double, const double*, const double*, double, const double*,
                            double*, hipStream_t);
template void axpby<float>(long, double, const double*, const float*, double, const double*,
                           float*, hipStream_t);
template void fill<double>(long, double, double*, hipStream_t);
template void fill<float>(long, float, float*, hipStream_t);
template void fill<int>(long, int, int*, hipStream_t);
template void convert<double, float>(long, const double*, float*, hipStream_t);
template void convert<float, double>(long, const float*, double*, hipStream_t);

}  // namespace mw

namespace mw {

Prof& prof() {
  static Prof p;
  return p;
}

thread_local int Prof::muted = 0;

void Prof::begin(int c, hipStream_t st) {
  if (!on || muted) return;
  auto& v = ev[c];
  if ((int)v.size() < used[c] + 2) {
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    v.push_back(a);
    v.push_back(b);
  }
  (void)hipEventRecord(v[used[c]], st);
}

void Prof::end(int c, hipStream_t st) {
  if (!on || muted) return;
  (void)hipEventRecord(ev[c][used[c] + 1], st);
  used[c] += 2;
}

void Prof::collect() {
  for (int c = 0; c < PROF_NCLASSES; c++) {
    for (int i = 0; i < used[c]; i += 2) {
      float ms = 0.f;
      (void)hipEventSynchronize(ev[c][i + 1]);
      (void)hipEventElapsedTime(&ms, ev[c][i], ev[c][i + 1]);
      total_ms[c] += ms;
      count[c]++;
    }
    used[c] = 0;
  }
}

void Prof::reset() {
  for (int c = 0; c < PROF_NCLASSES; c++) {
    total_ms[c] = 0;
    count[c] = 0;
    used[c] = 0;
  }
}

Prof::~Prof() {
  for (auto& v : ev)
    for (auto e : v) (void)hipEventDestroy(e);
}

}  // namespace mw
