// Learned Schur-complement correction network: forward, backward and Adam
// kernels for gfx950.
//
// Layout: activations are NHWC on the (nelz+1) x (nelx+1) pressure-node grid,
// 32 bfloat16 channels (64 bytes) per pixel; tap t = 3*(dy+1) + (dx+1) reads
// the pixel at (y+dy, x+dx) with zero padding.  The two 32 -> 32 layers (and
// their input gradients) run on MFMA (v_mfma_f32_16x16x16_bf16): one block
// computes an 8 x 32 pixel tile for all 32 output channels, the 10 x 34 input
// halo staged in LDS, the 9 x 32 x 32 weights held as MFMA B fragments in
// registers.  The small layers (4 -> 32, 32 -> 1) and the weight gradients are
// FP32 VALU kernels.  Reductions over pixels go through fixed per-block
// partials, so every result is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "mw_cnn.h"
#include "mw_common.h"

namespace mw {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) {  // round to nearest even
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float(h << 16); }
__device__ __forceinline__ bool bf_pos(uint32_t h) { return (h & 0x8000u) == 0 && h != 0; }

constexpr int TY = 8, TX = 32;          // output tile (pixels)
constexpr int HY = TY + 2, HX = TX + 2;  // input halo tile
constexpr int LP = 40;                   // LDS pixel stride in bf16 (80 B: no bank conflicts)

static inline int tiles_of(int H, int W) { return ((H + TY - 1) / TY) * ((W + TX - 1) / TX); }

// ------------------------------------------------------------ weights ------
__global__ void k_refresh_bf16(const float* __restrict__ p, uint16_t* __restrict__ wbf) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= CNN_NBF) return;
  float v;
  if (i < CNN_BF_W2) v = p[CNN_OFF_W1 + i];
  else if (i < CNN_BF_W3) v = p[CNN_OFF_W2 + (i - CNN_BF_W2)];
  else if (i < CNN_BF_W4) v = p[CNN_OFF_W3 + (i - CNN_BF_W3)];
  else if (i < CNN_BF_W2T) v = p[CNN_OFF_W4 + (i - CNN_BF_W4)];
  else {
    // WT[t][n][k] = W[8-t][k][n]
    int j = i < CNN_BF_W3T ? i - CNN_BF_W2T : i - CNN_BF_W3T;
    int off = i < CNN_BF_W3T ? CNN_OFF_W2 : CNN_OFF_W3;
    int t = j / 1024, n = (j / 32) % 32, k = j % 32;
    v = p[off + ((8 - t) * 32 + k) * 32 + n];
  }
  wbf[i] = f2bf(v);
}

__global__ void k_adam(int n, float* __restrict__ p, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.0f - b1) * gi;
  float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ------------------------------------------------------- small layers ------
// layer 1: X0 (4 FP32 channels) -> A1 = relu(conv + b1) (32 bf16 channels)
__global__ void __launch_bounds__(256) k_conv_in(int H, int W, const float* __restrict__ X0,
                                                 const uint16_t* __restrict__ w1,
                                                 const float* __restrict__ b1,
                                                 uint16_t* __restrict__ A1) {
  __shared__ float sw[9 * 32 * 4];
  for (int i = threadIdx.x; i < 9 * 32 * 4; i += 256) sw[i] = bf2f(w1[i]);
  __syncthreads();
  const long npix = (long)H * W;
  const int b = blockIdx.y;
  long p = blockIdx.x * 256L + threadIdx.x;
  if (p >= npix) return;
  int y = (int)(p / W), x = (int)(p % W);
  float acc[32];
#pragma unroll
  for (int c = 0; c < 32; c++) acc[c] = b1[c];
#pragma unroll
  for (int t = 0; t < 9; t++) {
    int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
    if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
    float4 v = reinterpret_cast<const float4*>(X0)[b * npix + (long)yy * W + xx];
#pragma unroll
    for (int c = 0; c < 32; c++) {
      const float* w = sw + (t * 32 + c) * 4;
      acc[c] += w[0] * v.x + w[1] * v.y + w[2] * v.z + w[3] * v.w;
    }
  }
  uint32_t pk[16];
#pragma unroll
  for (int c = 0; c < 16; c++)
    pk[c] = (uint32_t)f2bf(fmaxf(acc[2 * c], 0.0f)) |
            ((uint32_t)f2bf(fmaxf(acc[2 * c + 1], 0.0f)) << 16);
  uint4* dst = reinterpret_cast<uint4*>(A1 + (b * npix + p) * 32);
#pragma unroll
  for (int q = 0; q < 4; q++) dst[q] = make_uint4(pk[4 * q], pk[4 * q + 1], pk[4 * q + 2], pk[4 * q + 3]);
}

__device__ __forceinline__ void load_px32(const uint16_t* src, float* f) {
  const uint4* s = reinterpret_cast<const uint4*>(src);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    uint4 v = s[q];
    uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 4; j++) {
      f[8 * q + 2 * j] = bf2f(w[j] & 0xFFFFu);
      f[8 * q + 2 * j + 1] = bf2f(w[j] >> 16);
    }
  }
}

// layer 4: A3 (32 bf16) -> C = conv + b4 (1 FP32 channel)
__global__ void __launch_bounds__(256) k_conv_out(int H, int W, const uint16_t* __restrict__ A3,
                                                  const uint16_t* __restrict__ w4,
                                                  const float* __restrict__ b4,
                                                  float* __restrict__ C) {
  __shared__ float sw[9 * 32];
  for (int i = threadIdx.x; i < 9 * 32; i += 256) sw[i] = bf2f(w4[i]);
  __syncthreads();
  const long npix = (long)H * W;
  const int b = blockIdx.y;
  long p = blockIdx.x * 256L + threadIdx.x;
  if (p >= npix) return;
  int y = (int)(p / W), x = (int)(p % W);
  float acc = b4[0];
  for (int t = 0; t < 9; t++) {
    int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
    if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
    float a[32];
    load_px32(A3 + (b * npix + (long)yy * W + xx) * 32, a);
#pragma unroll
    for (int c = 0; c < 32; c++) acc += sw[t * 32 + c] * a[c];
  }
  C[b * npix + p] = acc;
}

// input gradient of layer 4, masked by relu'(layer 3): dZ3 = (conv^T(dC, W4)) * (A3 > 0)
__global__ void __launch_bounds__(256) k_convT_out(int H, int W, const float* __restrict__ dC,
                                                   const uint16_t* __restrict__ w4,
                                                   const uint16_t* __restrict__ A3,
                                                   uint16_t* __restrict__ dZ) {
  __shared__ float sw[9 * 32];
  for (int i = threadIdx.x; i < 9 * 32; i += 256) sw[i] = bf2f(w4[i]);
  __syncthreads();
  const long npix = (long)H * W;
  const int b = blockIdx.y;
  long p = blockIdx.x * 256L + threadIdx.x;
  if (p >= npix) return;
  int y = (int)(p / W), x = (int)(p % W);
  float acc[32];
#pragma unroll
  for (int c = 0; c < 32; c++) acc[c] = 0.0f;
#pragma unroll
  for (int t = 0; t < 9; t++) {
    int yy = y - (t / 3 - 1), xx = x - (t % 3 - 1);
    if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
    float d = dC[b * npix + (long)yy * W + xx];
#pragma unroll
    for (int c = 0; c < 32; c++) acc[c] += sw[t * 32 + c] * d;
  }
  const uint4* ms = reinterpret_cast<const uint4*>(A3 + (b * npix + p) * 32);
  uint4* dst = reinterpret_cast<uint4*>(dZ + (b * npix + p) * 32);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    uint4 m = ms[q];
    uint32_t mw[4] = {m.x, m.y, m.z, m.w}, o[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int c = 8 * q + 2 * j;
      uint32_t lo = bf_pos(mw[j] & 0xFFFFu) ? f2bf(acc[c]) : 0u;
      uint32_t hi = bf_pos(mw[j] >> 16) ? f2bf(acc[c + 1]) : 0u;
      o[j] = lo | (hi << 16);
    }
    dst[q] = make_uint4(o[0], o[1], o[2], o[3]);
  }
}

// ------------------------------------------------------- MFMA 32 -> 32 -----
// MODE 0: out = bf16(relu(conv(in, Wf) + bias))
// MODE 1: out = bf16(conv(in, Wf)) where mask > 0, else 0 (input gradient)
// Wf[t][n][k]: n output channel, k input channel.
__device__ __forceinline__ uint32_t mask2(uint32_t v, uint32_t m) {
  uint32_t lo = bf_pos(m & 0xFFFFu) ? (v & 0xFFFFu) : 0u;
  uint32_t hi = bf_pos(m >> 16) ? (v & 0xFFFF0000u) : 0u;
  return lo | hi;
}

template <int MODE>
__global__ void __launch_bounds__(256) k_conv_mid(int H, int W, int tilesX,
                                                  const uint16_t* __restrict__ in,
                                                  const uint16_t* __restrict__ wf,
                                                  const float* __restrict__ bias,
                                                  const uint16_t* __restrict__ mk,
                                                  uint16_t* __restrict__ out) {
  __shared__ __align__(16) uint16_t sIn[HY * HX * LP];
  __shared__ __align__(16) uint16_t sOut[TY * TX * LP];
  const int b = blockIdx.y;
  const int ty0 = (blockIdx.x / tilesX) * TY, tx0 = (blockIdx.x % tilesX) * TX;
  const long npix = (long)H * W;
  const uint16_t* inb = in + (long)b * npix * 32;
  for (int c = threadIdx.x; c < HY * HX * 4; c += 256) {
    int pix = c >> 2, part = c & 3;
    int ly = pix / HX, lx = pix % HX;
    int gy = ty0 + ly - 1, gx = tx0 + lx - 1;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (gy >= 0 && gy < H && gx >= 0 && gx < W)
      v = *reinterpret_cast<const uint4*>(inb + ((long)gy * W + gx) * 32 + part * 8);
    *reinterpret_cast<uint4*>(sIn + pix * LP + part * 8) = v;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, kg = lane >> 4;
  // B fragments: lane holds Wf[t][n = 16*nt + li][k = 16*kh + 4*kg .. +3]
  bf16x4 bfr[9][2][2];
#pragma unroll
  for (int t = 0; t < 9; t++)
#pragma unroll
    for (int kh = 0; kh < 2; kh++)
#pragma unroll
      for (int nt = 0; nt < 2; nt++)
        bfr[t][kh][nt] = *reinterpret_cast<const bf16x4*>(
            wf + ((t * 32 + nt * 16 + li) * 32 + kh * 16 + kg * 4));
  __syncthreads();
  f32x4 acc[4][2];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < 2; nt++) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  // wave w computes output rows 2w and 2w+1 (two 16-pixel M tiles each)
#pragma unroll
  for (int t = 0; t < 9; t++) {
    const int dy = t / 3, dx = t % 3;
#pragma unroll
    for (int kh = 0; kh < 2; kh++) {
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        const int r = wave * 2 + (mt >> 1), c = (mt & 1) * 16 + li;
        // A fragment: lane holds in[pixel (r, c) + tap][k = 16*kh + 4*kg .. +3]
        bf16x4 a = *reinterpret_cast<const bf16x4*>(sIn + ((r + dy) * HX + c + dx) * LP +
                                                     kh * 16 + kg * 4);
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
          acc[mt][nt] =
              __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bfr[t][kh][nt], acc[mt][nt], 0, 0, 0);
      }
    }
  }
  // D fragment: lane holds out[pixel 4*kg + i of the M tile][n = 16*nt + li]
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    const int r = wave * 2 + (mt >> 1);
#pragma unroll
    for (int nt = 0; nt < 2; nt++) {
      const int n = nt * 16 + li;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int c = (mt & 1) * 16 + kg * 4 + i;
        float v = acc[mt][nt][i];
        if (MODE == 0) v = fmaxf(v + bias[n], 0.0f);
        sOut[(r * TX + c) * LP + n] = f2bf(v);
      }
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < TY * TX * 4; c += 256) {
    int pix = c >> 2, part = c & 3;
    int gy = ty0 + pix / TX, gx = tx0 + pix % TX;
    if (gy >= H || gx >= W) continue;
    uint4 v = *reinterpret_cast<const uint4*>(sOut + pix * LP + part * 8);
    long gi = ((long)b * npix + (long)gy * W + gx) * 32 + part * 8;
    if (MODE == 1) {
      uint4 m = *reinterpret_cast<const uint4*>(mk + gi);
      v = make_uint4(mask2(v.x, m.x), mask2(v.y, m.y), mask2(v.z, m.z), mask2(v.w, m.w));
    }
    *reinterpret_cast<uint4*>(out + gi) = v;
  }
}

// ------------------------------------------------------ weight gradients ---
// dW[t][co][ci] = sum_{b,p} D[p][co] In[p + off_t][ci], db[co] = sum D[p][co];
// per-block partials over a grid-stride loop of 8 x 32 tiles.
// 32 x 32 layers: thread -> (co = tid/8, ci = 4*(tid%8) .. +3), 9 taps.
__global__ void __launch_bounds__(256) k_dw32(int B, int H, int W, int tilesX, int ntiles,
                                              const uint16_t* __restrict__ D,
                                              const uint16_t* __restrict__ In,
                                              float* __restrict__ part) {
  __shared__ __align__(16) uint16_t sIn[HY * HX * LP];
  __shared__ __align__(16) uint16_t sD[TY * TX * LP];
  const long npix = (long)H * W;
  const int co = threadIdx.x >> 3, ci0 = (threadIdx.x & 7) * 4;
  float acc[9][4];
#pragma unroll
  for (int t = 0; t < 9; t++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[t][j] = 0.0f;
  float db = 0.0f;
  for (int g = blockIdx.x; g < B * ntiles; g += gridDim.x) {
    const int b = g / ntiles, tile = g % ntiles;
    const int ty0 = (tile / tilesX) * TY, tx0 = (tile % tilesX) * TX;
    __syncthreads();
    for (int c = threadIdx.x; c < HY * HX * 4; c += 256) {
      int pix = c >> 2, pt = c & 3;
      int gy = ty0 + pix / HX - 1, gx = tx0 + pix % HX - 1;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (gy >= 0 && gy < H && gx >= 0 && gx < W)
        v = *reinterpret_cast<const uint4*>(In + ((long)b * npix + (long)gy * W + gx) * 32 + pt * 8);
      *reinterpret_cast<uint4*>(sIn + pix * LP + pt * 8) = v;
    }
    for (int c = threadIdx.x; c < TY * TX * 4; c += 256) {
      int pix = c >> 2, pt = c & 3;
      int gy = ty0 + pix / TX, gx = tx0 + pix % TX;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (gy < H && gx < W)
        v = *reinterpret_cast<const uint4*>(D + ((long)b * npix + (long)gy * W + gx) * 32 + pt * 8);
      *reinterpret_cast<uint4*>(sD + pix * LP + pt * 8) = v;
    }
    __syncthreads();
    for (int pix = 0; pix < TY * TX; pix++) {
      const float d = bf2f(sD[pix * LP + co]);
      db += d;
      const int r = pix / TX, c = pix % TX;
#pragma unroll
      for (int t = 0; t < 9; t++) {
        const uint2 a = *reinterpret_cast<const uint2*>(sIn + ((r + t / 3) * HX + c + t % 3) * LP + ci0);
        acc[t][0] += d * bf2f(a.x & 0xFFFFu);
        acc[t][1] += d * bf2f(a.x >> 16);
        acc[t][2] += d * bf2f(a.y & 0xFFFFu);
        acc[t][3] += d * bf2f(a.y >> 16);
      }
    }
  }
  float* pb = part + (long)blockIdx.x * (9 * 32 * 32 + 32);
#pragma unroll
  for (int t = 0; t < 9; t++)
#pragma unroll
    for (int j = 0; j < 4; j++) pb[(t * 32 + co) * 32 + ci0 + j] = acc[t][j];
  if ((threadIdx.x & 7) == 0) pb[9 * 32 * 32 + co] = db;
}

// small layers: generic thread -> output mapping; D and In read as FP32
template <int CIN, int COUT, class TD, class TI>
__global__ void __launch_bounds__(256) k_dw_small(int B, int H, int W, int tilesX, int ntiles,
                                                  const TD* __restrict__ D,
                                                  const TI* __restrict__ In,
                                                  float* __restrict__ part) {
  constexpr int NW = 9 * COUT * CIN, NOUT = NW + COUT, PER = (NOUT + 255) / 256;
  __shared__ float sIn[HY * HX * CIN];
  __shared__ float sD[TY * TX * COUT];
  const long npix = (long)H * W;
  float acc[PER];
#pragma unroll
  for (int j = 0; j < PER; j++) acc[j] = 0.0f;
  for (int g = blockIdx.x; g < B * ntiles; g += gridDim.x) {
    const int b = g / ntiles, tile = g % ntiles;
    const int ty0 = (tile / tilesX) * TY, tx0 = (tile % tilesX) * TX;
    __syncthreads();
    for (int c = threadIdx.x; c < HY * HX * CIN; c += 256) {
      int pix = c / CIN, k = c % CIN;
      int gy = ty0 + pix / HX - 1, gx = tx0 + pix % HX - 1;
      float v = 0.0f;
      if (gy >= 0 && gy < H && gx >= 0 && gx < W) {
        TI raw = In[((long)b * npix + (long)gy * W + gx) * CIN + k];
        if constexpr (sizeof(TI) == 2) v = bf2f(raw); else v = (float)raw;
      }
      sIn[c] = v;
    }
    for (int c = threadIdx.x; c < TY * TX * COUT; c += 256) {
      int pix = c / COUT, k = c % COUT;
      int gy = ty0 + pix / TX, gx = tx0 + pix % TX;
      float v = 0.0f;
      if (gy < H && gx < W) {
        TD raw = D[((long)b * npix + (long)gy * W + gx) * COUT + k];
        if constexpr (sizeof(TD) == 2) v = bf2f(raw); else v = (float)raw;
      }
      sD[c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; j++) {
      const int o = threadIdx.x + 256 * j;
      if (o >= NOUT) break;
      float s = 0.0f;
      if (o < NW) {
        const int t = o / (COUT * CIN), co = (o / CIN) % COUT, ci = o % CIN;
        const int dy = t / 3, dx = t % 3;
        for (int pix = 0; pix < TY * TX; pix++) {
          const int r = pix / TX, c = pix % TX;
          s += sD[pix * COUT + co] * sIn[((r + dy) * HX + c + dx) * CIN + ci];
        }
      } else {
        const int co = o - NW;
        for (int pix = 0; pix < TY * TX; pix++) s += sD[pix * COUT + co];
      }
      acc[j] += s;
    }
  }
  float* pb = part + (long)blockIdx.x * NOUT;
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const int o = threadIdx.x + 256 * j;
    if (o < NOUT) pb[o] = acc[j];
  }
}

__global__ void k_reduce_part(int nblk, int n, const float* __restrict__ part,
                              float* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int k = 0; k < nblk; k++) s += part[(long)k * n + i];
  out[i] = s;
}

// ------------------------------------------------------------- host --------
int Cnn::init(int H_, int W_, int maxB_) {
  if (H_ < 1 || W_ < 1 || maxB_ < 1) return fail(MW_EINVAL, "cnn: bad image size or batch");
  H = H_; W = W_; maxB = maxB_;
  npix = (long)H * W;
  t_adam = 0;
  RC(params.alloc(CNN_NPARAM)); RC(grads.alloc(CNN_NPARAM));
  RC(m.alloc(CNN_NPARAM)); RC(v.alloc(CNN_NPARAM));
  RC(part.alloc((size_t)CNN_PART_BLOCKS * (9 * 32 * 32 + 32)));
  RC(wbf.alloc(CNN_NBF));
  RC(X0.alloc((size_t)maxB * npix * 4));
  for (DBuf<uint16_t>* b : {&A1, &A2, &A3, &dZa, &dZb}) RC(b->alloc((size_t)maxB * npix * 32));
  MW_HIP(hipMemset(m.p, 0, CNN_NPARAM * 4));
  MW_HIP(hipMemset(v.p, 0, CNN_NPARAM * 4));
  MW_HIP(hipMemset(grads.p, 0, CNN_NPARAM * 4));
  return MW_OK;
}

void Cnn::refresh_bf16(hipStream_t st) {
  k_refresh_bf16<<<nblocks(CNN_NBF, 256), 256, 0, st>>>(params.p, wbf.p);
}

int Cnn::set_params(const float* host, hipStream_t st) {
  MW_HIP(hipMemcpyAsync(params.p, host, CNN_NPARAM * 4, hipMemcpyHostToDevice, st));
  MW_HIP(hipMemsetAsync(m.p, 0, CNN_NPARAM * 4, st));
  MW_HIP(hipMemsetAsync(v.p, 0, CNN_NPARAM * 4, st));
  t_adam = 0;
  refresh_bf16(st);
  MW_CHECK_LAUNCH();
  MW_HIP(hipStreamSynchronize(st));
  return MW_OK;
}

int Cnn::get(const DBuf<float>& b, float* host, hipStream_t st) {
  MW_HIP(hipMemcpyAsync(host, b.p, CNN_NPARAM * 4, hipMemcpyDeviceToHost, st));
  MW_HIP(hipStreamSynchronize(st));
  return MW_OK;
}

int Cnn::forward(int B, const float* x0, float* C, hipStream_t st) {
  if (B < 1 || B > maxB) return fail(MW_EINVAL, "cnn_forward: batch out of range");
  if (x0 != X0.p)
    MW_HIP(hipMemcpyAsync(X0.p, x0, (size_t)B * npix * 16, hipMemcpyDeviceToDevice, st));
  dim3 gp(nblocks(npix, 256), B), gt(tiles_of(H, W), B);
  int tilesX = (W + TX - 1) / TX;
  k_conv_in<<<gp, 256, 0, st>>>(H, W, X0.p, wbf.p + CNN_BF_W1, params.p + CNN_OFF_W1 + 1152, A1.p);
  k_conv_mid<0><<<gt, 256, 0, st>>>(H, W, tilesX, A1.p, wbf.p + CNN_BF_W2,
                                    params.p + CNN_OFF_W2 + 9216, nullptr, A2.p);
  k_conv_mid<0><<<gt, 256, 0, st>>>(H, W, tilesX, A2.p, wbf.p + CNN_BF_W3,
                                    params.p + CNN_OFF_W3 + 9216, nullptr, A3.p);
  k_conv_out<<<gp, 256, 0, st>>>(H, W, A3.p, wbf.p + CNN_BF_W4, params.p + CNN_OFF_W4 + 288, C);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

int Cnn::backward(int B, const float* dC, hipStream_t st) {
  if (B < 1 || B > maxB) return fail(MW_EINVAL, "cnn_backward: batch out of range");
  dim3 gp(nblocks(npix, 256), B), gt(tiles_of(H, W), B);
  const int tilesX = (W + TX - 1) / TX, ntiles = tiles_of(H, W);
  const int nblk = std::min(CNN_PART_BLOCKS, B * ntiles);
  // layer 4
  k_dw_small<32, 1, float, uint16_t><<<nblk, 256, 0, st>>>(B, H, W, tilesX, ntiles, dC, A3.p, part.p);
  k_reduce_part<<<nblocks(289, 256), 256, 0, st>>>(nblk, 289, part.p, grads.p + CNN_OFF_W4);
  k_convT_out<<<gp, 256, 0, st>>>(H, W, dC, wbf.p + CNN_BF_W4, A3.p, dZa.p);
  // layer 3
  k_dw32<<<nblk, 256, 0, st>>>(B, H, W, tilesX, ntiles, dZa.p, A2.p, part.p);
  k_reduce_part<<<nblocks(9248, 256), 256, 0, st>>>(nblk, 9248, part.p, grads.p + CNN_OFF_W3);
  k_conv_mid<1><<<gt, 256, 0, st>>>(H, W, tilesX, dZa.p, wbf.p + CNN_BF_W3T, nullptr, A2.p, dZb.p);
  // layer 2
  k_dw32<<<nblk, 256, 0, st>>>(B, H, W, tilesX, ntiles, dZb.p, A1.p, part.p);
  k_reduce_part<<<nblocks(9248, 256), 256, 0, st>>>(nblk, 9248, part.p, grads.p + CNN_OFF_W2);
  k_conv_mid<1><<<gt, 256, 0, st>>>(H, W, tilesX, dZb.p, wbf.p + CNN_BF_W2T, nullptr, A1.p, dZa.p);
  // layer 1
  k_dw_small<4, 32, uint16_t, float><<<nblk, 256, 0, st>>>(B, H, W, tilesX, ntiles, dZa.p, X0.p,
                                                            part.p);
  k_reduce_part<<<nblocks(1184, 256), 256, 0, st>>>(nblk, 1184, part.p, grads.p + CNN_OFF_W1);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

void Cnn::adam(double lr, double beta1, double beta2, double eps, hipStream_t st) {
  t_adam++;
  double c1 = 1.0 - std::pow(beta1, (double)t_adam), c2 = 1.0 - std::pow(beta2, (double)t_adam);
  k_adam<<<nblocks(CNN_NPARAM, 256), 256, 0, st>>>(CNN_NPARAM, params.p, grads.p, m.p, v.p,
                                                   (float)lr, (float)beta1, (float)beta2,
                                                   (float)eps, (float)c1, (float)c2);
  refresh_bf16(st);
}

}  // namespace mw

using namespace mw;

extern "C" {

int mw_cnn_create(int H, int W, int max_batch, void** net) {
  if (!net) return fail(MW_EINVAL, "cnn_create: null handle");
  Cnn* c = new Cnn();
  int rc = c->init(H, W, max_batch);
  if (rc) {
    delete c;
    return rc;
  }
  *net = c;
  return MW_OK;
}

int mw_cnn_destroy(void* net) {
  delete (Cnn*)net;
  return MW_OK;
}

int mw_cnn_set_params(void* net, const float* params_host, void* stream) {
  if (!net || !params_host) return fail(MW_EINVAL, "cnn_set_params: null argument");
  return ((Cnn*)net)->set_params(params_host, (hipStream_t)stream);
}

int mw_cnn_get_params(void* net, float* params_host, void* stream) {
  if (!net || !params_host) return fail(MW_EINVAL, "cnn_get_params: null argument");
  return ((Cnn*)net)->get(((Cnn*)net)->params, params_host, (hipStream_t)stream);
}

int mw_cnn_get_grads(void* net, float* grads_host, void* stream) {
  if (!net || !grads_host) return fail(MW_EINVAL, "cnn_get_grads: null argument");
  return ((Cnn*)net)->get(((Cnn*)net)->grads, grads_host, (hipStream_t)stream);
}

int mw_cnn_forward(void* net, int batch, const float* X0, float* C, void* stream) {
  if (!net || !X0 || !C) return fail(MW_EINVAL, "cnn_forward: null argument");
  return ((Cnn*)net)->forward(batch, X0, C, (hipStream_t)stream);
}

int mw_cnn_backward(void* net, int batch, const float* dC, void* stream) {
  if (!net || !dC) return fail(MW_EINVAL, "cnn_backward: null argument");
  return ((Cnn*)net)->backward(batch, dC, (hipStream_t)stream);
}

int mw_cnn_adam(void* net, double lr, void* stream) {
  if (!net) return fail(MW_EINVAL, "cnn_adam: null handle");
  ((Cnn*)net)->adam(lr, 0.9, 0.999, 1e-8, (hipStream_t)stream);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

}  // extern "C"
