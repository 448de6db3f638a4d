// Fitness surrogate MLP kernels for gfx950: BF16 matrix-core GEMMs with FP32
// accumulation and fused bias / GELU epilogues, GEMV kernels for the scalar
// output layer, column reductions, and Adam.
//
// The GEMM computes C[m, n] = sum_k A(m, k) B(k, n) for A, B given by element
// strides, so one kernel serves the forward product X W^T, the input gradient
// dZ W and the weight gradient dZ^T X.  A 256-thread workgroup owns a 64x64
// tile of C; each of its four waves owns a 32x32 quarter and accumulates it
// with v_mfma_f32_32x32x8_bf16 over K in steps of 32 staged through LDS.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include <mutex>

#include "mlp.h"

namespace emt6ro {
namespace mlp {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> BF16, round to nearest even (NaN kept quiet).
__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// tanh(u) = sign(u) (1 - e) / (1 + e), e = exp(-2 |u|), from the hardware
// exp2 and reciprocal (absolute error a few 1e-8, far below the bf16
// rounding of every value it feeds); the GEMM epilogues evaluate it once per
// output element.
__device__ __forceinline__ float fast_tanh(float u) {
  const float e = __builtin_amdgcn_exp2f(-2.8853900817779268f * fabsf(u));  // exp(-2|u|)
  const float t = (1.f - e) * __builtin_amdgcn_rcpf(1.f + e);
  return copysignf(t, u);
}

// GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))).
__device__ __forceinline__ float gelu(float x) {
  const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
  return 0.5f * x * (1.f + fast_tanh(u));
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
  const float t = fast_tanh(u);
  return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * 0.7978845608028654f * (1.f + 0.134145f * x * x);
}

constexpr int kTile = 64, kTK = 32, kLdsK = kTK + 4;  // LDS row: 32 bf16 + pad

// Epilogue operations.
enum Epi : int {
  kEpiForward = 0,   // z = acc + bias[n]; store z (bf16) and gelu(z) (bf16) or z if no act
  kEpiBackGelu = 1,  // out = acc * gelu'(z[m, n]) (bf16)
  kEpiStoreF32 = 2,  // out = acc (fp32)
};

struct GemmArgs {
  const uint16_t *A, *B;
  int64_t sam, sak, sbk, sbn;  // element strides
  int M, N, K;
  int epi;
  const float *bias;           // kEpiForward
  int act;                     // kEpiForward: apply GELU
  uint16_t *z_out;             // kEpiForward: pre-activation (ldc)
  uint16_t *h_out;             // kEpiForward / kEpiBackGelu output (ldc)
  const uint16_t *z_in;        // kEpiBackGelu: pre-activation of this layer's input
  float *f_out;                // kEpiStoreF32
  int64_t ldc;
  int k_chunk;                 // reduction slice of one blockIdx.z (multiple of kTK)
  int64_t split_stride;        // kEpiStoreF32 with split K: floats between slices
  uint16_t *ht_out;            // optional transposed copy of h_out: ht_out[n * ldt + m]
  int64_t ldt;
};

__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
  __shared__ uint16_t As[kTile][kLdsK];
  __shared__ uint16_t Bs[kTile][kLdsK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x16 acc;
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = 0.f;
  const bool a_kfast = g.sak == 1, b_kfast = g.sbk == 1;
  // Split K: slice blockIdx.z covers [kb, ke) of the reduction.
  const int kb = blockIdx.z * g.k_chunk;
  const int ke = min(g.K, kb + g.k_chunk);
  for (int k0 = kb; k0 < ke; k0 += kTK) {
    // Stage a 64 x 32 tile of A and of B (rows m / n, k contiguous in LDS).
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = tid + e * 256;  // 0..2047
      const int r = a_kfast ? idx / kTK : idx % kTile;
      const int c = a_kfast ? idx % kTK : idx / kTile;
      const int m = m0 + r, k = k0 + c;
      As[r][c] = (m < g.M && k < ke) ? g.A[m * g.sam + k * g.sak] : (uint16_t)0;
      const int rb = b_kfast ? idx / kTK : idx % kTile;
      const int cb = b_kfast ? idx % kTK : idx / kTile;
      const int n = n0 + rb, kk = k0 + cb;
      Bs[rb][cb] = (n < g.N && kk < ke) ? g.B[kk * g.sbk + n * g.sbn] : (uint16_t)0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTK; kk += 8) {
      const int kq = kk + (lane >> 5) * 4;
      bf16x4 a, b;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[i] = (short)As[wm + (lane & 31)][kq + i];
        b[i] = (short)Bs[wn + (lane & 31)][kq + i];
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // acc[j]: row 8 (j / 4) + 4 (lane / 32) + j % 4, column lane % 32.
  const int n = n0 + wn + (lane & 31);
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int m = m0 + wm + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
    if (m >= g.M || n >= g.N) continue;
    const int64_t o = (int64_t)m * g.ldc + n;
    const float v = acc[j];
    if (g.epi == kEpiForward) {
      const float z = v + g.bias[n];
      const uint16_t zb = to_bf16(z);
      if (g.z_out) g.z_out[o] = zb;
      const uint16_t hb = g.act ? to_bf16(gelu(from_bf16(zb))) : zb;
      g.h_out[o] = hb;
      if (g.ht_out) g.ht_out[(int64_t)n * g.ldt + m] = hb;
    } else if (g.epi == kEpiBackGelu) {
      const uint16_t hb = to_bf16(v * gelu_grad(from_bf16(g.z_in[o])));
      g.h_out[o] = hb;
      if (g.ht_out) g.ht_out[(int64_t)n * g.ldt + m] = hb;
    } else {
      g.f_out[(int64_t)blockIdx.z * g.split_stride + o] = v;
    }
  }
}

// Epilogue of one accumulator element (row m, column n, value v); returns
// the bf16 value stored to h_out (epilogues 0 and 1).
__device__ __forceinline__ uint16_t gemm_store(const GemmArgs &g, int m, int n, float v) {
  if (m >= g.M || n >= g.N) return 0;
  const int64_t o = (int64_t)m * g.ldc + n;
  if (g.epi == kEpiForward) {
    const float z = v + g.bias[n];
    const uint16_t zb = to_bf16(z);
    if (g.z_out) g.z_out[o] = zb;
    const uint16_t hb = g.act ? to_bf16(gelu(from_bf16(zb))) : zb;
    g.h_out[o] = hb;
    return hb;
  } else if (g.epi == kEpiBackGelu) {
    const uint16_t hb = to_bf16(v * gelu_grad(from_bf16(g.z_in[o])));
    g.h_out[o] = hb;
    return hb;
  } else {
    g.f_out[(int64_t)blockIdx.z * g.split_stride + o] = v;
  }
  return 0;
}

// The same product on TM x TN tiles (128 x 128, 128 x 64 or 64 x 64) with 16-byte
// loads, for M, N, K and the strides multiples of 8.  The four waves split
// the tile 2 x 2; each owns (TM / 2) x (TN / 2) as blocks of
// v_mfma_f32_32x32x16_bf16 (lane l holds row l % 32 and the 8 consecutive k
// from 8 (l / 32) of each operand).  AK / BK: operand A (resp. B) is
// contiguous along k; otherwise along m (resp. n), and its 8-element chunks
// are transposed into LDS.  The k-tile is KK wide; the next one is fetched
// into registers while the current one is multiplied.  The order of the
// products over k does not depend on TN or KK.
typedef short bf16x8 __attribute__((ext_vector_type(8)));
#ifndef EMT6RO_GEMM_KK
#define EMT6RO_GEMM_KK 128
#endif
constexpr int kT2 = 128, kK2 = EMT6RO_GEMM_KK;

template <bool AK, int ROWS, int KK>
__device__ __forceinline__ void load_chunks(const uint16_t *P, int64_t s_mn, int64_t s_k, int mn0,
                                            int k0, int MN, int ke, uint4 *v) {
  constexpr int NC = ROWS * KK / 8 / 256;   // 8-element chunks per thread
#pragma unroll
  for (int r = 0; r < NC; ++r) {
    const int c = (int)threadIdx.x + r * 256;
    int mn, k;
    if (AK) { mn = mn0 + c / (KK / 8); k = k0 + (c % (KK / 8)) * 8; }
    else { k = k0 + c / (ROWS / 8); mn = mn0 + (c % (ROWS / 8)) * 8; }
    const bool ok = mn < MN && k < ke;
    v[r] = ok ? *reinterpret_cast<const uint4 *>(P + mn * s_mn + k * s_k) : make_uint4(0, 0, 0, 0);
  }
}

template <bool AK, int ROWS, int KK>
__device__ __forceinline__ void store_chunks(uint16_t (*T)[KK + 8], const uint4 *v) {
  constexpr int NC = ROWS * KK / 8 / 256;
#pragma unroll
  for (int r = 0; r < NC; ++r) {
    const int c = (int)threadIdx.x + r * 256;
    if (AK) {
      *reinterpret_cast<uint4 *>(&T[c / (KK / 8)][(c % (KK / 8)) * 8]) = v[r];
    } else {
      const int k = c / (ROWS / 8), mn = (c % (ROWS / 8)) * 8;
      const uint16_t *e = reinterpret_cast<const uint16_t *>(&v[r]);
#pragma unroll
      for (int i = 0; i < 8; ++i) T[mn + i][k] = e[i];
    }
  }
}

template <bool AK, bool BK, int TM, int TN>
__global__ __launch_bounds__(256) void gemm128_kernel(GemmArgs g) {
  constexpr int KK = kK2, XB = TM / 64, YB = TN / 64;
  constexpr int NA = TM * KK / 8 / 256, NB = TN * KK / 8 / 256;
  __shared__ __align__(16) uint16_t As[TM][KK + 8];
  __shared__ __align__(16) uint16_t Bs[TN][KK + 8];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int wm = (wave >> 1) * (TM / 2), wn = (wave & 1) * (TN / 2);
  const int kb = blockIdx.z * g.k_chunk;
  const int ke = min(g.K, kb + g.k_chunk);
  f32x16 acc[XB][YB];
#pragma unroll
  for (int a = 0; a < XB; ++a)
#pragma unroll
    for (int b = 0; b < YB; ++b)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[a][b][i] = 0.f;
  // A(m, k) = A[m sam + k sak]; B(k, n) = B[k sbk + n sbn].
  uint4 ra[NA], rb[NB];
  load_chunks<AK, TM, KK>(g.A, g.sam, g.sak, m0, kb, g.M, ke, ra);
  load_chunks<BK, TN, KK>(g.B, g.sbn, g.sbk, n0, kb, g.N, ke, rb);
  for (int k0 = kb; k0 < ke; k0 += KK) {
    __syncthreads();
    store_chunks<AK, TM, KK>(As, ra);
    store_chunks<BK, TN, KK>(Bs, rb);
    __syncthreads();
    if (k0 + KK < ke) {
      load_chunks<AK, TM, KK>(g.A, g.sam, g.sak, m0, k0 + KK, g.M, ke, ra);
      load_chunks<BK, TN, KK>(g.B, g.sbn, g.sbk, n0, k0 + KK, g.N, ke, rb);
    }
#pragma unroll
    for (int kk = 0; kk < KK; kk += 16) {
      const int kq = kk + (lane >> 5) * 8;
      bf16x8 a[XB], b[YB];
#pragma unroll
      for (int x = 0; x < XB; ++x)
        a[x] = *reinterpret_cast<const bf16x8 *>(&As[wm + x * 32 + (lane & 31)][kq]);
#pragma unroll
      for (int y = 0; y < YB; ++y)
        b[y] = *reinterpret_cast<const bf16x8 *>(&Bs[wn + y * 32 + (lane & 31)][kq]);
#pragma unroll
      for (int x = 0; x < XB; ++x)
#pragma unroll
        for (int y = 0; y < YB; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[x], b[y], acc[x][y], 0, 0, 0);
    }
  }
  // acc[x][y][j]: row 8 (j / 4) + 4 (lane / 32) + j % 4 of block x, column
  // lane % 32 of block y; the 4 rows of each j / 4 go to the transposed copy
  // as one 8-byte store.
#pragma unroll
  for (int x = 0; x < XB; ++x)
#pragma unroll
    for (int y = 0; y < YB; ++y)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int mb = m0 + wm + x * 32 + 8 * q + 4 * (lane >> 5);
        const int n = n0 + wn + y * 32 + (lane & 31);
        uint16_t hv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) hv[i] = gemm_store(g, mb + i, n, acc[x][y][4 * q + i]);
        if (g.ht_out && mb < g.M && n < g.N) {
          const uint2 w2 = make_uint2((uint32_t)hv[0] | ((uint32_t)hv[1] << 16),
                                      (uint32_t)hv[2] | ((uint32_t)hv[3] << 16));
          *reinterpret_cast<uint2 *>(g.ht_out + (int64_t)n * g.ldt + mb) = w2;
        }
      }
}

// y[b] = sum_k h[b, k] w[k] + bias   (one wave per row)
__global__ __launch_bounds__(256) void gemv_out_kernel(const uint16_t *h, const float *w,
                                                       const float *bias, float *y, int B, int K) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= B) return;
  float s = 0.f;
  for (int k = lane; k < K; k += 64) s += from_bf16(h[(int64_t)row * K + k]) * w[k];
  for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
  if (lane == 0) y[row] = s + bias[0];
}

// Loss gradient dy = (y - t) / B and per-row squared error.
__global__ void loss_kernel(const float *y, const float *t, float *dy, float *sq, int B) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  const float d = y[i] - t[i];
  dy[i] = d / (float)B;
  sq[i] = d * d;
}

// dz[b, k] = dy[b] w[k] gelu'(z[b, k])  (bf16)
// A 64 x 64 tile of a row-major bf16 result, written to out (row-major,
// ld cols) and through LDS to out_t (its transpose, ld rows), both with
// coalesced stores.  value(r, c) gives the element; 256 threads.
template <typename F>
__device__ __forceinline__ void tile_store_both(int r0, int c0, int rows, int cols,
                                                uint16_t *out, uint16_t *out_t, const F &value) {
  __shared__ uint16_t tile[64][66];
  const int tid = threadIdx.x;
#pragma unroll 4
  for (int q = tid; q < 64 * 64; q += 256) {
    const int r = q >> 6, c = q & 63;
    if (r0 + r < rows && c0 + c < cols) {
      const uint16_t v = value(r0 + r, c0 + c);
      out[(int64_t)(r0 + r) * cols + c0 + c] = v;
      tile[r][c] = v;
    }
  }
  if (!out_t) return;
  __syncthreads();
#pragma unroll 4
  for (int q = tid; q < 64 * 64; q += 256) {
    const int c = q >> 6, r = q & 63;
    if (r0 + r < rows && c0 + c < cols) out_t[(int64_t)(c0 + c) * rows + r0 + r] = tile[r][c];
  }
}

// dz = dy w4 gelu'(z3) (and its transpose), one 64 x 64 tile per block.
__global__ __launch_bounds__(256) void out_back_kernel(const float *dy, const float *w,
                                                       const uint16_t *z, uint16_t *dz,
                                                       uint16_t *dzt, int B, int K) {
  tile_store_both(blockIdx.y * 64, blockIdx.x * 64, B, K, dz, dzt, [&](int b, int k) {
    return to_bf16(dy[b] * w[k] * gelu_grad(from_bf16(z[(int64_t)b * K + k])));
  });
}

// yt[c * rows + r] = x[r * cols + c] (bf16).
__global__ void transpose_bf16_kernel(const uint16_t *x, int rows, int cols, uint16_t *yt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)rows * cols) return;
  const int r = (int)(i / cols), c = (int)(i % cols);
  yt[(int64_t)c * rows + r] = x[i];
}

// out[i] = sum over s < nsplit of part[s * stride + i] (fixed order).
__global__ void sum_slices_kernel(const float *part, int64_t stride, int nsplit, float *out,
                                  int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = part[i];
  for (int z = 1; z < nsplit; ++z) s += part[(int64_t)z * stride + i];
  out[i] = s;
}

// Column sums, stage 1: block (bx, by) sums rows [by * kColRows, ...) of
// columns [bx * 64, ...): out[k] = sum_b x[b, k] (x bf16), or sum_b dy[b] x[b, k].
constexpr int kColRows = 128;
__global__ __launch_bounds__(256) void colsum_part_kernel(const uint16_t *x, const float *rowscale,
                                                          float *part, int B, int K) {
  __shared__ float red[4][64];
  const int k = blockIdx.x * 64 + (threadIdx.x & 63), r0 = threadIdx.x >> 6;
  const int b0 = blockIdx.y * kColRows, b1 = min(B, b0 + kColRows);
  float s = 0.f;
  if (k < K)
    for (int b = b0 + r0; b < b1; b += 4) {
      const float v = from_bf16(x[(int64_t)b * K + k]);
      s += rowscale ? rowscale[b] * v : v;
    }
  red[r0][threadIdx.x & 63] = s;
  __syncthreads();
  if (r0 == 0 && k < K)
    part[(int64_t)blockIdx.y * K + k] =
        ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

// Batched forms of the two kernels above for the training step: blockIdx.z
// (column sums) or blockIdx.y (slice sums) picks one of up to kMaxJobs jobs,
// each summed exactly as the single-job kernels do.
__global__ __launch_bounds__(256) void colsum_part_multi_kernel(ColsumJobs j, int B) {
  const ColsumJob &c = j.job[blockIdx.z];
  if ((int)blockIdx.x * 64 >= c.K) return;
  __shared__ float red[4][64];
  const int K = c.K;
  const int k = blockIdx.x * 64 + (threadIdx.x & 63), r0 = threadIdx.x >> 6;
  const int b0 = blockIdx.y * kColRows, b1 = min(B, b0 + kColRows);
  float sum = 0.f;
  if (k < K)
    for (int b = b0 + r0; b < b1; b += 4) {
      const float v = from_bf16(c.x[(int64_t)b * K + k]);
      sum += c.rowscale ? c.rowscale[b] * v : v;
    }
  red[r0][threadIdx.x & 63] = sum;
  __syncthreads();
  if (r0 == 0 && k < K)
    c.part[(int64_t)blockIdx.y * K + k] =
        ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

__global__ void sum_slices_multi_kernel(SliceJobs j) {
  const SliceJob &c = j.job[blockIdx.y];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.n) return;
  float sum = c.part[i];
  for (int z = 1; z < c.nsplit; ++z) sum += c.part[(int64_t)z * c.stride + i];
  c.out[i] = sum;
}

// Adam (Kingma & Ba, 2015) with bias correction; refreshes the bf16 copy.
__global__ void adam_kernel(float *w, float *m, float *v, const float *g, uint16_t *wb, int n,
                            float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = to_bf16(wi);
}

// Adam over the whole flat parameter vector in one launch; elements of the
// kAdamBf16 segments [off, off + n) also refresh their bf16 copies.
__global__ void adam_all_kernel(float *w, float *m, float *v, const float *g, AdamBf16 seg,
                                int n, float lr, float b1, float b2, float eps, float c1,
                                float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
#pragma unroll
  for (int k = 0; k < kAdamBf16; ++k)
    if (i >= seg.off[k] && i < seg.off[k] + seg.n[k]) {
      const int e = i - seg.off[k];
      const uint16_t b = to_bf16(wi);
      seg.wb[k][e] = b;
      if (seg.wbt[k]) {
        const int r = e / seg.cols[k], c = e - r * seg.cols[k];
        seg.wbt[k][(int64_t)c * (seg.n[k] / seg.cols[k]) + r] = b;
      }
    }
}

__global__ void f32_to_bf16_kernel(const float *x, uint16_t *y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = to_bf16(x[i]);
}

// Minibatch row index: splitmix64 of (seed, step, row) reduced onto [0, n).
__device__ __forceinline__ int32_t pick_row(uint64_t seed, uint32_t step, uint32_t b, int n) {
  uint64_t z = seed + 0x9e3779b97f4a7c15ull * (((uint64_t)step << 32 | b) + 1);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (int32_t)(z % (uint64_t)n);
}

// Rows of a dataset (bf16, ld columns) gathered into a minibatch; the rows
// are drawn on the device, so a training step needs no host round trip.
// One 64-row x 64-column tile of the minibatch (and its transpose).
__global__ __launch_bounds__(256) void gather_kernel(const uint16_t *X, const float *t,
                                                     uint64_t seed, uint32_t step, int n,
                                                     uint16_t *Xb, uint16_t *Xbt, float *tb,
                                                     int B, int ld) {
  const int b0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  __shared__ int32_t src[64];
  if (threadIdx.x < 64 && b0 + (int)threadIdx.x < B) {
    const int b = b0 + threadIdx.x;
    src[threadIdx.x] = pick_row(seed, step, (uint32_t)b, n);
    if (blockIdx.x == 0) tb[b] = t[src[threadIdx.x]];
  }
  __syncthreads();
  tile_store_both(b0, c0, B, ld, Xb, Xbt, [&](int b, int c) {
    return X[(int64_t)src[b - b0] * ld + c];
  });
}

// out[0] = mean of sq[0..B) (one workgroup).
__global__ __launch_bounds__(256) void mean_kernel(const float *sq, int B, float *out) {
  __shared__ float part[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < B; i += 256) s += sq[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0] / (float)B;
}

}  // namespace mlp

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
namespace {
inline int blocks(int64_t n, int t) { return (int)((n + t - 1) / t); }
}  // namespace

#ifndef EMT6RO_GEMM_MIN_GROUPS
#define EMT6RO_GEMM_MIN_GROUPS 512
#endif
constexpr int kMinGroups = EMT6RO_GEMM_MIN_GROUPS;

template <int TM, int TN>
void launch_wide(bool a_k, bool b_k, dim3 grid, const mlp::GemmArgs &g, hipStream_t s) {
  if (a_k && b_k) hipLaunchKernelGGL((mlp::gemm128_kernel<true, true, TM, TN>), grid, dim3(256), 0, s, g);
  else if (a_k) hipLaunchKernelGGL((mlp::gemm128_kernel<true, false, TM, TN>), grid, dim3(256), 0, s, g);
  else if (b_k) hipLaunchKernelGGL((mlp::gemm128_kernel<false, true, TM, TN>), grid, dim3(256), 0, s, g);
  else hipLaunchKernelGGL((mlp::gemm128_kernel<false, false, TM, TN>), grid, dim3(256), 0, s, g);
}

hipError_t mlp_gemm(const MlpGemm &q, hipStream_t s) {
  mlp::GemmArgs g{};
  g.A = q.A; g.B = q.B;
  g.sam = q.sam; g.sak = q.sak; g.sbk = q.sbk; g.sbn = q.sbn;
  g.M = q.M; g.N = q.N; g.K = q.K;
  g.epi = q.epi; g.bias = q.bias; g.act = q.act;
  g.z_out = q.z_out; g.h_out = q.h_out; g.z_in = q.z_in; g.f_out = q.f_out; g.ldc = q.ldc;
  g.ht_out = q.ht_out; g.ldt = q.ldt;
  const bool a_k = q.sak == 1, b_k = q.sbk == 1;
  const int64_t sa = a_k ? q.sam : q.sak, sb = b_k ? q.sbn : q.sbk;
  const bool wide = q.M % 8 == 0 && q.N % 8 == 0 && q.K % 8 == 0 && sa % 8 == 0 &&
                    sb % 8 == 0 && (a_k || q.sam == 1) && (b_k || q.sbn == 1) &&
                    ((uintptr_t)q.A & 15) == 0 && ((uintptr_t)q.B & 15) == 0;
  // Tiles: on the wide path the largest of 128 x 128, 128 x 64 and 64 x 64
  // that still gives kMinGroups workgroups (two per CU), else 64 x 64; the
  // 64 x 64 kernel of the narrow path otherwise.
  int tm = wide ? mlp::kT2 : mlp::kTile;
  int tn = wide ? mlp::kT2 : mlp::kTile;
  auto ntiles = [&](int tn_) { return ((q.N + tn_ - 1) / tn_) * ((q.M + tm - 1) / tm); };
  // Split a long reduction over blockIdx.z when the output has few tiles
  // (the weight gradients: 512 x 512 outputs, 4096-row batches); the
  // slices land in `work` and are summed in a fixed order.
  const int kq = wide ? mlp::kK2 : mlp::kTK;
  auto nsplit = [&](int tiles) {
    int sp = 1;
    if (q.epi == mlp::kEpiStoreF32 && q.work && q.work_floats > 0)
      while (sp < 16 && tiles * sp * 2 <= 1024 && q.K / (sp * 2) >= 4 * kq &&
             (int64_t)(sp * 2) * q.M * q.ldc <= q.work_floats)
        sp *= 2;
    return sp;
  };
  int split = nsplit(ntiles(tn));
  if (wide && ntiles(tn) * split < kMinGroups) {
    tn = 64;
    split = nsplit(ntiles(tn));
  }
  if (wide && ntiles(tn) * split < kMinGroups) {
    tm = 64;
    split = nsplit(ntiles(tn));
  }
  g.k_chunk = ((q.K + split - 1) / split + kq - 1) / kq * kq;
  split = (q.K + g.k_chunk - 1) / g.k_chunk;
  if (split > 1) {
    g.f_out = q.work;
    g.split_stride = (int64_t)q.M * q.ldc;
  }
  if (wide) {
    dim3 grid2((q.N + tn - 1) / tn, (q.M + tm - 1) / tm, split);
    if (tm == 64)
      launch_wide<64, 64>(a_k, b_k, grid2, g, s);
    else if (tn == 64)
      launch_wide<128, 64>(a_k, b_k, grid2, g, s);
    else
      launch_wide<128, 128>(a_k, b_k, grid2, g, s);
  } else {
    dim3 grid((q.N + mlp::kTile - 1) / mlp::kTile, (q.M + mlp::kTile - 1) / mlp::kTile, split);
    hipLaunchKernelGGL(mlp::gemm_bf16_kernel, grid, dim3(256), 0, s, g);
  }
  if (q.split_out) *q.split_out = split;
  if (split > 1 && !q.defer_sum) {
    const int64_t n = (int64_t)q.M * q.ldc;
    hipLaunchKernelGGL(mlp::sum_slices_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, q.work,
                       n, split, q.f_out, n);
  }
  return hipGetLastError();
}

hipError_t mlp_gemv_out(const uint16_t *h, const float *w, const float *bias, float *y, int B,
                        int K, hipStream_t s) {
  hipLaunchKernelGGL(mlp::gemv_out_kernel, dim3(blocks(B, 4)), dim3(256), 0, s, h, w, bias, y, B, K);
  return hipGetLastError();
}

hipError_t mlp_loss(const float *y, const float *t, float *dy, float *sq, int B, hipStream_t s) {
  hipLaunchKernelGGL(mlp::loss_kernel, dim3(blocks(B, 256)), dim3(256), 0, s, y, t, dy, sq, B);
  return hipGetLastError();
}

hipError_t mlp_out_back(const float *dy, const float *w, const uint16_t *z, uint16_t *dz,
                        uint16_t *dzt, int B, int K, hipStream_t s) {
  hipLaunchKernelGGL(mlp::out_back_kernel, dim3(blocks(K, 64), blocks(B, 64)), dim3(256), 0, s, dy,
                     w, z, dz, dzt, B, K);
  return hipGetLastError();
}

hipError_t mlp_transpose_bf16(const uint16_t *x, int rows, int cols, uint16_t *yt, hipStream_t s) {
  hipLaunchKernelGGL(mlp::transpose_bf16_kernel, dim3(blocks((int64_t)rows * cols, 256)), dim3(256),
                     0, s, x, rows, cols, yt);
  return hipGetLastError();
}

hipError_t mlp_colsum(const uint16_t *x, const float *rowscale, float *out, int B, int K,
                      float *work, hipStream_t s) {
  const int nb = (B + mlp::kColRows - 1) / mlp::kColRows;
  hipLaunchKernelGGL(mlp::colsum_part_kernel, dim3(blocks(K, 64), nb), dim3(256), 0, s, x,
                     rowscale, work, B, K);
  hipLaunchKernelGGL(mlp::sum_slices_kernel, dim3(blocks(K, 256)), dim3(256), 0, s, work,
                     (int64_t)K, nb, out, (int64_t)K);
  return hipGetLastError();
}

hipError_t mlp_colsum_multi(const ColsumJobs &jobs, int njobs, int B, hipStream_t s) {
  if (njobs <= 0 || njobs > kMaxJobs) return hipErrorInvalidValue;
  int kmax = 0;
  SliceJobs sj{};
  const int nb = (B + mlp::kColRows - 1) / mlp::kColRows;
  for (int i = 0; i < njobs; ++i) {
    kmax = std::max(kmax, jobs.job[i].K);
    sj.job[i] = SliceJob{jobs.job[i].part, (int64_t)jobs.job[i].K, nb, jobs.job[i].out,
                         (int64_t)jobs.job[i].K};
  }
  hipLaunchKernelGGL(mlp::colsum_part_multi_kernel, dim3(blocks(kmax, 64), nb, njobs), dim3(256), 0,
                     s, jobs, B);
  return mlp_sum_slices_multi(sj, njobs, s);
}

hipError_t mlp_sum_slices_multi(const SliceJobs &jobs, int njobs, hipStream_t s) {
  if (njobs <= 0 || njobs > kMaxJobs) return hipErrorInvalidValue;
  int64_t nmax = 0;
  for (int i = 0; i < njobs; ++i) nmax = std::max(nmax, jobs.job[i].n);
  hipLaunchKernelGGL(mlp::sum_slices_multi_kernel, dim3(blocks(nmax, 256), njobs), dim3(256), 0, s,
                     jobs);
  return hipGetLastError();
}

hipError_t mlp_adam_all(float *w, float *m, float *v, const float *g, const AdamBf16 &seg, int n,
                        float lr, float b1, float b2, float eps, int t, hipStream_t s) {
  float c1 = 1.f, c2 = 1.f;
  for (int i = 0; i < t; ++i) { c1 *= b1; c2 *= b2; }
  c1 = 1.f - c1;
  c2 = 1.f - c2;
  hipLaunchKernelGGL(mlp::adam_all_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, w, m, v, g, seg, n,
                     lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t mlp_adam(float *w, float *m, float *v, const float *g, uint16_t *wb, int n, float lr,
                    float b1, float b2, float eps, int t, hipStream_t s) {
  float c1 = 1.f, c2 = 1.f;
  for (int i = 0; i < t; ++i) { c1 *= b1; c2 *= b2; }
  c1 = 1.f - c1;
  c2 = 1.f - c2;
  hipLaunchKernelGGL(mlp::adam_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, w, m, v, g, wb, n, lr,
                     b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t mlp_to_bf16(const float *x, uint16_t *y, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(mlp::f32_to_bf16_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, x, y, n);
  return hipGetLastError();
}

hipError_t mlp_gather(const uint16_t *X, const float *t, uint64_t seed, uint32_t step, int n,
                      uint16_t *Xb, uint16_t *Xbt, float *tb, int B, int ld, hipStream_t s) {
  hipLaunchKernelGGL(mlp::gather_kernel, dim3(blocks(ld, 64), blocks(B, 64)), dim3(256), 0, s, X, t,
                     seed, step, n, Xb, Xbt, tb, B, ld);
  return hipGetLastError();
}

hipError_t mlp_mean(const float *sq, int B, float *out, hipStream_t s) {
  hipLaunchKernelGGL(mlp::mean_kernel, dim3(1), dim3(256), 0, s, sq, B, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Forward pass that allocates nothing (emt6ro_mlp_predict_noalloc): its
// activations live in a fixed device array, three bf16 buffers of
// kNoAllocRows x kNoAllocMaxCols (the padded input and two hidden layers,
// used in turn), 12 MiB that each device holds until the process ends.
// Every call on a device shares that device's copy, so calls that overlap on
// one device race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) uint16_t scratch[3 * (size_t)kNoAllocRows * kNoAllocMaxCols];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint16_t *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

namespace mlp {
// Rows of X (in columns, fp32) -> bf16 rows of kin columns, zero padded;
// the same rounding as f32_to_bf16_kernel.
__global__ void pad_to_bf16_kernel(const float *X, int64_t rows, int in, int kin, uint16_t *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * kin) return;
  const int64_t r = i / kin;
  const int c = (int)(i - r * kin);
  out[i] = c < in ? to_bf16(X[r * in + c]) : (uint16_t)0;
}
}  // namespace mlp

hipError_t mlp_noalloc_scratch(int device, uint16_t **out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    // The array's device address on this device (the caller made it current);
    // the host-side name `scratch` is not a device address.
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<uint16_t *>(p);
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

hipError_t mlp_forward_noalloc(const MlpForward &net, int device, const float *X, int n, float *y,
                               hipStream_t s) {
  if (net.in > kNoAllocMaxCols || net.kin > kNoAllocMaxCols || net.H > kNoAllocMaxCols)
    return hipErrorInvalidValue;
  uint16_t *base = nullptr;
  hipError_t e = mlp_noalloc_scratch(device, &base);
  if (e != hipSuccess) return e;
  const size_t buf = (size_t)kNoAllocRows * kNoAllocMaxCols;
  uint16_t *xb = base, *h[2] = {base + buf, base + 2 * buf};
  for (int r0 = 0; r0 < n; r0 += kNoAllocRows) {
    const int b = n - r0 < kNoAllocRows ? n - r0 : kNoAllocRows;
    hipLaunchKernelGGL(mlp::pad_to_bf16_kernel, dim3(blocks((int64_t)b * net.kin, 256)), dim3(256),
                       0, s, X + (int64_t)r0 * net.in, (int64_t)b, net.in, net.kin, xb);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const uint16_t *in = xb;
    int K = net.kin;
    for (int l = 0; l < 3; ++l) {
      MlpGemm q{};
      q.A = in; q.sam = K; q.sak = 1;
      q.B = net.wb[l]; q.sbk = 1; q.sbn = K;
      q.M = b; q.N = net.H; q.K = K;
      q.epi = 0; q.bias = net.bias[l]; q.act = 1;
      q.z_out = nullptr; q.h_out = h[l & 1]; q.ldc = net.H;
      if ((e = mlp_gemm(q, s)) != hipSuccess) return e;
      in = h[l & 1];
      K = net.H;
    }
    if ((e = mlp_gemv_out(in, net.w4, net.b4, y + r0, b, net.H, s)) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace emt6ro
