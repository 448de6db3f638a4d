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
  float *colpart;              // kEpiBackGelu: column sums of h_out per 32-row block
  int stage;                   // 64 x 64 tiles: stage epilogues 0 and 1 through LDS
  const float *w4;             // staged kEpiForward: row dot products of the tile with w4
  float *ypart;                //   ypart[(n0 / 64) * M + m]
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
  float cs = 0.f;  // column sum of the stored h_out values (kEpiBackGelu)
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
      cs += from_bf16(hb);
    } else {
      g.f_out[(int64_t)blockIdx.z * g.split_stride + o] = v;
    }
  }
  // Per 32-row block column sums, as gemm128_kernel writes them.
  if (g.colpart) {
    const float c = cs + __shfl_xor(cs, 32, 64);
    if (lane < 32 && n < g.N && m0 + wm < g.M)
      g.colpart[(int64_t)((m0 + wm) >> 5) * g.ldc + n] = c;
  }
}

// Epilogue of one accumulator element (row m, column n, value v) of split-K
// slice bz; returns the bf16 value stored to h_out (epilogues 0 and 1).
__device__ __forceinline__ uint16_t gemm_store(const GemmArgs &g, int m, int n, float v, int bz) {
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
    g.f_out[(int64_t)bz * g.split_stride + o] = v;
  }
  return 0;
}

// The values of one element for epilogues 0 and 1 (as gemm_store computes
// them): zb the pre-activation (epilogue 0), hb the stored output.
__device__ __forceinline__ void epi_values(const GemmArgs &g, int m, int n, float v, uint16_t &zb,
                                           uint16_t &hb) {
  if (g.epi == kEpiForward) {
    zb = to_bf16(v + g.bias[n]);
    hb = g.act ? to_bf16(gelu(from_bf16(zb))) : zb;
  } else {
    zb = 0;
    hb = to_bf16(v * gelu_grad(from_bf16(g.z_in[(int64_t)m * g.ldc + n])));
  }
}

// Workgroups are dealt to the GPU's 8 XCDs (each with its own L2) in turn by
// linear id.  The tile of a workgroup is chosen so that each XCD takes a
// contiguous run of tiles, x fastest, then y, then z: the rows of A it reads
// (a forward or input-gradient product: 8 row tiles) or its split-K slice (a
// weight-gradient product) then stay in that XCD's L2, where the plain order
// would give every XCD one column tile and make each read all of A.
constexpr int kXcds = 8;
__device__ __forceinline__ void xcd_tile(int &bx, int &by, int &bz) {
  const int nx = gridDim.x, ny = gridDim.y, total = nx * ny * (int)gridDim.z;
  int l = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
  if (total % kXcds == 0) l = (l % kXcds) * (total / kXcds) + l / kXcds;
  bx = l % nx;
  by = (l / nx) % ny;
  bz = l / (nx * ny);
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
  int bx, by, bz;
  xcd_tile(bx, by, bz);
  const int m0 = by * TM, n0 = bx * TN;
  const int wm = (wave >> 1) * (TM / 2), wn = (wave & 1) * (TN / 2);
  const int kb = bz * g.k_chunk;
  const int ke = min(g.K, kb + g.k_chunk);
  f32x16 acc[XB][YB];
#pragma unroll
  for (int a = 0; a < XB; ++a)
#pragma unroll
    for (int b = 0; b < YB; ++b)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[a][b][i] = 0.f;
  // Output-layer partial sums: this tile's 64 entries of w4, fetched now.
  __shared__ float w4s[64];
  float w4v = 0.f;
  if (XB == 1 && YB == 1 && g.ypart && tid < 64 && n0 + tid < g.N) w4v = g.w4[n0 + tid];
  // Staged input-gradient epilogue: its 64 x 64 tile of z_in is fetched now,
  // 16 bytes a load, so that its latency hides behind the products.
  const bool zpre = XB == 1 && YB == 1 && g.stage && g.epi == kEpiBackGelu;
  uint4 zr[2] = {make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)};
  if (zpre) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int e = tid + 256 * p, m = m0 + (e >> 3), nc = n0 + (e & 7) * 8;
      if (m < g.M && nc < g.N)
        zr[p] = *reinterpret_cast<const uint4 *>(g.z_in + (int64_t)m * g.ldc + nc);
    }
  }
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
  // lane % 32 of block y.
  if (XB == 1 && YB == 1 && g.stage) {
    // 64 x 64 tiles, epilogues 0 and 1: the bf16 results go through LDS
    // (the operand tiles, free after the last product) so that h, z and the
    // transposed copy leave as 16-byte stores of 8 consecutive elements.
    __syncthreads();
    uint16_t(*Hs)[KK + 8] = As;
    uint16_t(*Zs)[KK + 8] = Bs;
    if (zpre) {
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int e = tid + 256 * p;
        *reinterpret_cast<uint4 *>(&Zs[e >> 3][(e & 7) * 8]) = zr[p];
      }
      __syncthreads();
    }
    const int nl = wn + (lane & 31), n = n0 + nl;
    float cs0 = 0.f;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int ml = wm + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3), m = m0 + ml;
      uint16_t zb = 0, hb = 0;
      if (m < g.M && n < g.N) {
        if (zpre)  // epi_values' input-gradient value, z from the fetched tile
          hb = to_bf16(acc[0][0][j] * gelu_grad(from_bf16(Zs[ml][nl])));
        else
          epi_values(g, m, n, acc[0][0][j], zb, hb);
      }
      Hs[ml][nl] = hb;
      if (!zpre) Zs[ml][nl] = zb;
      cs0 += from_bf16(hb);
    }
    if (g.ypart && tid < 64) w4s[tid] = w4v;
    __syncthreads();
    if (g.ypart) {
      // Row r of the tile against w4: four threads take 16 columns each,
      // added in a fixed order ((q0 + q1) + (q2 + q3)).
      const int r = tid >> 2, qq = tid & 3;
      float p = 0.f;
#pragma unroll
      for (int c = 16 * qq; c < 16 * qq + 16; ++c)
        if (n0 + c < g.N) p += from_bf16(Hs[r][c]) * w4s[c];
      p += __shfl_xor(p, 1, 64);
      p += __shfl_xor(p, 2, 64);
      if (qq == 0 && m0 + r < g.M) g.ypart[(int64_t)bx * g.M + m0 + r] = p;
    }
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int e = tid + 256 * p, r = e >> 3, cj = (e & 7) * 8;
      const int m = m0 + r, nc = n0 + cj;
      if (m < g.M && nc < g.N) {
        const int64_t o = (int64_t)m * g.ldc + nc;
        *reinterpret_cast<uint4 *>(g.h_out + o) = *reinterpret_cast<const uint4 *>(&Hs[r][cj]);
        if (g.z_out)
          *reinterpret_cast<uint4 *>(g.z_out + o) = *reinterpret_cast<const uint4 *>(&Zs[r][cj]);
      }
    }
    if (g.ht_out) {
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int e = tid + 256 * p, c = e >> 3, rj = (e & 7) * 8;
        const int nc = n0 + c, m = m0 + rj;
        if (nc < g.N && m < g.M) {
          uint32_t w[4];
#pragma unroll
          for (int k = 0; k < 4; ++k)
            w[k] = (uint32_t)Hs[rj + 2 * k][c] | ((uint32_t)Hs[rj + 2 * k + 1][c] << 16);
          *reinterpret_cast<uint4 *>(g.ht_out + (int64_t)nc * g.ldt + m) =
              make_uint4(w[0], w[1], w[2], w[3]);
        }
      }
    }
    if (g.colpart) {
      const float c = cs0 + __shfl_xor(cs0, 32, 64);
      if (lane < 32 && n < g.N && m0 + wm < g.M)
        g.colpart[(int64_t)((m0 + wm) >> 5) * g.ldc + n] = c;
    }
    return;
  }
  // Otherwise element by element; the 4 rows of each j / 4 go to the
  // transposed copy as one 8-byte store.
  float cs[XB][YB];
#pragma unroll
  for (int x = 0; x < XB; ++x)
#pragma unroll
    for (int y = 0; y < YB; ++y) cs[x][y] = 0.f;
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
        for (int i = 0; i < 4; ++i) hv[i] = gemm_store(g, mb + i, n, acc[x][y][4 * q + i], bz);
        if (g.ht_out && mb < g.M && n < g.N) {
          const uint2 w2 = make_uint2((uint32_t)hv[0] | ((uint32_t)hv[1] << 16),
                                      (uint32_t)hv[2] | ((uint32_t)hv[3] << 16));
          *reinterpret_cast<uint2 *>(g.ht_out + (int64_t)n * g.ldt + mb) = w2;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) cs[x][y] += from_bf16(hv[i]);
      }
  // Column sums of the stored bf16 values over each 32-row block (rows past
  // M stored nothing and add 0): lanes l and l + 32 hold the same column.
  if (g.colpart) {
#pragma unroll
    for (int x = 0; x < XB; ++x)
#pragma unroll
      for (int y = 0; y < YB; ++y) {
        const float c = cs[x][y] + __shfl_xor(cs[x][y], 32, 64);
        const int n = n0 + wn + y * 32 + (lane & 31), mblk = (m0 + wm + x * 32) >> 5;
        if (lane < 32 && n < g.N && m0 + wm + x * 32 < g.M) g.colpart[(int64_t)mblk * g.ldc + n] = c;
      }
  }
}

// y[b] = sum_k h[b, k] w[k] + bias   (one wave per row); with targets t,
// also the loss gradient dy = (y - t) / B and the squared error per row.
__global__ __launch_bounds__(256) void gemv_out_kernel(const uint16_t *h, const float *w,
                                                       const float *bias, float *y, int B, int K,
                                                       const float *t, float *dy, float *sq) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= B) return;
  float s = 0.f;
  for (int k = lane; k < K; k += 64) s += from_bf16(h[(int64_t)row * K + k]) * w[k];
  for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
  if (lane == 0) {
    const float yv = s + bias[0];
    y[row] = yv;
    if (t) {
      const float d = yv - t[row];
      dy[row] = d / (float)B;
      sq[row] = d * d;
    }
  }
}

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
// The output layer's backward pass, one 64 x 64 tile per block: dz = dy w
// gelu'(z) (bf16), row-major and (through LDS) transposed, and the sums over
// the tile's 64 rows of dz, of dy h and of dy that make the gradients of
// b3, w4 and b4.  Thread t takes column t % 64 and rows t / 64 + 4 j; its
// four partial sums per column are added in order.
__global__ __launch_bounds__(256) void out_back_kernel(float *dy, const float *w,
                                                       const uint16_t *z, const uint16_t *h,
                                                       uint16_t *dz, uint16_t *dzt, int B, int K,
                                                       float *pdz, float *pdw, float *pdb,
                                                       OutLoss out) {
  __shared__ uint16_t tile[64][66];
  __shared__ float red[3][4][64];
  __shared__ float dys[64];
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64, tid = threadIdx.x;
  const int c = tid & 63, q = tid >> 6;
  if (out.ypart) {
    // The output and the loss terms of the block's 64 rows (the partial dot
    // products added in column-tile order); column block 0 stores them.
    if (tid < 64) {
      const int b = r0 + tid;
      float dv = 0.f;
      if (b < B) {
        float yv = 0.f;
        for (int t0 = 0; t0 < out.nyp; t0 += 8) {
          float v[8];
#pragma unroll
          for (int k = 0; k < 8; ++k)
            v[k] = t0 + k < out.nyp ? out.ypart[(int64_t)(t0 + k) * B + b] : 0.f;
#pragma unroll
          for (int k = 0; k < 8; ++k)
            if (t0 + k < out.nyp) yv += v[k];
        }
        yv += out.b4[0];
        const float d = yv - out.t[b];
        dv = d / (float)B;
        if (blockIdx.x == 0) {
          out.y[b] = yv;
          dy[b] = dv;
          out.sq[b] = d * d;
        }
      }
      dys[tid] = dv;
    }
    __syncthreads();
  }
  const bool colok = c0 + c < K;
  float sdz = 0.f, sdw = 0.f, sdb = 0.f;
#pragma unroll 4
  for (int r = q; r < 64; r += 4) {
    const int b = r0 + r;
    if (b < B && colok) {
      const int64_t o = (int64_t)b * K + c0 + c;
      const float d = out.ypart ? dys[r] : dy[b];
      const uint16_t v = to_bf16(d * w[c0 + c] * gelu_grad(from_bf16(z[o])));
      dz[o] = v;
      tile[r][c] = v;
      sdz += from_bf16(v);
      sdw += d * from_bf16(h[o]);
      sdb += d;
    }
  }
  red[0][q][c] = sdz;
  red[1][q][c] = sdw;
  red[2][q][c] = sdb;
  __syncthreads();
#pragma unroll 4
  for (int e = tid; e < 64 * 64; e += 256) {
    const int cc = e >> 6, rr = e & 63;
    if (r0 + rr < B && c0 + cc < K) dzt[(int64_t)(c0 + cc) * B + r0 + rr] = tile[rr][cc];
  }
  if (q == 0 && colok) {
    pdz[(int64_t)blockIdx.y * K + c0 + c] =
        ((red[0][0][c] + red[0][1][c]) + red[0][2][c]) + red[0][3][c];
    pdw[(int64_t)blockIdx.y * K + c0 + c] =
        ((red[1][0][c] + red[1][1][c]) + red[1][2][c]) + red[1][3][c];
  }
  if (blockIdx.x == 0 && tid == 0)
    pdb[blockIdx.y] = ((red[2][0][0] + red[2][1][0]) + red[2][2][0]) + red[2][3][0];
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

// Slice sums of up to kMaxJobs jobs (blockIdx.y picks one), each in the
// fixed order of sum_slices_kernel: the split-K slices of the weight
// gradients and the per-row-block column sums of the bias gradients.
constexpr int kWideSlices = 32;
__global__ void sum_slices_multi_kernel(SliceJobs j) {
  if (blockIdx.y == gridDim.y - 1 && j.loss.sq) {
    // The loss job: block 0 adds sq in a fixed order, stores the mean in
    // this step's slot and advances the clock (the products rounded on their
    // own, as the host computes them).
    if (blockIdx.x != 0) return;
    const LossTick &L = j.loss;
    __shared__ float wsum[4];
    // Thread t adds rows t, t + 256, ..., eight loads in flight at a time;
    // then a butterfly per wave and the four waves in order.
    float sum = 0.f;
    for (int b0 = 0; b0 < L.B; b0 += 8 * 256) {
      float v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int b = b0 + k * 256 + (int)threadIdx.x;
        v[k] = b < L.B ? L.sq[b] : 0.f;
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) sum += v[k];
    }
    for (int m = 32; m >= 1; m >>= 1) sum += __shfl_xor(sum, m, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
      sum = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
      const int it = L.clk->it;
      L.out[it] = sum / (float)L.B;
      L.clk->it = it + 1;
      L.clk->p1 = __fmul_rn(L.clk->p1, L.b1);
      L.clk->p2 = __fmul_rn(L.clk->p2, L.b2);
    }
    return;
  }
  const SliceJob &c = j.job[blockIdx.y];
  if (c.nsplit >= kWideSlices) {
    // Many slices (the bias gradients' row blocks): one wave per output;
    // lane l adds slices l, l + 64, ... in order, then the lanes are added
    // by a fixed butterfly.
    const int64_t i = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= c.n) return;  // the whole wave
    float sum = 0.f;
    for (int z = lane; z < c.nsplit; z += 64) sum += c.part[(int64_t)z * c.stride + i];
    for (int m = 32; m >= 1; m >>= 1) sum += __shfl_xor(sum, m, 64);
    if (lane == 0) c.out[i] = sum;
    return;
  }
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.n) return;
  float sum = c.part[i];
  for (int z = 1; z < c.nsplit; ++z) sum += c.part[(int64_t)z * c.stride + i];
  c.out[i] = sum;
}

// Adam (Kingma & Ba, 2015) with bias correction over the whole flat
// parameter vector in one launch; elements of the kAdamBf16 segments
// [off, off + n) also refresh their bf16 copies.  The clock holds b1^t and
// b2^t for this step t: c1 = 1 - b1^t, c2 = 1 - b2^t.
__global__ void adam_all_kernel(float *w, float *m, float *v, const float *g, AdamBf16 seg,
                                int n, float lr, float b1, float b2, float eps,
                                const TrainClock *clk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float c1 = 1.f - clk->p1, c2 = 1.f - clk->p2;
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
                                                     uint64_t seed, const TrainClock *clk, int n,
                                                     uint16_t *Xb, uint16_t *Xbt, float *tb,
                                                     int B, int ld) {
  const int b0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const uint32_t step = (uint32_t)clk->it;
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
  g.colpart = q.epi == mlp::kEpiBackGelu ? q.colpart : nullptr;
  // The staged epilogue stores 8 elements (16 bytes) at a time.
  auto al16 = [](const void *p) { return ((uintptr_t)p & 15) == 0; };
  g.stage = q.epi != mlp::kEpiStoreF32 && q.ldc % 8 == 0 && al16(q.h_out) &&
            (!q.z_out || al16(q.z_out)) && (!q.ht_out || (q.ldt % 8 == 0 && al16(q.ht_out)));
  g.w4 = q.w4;
  g.ypart = nullptr;
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
  // The output-layer partial sums need the staged 64 x 64 epilogue.
  if (q.ypart && q.epi == mlp::kEpiForward && wide && g.stage && tm == 64 && tn == 64 && split == 1)
    g.ypart = q.ypart;
  if (q.ypart_done) *q.ypart_done = g.ypart != nullptr;
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
                        int K, hipStream_t s, const float *t, float *dy, float *sq) {
  hipLaunchKernelGGL(mlp::gemv_out_kernel, dim3(blocks(B, 4)), dim3(256), 0, s, h, w, bias, y, B, K,
                     t, dy, sq);
  return hipGetLastError();
}

hipError_t mlp_out_back(float *dy, const float *w, const uint16_t *z, const uint16_t *h,
                        uint16_t *dz, uint16_t *dzt, int B, int K, float *pdz, float *pdw,
                        float *pdb, const OutLoss &out, hipStream_t s) {
  hipLaunchKernelGGL(mlp::out_back_kernel, dim3(blocks(K, 64), blocks(B, 64)), dim3(256), 0, s, dy,
                     w, z, h, dz, dzt, B, K, pdz, pdw, pdb, out);
  return hipGetLastError();
}

hipError_t mlp_transpose_bf16(const uint16_t *x, int rows, int cols, uint16_t *yt, hipStream_t s) {
  hipLaunchKernelGGL(mlp::transpose_bf16_kernel, dim3(blocks((int64_t)rows * cols, 256)), dim3(256),
                     0, s, x, rows, cols, yt);
  return hipGetLastError();
}

hipError_t mlp_sum_slices_multi(const SliceJobs &jobs, int njobs, hipStream_t s) {
  if (njobs < 0 || njobs > kMaxJobs || (njobs == 0 && !jobs.loss.sq)) return hipErrorInvalidValue;
  int bx = 1;
  for (int i = 0; i < njobs; ++i) {
    const SliceJob &c = jobs.job[i];
    bx = std::max(bx, c.nsplit >= mlp::kWideSlices ? blocks(c.n, 4) : blocks(c.n, 256));
  }
  hipLaunchKernelGGL(mlp::sum_slices_multi_kernel, dim3(bx, njobs + (jobs.loss.sq ? 1 : 0)),
                     dim3(256), 0, s, jobs);
  return hipGetLastError();
}

hipError_t mlp_adam_all(float *w, float *m, float *v, const float *g, const AdamBf16 &seg, int n,
                        float lr, float b1, float b2, float eps, const TrainClock *clk,
                        hipStream_t s) {
  hipLaunchKernelGGL(mlp::adam_all_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, w, m, v, g, seg, n,
                     lr, b1, b2, eps, clk);
  return hipGetLastError();
}

hipError_t mlp_to_bf16(const float *x, uint16_t *y, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(mlp::f32_to_bf16_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, x, y, n);
  return hipGetLastError();
}

hipError_t mlp_gather(const uint16_t *X, const float *t, uint64_t seed, const TrainClock *clk,
                      int n, uint16_t *Xb, uint16_t *Xbt, float *tb, int B, int ld,
                      hipStream_t s) {
  hipLaunchKernelGGL(mlp::gather_kernel, dim3(blocks(ld, 64), blocks(B, 64)), dim3(256), 0, s, X, t,
                     seed, clk, n, Xb, Xbt, tb, B, ld);
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
