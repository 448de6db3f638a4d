// Fitness surrogate MLP: kernel launchers (mlp_kernels.hip) and the host
// model (mlp.cpp).
#ifndef EMT6RO_AMD_MLP_H_
#define EMT6RO_AMD_MLP_H_

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

namespace emt6ro {

// C[m, n] = sum_k A(m, k) B(k, n), A(m, k) = A[m*sam + k*sak],
// B(k, n) = B[k*sbk + n*sbn], bf16 in, fp32 accumulate.  Epilogue `epi`:
// 0 z = acc + bias[n] -> z_out (bf16), h_out = act ? gelu(z) : z (bf16);
// 1 h_out = acc * gelu'(z_in) (bf16); 2 f_out = acc (fp32).  Output row
// stride ldc.
struct MlpGemm {
  const uint16_t *A, *B;
  int64_t sam, sak, sbk, sbn;
  int M, N, K;
  int epi;
  const float *bias;
  int act;
  uint16_t *z_out, *h_out;
  const uint16_t *z_in;
  float *f_out;
  int64_t ldc;
  // epi 2 only: scratch for split-K slices (may be null: no split).
  float *work;
  int64_t work_floats;
  // Split-K: the number of slices goes to *split_out (if set); with
  // defer_sum the slices stay in `work` for a later mlp_sum_slices_multi.
  int *split_out;
  int defer_sum;
  // epi 0 / 1: optional transposed copy of h_out, ht_out[n * ldt + m].
  uint16_t *ht_out;
  int64_t ldt;
};

// Batched reductions of the training step (one launch each).
constexpr int kMaxJobs = 8;
struct ColsumJob {            // out[k] = sum_b x[b, k] (times rowscale[b] if set)
  const uint16_t *x;
  const float *rowscale;
  float *out, *part;          // part: ceil(B / 128) x K floats of scratch
  int K;
};
struct ColsumJobs { ColsumJob job[kMaxJobs]; };
struct SliceJob {             // out[i] = sum_z part[z * stride + i], z < nsplit, in order
  const float *part;
  int64_t stride;
  int nsplit;
  float *out;
  int64_t n;
};
struct SliceJobs { SliceJob job[kMaxJobs]; };
constexpr int kAdamBf16 = 3;
struct AdamBf16 {             // flat ranges whose bf16 copies Adam refreshes
  uint16_t *wb[kAdamBf16];
  uint16_t *wbt[kAdamBf16];   // optional transposed copies (cols x rows)
  int off[kAdamBf16], n[kAdamBf16], cols[kAdamBf16];
};

hipError_t mlp_gemm(const MlpGemm &q, hipStream_t s);
hipError_t mlp_gemv_out(const uint16_t *h, const float *w, const float *bias, float *y, int B,
                        int K, hipStream_t s);
hipError_t mlp_loss(const float *y, const float *t, float *dy, float *sq, int B, hipStream_t s);
hipError_t mlp_out_back(const float *dy, const float *w, const uint16_t *z, uint16_t *dz,
                        uint16_t *dzt, int B, int K, hipStream_t s);
// yt[c * rows + r] = x[r * cols + c]
hipError_t mlp_transpose_bf16(const uint16_t *x, int rows, int cols, uint16_t *yt, hipStream_t s);
// `work` holds ceil(B / 128) * K floats.
hipError_t mlp_colsum(const uint16_t *x, const float *rowscale, float *out, int B, int K,
                      float *work, hipStream_t s);
hipError_t mlp_adam(float *w, float *m, float *v, const float *g, uint16_t *wb, int n, float lr,
                    float b1, float b2, float eps, int t, hipStream_t s);
hipError_t mlp_adam_all(float *w, float *m, float *v, const float *g, const AdamBf16 &seg, int n,
                        float lr, float b1, float b2, float eps, int t, hipStream_t s);
hipError_t mlp_colsum_multi(const ColsumJobs &jobs, int njobs, int B, hipStream_t s);
hipError_t mlp_sum_slices_multi(const SliceJobs &jobs, int njobs, hipStream_t s);
hipError_t mlp_to_bf16(const float *x, uint16_t *y, int64_t n, hipStream_t s);
// Minibatch of B rows drawn on the device from (seed, step).
hipError_t mlp_gather(const uint16_t *X, const float *t, uint64_t seed, uint32_t step, int n,
                      uint16_t *Xb, uint16_t *Xbt, float *tb, int B, int ld, hipStream_t s);
hipError_t mlp_mean(const float *sq, int B, float *out, hipStream_t s);

// Forward pass that allocates nothing: activations in a fixed device array
// of three kNoAllocRows x kNoAllocMaxCols bf16 buffers (mlp_kernels.hip).
constexpr int kNoAllocRows = 4096;      // rows per pass (the predict chunk)
constexpr int kNoAllocMaxCols = 512;    // in, padded in and hidden at most
struct MlpForward {
  const uint16_t *wb[3];    // bf16 weights of the hidden layers
  const float *bias[3];
  const float *w4, *b4;     // output layer (fp32)
  int in, kin, H;
};
// y[0..n) = network(X rows of `in` fp32 columns); X and y in device memory.
hipError_t mlp_forward_noalloc(const MlpForward &net, int device, const float *X, int n, float *y,
                               hipStream_t s);

}  // namespace emt6ro

#endif  // EMT6RO_AMD_MLP_H_
