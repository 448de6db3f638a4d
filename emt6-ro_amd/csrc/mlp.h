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
};

hipError_t mlp_gemm(const MlpGemm &q, hipStream_t s);
hipError_t mlp_gemv_out(const uint16_t *h, const float *w, const float *bias, float *y, int B,
                        int K, hipStream_t s);
hipError_t mlp_loss(const float *y, const float *t, float *dy, float *sq, int B, hipStream_t s);
hipError_t mlp_out_back(const float *dy, const float *w, const uint16_t *z, uint16_t *dz, int B,
                        int K, hipStream_t s);
// `work` holds ceil(B / 128) * K floats.
hipError_t mlp_colsum(const uint16_t *x, const float *rowscale, float *out, int B, int K,
                      float *work, hipStream_t s);
hipError_t mlp_adam(float *w, float *m, float *v, const float *g, uint16_t *wb, int n, float lr,
                    float b1, float b2, float eps, int t, hipStream_t s);
hipError_t mlp_to_bf16(const float *x, uint16_t *y, int64_t n, hipStream_t s);
// Minibatch of B rows drawn on the device from (seed, step).
hipError_t mlp_gather(const uint16_t *X, const float *t, uint64_t seed, uint32_t step, int n,
                      uint16_t *Xb, float *tb, int B, int ld, hipStream_t s);
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
