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
  // epi 1: optional column sums of h_out (as stored, bf16) over each block
  // of 32 rows, colpart[(m / 32) * ldc + n]; mlp_sum_slices_multi finishes
  // them (the bias gradients).
  float *colpart;
  // epi 0, 64 x 64 tiles (training, last hidden layer): optional per column
  // tile dot products of h_out's rows with w4, ypart[(n / 64) * M + m]; the
  // output layer's backward pass adds them.  *ypart_done tells whether they
  // were written (else the caller computes the output with mlp_gemv_out).
  const float *w4;
  float *ypart;
  int *ypart_done;
};

// Batched reductions of the training step (one launch each).
// Training-step state kept on the device, so that the kernels of a step take
// the same arguments every step (and a captured graph of steps can be
// replayed): `it` counts the steps of the current training call (minibatch
// stream and loss slot), p1 = b1^t and p2 = b2^t for the Adam step t being
// taken (the bias corrections, as running products; advanced before Adam).
struct TrainClock {
  int32_t it, pad;
  float p1, p2;
};
// The loss and clock job of a training step's slice sums: out[clk->it] =
// mean of sq[0..B); then it + 1, p1 b1, p2 b2.
struct LossTick {
  const float *sq;
  int B;
  float *out;
  TrainClock *clk;
  float b1, b2;
};
constexpr int kMaxJobs = 8;
struct SliceJob {             // out[i] = sum_z part[z * stride + i], z < nsplit, in order
  const float *part;
  int64_t stride;
  int nsplit;
  float *out;
  int64_t n;
};
struct SliceJobs {
  SliceJob job[kMaxJobs];
  LossTick loss;  // optional (loss.sq null: none)
};
constexpr int kAdamBf16 = 3;
struct AdamBf16 {             // flat ranges whose bf16 copies Adam refreshes
  uint16_t *wb[kAdamBf16];
  uint16_t *wbt[kAdamBf16];   // optional transposed copies (cols x rows)
  int off[kAdamBf16], n[kAdamBf16], cols[kAdamBf16];
};

hipError_t mlp_gemm(const MlpGemm &q, hipStream_t s);
// y = h w + bias per row; with targets t (else null), also dy = (y - t) / B
// and sq = (y - t)^2.
hipError_t mlp_gemv_out(const uint16_t *h, const float *w, const float *bias, float *y, int B,
                        int K, hipStream_t s, const float *t = nullptr, float *dy = nullptr,
                        float *sq = nullptr);
// Output layer backward: dz = dy w gelu'(z) (bf16) and its transpose dzt,
// with column sums over each block of 64 rows, [b / 64][k]: of dz (pdz),
// of dy[b] h[b, k] (pdw, the output weights' gradient) and of dy (pdb, one
// column).
// With out.ypart set, the output y = sum of the nyp partial dot products +
// b4 and the loss terms (y, dy = (y - t) / B, sq = (y - t)^2 for rows of
// column block 0) are computed here first; otherwise dy is read.
struct OutLoss {
  const float *ypart;
  int nyp;
  const float *b4, *t;
  float *y, *sq;
};
hipError_t mlp_out_back(float *dy, const float *w, const uint16_t *z, const uint16_t *h,
                        uint16_t *dz, uint16_t *dzt, int B, int K, float *pdz, float *pdw,
                        float *pdb, const OutLoss &out, hipStream_t s);
// yt[c * rows + r] = x[r * cols + c]
hipError_t mlp_transpose_bf16(const uint16_t *x, int rows, int cols, uint16_t *yt, hipStream_t s);
// Adam with c1 = 1 - p1, c2 = 1 - p2 from `clk`.
hipError_t mlp_adam_all(float *w, float *m, float *v, const float *g, const AdamBf16 &seg, int n,
                        float lr, float b1, float b2, float eps, const TrainClock *clk,
                        hipStream_t s);
hipError_t mlp_sum_slices_multi(const SliceJobs &jobs, int njobs, hipStream_t s);
hipError_t mlp_to_bf16(const float *x, uint16_t *y, int64_t n, hipStream_t s);
// Minibatch of B rows drawn on the device from (seed, clk->it).
hipError_t mlp_gather(const uint16_t *X, const float *t, uint64_t seed, const TrainClock *clk,
                      int n, uint16_t *Xb, uint16_t *Xbt, float *tb, int B, int ld,
                      hipStream_t s);

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
