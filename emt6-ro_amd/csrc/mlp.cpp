// Fitness surrogate: an MLP in -> H -> H -> H -> 1 with GELU, BF16 weights and
// activations on the matrix cores, FP32 accumulation, FP32 master weights and
// Adam, trained by the kernels of mlp_kernels.hip.  C ABI at the end
// (include/emt6ro_amd.h, emt6ro_mlp_*).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/emt6ro_amd.h"
#include "host.h"
#include "mlp.h"

namespace emt6ro {

#define EMT_HIPM(call)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw Error(cat("HIP error ", hipGetErrorString(e_), " at ", #call), -3);          \
  } while (0)

class Mlp {
 public:
  static constexpr int kLayers = 3;  // hidden layers

  Mlp(int in, int hidden, uint64_t seed, int device)
      : in_(in), kin_((in + 31) / 32 * 32), H_(hidden), device_(device) {
    EMT_ENFORCE(in > 0 && hidden > 0 && hidden % 64 == 0, -1,
                "MLP needs in > 0 and hidden a multiple of 64");
    EMT_HIPM(hipSetDevice(device_));
    EMT_HIPM(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    // Parameters: W1 [H, kin] (padded columns zero), b1, W2 [H, H], b2,
    // W3 [H, H], b3, w4 [H], b4 [1].
    sizes_ = {(int64_t)H_ * kin_, H_, (int64_t)H_ * H_, H_, (int64_t)H_ * H_, H_, H_, 1};
    int64_t tot = 0;
    for (int64_t n : sizes_) { offs_.push_back(tot); tot += n; }
    total_ = tot;
    w_ = dalloc<float>(tot); m_ = dalloc<float>(tot); v_ = dalloc<float>(tot); g_ = dalloc<float>(tot);
    EMT_HIPM(hipMemsetAsync(m_, 0, tot * 4, s_));
    EMT_HIPM(hipMemsetAsync(v_, 0, tot * 4, s_));
    EMT_HIPM(hipMemsetAsync(g_, 0, tot * 4, s_));
    for (int l = 0; l < kLayers; ++l) wb_[l] = dalloc<uint16_t>(sizes_[2 * l]);
    // Transposed bf16 copies of W2 and W3 ([in x out]), the k-contiguous
    // operand of the input-gradient products.
    for (int l = 1; l < kLayers; ++l) wbt_[l] = dalloc<uint16_t>(sizes_[2 * l]);
    // Split-K slices of the weight-gradient products: up to 16 slices of the
    // largest weight matrix, for each hidden layer (summed together at the
    // end of the backward pass).
    kwork_floats_ = 16 * std::max<int64_t>((int64_t)H_ * kin_, (int64_t)H_ * H_);
    kwork_ = dalloc<float>(kLayers * kwork_floats_);
    clock_ = dalloc<TrainClock>(1);
    // PyTorch-style init: U(-1/sqrt(fan_in), 1/sqrt(fan_in)).
    std::mt19937_64 rng(seed);
    std::vector<float> h(tot, 0.f);
    const int fan[4] = {in_, H_, H_, H_};
    for (int l = 0; l < 4; ++l) {
      const float bnd = 1.f / std::sqrt((float)fan[l]);
      std::uniform_real_distribution<float> u(-bnd, bnd);
      float *W = &h[offs_[2 * l]], *b = &h[offs_[2 * l + 1]];
      const int rows = l < 3 ? H_ : 1, cols = l == 0 ? in_ : H_, ld = l == 0 ? kin_ : H_;
      for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) W[(int64_t)r * ld + c] = u(rng);
      for (int r = 0; r < rows; ++r) b[r] = u(rng);
    }
    upload_params(h);
  }

  ~Mlp() {
    (void)hipSetDevice(device_);
    if (s_) (void)hipStreamSynchronize(s_);
    for (void *p : allocs_) (void)hipFree(p);
    if (s_) (void)hipStreamDestroy(s_);
  }

  int64_t num_params() const { return (int64_t)H_ * in_ + H_ + 2 * ((int64_t)H_ * H_ + H_) + H_ + 1; }

  // Flat parameters without the padding columns: W1, b1, W2, b2, W3, b3, w4, b4.
  void get_params(float *out) {
    EMT_HIPM(hipSetDevice(device_));
    std::vector<float> h(total_);
    EMT_HIPM(hipMemcpyAsync(h.data(), w_, total_ * 4, hipMemcpyDeviceToHost, s_));
    EMT_HIPM(hipStreamSynchronize(s_));
    unpad(h, out);
  }

  void set_params(const float *in) {
    EMT_HIPM(hipSetDevice(device_));
    std::vector<float> h(total_, 0.f);
    pad(in, h);
    upload_params(h);
    EMT_HIPM(hipMemsetAsync(m_, 0, total_ * 4, s_));
    EMT_HIPM(hipMemsetAsync(v_, 0, total_ * 4, s_));
    step_ = 0;
  }

  void set_data(const float *X, const float *t, int n) {
    EMT_HIPM(hipSetDevice(device_));
    EMT_ENFORCE(n > 0, -1, "empty data set");
    release(data_x_);
    release(data_t_);
    n_data_ = 0;
    std::vector<float> xp((size_t)n * kin_, 0.f);
    for (int r = 0; r < n; ++r) std::memcpy(&xp[(size_t)r * kin_], &X[(size_t)r * in_], in_ * 4);
    float *tmp = dalloc<float>(xp.size());
    data_x_ = dalloc<uint16_t>(xp.size());
    data_t_ = dalloc<float>(n);
    EMT_HIPM(hipMemcpyAsync(tmp, xp.data(), xp.size() * 4, hipMemcpyHostToDevice, s_));
    EMT_HIPM(mlp_to_bf16(tmp, data_x_, (int64_t)xp.size(), s_));
    EMT_HIPM(hipMemcpyAsync(data_t_, t, (size_t)n * 4, hipMemcpyHostToDevice, s_));
    EMT_HIPM(hipStreamSynchronize(s_));
    release(tmp);
    n_data_ = n;
  }

  // Adam on minibatches drawn from the data set on the device; returns each
  // step's mean squared error.  The steps are enqueued back to back and the
  // host waits once, at the end.  With graph_steps_ > 0 the kernels of that
  // many steps are captured once into a HIP graph, which is replayed; the
  // remaining steps are enqueued directly.  Both orders run the same kernels
  // on the same arguments: what changes from step to step (minibatch stream,
  // loss slot, Adam bias corrections) is read from the device clock.
  std::vector<float> train(int steps, int batch, float lr, uint64_t seed) {
    float *dl = train_begin(steps, batch);
    if (!dl) return {};
    int done = 0;
    const int g = graph_steps_;
    if (g > 0 && steps >= g) {
      hipGraphExec_t exec = capture_steps(g, batch, lr, seed, dl);
      hipError_t e = hipSuccess;
      for (; e == hipSuccess && done + g <= steps; done += g) e = hipGraphLaunch(exec, s_);
      if (e == hipSuccess) e = hipStreamSynchronize(s_);
      (void)hipGraphExecDestroy(exec);
      EMT_HIPM(e);
    }
    for (; done < steps; ++done) enqueue_step(batch, lr, seed, dl);
    return train_end(dl, steps);
  }

  // The kernels of g training steps captured from this network's stream
  // into an executable graph (the caller launches and destroys it).
  hipGraphExec_t capture_steps(int g, int batch, float lr, uint64_t seed, float *dl) {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    EMT_HIPM(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
    try {
      for (int k = 0; k < g; ++k) enqueue_step(batch, lr, seed, dl);
    } catch (...) {
      (void)hipStreamEndCapture(s_, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    EMT_HIPM(hipStreamEndCapture(s_, &graph));
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    EMT_HIPM(e);
    return exec;
  }

  hipStream_t stream() const { return s_; }
  int graph_steps() const { return graph_steps_; }

  // train() in parts, so that several networks can interleave their steps
  // on their own streams (emt6ro_mlp_train_many): train_begin() checks the
  // arguments, sizes the batch buffers, allocates the loss slots and sets the
  // device clock (null when there are no steps); enqueue_step() adds one
  // step; train_end() waits for this network's stream and returns the losses.
  float *train_begin(int steps, int batch) {
    EMT_HIPM(hipSetDevice(device_));
    EMT_ENFORCE(n_data_ > 0, -1, "set_data first");
    EMT_ENFORCE(steps >= 0 && batch > 0, -1, "steps >= 0 and batch > 0");
    ensure_batch(batch);
    if (steps == 0) return nullptr;
    float *dl = dalloc<float>(steps);
    // b1^t and b2^t for the t Adam steps taken so far; each step advances
    // them before its Adam update.
    TrainClock c{0, 0, 1.f, 1.f};
    for (int i = 0; i < step_; ++i) { c.p1 *= kB1; c.p2 *= kB2; }
    EMT_HIPM(hipMemcpyAsync(clock_, &c, sizeof c, hipMemcpyHostToDevice, s_));
    return dl;
  }

  std::vector<float> train_end(float *dl, int steps) {
    std::vector<float> losses(steps);
    step_ += steps;
    EMT_HIPM(hipMemcpyAsync(losses.data(), dl, (size_t)steps * 4, hipMemcpyDeviceToHost, s_));
    EMT_HIPM(hipStreamSynchronize(s_));
    release(dl);
    return losses;
  }

  int device() const { return device_; }

  // One training step: minibatch, forward, backward (whose slice sums also
  // store the loss and advance the device clock), Adam.
  void enqueue_step(int batch, float lr, uint64_t seed, float *dl) {
    EMT_HIPM(mlp_gather(data_x_, data_t_, seed, clock_, n_data_, x_, xt_, t_, batch, kin_, s_));
    forward(batch, true);
    backward(batch, dl);
    AdamBf16 seg{};
    for (int l = 0; l < kLayers; ++l) {
      seg.wb[l] = wb_[l];
      seg.wbt[l] = wbt_[l];
      seg.off[l] = (int)offs_[2 * l];
      seg.n[l] = (int)sizes_[2 * l];
      seg.cols[l] = l == 0 ? kin_ : H_;
    }
    EMT_HIPM(mlp_adam_all(w_, m_, v_, g_, seg, (int)total_, lr, kB1, kB2, 1e-8f, clock_, s_));
  }


  // Steps per captured graph in train() (0: none).
  void set_graph_steps(int g) {
    EMT_ENFORCE(g >= 0 && g <= 1024, -1, "graph steps must be in [0, 1024]");
    graph_steps_ = g;
  }

  void predict(const float *X, int n, float *y) {
    EMT_HIPM(hipSetDevice(device_));
    const int chunk = 4096;
    for (int r0 = 0; r0 < n; r0 += chunk) {
      const int b = std::min(chunk, n - r0);
      load_batch(X + (size_t)r0 * in_, nullptr, b);
      forward(b);
      EMT_HIPM(hipMemcpyAsync(y + r0, y_, b * 4, hipMemcpyDeviceToHost, s_));
    }
    EMT_HIPM(hipStreamSynchronize(s_));
  }

  // predict() on device rows, with no allocation: X_dev [n x in_] and
  // y_dev [n] are device memory of this network's device.
  void predict_noalloc(const float *X_dev, int n, float *y_dev) {
    EMT_HIPM(hipSetDevice(device_));
    EMT_ENFORCE(n >= 0, -1, "n must be >= 0; got ", n);
    EMT_ENFORCE(in_ <= kNoAllocMaxCols && kin_ <= kNoAllocMaxCols && H_ <= kNoAllocMaxCols, -1,
                "the allocation-free forward pass takes at most ", kNoAllocMaxCols,
                " input and hidden features; this network has ", in_, " and ", H_);
    if (n == 0) return;
    EMT_ENFORCE(X_dev && y_dev, -1, "null device pointer");
    MlpForward net{};
    for (int l = 0; l < kLayers; ++l) { net.wb[l] = wb_[l]; net.bias[l] = w_ + offs_[2 * l + 1]; }
    net.w4 = w_ + offs_[6];
    net.b4 = w_ + offs_[7];
    net.in = in_; net.kin = kin_; net.H = H_;
    EMT_HIPM(mlp_forward_noalloc(net, device_, X_dev, n, y_dev, s_));
    EMT_HIPM(hipStreamSynchronize(s_));
  }

  // Loss and gradient (flat, unpadded) of one batch, without an update.
  float grads(const float *X, const float *t, int n, float *gout) {
    EMT_HIPM(hipSetDevice(device_));
    load_batch(X, t, n);
    EMT_HIPM(mlp_transpose_bf16(x_, n, kin_, xt_, s_));
    forward(n, true);
    backward(n);
    std::vector<float> h(total_), sq(n);
    EMT_HIPM(hipMemcpyAsync(h.data(), g_, total_ * 4, hipMemcpyDeviceToHost, s_));
    EMT_HIPM(hipMemcpyAsync(sq.data(), sq_, n * 4, hipMemcpyDeviceToHost, s_));
    EMT_HIPM(hipStreamSynchronize(s_));
    unpad(h, gout);
    double s = 0;
    for (float v : sq) s += v;
    return (float)(s / (2.0 * n));
  }

 private:
  template <typename T>
  T *dalloc(size_t n) {
    void *p = nullptr;
    EMT_HIPM(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    allocs_.push_back(p);
    return static_cast<T *>(p);
  }
  template <typename T>
  void release(T *&p) {
    if (!p) return;
    auto it = std::find(allocs_.begin(), allocs_.end(), (void *)p);
    if (it != allocs_.end()) allocs_.erase(it);
    (void)hipFree(p);
    p = nullptr;
  }

  void pad(const float *in, std::vector<float> &h) {
    size_t o = 0;
    for (int r = 0; r < H_; ++r)
      for (int c = 0; c < in_; ++c) h[offs_[0] + (size_t)r * kin_ + c] = in[o++];
    for (int i = 1; i < 8; ++i)
      for (int64_t k = 0; k < sizes_[i]; ++k) h[offs_[i] + k] = in[o++];
  }
  void unpad(const std::vector<float> &h, float *out) {
    size_t o = 0;
    for (int r = 0; r < H_; ++r)
      for (int c = 0; c < in_; ++c) out[o++] = h[offs_[0] + (size_t)r * kin_ + c];
    for (int i = 1; i < 8; ++i)
      for (int64_t k = 0; k < sizes_[i]; ++k) out[o++] = h[offs_[i] + k];
  }

  void upload_params(const std::vector<float> &h) {
    EMT_HIPM(hipMemcpyAsync(w_, h.data(), total_ * 4, hipMemcpyHostToDevice, s_));
    for (int l = 0; l < kLayers; ++l) {
      EMT_HIPM(mlp_to_bf16(w_ + offs_[2 * l], wb_[l], sizes_[2 * l], s_));
      if (wbt_[l]) EMT_HIPM(mlp_transpose_bf16(wb_[l], H_, H_, wbt_[l], s_));
    }
    EMT_HIPM(hipStreamSynchronize(s_));
  }

  void ensure_batch(int B) {
    if (B <= bmax_) return;
    EMT_HIPM(hipStreamSynchronize(s_));
    release(x_);
    for (int l = 0; l < kLayers; ++l) { release(z_[l]); release(h_[l]); release(dz_[l]); }
    release(y_); release(t_); release(dy_); release(sq_); release(xs_); release(ypart_);
    release(cwork_); release(xt_);
    for (int l = 0; l < kLayers; ++l) { release(ht_[l]); release(dzt_[l]); }
    bmax_ = B;
    x_ = dalloc<uint16_t>((size_t)B * kin_);
    // Transposed copies ([features x batch]) of the layer inputs and of the
    // pre-activation gradients: the k-contiguous operands of the
    // weight-gradient products.
    xt_ = dalloc<uint16_t>((size_t)B * kin_);
    for (int l = 0; l < kLayers; ++l) {
      if (l < kLayers - 1) ht_[l] = dalloc<uint16_t>((size_t)B * H_);
      dzt_[l] = dalloc<uint16_t>((size_t)B * H_);
    }
    xs_ = dalloc<float>((size_t)B * kin_);
    for (int l = 0; l < kLayers; ++l) {
      z_[l] = dalloc<uint16_t>((size_t)B * H_);
      h_[l] = dalloc<uint16_t>((size_t)B * H_);
      dz_[l] = dalloc<uint16_t>((size_t)B * H_);
    }
    y_ = dalloc<float>(B); t_ = dalloc<float>(B); dy_ = dalloc<float>(B); sq_ = dalloc<float>(B);
    ypart_ = dalloc<float>((size_t)((H_ + 63) / 64) * B);
    // Bias-gradient column sums per row block: 3 of ceil(B / 64) x H (output
    // layer backward) and 2 of ceil(B / 32) x H (input-gradient epilogues).
    cwork_ = dalloc<float>((size_t)5 * ((B + 31) / 32) * H_);
  }

  // Host rows (in_ columns) -> padded bf16 input x_, targets -> t_.
  void load_batch(const float *X, const float *t, int n) {
    ensure_batch(n);
    std::vector<float> xp((size_t)n * kin_, 0.f);
    for (int r = 0; r < n; ++r) std::memcpy(&xp[(size_t)r * kin_], &X[(size_t)r * in_], in_ * 4);
    EMT_HIPM(hipMemcpyAsync(xs_, xp.data(), xp.size() * 4, hipMemcpyHostToDevice, s_));
    EMT_HIPM(mlp_to_bf16(xs_, x_, (int64_t)xp.size(), s_));
    if (t) EMT_HIPM(hipMemcpyAsync(t_, t, n * 4, hipMemcpyHostToDevice, s_));
    else EMT_HIPM(hipMemsetAsync(t_, 0, n * 4, s_));
    EMT_HIPM(hipStreamSynchronize(s_));
  }

  // keep_t (training): also write the transposed layer outputs the backward
  // pass needs, and leave the output layer to the backward pass when the
  // last hidden layer's product can give its partial sums (ypart_ready_).
  void forward(int B, bool keep_t = false) {
    ypart_ready_ = 0;
    const uint16_t *in = x_;
    int K = kin_;
    for (int l = 0; l < kLayers; ++l) {
      MlpGemm q{};
      q.A = in; q.sam = K; q.sak = 1;
      q.B = wb_[l]; q.sbk = 1; q.sbn = K;
      q.M = B; q.N = H_; q.K = K;
      q.epi = 0; q.bias = w_ + offs_[2 * l + 1]; q.act = 1;
      q.z_out = z_[l]; q.h_out = h_[l]; q.ldc = H_;
      if (keep_t && l < kLayers - 1) { q.ht_out = ht_[l]; q.ldt = B; }
      if (keep_t && l == kLayers - 1) {
        q.w4 = w_ + offs_[6];
        q.ypart = ypart_;
        q.ypart_done = &ypart_ready_;
      }
      EMT_HIPM(mlp_gemm(q, s_));
      in = h_[l];
      K = H_;
    }
    if (!ypart_ready_)
      EMT_HIPM(mlp_gemv_out(h_[2], w_ + offs_[6], w_ + offs_[7], y_, B, H_, s_,
                            keep_t ? t_ : nullptr, dy_, sq_));
  }

  // dl (training): the step's loss slots; the slice sums then also store
  // the loss and advance the clock.
  void backward(int B, float *dl = nullptr) {
    // Column-sum partials of the bias gradients (summed with the split-K
    // slices at the end): the output layer's kernel gives db3, dw4 and db4
    // per 64-row block, the input-gradient products db2 and db1 per 32-row
    // block.
    const int nb64 = (B + 63) / 64, nb32 = (B + 31) / 32;
    const size_t region = (size_t)nb32 * H_;
    float *p_db3 = cwork_, *p_dw4 = cwork_ + region, *p_db4 = cwork_ + 2 * region;
    float *p_db[kLayers - 1] = {cwork_ + 3 * region, cwork_ + 4 * region};
    // Output layer: dz3 = dy w4 gelu'(z3).
    OutLoss ol{};
    if (ypart_ready_) ol = OutLoss{ypart_, (H_ + 63) / 64, w_ + offs_[7], t_, y_, sq_};
    EMT_HIPM(mlp_out_back(dy_, w_ + offs_[6], z_[2], h_[2], dz_[2], dzt_[2], B, H_, p_db3, p_dw4,
                          p_db4, ol, s_));
    SliceJobs slices{};
    int nslices = 0;
    for (int l = kLayers - 1; l >= 0; --l) {
      const uint16_t *in_t = l == 0 ? xt_ : ht_[l - 1];
      const int K = l == 0 ? kin_ : H_;
      // dW_l = dz_l^T in  ([H, K], reduction over the batch), from the
      // transposed copies so both operands are contiguous along the batch;
      // split-K slices are summed for all layers in one launch below.
      MlpGemm q{};
      q.A = dzt_[l]; q.sam = B; q.sak = 1;
      q.B = in_t; q.sbk = 1; q.sbn = B;
      q.M = H_; q.N = K; q.K = B;
      q.epi = 2; q.f_out = g_ + offs_[2 * l]; q.ldc = K;
      q.work = kwork_ + l * kwork_floats_; q.work_floats = kwork_floats_;
      int split = 1;
      q.split_out = &split;
      q.defer_sum = 1;
      EMT_HIPM(mlp_gemm(q, s_));
      if (split > 1)
        slices.job[nslices++] = SliceJob{q.work, (int64_t)H_ * K, split, q.f_out, (int64_t)H_ * K};
      if (l > 0) {
        // dz_{l-1} = (dz_l W_l) * gelu'(z_{l-1}), W_l from its transposed
        // copy; the epilogue also sums dz_{l-1}'s columns (db_{l-1}).
        MlpGemm r{};
        r.A = dz_[l]; r.sam = H_; r.sak = 1;
        r.B = wbt_[l]; r.sbk = 1; r.sbn = H_;
        r.M = B; r.N = H_; r.K = H_;
        r.epi = 1; r.z_in = z_[l - 1]; r.h_out = dz_[l - 1]; r.ldc = H_;
        r.ht_out = dzt_[l - 1]; r.ldt = B;
        r.colpart = p_db[l - 1];
        EMT_HIPM(mlp_gemm(r, s_));
      }
    }
    slices.job[nslices++] = SliceJob{p_db3, H_, nb64, g_ + offs_[5], H_};
    slices.job[nslices++] = SliceJob{p_dw4, H_, nb64, g_ + offs_[6], H_};
    slices.job[nslices++] = SliceJob{p_db4, 1, nb64, g_ + offs_[7], 1};
    for (int l = 0; l < kLayers - 1; ++l)
      slices.job[nslices++] = SliceJob{p_db[l], H_, nb32, g_ + offs_[2 * l + 1], H_};
    if (dl) slices.loss = LossTick{sq_, B, dl, clock_, kB1, kB2};
    EMT_HIPM(mlp_sum_slices_multi(slices, nslices, s_));
  }

  static constexpr float kB1 = 0.9f, kB2 = 0.999f;
  int in_, kin_, H_, device_;
  hipStream_t s_ = nullptr;
  std::vector<int64_t> sizes_, offs_;
  int64_t total_ = 0;
  float *w_ = nullptr, *m_ = nullptr, *v_ = nullptr, *g_ = nullptr;
  uint16_t *wb_[kLayers] = {}, *wbt_[kLayers] = {};
  int bmax_ = 0;
  uint16_t *x_ = nullptr, *z_[kLayers] = {}, *h_[kLayers] = {}, *dz_[kLayers] = {};
  uint16_t *xt_ = nullptr, *ht_[kLayers] = {}, *dzt_[kLayers] = {};
  float *y_ = nullptr, *t_ = nullptr, *dy_ = nullptr, *sq_ = nullptr, *xs_ = nullptr;
  float *ypart_ = nullptr;  // output-layer partial sums, (H / 64) x B
  int ypart_ready_ = 0;
  float *cwork_ = nullptr, *kwork_ = nullptr;
  int64_t kwork_floats_ = 0;
  uint16_t *data_x_ = nullptr;
  float *data_t_ = nullptr;
  int n_data_ = 0;
  int step_ = 0;          // Adam steps taken
  int graph_steps_ = 0;
  TrainClock *clock_ = nullptr;
  std::vector<void *> allocs_;
};

}  // namespace emt6ro

struct emt6ro_mlp { std::unique_ptr<emt6ro::Mlp> m; };

namespace {
thread_local std::string g_mlp_error;
template <typename F>
int mlp_guarded(F &&f) {
  try {
    f();
    return EMT6RO_OK;
  } catch (const emt6ro::Error &e) {
    g_mlp_error = e.what();
    return e.code();
  } catch (const std::exception &e) {
    g_mlp_error = e.what();
    return EMT6RO_ERR_ARG;
  }
}
}  // namespace

extern "C" {

const char *emt6ro_mlp_last_error(void) { return g_mlp_error.c_str(); }

int emt6ro_mlp_create(int32_t in_features, int32_t hidden, uint64_t seed, int32_t device,
                      emt6ro_mlp **out) {
  return mlp_guarded([&] {
    if (!out) throw emt6ro::Error("Check failed: out; null pointer", EMT6RO_ERR_ARG);
    auto m = std::make_unique<emt6ro_mlp>();
    m->m = std::make_unique<emt6ro::Mlp>(in_features, hidden, seed, device);
    *out = m.release();
  });
}

int64_t emt6ro_mlp_num_params(const emt6ro_mlp *m) { return m ? m->m->num_params() : -1; }

int emt6ro_mlp_get_params(emt6ro_mlp *m, float *out) {
  return mlp_guarded([&] { m->m->get_params(out); });
}

int emt6ro_mlp_set_params(emt6ro_mlp *m, const float *in) {
  return mlp_guarded([&] { m->m->set_params(in); });
}

int emt6ro_mlp_set_data(emt6ro_mlp *m, const float *X, const float *y, int32_t n) {
  return mlp_guarded([&] { m->m->set_data(X, y, n); });
}

int emt6ro_mlp_train(emt6ro_mlp *m, int32_t steps, int32_t batch, float lr, uint64_t seed,
                     float *losses) {
  return mlp_guarded([&] {
    std::vector<float> l = m->m->train(steps, batch, lr, seed);
    if (losses) std::memcpy(losses, l.data(), l.size() * 4);
  });
}

int emt6ro_mlp_train_many(emt6ro_mlp *const *nets, int32_t count, int32_t steps, int32_t batch,
                          float lr, const uint64_t *seeds, float *losses) {
  return mlp_guarded([&] {
    if (!nets || !seeds || count < 1)
      throw emt6ro::Error("Check failed: nets, seeds and count >= 1", EMT6RO_ERR_ARG);
    for (int i = 0; i < count; ++i) {
      if (!nets[i]) throw emt6ro::Error("Check failed: null network", EMT6RO_ERR_ARG);
      for (int j = 0; j < i; ++j)
        if (nets[j] == nets[i]) throw emt6ro::Error("a network is listed twice", EMT6RO_ERR_ARG);
      if (nets[i]->m->device() != nets[0]->m->device())
        throw emt6ro::Error("the networks are on different devices", EMT6RO_ERR_ARG);
    }
    std::vector<float *> dl(count);
    for (int i = 0; i < count; ++i) dl[i] = nets[i]->m->train_begin(steps, batch);
    if (steps == 0) return;
    // Step k of every network before step k + 1 of any: each network's
    // kernels stay in order on its own stream, and the streams overlap.  With
    // graphs (the first network's setting), g steps of each network are
    // captured once and the graphs are launched in turn.
    int done = 0;
    const int g = nets[0]->m->graph_steps();
    if (g > 0 && steps >= g) {
      std::vector<hipGraphExec_t> ex;
      hipError_t e = hipSuccess;
      try {
        for (int i = 0; i < count; ++i)
          ex.push_back(nets[i]->m->capture_steps(g, batch, lr, seeds[i], dl[i]));
      } catch (...) {
        for (hipGraphExec_t x : ex) (void)hipGraphExecDestroy(x);
        throw;
      }
      for (; e == hipSuccess && done + g <= steps; done += g)
        for (int i = 0; i < count && e == hipSuccess; ++i)
          e = hipGraphLaunch(ex[i], nets[i]->m->stream());
      for (int i = 0; i < count; ++i) {
        const hipError_t e2 = hipStreamSynchronize(nets[i]->m->stream());
        if (e == hipSuccess) e = e2;
      }
      for (hipGraphExec_t x : ex) (void)hipGraphExecDestroy(x);
      if (e != hipSuccess)
        throw emt6ro::Error(emt6ro::cat("HIP error ", hipGetErrorString(e), " in train_many"), -3);
    }
    for (; done < steps; ++done)
      for (int i = 0; i < count; ++i) nets[i]->m->enqueue_step(batch, lr, seeds[i], dl[i]);
    for (int i = 0; i < count; ++i) {
      std::vector<float> l = nets[i]->m->train_end(dl[i], steps);
      if (losses) std::memcpy(losses + (size_t)i * steps, l.data(), l.size() * 4);
    }
  });
}

int emt6ro_mlp_set_train_graph(emt6ro_mlp *m, int32_t steps_per_graph) {
  return mlp_guarded([&] {
    if (!m) throw emt6ro::Error("Check failed: m; null network", EMT6RO_ERR_ARG);
    m->m->set_graph_steps(steps_per_graph);
  });
}

int emt6ro_mlp_predict(emt6ro_mlp *m, const float *X, int32_t n, float *y) {
  return mlp_guarded([&] { m->m->predict(X, n, y); });
}

int emt6ro_mlp_predict_noalloc(emt6ro_mlp *m, const float *X_dev, int32_t n, float *y_dev) {
  return mlp_guarded([&] {
    if (!m) throw emt6ro::Error("Check failed: m; null network", EMT6RO_ERR_ARG);
    m->m->predict_noalloc(X_dev, n, y_dev);
  });
}

int emt6ro_mlp_grads(emt6ro_mlp *m, const float *X, const float *y, int32_t n, float *grads,
                     float *loss) {
  return mlp_guarded([&] {
    const float l = m->m->grads(X, y, n, grads);
    if (loss) *loss = l;
  });
}

void emt6ro_mlp_destroy(emt6ro_mlp *m) { delete m; }

}  // extern "C"
