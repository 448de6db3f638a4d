// Batch engine: builds the canonical device state of a batch of simulations
// from tumour states, and runs it with the persistent kernel, the per-step
// kernels, or the per-step kernels replayed from a captured HIP graph.
#include "engine.h"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>

namespace emt6ro {

struct ImgDesc {
  int64_t sub_off, st_off, slot_off;
  int32_t words;  // slot words of buffer 0 (kSlotFields * cap)
  int32_t pad;
};

namespace {

// Copies each simulation's tumour image into its own state and header.
__global__ void replicate_kernel(const SimDesc *desc, const int32_t *tumor_of,
                                 const ImgDesc *img, const SimHdr *img_hdr,
                                 const float *img_sub, const uint8_t *img_st,
                                 const uint32_t *img_slots, float *sub, uint8_t *st,
                                 uint32_t *slots, SimHdr *hdr, int nsims) {
  const int sim = blockIdx.x;
  if (sim >= nsims) return;
  const SimDesc d = desc[sim];
  const int t = tumor_of[sim];
  const ImgDesc im = img[t];
  const int hw = d.H * d.W;
  for (int i = threadIdx.x; i < 3 * hw; i += blockDim.x) sub[d.sub_off + i] = img_sub[im.sub_off + i];
  const int stw = (int)(st_bytes(d.H, d.W) / 4);
  const uint32_t *s32 = reinterpret_cast<const uint32_t *>(img_st + im.st_off);
  uint32_t *d32 = reinterpret_cast<uint32_t *>(st + d.st_off);
  for (int i = threadIdx.x; i < stw; i += blockDim.x) d32[i] = s32[i];
  for (int i = threadIdx.x; i < im.words; i += blockDim.x)
    slots[d.slot_off + i] = img_slots[im.slot_off + i];
  if (threadIdx.x == 0) hdr[sim] = img_hdr[t];
}

template <typename T>
T *dalloc(size_t n) {
  void *p = nullptr;
  if (n == 0) n = 1;
  EMT_HIP(hipMalloc(&p, n * sizeof(T)));
  return static_cast<T *>(p);
}

template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace

Batch::Batch(const HostParams &p, const std::vector<const TumorState *> &tumors,
             const BatchSpec &spec, int device, uint64_t seed)
    : params_(p), dparams_(p.device()), device_(device),
      seed_((uint32_t)(seed ^ (seed >> 32))), nsims_((int)spec.tumor_index.size()),
      spec_(spec) {
  EMT_ENFORCE(nsims_ > 0, -1, "A batch needs at least one simulation");
  EMT_ENFORCE(!tumors.empty(), -1, "A batch needs at least one tumour");
  EMT_ENFORCE(spec.prot_index.size() == (size_t)nsims_ && spec.horizon.size() == (size_t)nsims_ &&
                  spec.sim_id.size() == (size_t)nsims_, -1, "Inconsistent batch arrays");
  EMT_ENFORCE(spec.prot_res > 0, -1, "Protocol resolution must be positive");
  EMT_ENFORCE(spec.n_prot > 0 && spec.n_slots > 0, -1, "Protocol table must be non-empty");
  // Per-tumour images (host only: every argument is checked before the
  // first device call).
  const int T = (int)tumors.size();
  std::vector<ImgDesc> img(T);
  int64_t isub = 0, ist = 0, islot = 0;
  for (int t = 0; t < T; ++t) {
    const TumorState &s = *tumors[t];
    EMT_ENFORCE(s.H >= 3 && s.W >= 3 && s.H <= 255 && s.W <= 255, -1, "Bad tumour dims");
    img[t] = ImgDesc{isub, ist, islot, kSlotFields * slot_cap(s.H, s.W), 0};
    isub += 3 * (int64_t)s.sites();
    ist += st_bytes(s.H, s.W);
    islot += (int64_t)kSlotFields * slot_cap(s.H, s.W);
  }
  std::vector<float> h_sub(isub);
  std::vector<uint8_t> h_st(ist, 0);
  std::vector<uint32_t> h_slots(islot, 0);
  img_hdr_.assign(T, SimHdr{});
  for (int t = 0; t < T; ++t) {
    const TumorState &s = *tumors[t];
    const size_t hw = s.sites();
    for (int k = 0; k < 3; ++k)
      std::memcpy(&h_sub[img[t].sub_off + k * hw], s.f[k].data(), hw * sizeof(float));
    std::memcpy(&h_st[img[t].st_off], s.occ.data(), hw);
    const int cap = slot_cap(s.H, s.W);
    const std::vector<int32_t> list = occupied_list_order(s);
    EMT_ENFORCE((int)list.size() <= cap, -1, "Too many occupied sites");
    uint32_t *S = &h_slots[img[t].slot_off];
    for (size_t i = 0; i < list.size(); ++i) {
      const int32_t x = list[i];
      auto put = [&](int field, float v) { std::memcpy(&S[(size_t)field * cap + i], &v, 4); };
      put(kTRep, s.f[TumorState::kTRep][x]);
      put(kR, s.f[TumorState::kIrr][x]);
      put(kDelay, s.f[TumorState::kDelay][x]);
      put(kProlif, s.f[TumorState::kProlif][x]);
      for (int k = 0; k < 5; ++k) put(kT0 + k, s.f[TumorState::kG1 + k][x]);
      S[(size_t)kPacked * cap + i] = pack_slot((uint32_t)x, s.mode[x], s.phase[x], true);
    }
    SimHdr h{};
    h.n = (int32_t)list.size();
    img_hdr_[t] = h;
  }
  // Per-simulation layout.
  desc_h_.resize(nsims_);
  std::vector<int32_t> tumor_of(nsims_);
  for (int i = 0; i < nsims_; ++i) {
    const int t = spec.tumor_index[i];
    EMT_ENFORCE(t >= 0 && t < T, -1, "Tumour index out of range: ", t);
    EMT_ENFORCE(spec.prot_index[i] >= 0 && spec.prot_index[i] < spec.n_prot, -1,
                "Protocol index out of range: ", spec.prot_index[i]);
    const TumorState &s = *tumors[t];
    SimDesc d{};
    d.sub_off = sub_total_;
    d.st_off = st_total_;
    d.slot_off = slot_total_;
    d.H = s.H;
    d.W = s.W;
    d.cap = slot_cap(s.H, s.W);
    d.prot = spec.prot_index[i];
    d.sim_id = spec.sim_id[i];
    d.horizon = spec.horizon[i];
    desc_h_[i] = d;
    tumor_of[i] = t;
    sub_total_ += 3 * (int64_t)s.sites();
    st_total_ += st_bytes(s.H, s.W);
    slot_total_ += 2 * (int64_t)kSlotFields * d.cap;
    hmax_ = std::max(hmax_, s.H);
    wmax_ = std::max(wmax_, s.W);
    capmax_ = std::max(capmax_, d.cap);
  }
  // Size classes by LDS row stride (36 / 56 / 88), longest work first within
  // each: lattice area times horizon.
  auto class_of = [](int H, int W) { const int m = std::max(H, W) + 2; return m <= 36 ? 0 : m <= 56 ? 1 : 2; };
  const int rs_of[3] = {36, 56, 88};
  order_h_.resize(nsims_);
  std::iota(order_h_.begin(), order_h_.end(), 0);
  std::stable_sort(order_h_.begin(), order_h_.end(), [&](int a, int b) {
    const int ka = class_of(desc_h_[a].H, desc_h_[a].W), kb = class_of(desc_h_[b].H, desc_h_[b].W);
    if (ka != kb) return ka < kb;
    const double ca = (double)desc_h_[a].H * desc_h_[a].W * desc_h_[a].horizon;
    const double cb = (double)desc_h_[b].H * desc_h_[b].W * desc_h_[b].horizon;
    return ca > cb;
  });
  for (int i = 0; i < nsims_; ++i) {
    const SimDesc &d = desc_h_[order_h_[i]];
    const int k = class_of(d.H, d.W);
    EMT_ENFORCE(std::max(d.H, d.W) + 2 <= 88, -1, "Lattices are limited to 86 x 86");
    if (classes_.empty() || classes_.back().rs != rs_of[k])
      classes_.push_back(Class{rs_of[k], i, 0, 0, 0});
    Class &c = classes_.back();
    c.count += 1;
    c.hmax = std::max(c.hmax, d.H);
    c.wmax = std::max(c.wmax, d.W);
  }
  EMT_HIP(hipSetDevice(device_));
  EMT_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  EMT_HIP(hipEventCreate(&ev0_));
  EMT_HIP(hipEventCreate(&ev1_));
  desc_ = dalloc<SimDesc>(nsims_);
  hdr_ = dalloc<SimHdr>(nsims_);
  sub_ = dalloc<float>(sub_total_);
  st_ = dalloc<uint8_t>(st_total_);
  slots_ = dalloc<uint32_t>(slot_total_);
  prot_ = dalloc<float>((size_t)spec.n_prot * spec.n_slots);
  order_ = dalloc<int32_t>(nsims_);
  order_next_ = dalloc<int32_t>(nsims_);
  order_win_ = dalloc<int32_t>(nsims_);
  wcount_ = dalloc<int32_t>(classes_.size() * kWindowKinds);
  win_frac_.assign(classes_.size() * kWindowKinds, 1.0);
  target_ = dalloc<uint32_t>(nsims_);
  queue_ = dalloc<int32_t>(1);
  results_ = dalloc<uint32_t>(nsims_);
  EMT_HIP(hipMemcpyAsync(desc_, desc_h_.data(), nsims_ * sizeof(SimDesc), hipMemcpyHostToDevice, stream_));
  EMT_HIP(hipMemcpyAsync(order_, order_h_.data(), nsims_ * sizeof(int32_t), hipMemcpyHostToDevice, stream_));
  EMT_HIP(hipMemsetAsync(st_, 0, st_total_, stream_));
  EMT_HIP(hipMemsetAsync(prot_, 0, (size_t)spec.n_prot * spec.n_slots * sizeof(float), stream_));
  // Images on the device (kept for reset).
  float *d_isub = dalloc<float>(isub);
  uint8_t *d_ist = dalloc<uint8_t>(ist);
  uint32_t *d_islot = dalloc<uint32_t>(islot);
  ImgDesc *d_img = dalloc<ImgDesc>(T);
  SimHdr *d_ihdr = dalloc<SimHdr>(T);
  int32_t *d_tumor_of = dalloc<int32_t>(nsims_);
  EMT_HIP(hipMemcpyAsync(d_isub, h_sub.data(), isub * sizeof(float), hipMemcpyHostToDevice, stream_));
  EMT_HIP(hipMemcpyAsync(d_ist, h_st.data(), ist, hipMemcpyHostToDevice, stream_));
  EMT_HIP(hipMemcpyAsync(d_islot, h_slots.data(), islot * sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
  EMT_HIP(hipMemcpyAsync(d_img, img.data(), T * sizeof(ImgDesc), hipMemcpyHostToDevice, stream_));
  EMT_HIP(hipMemcpyAsync(d_ihdr, img_hdr_.data(), T * sizeof(SimHdr), hipMemcpyHostToDevice, stream_));
  EMT_HIP(hipMemcpyAsync(d_tumor_of, tumor_of.data(), nsims_ * sizeof(int32_t), hipMemcpyHostToDevice, stream_));
  img_sub_ = {d_isub};
  img_st_ = {d_ist};
  img_slots_ = {d_islot};
  // Keep the small tables with the images.
  img_desc_ = d_img;
  img_hdr_d_ = d_ihdr;
  tumor_of_ = d_tumor_of;
  sync();
  reset(0);
}

Batch::~Batch() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  dfree(desc_); dfree(hdr_); dfree(sub_); dfree(st_); dfree(slots_); dfree(prot_);
  dfree(order_); dfree(order_next_); dfree(order_win_); dfree(wcount_); dfree(target_); dfree(queue_); dfree(results_); dfree(scratch_); dfree(ovf_);
  for (auto &p : img_sub_) dfree(p);
  for (auto &p : img_st_) dfree(p);
  for (auto &p : img_slots_) dfree(p);
  dfree(img_desc_); dfree(img_hdr_d_); dfree(tumor_of_);
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Batch::sync() { EMT_HIP(hipStreamSynchronize(stream_)); }

void Batch::set_protocols(const float *doses) {
  EMT_HIP(hipSetDevice(device_));
  EMT_HIP(hipMemcpyAsync(prot_, doses, (size_t)spec_.n_prot * spec_.n_slots * sizeof(float),
                         hipMemcpyHostToDevice, stream_));
  sync();
}

void Batch::reset(uint32_t epoch) {
  EMT_HIP(hipSetDevice(device_));
  epoch_ = epoch;
  std::fill(win_frac_.begin(), win_frac_.end(), 1.0);
  hipLaunchKernelGGL(replicate_kernel, dim3(nsims_), dim3(256), 0, stream_, desc_, tumor_of_,
                     img_desc_, img_hdr_d_, img_sub_[0], img_st_[0], img_slots_[0], sub_, st_,
                     slots_, hdr_, nsims_);
  EMT_HIP(hipGetLastError());
  sync();
}

void Batch::ensure_scratch(int wgs) {
  if (wgs <= scratch_wgs_) return;
  dfree(scratch_);
  dfree(ovf_);
  const int64_t sstride = 32 * (int64_t)hmax_;
  const int64_t ostride = capmax_;
  scratch_ = dalloc<int32_t>((size_t)wgs * sstride);
  ovf_ = dalloc<uint64_t>((size_t)wgs * ostride);
  scratch_wgs_ = wgs;
}

// ---------------------------------------------------------------------------
// Instrumentation records.
// ---------------------------------------------------------------------------
namespace {
std::mutex g_instr_mu;
std::map<int, unsigned long long *> g_instr_rec;
std::atomic<int> g_instr_on{-1};  // -1: not read from the environment yet
}  // namespace

bool instr_enabled() {
  int v = g_instr_on.load();
  if (v < 0) {
    const char *e = std::getenv("EMT6RO_INSTRUMENT");
    v = (e && *e && std::strcmp(e, "0") != 0) ? 1 : 0;
    int expect = -1;
    g_instr_on.compare_exchange_strong(expect, v);
    v = g_instr_on.load();
  }
  return v == 1;
}

void instr_set_enabled(bool on) { g_instr_on.store(on ? 1 : 0); }

unsigned long long *instr_device_record(int device) {
  if (!instr_enabled()) return nullptr;
  std::lock_guard<std::mutex> lk(g_instr_mu);
  auto it = g_instr_rec.find(device);
  if (it != g_instr_rec.end()) return it->second;
  void *p = nullptr;
  EMT_HIP(hipSetDevice(device));
  EMT_HIP(hipMalloc(&p, sizeof(unsigned long long) * dev::kInstrN));
  EMT_HIP(hipMemset(p, 0, sizeof(unsigned long long) * dev::kInstrN));
  EMT_HIP(hipDeviceSynchronize());
  g_instr_rec[device] = static_cast<unsigned long long *>(p);
  return g_instr_rec[device];
}

void instr_collect(unsigned long long *out, bool reset) {
  std::lock_guard<std::mutex> lk(g_instr_mu);
  for (int i = 0; i < dev::kInstrN; ++i) out[i] = 0;
  std::vector<unsigned long long> h(dev::kInstrN);
  for (auto &kv : g_instr_rec) {
    EMT_HIP(hipSetDevice(kv.first));
    EMT_HIP(hipDeviceSynchronize());
    EMT_HIP(hipMemcpy(h.data(), kv.second, sizeof(unsigned long long) * dev::kInstrN,
                      hipMemcpyDeviceToHost));
    for (int i = 0; i < dev::kInstrN; ++i) out[i] += h[i];
    if (reset) {
      EMT_HIP(hipMemset(kv.second, 0, sizeof(unsigned long long) * dev::kInstrN));
      EMT_HIP(hipDeviceSynchronize());
    }
  }
}

dev::RunArgs Batch::args(uint32_t nsteps) const {
  dev::RunArgs a{};
  a.p = dparams_;
  a.desc = desc_;
  a.hdr = hdr_;
  a.sub_pool = sub_;
  a.st_pool = st_;
  a.slot_pool = slots_;
  a.prot_data = prot_;
  a.n_slots = spec_.n_slots;
  a.prot_res = spec_.prot_res;
  a.prot_length = spec_.prot_length;
  a.seed = seed_;
  a.epoch = epoch_;
  a.nsteps = nsteps;
  a.order = order_;
  a.nsims = nsims_;
  a.queue = queue_;
  a.scratch = scratch_;
  a.scratch_stride = 32 * (int64_t)hmax_;
  a.ovf = ovf_;
  a.ovf_stride = capmax_;
  a.pv_max = 0;
  a.instr = instr_device_record(device_);
  return a;
}

int Batch::choose_variant(int hmax, int wmax, int kind) {
  int cnt = 0;
  const VariantInfo *v = variant_table(&cnt);
  int best = -1;
  for (int i = 0; i < cnt; ++i) {
    if (!variant_fits(v[i], hmax, wmax) || (kind >= 0 && v[i].kind != kind)) continue;
    if (kind < 0 && persistent_occupancy(i, persistent_lds_bytes(v[i], hmax, wmax)) == 0)
      continue;
    if (best < 0 || v[i].rs < v[best].rs ||
        (v[i].rs == v[best].rs && v[i].kind > v[best].kind))
      best = i;
  }
  if (kind >= 0 || best < 0) return best;
  // Default without tuning: the variants that tuned fastest on MI355X for
  // lattices of each row stride (36: 33x33, 56: 53x53, 88: 81x81), in
  // order of preference, at the smallest row stride that fits.
  static const int kPreferred[] = {49, 36, 35, 39, 32, 43, 42, 41, 34, 25, 24, 14, 17, 19, 11};
  for (int id : kPreferred) {
    if (id >= cnt || v[id].rs != v[best].rs || !variant_fits(v[id], hmax, wmax)) continue;
    if (persistent_occupancy(id, persistent_lds_bytes(v[id], hmax, wmax)) == 0) continue;
    return id;
  }
  return best;
}

void Batch::set_window(int rows) {
  EMT_ENFORCE(rows == -1 || rows == 0 || rows >= 8, -1,
              "Window rows must be -1 (automatic), 0 (off) or at least 8; got ", rows);
  window_rows_ = rows;
}

void Batch::set_preferred_variant(int rs, int variant) {
  const int k = rs <= 36 ? 0 : rs <= 56 ? 1 : 2;
  preferred_[k] = variant;
}

void Batch::run(uint32_t nsteps, int path, int variant, uint32_t steps_per_launch) {
  EMT_HIP(hipSetDevice(device_));
  int cnt = 0;
  const VariantInfo *vt = variant_table(&cnt);
  EMT_ENFORCE(variant < cnt, -1, "Kernel variant ", variant, " does not exist");
  int64_t launches = 0;
  if (path == kPathPersistent) {
    int dev_cus = 0;
    EMT_HIP(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, device_));
    struct Plan { int v, grid; size_t lds; };
    std::vector<Plan> plans;
    int max_grid = 0;
    for (const Class &c : classes_) {
      const int k = c.rs <= 36 ? 0 : c.rs <= 56 ? 1 : 2;
      int v = variant >= 0 ? variant : preferred_[k];
      if (v < 0 || !variant_fits(vt[v], c.hmax, c.wmax)) v = choose_variant(c.hmax, c.wmax);
      EMT_ENFORCE(v >= 0 && variant_fits(vt[v], c.hmax, c.wmax), -1, "Kernel variant ", variant,
                  " does not fit lattices of ", c.hmax, " x ", c.wmax);
      const size_t lds = persistent_lds_bytes(vt[v], c.hmax, c.wmax);
      const int occ = persistent_occupancy(v, lds);
      EMT_ENFORCE(occ > 0, -1, "Persistent kernel does not fit: ", lds, " bytes of LDS");
      const int grid = std::min(c.count, occ * dev_cus);
      plans.push_back(Plan{v, grid, lds});
      max_grid = std::max(max_grid, grid);
    }
    // Window pass: for classes wider than row stride 36 (and the default
    // variant choice), simulations whose ROI fits a window first run in the
    // window kernel; the class's own launch then takes every simulation to
    // the same per-simulation target.
    // Window passes: for classes wider than a window kernel's row stride
    // (36: 128 threads, eight workgroups per CU; 56: 256 threads, four per
    // CU, only for classes whose whole-lattice kernel needs the 88 stride)
    // and the default variant choice, the simulations whose ROI fits run
    // first in the window kernels, narrowest first; the class's own launch
    // then takes every simulation to the same per-simulation target.
    struct WinPlan { int rs, rows, grid; size_t lds; };
    static const int kWinRs[kWindowKinds] = {36, 56};
    static const int kWinPerCu[kWindowKinds] = {8, 4};
    std::vector<WinPlan> wins(classes_.size() * kWindowKinds, WinPlan{0, 0, 0, 0});
    if (variant < 0 && window_rows_ != 0) {
      for (size_t ci = 0; ci < classes_.size(); ++ci) {
        const Class &c = classes_[ci];
        for (int wk = 0; wk < kWindowKinds; ++wk) {
          const int rs = kWinRs[wk];
          if (c.rs <= rs || (wk > 0 && window_rows_ > 0)) continue;
          int rows = window_rows_;
          if (rows < 0) {
            const size_t fixed = window_lds_bytes(rs, 0, c.hmax, c.wmax);
            const size_t budget = (size_t)163840 / kWinPerCu[wk];
            rows = budget > fixed ? (int)((budget - fixed) / (3 * rs * sizeof(float))) : 0;
          }
          rows = std::min(rows, c.hmax + 2);
          if (rows < 8) continue;
          const size_t lds = window_lds_bytes(rs, rows, c.hmax, c.wmax);
          const int occ = window_occupancy(rs, lds);
          if (occ <= 0) continue;
          WinPlan &w = wins[ci * kWindowKinds + wk];
          w = WinPlan{rs, rows, std::min(c.count, occ * dev_cus), lds};
          max_grid = std::max(max_grid, w.grid);
        }
      }
    }
    ensure_scratch(max_grid);
    EMT_HIP(hipEventRecord(ev0_, stream_));
    uint32_t done = 0;
    const uint32_t chunk = steps_per_launch ? steps_per_launch : nsteps;
    while (done < nsteps) {
      const uint32_t k = std::min(chunk, nsteps - done);
      for (size_t ci = 0; ci < classes_.size(); ++ci) {
        const Class &c = classes_[ci];
        dev::RunArgs a = args(k);
        bool any = false;
        for (int wk = 0; wk < kWindowKinds; ++wk) any = any || wins[ci * kWindowKinds + wk].rows > 0;
        if (any) {
          EMT_HIP(launch_set_target(order_ + c.offset, c.count, hdr_, desc_, target_, k, stream_));
          for (int wk = 0; wk < kWindowKinds; ++wk) {
            const size_t wi = ci * kWindowKinds + wk;
            const WinPlan &w = wins[wi];
            if (w.rows == 0) continue;
            EMT_HIP(launch_order_window(order_ + c.offset, order_win_ + c.offset, c.count, hdr_,
                                        desc_, target_, w.rows, w.rs, wcount_ + wi, stream_));
            launches += 1;
            if (win_frac_[wi] < kWindowMin) continue;
            dev::RunArgs aw = args(k);
            aw.order = order_win_ + c.offset;
            aw.nsims = c.count;
            aw.nsims_dev = wcount_ + wi;
            aw.target = target_;
            aw.pv_max = w.rows * w.rs;
            EMT_HIP(launch_fill_i32(queue_, 0, 1, stream_));
            EMT_HIP(launch_window(w.rs, aw, w.grid, w.lds, stream_));
            launches += 2;
          }
          a.target = target_;
          launches += 1;
        }
        // Longest first by the ROI each simulation reached.
        EMT_HIP(launch_order_by_cost(order_ + c.offset, order_next_ + c.offset, c.count, hdr_,
                                     desc_, stream_));
        a.order = order_next_ + c.offset;
        a.nsims = c.count;
        a.pv_max = vt[plans[ci].v].rs * (c.hmax + 2);
        EMT_HIP(launch_fill_i32(queue_, 0, 1, stream_));
        EMT_HIP(launch_persistent_lds(plans[ci].v, a, plans[ci].grid, plans[ci].lds, stream_));
        launches += 3;
      }
      done += k;
    }
    EMT_HIP(hipEventRecord(ev1_, stream_));
    // Eligible fraction of each windowed class, for the next run's choice.
    bool any_win = false;
    for (const WinPlan &w : wins) any_win = any_win || w.rows > 0;
    if (any_win) {
      std::vector<int32_t> wc(wins.size(), 0);
      EMT_HIP(hipMemcpyAsync(wc.data(), wcount_, wc.size() * sizeof(int32_t),
                             hipMemcpyDeviceToHost, stream_));
      sync();
      for (size_t wi = 0; wi < wins.size(); ++wi)
        if (wins[wi].rows > 0 && classes_[wi / kWindowKinds].count > 0)
          win_frac_[wi] = (double)wc[wi] / classes_[wi / kWindowKinds].count;
    }
  } else if (path == kPathStep || path == kPathStepGraph) {
    // The per-step kernels use the banded scheme.
    if (variant < 0 || vt[variant].kind != 0) variant = choose_variant(hmax_, wmax_, 0);
    EMT_ENFORCE(variant >= 0 && variant_fits(vt[variant], hmax_, wmax_), -1, "Kernel variant ",
                variant, " does not fit lattices of ", hmax_, " x ", wmax_);
    ensure_scratch(nsims_);
    dev::RunArgs a = args(1);
    a.pv_max = vt[variant].rs * (hmax_ + 2);
    EMT_HIP(hipEventRecord(ev0_, stream_));
    uint32_t done = 0;
    if (path == kPathStepGraph) {
      const int K = 128;
      if (!graph_exec_ || graph_variant_ != variant || graph_steps_ != K ||
          graph_epoch_ != a.epoch || graph_instr_ != a.instr) {
        if (graph_exec_) EMT_HIP(hipGraphExecDestroy(graph_exec_));
        graph_exec_ = nullptr;
        hipGraph_t g = nullptr;
        EMT_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        for (int s = 0; s < K; ++s) EMT_HIP(launch_step(variant, a, nsims_, hmax_, wmax_, stream_));
        EMT_HIP(hipStreamEndCapture(stream_, &g));
        EMT_HIP(hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0));
        EMT_HIP(hipGraphDestroy(g));
        graph_variant_ = variant;
        graph_steps_ = K;
        graph_epoch_ = a.epoch;
        graph_instr_ = a.instr;
      }
      while (nsteps - done >= (uint32_t)K) {
        EMT_HIP(hipGraphLaunch(graph_exec_, stream_));
        launches += 3 * K;
        done += K;
      }
    }
    while (done < nsteps) {
      EMT_HIP(launch_step(variant, a, nsims_, hmax_, wmax_, stream_));
      launches += 3;
      ++done;
    }
    EMT_HIP(hipEventRecord(ev1_, stream_));
  } else {
    throw Error(cat("Check failed: valid path; unknown run path ", path), -1);
  }
  sync();
  float ms = 0.f;
  EMT_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
  last_ms_ = ms;
  last_launches_ = launches;
}

void Batch::results(uint32_t *dst) {
  EMT_HIP(hipSetDevice(device_));
  EMT_HIP(launch_count_living(desc_, st_, results_, nsims_, stream_));
  EMT_HIP(hipMemcpyAsync(dst, results_, nsims_ * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
  sync();
}

void Batch::counters(uint64_t *dst) {
  EMT_HIP(hipSetDevice(device_));
  std::vector<SimHdr> h(nsims_);
  EMT_HIP(hipMemcpyAsync(h.data(), hdr_, nsims_ * sizeof(SimHdr), hipMemcpyDeviceToHost, stream_));
  sync();
  for (int i = 0; i < nsims_; ++i) {
    dst[5 * i + 0] = h[i].step;
    dst[5 * i + 1] = (uint64_t)h[i].n;
    dst[5 * i + 2] = h[i].divisions;
    dst[5 * i + 3] = h[i].status;
    dst[5 * i + 4] = h[i].site_substeps;
  }
}

TumorState Batch::state(int sim) {
  EMT_ENFORCE(sim >= 0 && sim < nsims_, -1, "Cannot read data from sample ", sim,
              ". Batch size: ", nsims_);
  EMT_HIP(hipSetDevice(device_));
  const SimDesc &d = desc_h_[sim];
  TumorState s(d.H, d.W);
  const size_t hw = s.sites();
  std::vector<float> sub(3 * hw);
  std::vector<uint32_t> slots((size_t)2 * kSlotFields * d.cap);
  SimHdr h{};
  EMT_HIP(hipMemcpyAsync(sub.data(), sub_ + d.sub_off, 3 * hw * sizeof(float), hipMemcpyDeviceToHost, stream_));
  EMT_HIP(hipMemcpyAsync(s.occ.data(), st_ + d.st_off, hw, hipMemcpyDeviceToHost, stream_));
  EMT_HIP(hipMemcpyAsync(slots.data(), slots_ + d.slot_off, slots.size() * 4, hipMemcpyDeviceToHost, stream_));
  EMT_HIP(hipMemcpyAsync(&h, hdr_ + sim, sizeof(SimHdr), hipMemcpyDeviceToHost, stream_));
  sync();
  for (int k = 0; k < 3; ++k) std::memcpy(s.f[k].data(), &sub[k * hw], hw * sizeof(float));
  const uint32_t *S = &slots[(size_t)h.cur * kSlotFields * d.cap];
  for (int i = 0; i < h.n; ++i) {
    const uint32_t w = S[(size_t)kPacked * d.cap + i];
    if (!slot_live(w)) continue;
    const uint32_t x = slot_site(w);
    auto get = [&](int field) { float v; std::memcpy(&v, &S[(size_t)field * d.cap + i], 4); return v; };
    s.f[TumorState::kTRep][x] = get(kTRep);
    s.f[TumorState::kIrr][x] = get(kR);
    s.f[TumorState::kDelay][x] = get(kDelay);
    s.f[TumorState::kProlif][x] = get(kProlif);
    for (int k = 0; k < 5; ++k) s.f[TumorState::kG1 + k][x] = get(kT0 + k);
    s.mode[x] = (uint8_t)slot_mode(w);
    s.phase[x] = (uint8_t)slot_phase(w);
  }
  return s;
}

}  // namespace emt6ro
