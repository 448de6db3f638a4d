// Batch engine: the device state of many simulations and the runs over it.
#ifndef EMT6RO_AMD_ENGINE_H_
#define EMT6RO_AMD_ENGINE_H_

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>
#include "host.h"
#include "kernels.h"
#include "sim_phases.h"

namespace emt6ro {

#define EMT_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw ::emt6ro::Error(::emt6ro::cat("HIP error ", hipGetErrorString(e_), " at ", \
                                          #call),                                       \
                            -3);                                                        \
  } while (0)

enum RunPath { kPathPersistent = 0, kPathStep = 1, kPathStepGraph = 2 };

// Instrumentation counters (dev::kInstr* layout), one record per device,
// allocated on first use.  Runs started while instrumentation is enabled add
// to the record of their device.  Enabled at load time when the environment
// variable EMT6RO_INSTRUMENT is set to anything but "" or "0".
bool instr_enabled();
void instr_set_enabled(bool on);
unsigned long long *instr_device_record(int device);  // null while disabled
// Sum of every device's record (kInstrN values); zeroes them when `reset`.
void instr_collect(unsigned long long *out, bool reset);

struct BatchSpec {
  std::vector<int32_t> tumor_index, prot_index;
  std::vector<uint32_t> horizon, sim_id;
  int32_t n_prot = 0, n_slots = 0;
  uint32_t prot_res = 1, prot_length = 0;
};

class Batch {
 public:
  Batch(const HostParams &p, const std::vector<const TumorState *> &tumors,
        const BatchSpec &spec, int device, uint64_t seed);
  ~Batch();
  Batch(const Batch &) = delete;
  Batch &operator=(const Batch &) = delete;

  int nsims() const { return nsims_; }
  int device() const { return device_; }
  void set_protocols(const float *doses);  // n_prot * n_slots
  void reset(uint32_t epoch);
  void run(uint32_t nsteps, int path, int variant, uint32_t steps_per_launch);
  void results(uint32_t *dst);
  void counters(uint64_t *dst);
  TumorState state(int sim);
  double last_ms() const { return last_ms_; }
  int64_t last_launches() const { return last_launches_; }
  // Kernel variant for lattices up to hmax x wmax: among the fitting
  // variants of `kind` (-1: any whose persistent kernel fits in LDS) with
  // the smallest LDS row stride, the preferred one (kind -1: the fastest
  // measured), else the highest kind, then table order.
  static int choose_variant(int hmax, int wmax, int kind = -1);
  // Preferred variant per LDS row stride class (-1: choose_variant).
  void set_preferred_variant(int rs, int variant);
  int class_count() const { return (int)classes_.size(); }
  // Window pass of the persistent path (lattices wider than row stride 36):
  // rows of the window (0: off, -1: the most that keep eight workgroups on
  // a CU).
  void set_window(int rows);

 private:
  void sync();
  void ensure_scratch(int wgs);
  dev::RunArgs args(uint32_t nsteps) const;
  // Simulations grouped by the LDS row stride their lattice needs; each
  // group is one persistent launch, its simulations longest-first.
  struct Class {
    int rs, offset, count, hmax, wmax;
  };
  std::vector<Class> classes_;
  int preferred_[3] = {-1, -1, -1};

  HostParams params_;
  DevParams dparams_;
  int device_;
  uint32_t seed_, epoch_ = 0;
  int nsims_;
  int hmax_ = 0, wmax_ = 0, capmax_ = 0;
  BatchSpec spec_;
  std::vector<SimDesc> desc_h_;
  std::vector<int32_t> order_h_;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  // device memory
  SimDesc *desc_ = nullptr;
  SimHdr *hdr_ = nullptr;
  float *sub_ = nullptr;
  uint8_t *st_ = nullptr;
  uint32_t *slots_ = nullptr;
  float *prot_ = nullptr;
  int32_t *order_ = nullptr;
  int32_t *order_next_ = nullptr;  // the order of the next launch, sorted by cost
  int32_t *order_win_ = nullptr;   // the window kernel's queue
  int32_t *wcount_ = nullptr;      // its length, per class
  // Fraction of each class the window queue held at the end of the last
  // run; below kWindowMin the window kernel is skipped (its launch would be
  // mostly tail).  Reset to 1 by reset().
  std::vector<double> win_frac_;
  static constexpr double kWindowMin = 0.4;
  static constexpr int kWindowKinds = 2;   // window row strides 36 and 56
  uint32_t *target_ = nullptr;     // per-simulation target step of a launch
  int window_rows_ = -1;
  int32_t *queue_ = nullptr;
  uint32_t *results_ = nullptr;
  int32_t *scratch_ = nullptr;
  uint64_t *ovf_ = nullptr;
  int scratch_wgs_ = 0;
  // initial images, one per tumour, and the per-simulation copy plan
  std::vector<float *> img_sub_;
  std::vector<uint8_t *> img_st_;
  std::vector<uint32_t *> img_slots_;
  std::vector<SimHdr> img_hdr_;
  struct ImgDesc *img_desc_ = nullptr;
  SimHdr *img_hdr_d_ = nullptr;
  int32_t *tumor_of_ = nullptr;
  int64_t sub_total_ = 0, st_total_ = 0, slot_total_ = 0;
  double last_ms_ = 0.0;
  int64_t last_launches_ = 0;
  // captured per-step graph and the launch arguments that can change between
  // runs (epoch, instrumentation record): a change recaptures it
  hipGraphExec_t graph_exec_ = nullptr;
  int graph_steps_ = 0, graph_variant_ = -1;
  uint32_t graph_epoch_ = 0;
  unsigned long long *graph_instr_ = nullptr;
};

}  // namespace emt6ro

#endif  // EMT6RO_AMD_ENGINE_H_
