// Host-side model objects: parameters, tumour states and their file formats.
#ifndef EMT6RO_AMD_HOST_H_
#define EMT6RO_AMD_HOST_H_

#include <cstdint>
#include <stdexcept>
#include <sstream>
#include <string>
#include <vector>
#include "model.h"

namespace emt6ro {

// "Check failed: <cond>; <message>", the text of the reference's ENFORCE
// (common/error.h:25-29).
class Error : public std::runtime_error {
 public:
  Error(const std::string &msg, int code) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

template <typename... Args>
std::string cat(const Args &...args) {
  std::ostringstream ss;
  (ss << ... << args);
  return ss.str();
}

#define EMT_ENFORCE(COND, CODE, ...)                                                    \
  do {                                                                                  \
    if (!(COND))                                                                        \
      throw ::emt6ro::Error(std::string("Check failed: " #COND "; ") +                 \
                                ::emt6ro::cat(__VA_ARGS__), CODE);                      \
  } while (0)

// Parameters (parameters/parameters.h:10-72), flattened in the order of the
// C ABI's parameter vector.
struct HostParams {
  int32_t lat_h = 0, lat_w = 0;
  float time_step = 0.f;
  float diff_coeffs[3] = {0.f, 0.f, 0.f};
  float diff_time_step = 0.f;
  float metab[4][3] = {};
  float ext[3] = {0.f, 0.f, 0.f};
  float death_gi = 0.f, quiescence_gi = 0.f, death_gi_production = 0.f;
  float cycle_mean[5] = {}, cycle_sd[5] = {};
  float delay_coeff = 0.f, delay_exp = 0.f, surv_coeff = 0.f, surv_exp = 0.f;
  float repair_half_time = 0.f;

  static constexpr int kNumValues = 40;
  void to_values(float *v) const;
  static HostParams from_values(const float *v);
  DevParams device() const;
};

HostParams load_params_json(const std::string &path);

// A tumour state: per-site arrays of the bordered H x W lattice
// (HostGrid<Site>, common/grid.h:150-170; Site, site/site.h:34-83).
struct TumorState {
  enum Field {
    kCho = 0, kOx, kGi, kOcc, kTRep, kIrr, kDelay, kProlif,
    kG1, kS, kG2, kM, kD, kMode, kPhase, kNumFields
  };
  int32_t H = 0, W = 0;
  std::vector<float> f[13];       // float fields, indexed by Field (< kMode)
  std::vector<uint8_t> occ, mode, phase;

  TumorState() = default;
  TumorState(int32_t h, int32_t w);
  static bool is_float(int field) { return field != kOcc && field != kMode && field != kPhase; }
  size_t sites() const { return (size_t)H * W; }
};

// loadFromFile / saveToFile (state/state.cc:9-82).
TumorState load_state_text(const std::string &path, const HostParams &p);
void save_state_text(const TumorState &s, const std::string &path);
// All-vacant lattice with a MOCKED border, every site at the external levels.
TumorState empty_state(int32_t H, int32_t W, const HostParams &p);

// FP32 exp with the fixed operation order the kernels use (device.h).
float expf_fixed_host(float x);
float delay_time_host(float R, const HostParams &p);

// The reference's occupied-list order (findOccupied, simulation.cu:41-64):
// column groups g = 0..31 (columns g, g+32, ...), rows 0..H-2 within a group.
std::vector<int32_t> occupied_list_order(const TumorState &s);

}  // namespace emt6ro

#endif  // EMT6RO_AMD_HOST_H_
