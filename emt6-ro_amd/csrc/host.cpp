// Parameters (JSON), tumour states (text format) and host arithmetic shared
// with the kernels.  Built with -ffp-contract=off like the kernels.
#include "host.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>

namespace emt6ro {

// ---------------------------------------------------------------------------
// A small JSON reader: objects, arrays, numbers, strings, true/false/null.
// Numbers are parsed as double and narrowed to float on use, as nlohmann's
// get<float>() does (parameters.cc:27-63).
// ---------------------------------------------------------------------------
namespace {

struct JValue {
  enum Kind { kNull, kBool, kNum, kStr, kArr, kObj } kind = kNull;
  double num = 0.0;
  bool b = false;
  std::string str;
  std::vector<JValue> arr;
  std::map<std::string, JValue> obj;

  const JValue &operator[](const std::string &key) const {
    auto it = obj.find(key);
    if (kind != kObj || it == obj.end())
      throw Error(cat("Check failed: key present; missing JSON key: ", key), -2);
    return it->second;
  }
  float f() const {
    if (kind != kNum) throw Error("Check failed: number; JSON value is not a number", -2);
    return (float)num;
  }
};

class JParser {
 public:
  explicit JParser(const std::string &s) : s_(s) {}
  JValue parse() {
    JValue v = value();
    ws();
    if (i_ != s_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string &s_;
  size_t i_ = 0;

  [[noreturn]] void fail(const char *what) {
    throw Error(cat("Check failed: valid JSON; ", what, " at offset ", i_), -2);
  }
  void ws() {
    while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\n' || s_[i_] == '\r' || s_[i_] == '\t'))
      ++i_;
  }
  JValue value() {
    ws();
    if (i_ >= s_.size()) fail("unexpected end");
    const char c = s_[i_];
    JValue v;
    if (c == '{') {
      v.kind = JValue::kObj;
      ++i_;
      ws();
      if (i_ < s_.size() && s_[i_] == '}') { ++i_; return v; }
      for (;;) {
        ws();
        if (i_ >= s_.size() || s_[i_] != '"') fail("expected key");
        std::string k = string();
        ws();
        if (i_ >= s_.size() || s_[i_] != ':') fail("expected ':'");
        ++i_;
        v.obj[k] = value();
        ws();
        if (i_ < s_.size() && s_[i_] == ',') { ++i_; continue; }
        if (i_ < s_.size() && s_[i_] == '}') { ++i_; return v; }
        fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      v.kind = JValue::kArr;
      ++i_;
      ws();
      if (i_ < s_.size() && s_[i_] == ']') { ++i_; return v; }
      for (;;) {
        v.arr.push_back(value());
        ws();
        if (i_ < s_.size() && s_[i_] == ',') { ++i_; continue; }
        if (i_ < s_.size() && s_[i_] == ']') { ++i_; return v; }
        fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.kind = JValue::kStr;
      v.str = string();
      return v;
    }
    if (s_.compare(i_, 4, "true") == 0) { i_ += 4; v.kind = JValue::kBool; v.b = true; return v; }
    if (s_.compare(i_, 5, "false") == 0) { i_ += 5; v.kind = JValue::kBool; return v; }
    if (s_.compare(i_, 4, "null") == 0) { i_ += 4; return v; }
    const char *start = s_.c_str() + i_;
    char *end = nullptr;
    v.num = std::strtod(start, &end);
    if (end == start) fail("bad value");
    i_ += (size_t)(end - start);
    v.kind = JValue::kNum;
    return v;
  }
  std::string string() {
    ++i_;  // opening quote
    std::string out;
    while (i_ < s_.size() && s_[i_] != '"') {
      if (s_[i_] == '\\' && i_ + 1 < s_.size()) {
        ++i_;
        const char e = s_[i_];
        out.push_back(e == 'n' ? '\n' : e == 't' ? '\t' : e);
      } else {
        out.push_back(s_[i_]);
      }
      ++i_;
    }
    if (i_ >= s_.size()) fail("unterminated string");
    ++i_;
    return out;
  }
};

void read_sub(const JValue &o, float out[3]) {
  out[0] = o["cho"].f();
  out[1] = o["ox"].f();
  out[2] = o["gi"].f();
}

}  // namespace

// Parameters::loadFromJSONFile (parameters.cc:66-80).
HostParams load_params_json(const std::string &path) {
  std::ifstream ifs(path);
  EMT_ENFORCE(ifs.good(), -2, "Could not open the file: ", path);
  std::stringstream ss;
  ss << ifs.rdbuf();
  const std::string text = ss.str();
  const JValue root = JParser(text).parse();
  HostParams p;
  p.lat_h = (int16_t)root["lattice_dims"]["height"].num;
  p.lat_w = (int16_t)root["lattice_dims"]["width"].num;
  p.time_step = root["time_step"].f();
  read_sub(root["diffusion_params"]["coeffs"], p.diff_coeffs);
  p.diff_time_step = root["diffusion_params"]["time_step"].f();
  const JValue &m = root["metabolism"];
  read_sub(m["aerobic_proliferation"], p.metab[0]);
  read_sub(m["anaerobic_proliferation"], p.metab[1]);
  read_sub(m["aerobic_quiescence"], p.metab[2]);
  read_sub(m["anaerobic_quiescence"], p.metab[3]);
  read_sub(root["external_levels"], p.ext);
  p.death_gi = root["death_gi"].f();
  p.quiescence_gi = root["quiescence_gi"].f();
  p.death_gi_production = root["death_gi_production"].f();
  const char *phases[5] = {"g1", "s", "g2", "m", "d"};
  for (int k = 0; k < 5; ++k) {
    p.cycle_mean[k] = root["cycle_times"][phases[k]]["mean"].f();
    p.cycle_sd[k] = root["cycle_times"][phases[k]]["stddev"].f();
  }
  const JValue &cr = root["cell_repair"];
  p.delay_coeff = cr["delay_time"]["coeff"].f();
  p.delay_exp = cr["delay_time"]["exp_coeff"].f();
  p.surv_coeff = cr["survival_prob"]["coeff"].f();
  p.surv_exp = cr["survival_prob"]["exp_coeff"].f();
  p.repair_half_time = cr["repair_half_time"].f();
  return p;
}

void HostParams::to_values(float *v) const {
  int i = 0;
  v[i++] = (float)lat_h;
  v[i++] = (float)lat_w;
  v[i++] = time_step;
  for (int s = 0; s < 3; ++s) v[i++] = diff_coeffs[s];
  v[i++] = diff_time_step;
  for (int m = 0; m < 4; ++m)
    for (int s = 0; s < 3; ++s) v[i++] = metab[m][s];
  for (int s = 0; s < 3; ++s) v[i++] = ext[s];
  v[i++] = death_gi;
  v[i++] = quiescence_gi;
  v[i++] = death_gi_production;
  for (int k = 0; k < 5; ++k) {
    v[i++] = cycle_mean[k];
    v[i++] = cycle_sd[k];
  }
  v[i++] = delay_coeff;
  v[i++] = delay_exp;
  v[i++] = surv_coeff;
  v[i++] = surv_exp;
  v[i++] = repair_half_time;
}

HostParams HostParams::from_values(const float *v) {
  HostParams p;
  int i = 0;
  p.lat_h = (int32_t)v[i++];
  p.lat_w = (int32_t)v[i++];
  p.time_step = v[i++];
  for (int s = 0; s < 3; ++s) p.diff_coeffs[s] = v[i++];
  p.diff_time_step = v[i++];
  for (int m = 0; m < 4; ++m)
    for (int s = 0; s < 3; ++s) p.metab[m][s] = v[i++];
  for (int s = 0; s < 3; ++s) p.ext[s] = v[i++];
  p.death_gi = v[i++];
  p.quiescence_gi = v[i++];
  p.death_gi_production = v[i++];
  for (int k = 0; k < 5; ++k) {
    p.cycle_mean[k] = v[i++];
    p.cycle_sd[k] = v[i++];
  }
  p.delay_coeff = v[i++];
  p.delay_exp = v[i++];
  p.surv_coeff = v[i++];
  p.surv_exp = v[i++];
  p.repair_half_time = v[i++];
  return p;
}

// Derived constants, in the reference's float expressions: clock step
// time_step / 3600 (cell.cu:25, cell.h:107), G1/S margin 2*time_step / 3600.f
// (cell.cu:68), diffusion coefficient ((D * dt) * HS) / f with HS = (float)
// (2 sqrt2) and f = (float)(4 + HS) (diffusion.cu:119-120,163), substeps
// (int16)(time_step / dt) (simulation.cu:170, diffusion.h).
DevParams HostParams::device() const {
  DevParams d{};
  d.time_step = time_step;
  d.hour_step = time_step / 3600.f;
  d.death_margin = time_step / 3600.f;
  d.g1s_margin = (2.f * time_step) / 3600.f;
  const float HS = (float)(2.0 * 1.41421356237309504880);
  const float F = (float)(4.0 + (double)HS);
  for (int s = 0; s < 3; ++s) {
    d.ext[s] = ext[s];
    d.k[s] = ((diff_coeffs[s] * diff_time_step) * HS) / F;
  }
  d.substeps = (int32_t)(int16_t)(time_step / diff_time_step);
  for (int m = 0; m < 4; ++m)
    for (int s = 0; s < 3; ++s) d.metab[m][s] = metab[m][s];
  d.death_gi = death_gi;
  d.quiescence_gi = quiescence_gi;
  for (int k = 0; k < 5; ++k) {
    d.cycle_mean[k] = cycle_mean[k];
    d.cycle_sd[k] = cycle_sd[k];
  }
  d.delay_coeff = delay_coeff;
  d.delay_exp = delay_exp;
  d.surv_coeff = surv_coeff;
  d.surv_exp = surv_exp;
  d.repair_half_time = repair_half_time;
  return d;
}

// The kernels' expf (Cephes expf: Cody-Waite reduction, degree-5 polynomial),
// in the same operation order.
float expf_fixed_host(float x) {
  x = x > 88.0f ? 88.0f : x;
  x = x < -87.0f ? -87.0f : x;
  float z = std::floor(1.44269504088896341f * x + 0.5f);
  float r = x - z * 0.693359375f;
  r = r - z * -2.12194440e-4f;
  int n = (int)z;
  float rr = r * r;
  float p = 1.9875691500e-4f;
  p = p * r + 1.3981999507e-3f;
  p = p * r + 8.3334519073e-3f;
  p = p * r + 4.1665795894e-2f;
  p = p * r + 1.6666665459e-1f;
  p = p * r + 5.0000001201e-1f;
  p = p * rr;
  p = p + r;
  p = p + 1.0f;
  uint32_t bits = (uint32_t)(n + 127) << 23;
  float scale;
  std::memcpy(&scale, &bits, 4);
  return p * scale;
}

float delay_time_host(float R, const HostParams &p) {
  return p.delay_coeff * expf_fixed_host(p.delay_exp * R);
}

TumorState::TumorState(int32_t h, int32_t w) : H(h), W(w) {
  const size_t n = (size_t)h * w;
  for (auto &v : f) v.assign(n, 0.f);
  occ.assign(n, kVacant);
  mode.assign(n, 0);
  phase.assign(n, 0);
}

TumorState empty_state(int32_t H, int32_t W, const HostParams &p) {
  EMT_ENFORCE(H >= 3 && W >= 3 && H <= 255 && W <= 255, -1,
              "Lattice dims must be within 3..255, got ", H, " x ", W);
  TumorState s(H, W);
  for (int32_t r = 0; r < H; ++r)
    for (int32_t c = 0; c < W; ++c) {
      const size_t x = (size_t)r * W + c;
      for (int k = 0; k < 3; ++k) s.f[k][x] = p.ext[k];
      s.occ[x] = (r == 0 || r == H - 1 || c == 0 || c == W - 1) ? kMocked : kVacant;
    }
  return s;
}

namespace {

class Tokens {
 public:
  explicit Tokens(std::istream &is) : is_(is) {}
  bool next(std::string &tok) { return static_cast<bool>(is_ >> tok); }
  float f(const std::string &path) {
    std::string t;
    EMT_ENFORCE(next(t), -2, "Unexpected end of file: ", path);
    char *end = nullptr;
    float v = std::strtof(t.c_str(), &end);
    EMT_ENFORCE(end != t.c_str(), -2, "Bad number '", t, "' in ", path);
    return v;
  }
  int i(const std::string &path) {
    std::string t;
    EMT_ENFORCE(next(t), -2, "Unexpected end of file: ", path);
    char *end = nullptr;
    long v = std::strtol(t.c_str(), &end, 10);
    EMT_ENFORCE(end != t.c_str(), -2, "Bad integer '", t, "' in ", path);
    return (int)v;
  }

 private:
  std::istream &is_;
};

}  // namespace

// loadFromFile (state.cc:9-53).
TumorState load_state_text(const std::string &path, const HostParams &p) {
  std::ifstream file(path);
  EMT_ENFORCE(file.good(), -2, "Could not open the file: ", path);
  Tokens tk(file);
  const int h = tk.i(path), w = tk.i(path);
  TumorState s = empty_state(h + 2, w + 2, p);
  for (int r = 1; r <= h; ++r)
    for (int c = 1; c <= w; ++c) {
      const size_t x = (size_t)r * s.W + c;
      const int occ = tk.i(path);
      s.f[TumorState::kCho][x] = tk.f(path);
      s.f[TumorState::kOx][x] = tk.f(path);
      s.f[TumorState::kGi][x] = tk.f(path);
      if (occ) {
        s.occ[x] = kOccupied;
        s.f[TumorState::kTRep][x] = tk.f(path);
        const float R = tk.f(path);
        s.f[TumorState::kIrr][x] = R;
        s.f[TumorState::kDelay][x] = R > 0.f ? delay_time_host(R, p) : 0.f;
        s.f[TumorState::kProlif][x] = tk.f(path);
        for (int k = 0; k < 5; ++k) s.f[TumorState::kG1 + k][x] = tk.f(path);
        s.mode[x] = (uint8_t)tk.i(path);
        s.phase[x] = (uint8_t)tk.i(path);
      } else {
        s.occ[x] = kVacant;
      }
    }
  return s;
}

// saveToFile (state.cc:55-82): default ostream formatting (6 significant
// digits), one value per line.
void save_state_text(const TumorState &s, const std::string &path) {
  std::ofstream file(path);
  EMT_ENFORCE(file.good(), -2, "Could not open the file: ", path);
  const int h = s.H - 2, w = s.W - 2;
  file << h << "\n" << w << "\n";
  for (int r = 1; r <= h; ++r)
    for (int c = 1; c <= w; ++c) {
      const size_t x = (size_t)r * s.W + c;
      const bool occ = s.occ[x] == kOccupied;
      file << (int)occ << "\n";
      file << s.f[TumorState::kCho][x] << "\n" << s.f[TumorState::kOx][x] << "\n"
           << s.f[TumorState::kGi][x] << "\n";
      if (occ) {
        file << s.f[TumorState::kTRep][x] << "\n";
        file << s.f[TumorState::kIrr][x] << "\n";
        file << s.f[TumorState::kProlif][x] << "\n";
        for (int k = 0; k < 5; ++k) file << s.f[TumorState::kG1 + k][x] << "\n";
        file << (int)s.mode[x] << "\n";
        file << (int)s.phase[x] << "\n";
      }
    }
}

std::vector<int32_t> occupied_list_order(const TumorState &s) {
  std::vector<int32_t> out;
  for (int g = 0; g < 32; ++g)
    for (int r = 0; r < s.H - 1; ++r)
      for (int c = g; c < s.W - 1; c += 32) {
        const int32_t x = r * s.W + c;
        if (s.occ[x] == kOccupied) out.push_back(x);
      }
  return out;
}

}  // namespace emt6ro
