// C ABI of the library (include/emt6ro_amd.h).  Exceptions stop here and
// become return codes plus a per-thread message.
#include "../../include/emt6ro_amd.h"

#include <cmath>
#include <cstring>
#include <future>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "engine.h"
#include "host.h"

struct emt6ro_params { emt6ro::HostParams p; };
struct emt6ro_state { emt6ro::TumorState s; };
struct emt6ro_batch { std::unique_ptr<emt6ro::Batch> b; };

namespace {

thread_local std::string g_last_error;

template <typename F>
int guarded(F &&f) {
  try {
    f();
    g_last_error.clear();
    return EMT6RO_OK;
  } catch (const emt6ro::Error &e) {
    g_last_error = e.what();
    return e.code();
  } catch (const std::bad_alloc &) {
    g_last_error = "Check failed: host allocation; out of host memory";
    return EMT6RO_ERR_ARG;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return EMT6RO_ERR_ARG;
  } catch (...) {
    // Nothing may unwind through the C ABI.
    g_last_error = "Check failed: no exception escapes the C ABI; unknown exception";
    return EMT6RO_ERR_ARG;
  }
}

void need(bool cond, const char *what) {
  if (!cond) throw emt6ro::Error(std::string("Check failed: argument; ") + what, EMT6RO_ERR_ARG);
}

}  // namespace

// _Experiment (python/emt6ro/simulation/backend.cc:16-136) on top of a Batch:
// simulation b = p*T*R + t*R + r; doses live in host memory and are uploaded
// before each run; run() is asynchronous like the reference's std::async.
struct emt6ro_experiment {
  emt6ro::HostParams params;
  int32_t T = 0, R = 0, P = 0;
  uint32_t sim_steps = 0, resolution = 0;
  int32_t n_slots = 0;
  std::vector<float> doses;
  std::unique_ptr<emt6ro::Batch> batch;
  std::future<std::vector<uint32_t>> pending;
  bool running = false;
  uint32_t epoch = 0;
  int path = emt6ro::kPathPersistent;
  std::mutex mu;
};

extern "C" {

const char *emt6ro_last_error(void) { return g_last_error.c_str(); }

const char *emt6ro_version(void) { return "emt6ro_amd 0.1.0 (gfx950)"; }

int emt6ro_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int emt6ro_params_load(const char *path, emt6ro_params **out) {
  return guarded([&] {
    need(path && out, "null pointer");
    auto p = std::make_unique<emt6ro_params>();
    p->p = emt6ro::load_params_json(path);
    *out = p.release();
  });
}

int emt6ro_params_from_values(const float *values, int n, emt6ro_params **out) {
  return guarded([&] {
    need(values && out, "null pointer");
    need(n == EMT6RO_NUM_PARAMS, "expected EMT6RO_NUM_PARAMS values");
    auto p = std::make_unique<emt6ro_params>();
    p->p = emt6ro::HostParams::from_values(values);
    *out = p.release();
  });
}

int emt6ro_params_values(const emt6ro_params *p, float *values, int n) {
  return guarded([&] {
    need(p && values, "null pointer");
    need(n == EMT6RO_NUM_PARAMS, "expected EMT6RO_NUM_PARAMS values");
    p->p.to_values(values);
  });
}

void emt6ro_params_free(emt6ro_params *p) { delete p; }

int emt6ro_state_load(const char *path, const emt6ro_params *p, emt6ro_state **out) {
  return guarded([&] {
    need(path && p && out, "null pointer");
    auto s = std::make_unique<emt6ro_state>();
    s->s = emt6ro::load_state_text(path, p->p);
    *out = s.release();
  });
}

int emt6ro_state_save(const emt6ro_state *s, const char *path) {
  return guarded([&] {
    need(s && path, "null pointer");
    emt6ro::save_state_text(s->s, path);
  });
}

int emt6ro_state_create(int32_t H, int32_t W, const emt6ro_params *p, emt6ro_state **out) {
  return guarded([&] {
    need(p && out, "null pointer");
    auto s = std::make_unique<emt6ro_state>();
    s->s = emt6ro::empty_state(H, W, p->p);
    *out = s.release();
  });
}

int emt6ro_state_dims(const emt6ro_state *s, int32_t *H, int32_t *W) {
  return guarded([&] {
    need(s && H && W, "null pointer");
    *H = s->s.H;
    *W = s->s.W;
  });
}

int emt6ro_state_get_field(const emt6ro_state *s, int field, void *dst) {
  return guarded([&] {
    need(s && dst, "null pointer");
    need(field >= 0 && field < EMT6RO_NUM_FIELDS, "field out of range");
    const auto &st = s->s;
    const size_t n = st.sites();
    if (field == EMT6RO_FIELD_OCCUPANCY) std::memcpy(dst, st.occ.data(), n);
    else if (field == EMT6RO_FIELD_MODE) std::memcpy(dst, st.mode.data(), n);
    else if (field == EMT6RO_FIELD_PHASE) std::memcpy(dst, st.phase.data(), n);
    else std::memcpy(dst, st.f[field].data(), n * sizeof(float));
  });
}

int emt6ro_state_set_field(emt6ro_state *s, int field, const void *src) {
  return guarded([&] {
    need(s && src, "null pointer");
    need(field >= 0 && field < EMT6RO_NUM_FIELDS, "field out of range");
    auto &st = s->s;
    const size_t n = st.sites();
    if (field == EMT6RO_FIELD_OCCUPANCY) {
      const uint8_t *v = static_cast<const uint8_t *>(src);
      for (size_t i = 0; i < n; ++i) need(v[i] <= emt6ro::kMocked, "occupancy values are 0, 1, 2");
      std::memcpy(st.occ.data(), src, n);
    } else if (field == EMT6RO_FIELD_MODE) {
      std::memcpy(st.mode.data(), src, n);
    } else if (field == EMT6RO_FIELD_PHASE) {
      std::memcpy(st.phase.data(), src, n);
    } else {
      std::memcpy(st.f[field].data(), src, n * sizeof(float));
    }
  });
}

void emt6ro_state_free(emt6ro_state *s) { delete s; }

int emt6ro_experiment_create(const emt6ro_params *p, const emt6ro_state *const *tumors,
                             int32_t num_tumors, int32_t runs, int32_t num_protocols,
                             uint32_t sim_steps, uint32_t resolution, int32_t device,
                             uint64_t seed, emt6ro_experiment **out) {
  return guarded([&] {
    need(p && tumors && out, "null pointer");
    need(num_tumors > 0 && runs > 0 && num_protocols > 0, "counts must be positive");
    need(resolution > 0, "protocol resolution must be positive");
    auto e = std::make_unique<emt6ro_experiment>();
    e->params = p->p;
    e->T = num_tumors;
    e->R = runs;
    e->P = num_protocols;
    e->sim_steps = sim_steps;
    e->resolution = resolution;
    // protocol_data_size = ceil(sim_steps / resolution) (backend.cc:48)
    e->n_slots = (int32_t)((sim_steps + resolution - 1) / resolution);
    if (e->n_slots < 1) e->n_slots = 1;
    e->doses.assign((size_t)e->P * e->n_slots, 0.f);
    std::vector<const emt6ro::TumorState *> ts;
    for (int t = 0; t < num_tumors; ++t) {
      need(tumors[t] != nullptr, "null tumour");
      ts.push_back(&tumors[t]->s);
      EMT_ENFORCE(tumors[t]->s.H == ts[0]->H && tumors[t]->s.W == ts[0]->W, EMT6RO_ERR_ARG,
                  "All tumours of an experiment must have the same dimensions");
    }
    const int B = e->P * e->T * e->R;
    emt6ro::BatchSpec spec;
    spec.tumor_index.resize(B);
    spec.prot_index.resize(B);
    spec.horizon.assign(B, 0xffffffffu);
    spec.sim_id.resize(B);
    for (int pi = 0; pi < e->P; ++pi)
      for (int t = 0; t < e->T; ++t)
        for (int r = 0; r < e->R; ++r) {
          const int b = pi * e->T * e->R + t * e->R + r;
          spec.tumor_index[b] = t;
          spec.prot_index[b] = pi;
          spec.sim_id[b] = (uint32_t)b;
        }
    spec.n_prot = e->P;
    spec.n_slots = e->n_slots;
    spec.prot_res = resolution;
    spec.prot_length = sim_steps;
    e->batch = std::make_unique<emt6ro::Batch>(e->params, ts, spec, device, seed);
    *out = e.release();
  });
}

int emt6ro_experiment_add_irradiations(emt6ro_experiment *e, int32_t num_protocols,
                                       const int32_t *offsets, const int32_t *steps,
                                       const float *doses) {
  return guarded([&] {
    need(e && offsets, "null pointer");
    std::lock_guard<std::mutex> lk(e->mu);
    const int32_t protocols_num = e->P;
    EMT_ENFORCE(num_protocols == protocols_num, EMT6RO_ERR_ARG,
                "Wrong number of protocols. Expected: ", protocols_num);
    EMT_ENFORCE(!e->running, EMT6RO_ERR_ARG, "Cannot add irradiations while the experiment runs.");
    for (int i = 0; i < e->P; ++i)
      for (int32_t k = offsets[i]; k < offsets[i + 1]; ++k) {
        need(steps && doses, "null pointer");
        const int32_t step = steps[k];
        EMT_ENFORCE(step >= 0 && (uint32_t)step / e->resolution < (uint32_t)e->n_slots,
                    EMT6RO_ERR_ARG, "Irradiation step ", step, " is outside the protocol of ",
                    e->sim_steps, " steps");
        // Protocol::closestDose (protocol.h:16-18)
        e->doses[(size_t)i * e->n_slots + (uint32_t)step / e->resolution] += doses[k];
      }
  });
}

int emt6ro_experiment_run(emt6ro_experiment *e, uint32_t nsteps) {
  return guarded([&] {
    need(e != nullptr, "null pointer");
    std::lock_guard<std::mutex> lk(e->mu);
    const bool running = e->running;
    EMT_ENFORCE(!running, EMT6RO_ERR_ARG, "Experiment already running.");
    e->batch->set_protocols(e->doses.data());
    e->running = true;
    emt6ro::Batch *b = e->batch.get();
    const int path = e->path;
    e->pending = std::async(std::launch::async, [b, nsteps, path]() {
      b->run(nsteps, path, -1, 0);
      std::vector<uint32_t> res(b->nsims());
      b->results(res.data());
      return res;
    });
  });
}

int emt6ro_experiment_results(emt6ro_experiment *e, uint32_t *dst) {
  return guarded([&] {
    need(e && dst, "null pointer");
    std::lock_guard<std::mutex> lk(e->mu);
    if (!e->running) throw emt6ro::Error("First you have to run the experiment.", EMT6RO_ERR_ARG);
    e->running = false;
    std::vector<uint32_t> res = e->pending.get();
    std::memcpy(dst, res.data(), res.size() * sizeof(uint32_t));
  });
}

int emt6ro_experiment_reset(emt6ro_experiment *e) {
  return guarded([&] {
    need(e != nullptr, "null pointer");
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->running) {
      e->pending.wait();
    }
    // A new epoch gives later runs fresh random numbers, as the reference's
    // generator states keep advancing across reset() (simulation.cu:212-216).
    e->epoch += 1;
    e->batch->reset(e->epoch);
    std::fill(e->doses.begin(), e->doses.end(), 0.f);
  });
}

int emt6ro_experiment_state(emt6ro_experiment *e, int32_t sample, emt6ro_state **out) {
  return guarded([&] {
    need(e && out, "null pointer");
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->running) e->pending.wait();
    auto s = std::make_unique<emt6ro_state>();
    s->s = e->batch->state(sample);
    *out = s.release();
  });
}

int emt6ro_experiment_batch_size(const emt6ro_experiment *e) {
  return e ? e->batch->nsims() : EMT6RO_ERR_ARG;
}

int emt6ro_experiment_set_path(emt6ro_experiment *e, int32_t path) {
  return guarded([&] {
    need(e != nullptr, "null pointer");
    need(path == emt6ro::kPathPersistent || path == emt6ro::kPathStep ||
             path == emt6ro::kPathStepGraph, "path must be 0, 1 or 2");
    std::lock_guard<std::mutex> lk(e->mu);
    e->path = path;
  });
}

void emt6ro_experiment_destroy(emt6ro_experiment *e) {
  if (!e) return;
  if (e->running) {
    try { e->pending.wait(); } catch (...) {}
  }
  delete e;
}

int emt6ro_batch_create(const emt6ro_params *p, const emt6ro_state *const *tumors,
                        int32_t num_tumors, int32_t num_sims, const int32_t *tumor_index,
                        const int32_t *prot_index, const uint32_t *horizon,
                        const uint32_t *sim_id, int32_t n_prot, int32_t n_slots,
                        const float *prot_doses, uint32_t prot_res, uint32_t prot_length,
                        int32_t device, uint64_t seed, emt6ro_batch **out) {
  return guarded([&] {
    need(p && tumors && tumor_index && prot_index && horizon && sim_id && prot_doses && out,
         "null pointer");
    need(num_tumors > 0 && num_sims > 0, "counts must be positive");
    std::vector<const emt6ro::TumorState *> ts;
    for (int t = 0; t < num_tumors; ++t) {
      need(tumors[t] != nullptr, "null tumour");
      ts.push_back(&tumors[t]->s);
    }
    emt6ro::BatchSpec spec;
    spec.tumor_index.assign(tumor_index, tumor_index + num_sims);
    spec.prot_index.assign(prot_index, prot_index + num_sims);
    spec.horizon.assign(horizon, horizon + num_sims);
    spec.sim_id.assign(sim_id, sim_id + num_sims);
    spec.n_prot = n_prot;
    spec.n_slots = n_slots;
    spec.prot_res = prot_res;
    spec.prot_length = prot_length;
    auto b = std::make_unique<emt6ro_batch>();
    b->b = std::make_unique<emt6ro::Batch>(p->p, ts, spec, device, seed);
    b->b->set_protocols(prot_doses);
    *out = b.release();
  });
}

int emt6ro_batch_run(emt6ro_batch *b, uint32_t nsteps, int32_t path, int32_t variant,
                     uint32_t steps_per_launch) {
  return guarded([&] {
    need(b != nullptr, "null pointer");
    b->b->run(nsteps, path, variant, steps_per_launch);
  });
}

int emt6ro_batch_results(emt6ro_batch *b, uint32_t *dst) {
  return guarded([&] {
    need(b && dst, "null pointer");
    b->b->results(dst);
  });
}

int emt6ro_batch_counters(emt6ro_batch *b, uint64_t *dst) {
  return guarded([&] {
    need(b && dst, "null pointer");
    b->b->counters(dst);
  });
}

int emt6ro_batch_state(emt6ro_batch *b, int32_t sim, emt6ro_state **out) {
  return guarded([&] {
    need(b && out, "null pointer");
    auto s = std::make_unique<emt6ro_state>();
    s->s = b->b->state(sim);
    *out = s.release();
  });
}

int emt6ro_batch_reset(emt6ro_batch *b, uint32_t epoch) {
  return guarded([&] {
    need(b != nullptr, "null pointer");
    b->b->reset(epoch);
  });
}

int emt6ro_batch_last_run(const emt6ro_batch *b, double *ms, int64_t *launches) {
  return guarded([&] {
    need(b && ms && launches, "null pointer");
    *ms = b->b->last_ms();
    *launches = b->b->last_launches();
  });
}

void emt6ro_batch_destroy(emt6ro_batch *b) { delete b; }

int emt6ro_batch_set_variant(emt6ro_batch *b, int32_t rs, int32_t variant) {
  return guarded([&] {
    need(b != nullptr, "null pointer");
    int cnt = 0;
    emt6ro::variant_table(&cnt);
    need(variant >= -1 && variant < cnt, "variant out of range");
    b->b->set_preferred_variant(rs, variant);
  });
}

int emt6ro_batch_set_window(emt6ro_batch *b, int32_t rows) {
  return guarded([&] {
    need(b != nullptr, "null pointer");
    b->b->set_window(rows);
  });
}

int emt6ro_batch_set_doses(emt6ro_batch *b, const float *prot_doses) {
  return guarded([&] {
    need(b && prot_doses, "null pointer");
    b->b->set_protocols(prot_doses);
  });
}

int emt6ro_variant_info(int32_t idx, int32_t *info) {
  int cnt = 0;
  const emt6ro::VariantInfo *v = emt6ro::variant_table(&cnt);
  if (idx < 0) return cnt;
  return guarded([&] {
    need(idx < cnt && info, "variant index out of range");
    info[0] = v[idx].id;
    info[1] = v[idx].kind;
    info[2] = v[idx].nt;
    info[3] = v[idx].l;
    info[4] = v[idx].kmax;
    info[5] = v[idx].kf;
    info[6] = v[idx].ps;
    info[7] = v[idx].rs;
  });
}

int emt6ro_variant_occupancy(int32_t variant, int32_t hmax, int32_t wmax) {
  int cnt = 0;
  const emt6ro::VariantInfo *v = emt6ro::variant_table(&cnt);
  if (variant < 0 || variant >= cnt || !emt6ro::variant_fits(v[variant], hmax, wmax)) return 0;
  return emt6ro::persistent_occupancy(variant,
                                      emt6ro::persistent_lds_bytes(v[variant], hmax, wmax));
}

int emt6ro_instr_read(uint64_t counters[4], double *sum_death_prob) {
  return guarded([&] {
    need(counters && sum_death_prob, "null pointer");
    std::vector<unsigned long long> rec(emt6ro::dev::kInstrN);
    emt6ro::instr_collect(rec.data(), false);
    for (int i = 0; i < 4; ++i) counters[i] = rec[i];
    *sum_death_prob = (double)rec[emt6ro::dev::kInstrDeathProbFx] / 4294967296.0;
  });
}

int emt6ro_instr_reset(void) {
  return guarded([&] {
    std::vector<unsigned long long> rec(emt6ro::dev::kInstrN);
    emt6ro::instr_collect(rec.data(), true);
  });
}

int emt6ro_instr_histograms(uint64_t *dst, int32_t n) {
  return guarded([&] {
    need(dst != nullptr || n == 0, "null pointer");
    std::vector<unsigned long long> rec(emt6ro::dev::kInstrN);
    emt6ro::instr_collect(rec.data(), false);
    for (int32_t i = 0; i < n; ++i)
      dst[i] = i < 2 * emt6ro::dev::kInstrHistN ? rec[emt6ro::dev::kInstrIrrHist + i] : 0;
  });
}

int emt6ro_instr_enable(int32_t on) {
  return guarded([&] { emt6ro::instr_set_enabled(on != 0); });
}

int emt6ro_instr_enabled(void) { return emt6ro::instr_enabled() ? 1 : 0; }

}  // extern "C"
