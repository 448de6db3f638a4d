/*
 * emt6ro_amd -- C ABI of the MI355X EMT6/Ro simulator (libemt6ro_amd.so).
 *
 * Plain C types only; every handle is opaque and every output buffer is
 * allocated by the caller.  Functions return 0 on success and a negative code
 * on failure; emt6ro_last_error() then gives the message of the calling
 * thread's last failure.  Where the reference raises an error through
 * ENFORCE (common/error.h:25-29) the message has the same text:
 * "Check failed: <condition>; <message>".
 *
 * The `python/emt6ro` package of the reference binds the same operations
 * through pybind11; each declaration cites the reference function it stands
 * in for.
 */
#ifndef EMT6RO_AMD_H_
#define EMT6RO_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMT6RO_OK 0
#define EMT6RO_ERR_ARG -1     /* invalid argument or state */
#define EMT6RO_ERR_IO -2      /* file could not be read or parsed */
#define EMT6RO_ERR_DEVICE -3  /* HIP runtime error */

typedef struct emt6ro_params emt6ro_params;
typedef struct emt6ro_state emt6ro_state;
typedef struct emt6ro_experiment emt6ro_experiment;
typedef struct emt6ro_batch emt6ro_batch;

/* Message of this thread's last failed call ("" if none). */
const char *emt6ro_last_error(void);
/* Library version string. */
const char *emt6ro_version(void);
/* Number of visible GPUs (0 when none or no driver); no reference
 * counterpart. */
int emt6ro_device_count(void);

/* ---- Parameters (parameters/parameters.h:10-72) ------------------------- */

/* Number of values in the flat parameter vector:
 *   [0] lattice height  [1] lattice width  [2] time_step
 *   [3..5] diffusion coeffs cho, ox, gi  [6] diffusion time_step
 *   [7..18] metabolism: aerobic prolif., anaerobic prolif., aerobic quiesc.,
 *           anaerobic quiesc., each cho, ox, gi
 *   [19..21] external levels cho, ox, gi
 *   [22] death_gi  [23] quiescence_gi  [24] death_gi_production
 *   [25..34] cycle times g1, s, g2, m, d, each mean, stddev
 *   [35] delay coeff  [36] delay exp_coeff  [37] survival coeff
 *   [38] survival exp_coeff  [39] repair_half_time */
#define EMT6RO_NUM_PARAMS 40

/* Parameters::loadFromJSONFile (parameters/parameters.cc:66-80);
 * Python: load_parameters (python/emt6ro/simulation/backend.cc:217). */
int emt6ro_params_load(const char *path, emt6ro_params **out);
int emt6ro_params_from_values(const float *values, int n, emt6ro_params **out);
int emt6ro_params_values(const emt6ro_params *p, float *values, int n);
void emt6ro_params_free(emt6ro_params *p);

/* ---- Tumour state (state/state.h:10-12; HostGrid<Site>, common/grid.h) -- */

/* Per-site fields of a tumour state; each is an H x W row-major array of the
 * bordered lattice (H = interior height + 2).  TumorState accessors in
 * python/emt6ro/simulation/backend.cc:138-179,198-205. */
enum emt6ro_field {
  EMT6RO_FIELD_CHO = 0,            /* float */
  EMT6RO_FIELD_OX = 1,             /* float */
  EMT6RO_FIELD_GI = 2,             /* float */
  EMT6RO_FIELD_OCCUPANCY = 3,      /* uint8: 0 vacant, 1 occupied, 2 mocked */
  EMT6RO_FIELD_TIME_IN_REPAIR = 4, /* float */
  EMT6RO_FIELD_IRRADIATION = 5,    /* float */
  EMT6RO_FIELD_REPAIR_DELAY = 6,   /* float */
  EMT6RO_FIELD_PROLIF_TIME = 7,    /* float */
  EMT6RO_FIELD_CYCLE_G1 = 8,       /* float, cumulative phase ends ... */
  EMT6RO_FIELD_CYCLE_S = 9,
  EMT6RO_FIELD_CYCLE_G2 = 10,
  EMT6RO_FIELD_CYCLE_M = 11,
  EMT6RO_FIELD_CYCLE_D = 12,
  EMT6RO_FIELD_MODE = 13,          /* uint8 metabolic mode */
  EMT6RO_FIELD_PHASE = 14,         /* uint8 cycle phase */
  EMT6RO_NUM_FIELDS = 15
};

/* loadFromFile (state/state.cc:9-53); Python: load_state (backend.cc:219). */
int emt6ro_state_load(const char *path, const emt6ro_params *p, emt6ro_state **out);
/* saveToFile (state/state.cc:55-82). */
int emt6ro_state_save(const emt6ro_state *s, const char *path);
/* An all-vacant state of bordered dims H x W with the border MOCKED and every
 * site at the external levels of `p`; fields are then set with
 * emt6ro_state_set_field. */
int emt6ro_state_create(int32_t H, int32_t W, const emt6ro_params *p, emt6ro_state **out);
int emt6ro_state_dims(const emt6ro_state *s, int32_t *H, int32_t *W);
int emt6ro_state_get_field(const emt6ro_state *s, int field, void *dst);
int emt6ro_state_set_field(emt6ro_state *s, int field, const void *src);
void emt6ro_state_free(emt6ro_state *s);

/* ---- Experiment (python/emt6ro/simulation/backend.cc:16-136) ------------ */

/* _Experiment(params, tumors, tests_num, protocols_num, sim_steps,
 * prot_resolution, device_id) (backend.cc:38-76).  Simulation b = p*T*R +
 * t*R + r runs tumour t under protocol p (replicate r).  `seed` keys the
 * counter-based random stream (the reference draws it from
 * std::random_device; pass any value, the same seed gives the same run). */
int emt6ro_experiment_create(const emt6ro_params *p, const emt6ro_state *const *tumors,
                             int32_t num_tumors, int32_t runs, int32_t num_protocols,
                             uint32_t sim_steps, uint32_t resolution, int32_t device,
                             uint64_t seed, emt6ro_experiment **out);
/* addIrradiations (backend.cc:91-101): protocol i's (step, dose) pairs are
 * steps[offsets[i] .. offsets[i+1]) and doses[...]; doses accumulate into
 * slot step / resolution.  `num_protocols` must equal the experiment's. */
int emt6ro_experiment_add_irradiations(emt6ro_experiment *e, int32_t num_protocols,
                                       const int32_t *offsets, const int32_t *steps,
                                       const float *doses);
/* run (backend.cc:103-113): starts nsteps steps and returns at once. */
int emt6ro_experiment_run(emt6ro_experiment *e, uint32_t nsteps);
/* results (backend.cc:115-121): waits for the run; dst holds P*T*R counts. */
int emt6ro_experiment_results(emt6ro_experiment *e, uint32_t *dst);
/* reset (backend.cc:84-89): initial states back, doses zeroed. */
int emt6ro_experiment_reset(emt6ro_experiment *e);
/* state (backend.cc:123-134): copy of simulation `sample`. */
int emt6ro_experiment_state(emt6ro_experiment *e, int32_t sample, emt6ro_state **out);
int emt6ro_experiment_batch_size(const emt6ro_experiment *e);
/* Kernel path used by later runs: 0 persistent (default), 1 per-step. */
int emt6ro_experiment_set_path(emt6ro_experiment *e, int32_t path);
void emt6ro_experiment_destroy(emt6ro_experiment *e);

/* ---- Batch: uneven simulations in one launch (no reference counterpart;
 * the reference runs one lattice size and one horizon per Simulation,
 * simulation.cu:123-137) ------------------------------------------------- */

/* Simulation i runs tumours[tumor_index[i]] under protocol
 * prot_index[i] (a row of n_prot x n_slots doses in `prot_doses`) up to
 * step horizon[i]; its random stream is sim_id[i].  Doses fire at steps
 * s < prot_length with s % prot_res == 0 (protocol.h:9-14). */
int emt6ro_batch_create(const emt6ro_params *p, const emt6ro_state *const *tumors,
                        int32_t num_tumors, int32_t num_sims, const int32_t *tumor_index,
                        const int32_t *prot_index, const uint32_t *horizon,
                        const uint32_t *sim_id, int32_t n_prot, int32_t n_slots,
                        const float *prot_doses, uint32_t prot_res, uint32_t prot_length,
                        int32_t device, uint64_t seed, emt6ro_batch **out);
/* Runs every simulation `nsteps` further (capped by its horizon) and waits.
 * path: 0 persistent kernel, 1 per-step kernels, 2 per-step kernels
 * replayed from a captured HIP graph.  variant: kernel shape from
 * emt6ro_variant_info, or -1 for the tuned / default choice.
 * steps_per_launch: persistent kernel steps per launch (0: all at once). */
int emt6ro_batch_run(emt6ro_batch *b, uint32_t nsteps, int32_t path, int32_t variant,
                     uint32_t steps_per_launch);
/* Living-cell counts (num_sims values). */
int emt6ro_batch_results(emt6ro_batch *b, uint32_t *dst);
/* Per-simulation counters: step, list length, divisions, status and
 * diffusion site updates, 5 values each (uint64). */
int emt6ro_batch_counters(emt6ro_batch *b, uint64_t *dst);
int emt6ro_batch_state(emt6ro_batch *b, int32_t sim, emt6ro_state **out);
/* Restores the initial states; `epoch` selects a fresh random stream. */
int emt6ro_batch_reset(emt6ro_batch *b, uint32_t epoch);
/* Milliseconds of the last run on the device (HIP events on its stream),
 * and the number of kernel launches it made. */
int emt6ro_batch_last_run(const emt6ro_batch *b, double *ms, int64_t *launches);
void emt6ro_batch_destroy(emt6ro_batch *b);

/* ---- Kernel variants and tuning ----------------------------------------- */

/* info[0..7] = id, diffusion scheme (0 banded in place, 1 ping-pong
 * site-parallel), workgroup size, band rows, items (or Jacobi sites) per
 * thread, final-substep sites per thread, ring slots per thread, LDS row
 * stride (fits lattices up to stride - 2 on a side).  Returns the number of
 * variants if idx < 0. */
int emt6ro_variant_info(int32_t idx, int32_t *info);
/* Variant the persistent path uses for simulations whose lattice needs LDS
 * row stride `rs` (36, 56 or 88); -1 restores the default choice. */
int emt6ro_batch_set_variant(emt6ro_batch *b, int32_t rs, int32_t variant);
/* Window pass of the persistent path, no reference counterpart (the
 * reference keeps every lattice whole, diffusion.cu:119-200): for lattices
 * wider than 34 sites, simulations whose ROI plus two rings fits a window of
 * `rows` x 36 padded sites (and, on lattices wider than 54 sites, an
 * automatic second window of 56 columns) run first in a kernel that keeps only that window
 * of the substrates in LDS, so more simulations fit a CU; the rest, and any
 * whose ROI outgrows the window, continue in the whole-lattice kernel.
 * Results do not depend on it.  rows: -1 automatic (the default), 0 off,
 * otherwise at least 8.  Only runs with variant -1 use it. */
int emt6ro_batch_set_window(emt6ro_batch *b, int32_t rows);
/* Replaces the batch's dose table (n_prot x n_slots, as at creation). */
int emt6ro_batch_set_doses(emt6ro_batch *b, const float *prot_doses);
/* Workgroups per CU the persistent kernel of `variant` reaches for lattices
 * up to hmax x wmax (0 if it does not fit). */
int emt6ro_variant_occupancy(int32_t variant, int32_t hmax, int32_t wmax);

/* ---- Fitness surrogate: MLP in -> hidden x3 (GELU) -> 1 on the matrix cores
 * (no reference counterpart: the reference's only surrogate is an external
 * pretrained Temporal Fusion Transformer, python_ga/example_run_model.py:
 * 24-184).  BF16 weights and activations, FP32 accumulation, FP32 master
 * weights and Adam.  Parameters are flat FP32 in the order W1 [hidden x in],
 * b1, W2, b2, W3, b3, w4 [hidden], b4 [1]. ----------------------------------- */
typedef struct emt6ro_mlp emt6ro_mlp;
const char *emt6ro_mlp_last_error(void);
/* hidden must be a multiple of 64. */
int emt6ro_mlp_create(int32_t in_features, int32_t hidden, uint64_t seed, int32_t device,
                      emt6ro_mlp **out);
int64_t emt6ro_mlp_num_params(const emt6ro_mlp *m);
int emt6ro_mlp_get_params(emt6ro_mlp *m, float *out);
int emt6ro_mlp_set_params(emt6ro_mlp *m, const float *in);
/* Training set: X [n x in_features] (row-major), targets y [n]. */
int emt6ro_mlp_set_data(emt6ro_mlp *m, const float *X, const float *y, int32_t n);
/* `steps` Adam steps on random minibatches of `batch` rows; losses[steps]
 * (may be NULL) receives each step's mean squared error. */
int emt6ro_mlp_train(emt6ro_mlp *m, int32_t steps, int32_t batch, float lr, uint64_t seed,
                     float *losses);
/* `steps` Adam steps of each of `count` distinct networks on one device,
 * network i drawing its minibatches with seeds[i]: step k of every network
 * is enqueued (on the network's own stream) before step k + 1 of any, so
 * the GPU overlaps the networks' kernels; with emt6ro_mlp_set_train_graph
 * (nets[0], g > 0), g steps of each network are captured once into a graph
 * of its own and the graphs are launched in turn.  Each network ends
 * exactly as emt6ro_mlp_train with its seed would leave it.  losses (count x steps,
 * may be NULL) receives each step's mean squared error.  No reference
 * counterpart (the surrogate has none); used to train an ensemble. */
int emt6ro_mlp_train_many(emt6ro_mlp *const *nets, int32_t count, int32_t steps, int32_t batch,
                          float lr, const uint64_t *seeds, float *losses);
/* How emt6ro_mlp_train enqueues its steps: steps_per_graph > 0 captures the
 * kernels of that many steps once per call into a HIP graph and replays it,
 * 0 (the default) launches every kernel directly.  Results are identical.
 * No reference counterpart (the surrogate has none). */
int emt6ro_mlp_set_train_graph(emt6ro_mlp *m, int32_t steps_per_graph);
int emt6ro_mlp_predict(emt6ro_mlp *m, const float *X, int32_t n, float *y);
/* emt6ro_mlp_predict for callers that cannot allocate device memory (code
 * under HIP graph capture, tight screening loops, harnesses that forbid
 * hipMalloc); no counterpart in the reference.  X_dev [n x in_features]
 * and y_dev [n] are device memory of the network's device; it calls no
 * allocator: activations live in a fixed 12 MiB device array per GPU (three
 * bf16 buffers of 4096 rows x 512 columns), so in_features and hidden must be
 * at most 512 (larger networks are rejected with an error); n is any size
 * (4096 rows per pass).  The array is shared by every call on that GPU:
 * calls that overlap on one GPU race.  Waits for the result, like
 * emt6ro_mlp_predict, whose output it matches bit for bit. */
int emt6ro_mlp_predict_noalloc(emt6ro_mlp *m, const float *X_dev, int32_t n, float *y_dev);
/* Loss (mean squared error / 2) and its gradient for one batch, no update. */
int emt6ro_mlp_grads(emt6ro_mlp *m, const float *X, const float *y, int32_t n, float *grads,
                     float *loss);
void emt6ro_mlp_destroy(emt6ro_mlp *m);

/* ---- Instrumentation (common/instr.h:33-58; backend.cc:221-231) -------- */

/* counters = repair decisions, repair kills, irradiation events, divisions
 * (InstrCounters, common/instr.h:46-52); histograms: 300 bins of the
 * irradiation then 300 of the time in repair at each repair decision.
 * Summed over the devices used; integer totals (the death-probability sum is
 * kept in 32.32 fixed point), so they do not depend on thread timing.
 * Counters read zero while instrumentation is off, as the reference's do in
 * its default build. */
int emt6ro_instr_read(uint64_t counters[4], double *sum_death_prob);
int emt6ro_instr_reset(void);
int emt6ro_instr_histograms(uint64_t *dst, int32_t n);
/* The reference compiles the counters in with -DEMT6RO_INSTRUMENT
 * (common/instr.h:18-21); here runs started while instrumentation is enabled
 * count.  Default: on when the environment variable EMT6RO_INSTRUMENT is set
 * to anything but "" or "0". */
int emt6ro_instr_enable(int32_t on);
int emt6ro_instr_enabled(void);

/* ---- Measurement (no reference counterpart) ------------------------------ */

/* LDS read bandwidth of `device`, the peak against which the simulator's LDS
 * traffic is reported: every lane of 16 workgroups per CU reads 32 KB of
 * LDS in lane-contiguous (conflict-free) order with loads of width_bytes (4
 * or 16).  out[0] bytes/s over the GPU, out[1] bytes per shader clock per CU,
 * out[2] the shader clock (Hz) under the probe, out[3] the CU count. */
int emt6ro_probe_lds(int32_t device, int32_t width_bytes, double out[4]);
/* What a barrier across every workgroup of a kernel costs against a kernel
 * boundary: one workgroup per CU (cooperative launch) runs 200 barriers,
 * each workgroup writing `words` (0..16384) words of its own between them
 * and checking another's after; mode 0 one arrival counter, mode 1 eight
 * group counters and group release words, both with agent-scope fences.
 * out[0] microseconds per barrier, out[1] microseconds per launch of a
 * kernel doing the same writes back to back.  Fails if a wait timed out (1
 * s) or stale data was read. */
int emt6ro_probe_grid_barrier(int32_t device, int32_t mode, int32_t words, double out[2]);

#ifdef __cplusplus
}
#endif

#endif /* EMT6RO_AMD_H_ */
