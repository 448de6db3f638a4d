/*
 * EMT6/Ro CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference simulator's algorithm, used to check
 * the HIP product path and as the timed CPU baseline.  Nothing in the product
 * (the emt6ro package, its C-ABI library or its kernels) may call into this
 * code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it, and only as the checker.
 *
 * Contents and what pins each part:
 *   - eo_philox4x32_10      Philox-4x32-10 (Salmon et al., SC'11); pinned by the
 *                           published Random123 known-answer vectors.
 *   - eo_expf / eo_logf /   FP32 elementary functions after the Cephes library
 *     eo_sincos2pi            (S. Moshier); pinned against libm in double.
 *   - cell rules            restatement of cell/cell.h:102-136, cell/cell.cu:9-100,
 *                           site/site.h:62-82; pinned by the reference's gtest
 *                           tables (cell-*-test.cc) and closed forms.
 *   - eo_find_roi           restatement of diffusion/diffusion.cu:16-100 with the
 *                           reference's float sqrt/ceil; pinned by brute force.
 *   - eo_diffuse_f64        FP64 restatement of diffusion/diffusion.cu:119-191
 *                           (the reference formula order, in double); pinned by
 *                           fixed-point, symmetry and brute-force checks.
 *   - eo_diffuse_f32 and    the FP32 simulator with the build's deterministic
 *     eo_run                 step rules (division arbitration by key, child
 *                           vacancy after all deaths, seeded counter-based
 *                           draws; stated beside each function below); the
 *                           FP32 diffusion is pinned to eo_diffuse_f64 within
 *                           1e-5, the bookkeeping by the tests listed above.
 */
#ifndef EMT6RO_ORACLE_H_
#define EMT6RO_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flat parameter block; field meaning follows parameters/parameters.h:10-72. */
typedef struct {
  int32_t H, W;               /* bordered lattice dims (interior + 2) */
  float time_step;            /* seconds per simulation step */
  float diff_coeffs[3];       /* cho, ox, gi diffusion coefficients */
  float diff_time_step;       /* seconds per diffusion substep */
  float metab[4][3];          /* per metabolic mode: cho, ox, gi consumed per step */
  float ext[3];               /* external substrate levels */
  float death_gi, quiescence_gi, death_gi_production;
  float cycle_mean[5], cycle_sd[5];
  float delay_coeff, delay_exp, surv_coeff, surv_exp, repair_half_time;
} eo_params;

/* 48-byte cell record; same field meaning as cell/cell.h:46-52. */
typedef struct {
  float t_rep, R, delay, prolif;
  float times[5];
  uint8_t mode, phase;
  uint16_t pad0;
  uint32_t pad1[2];
} eo_cell;

enum { EO_VACANT = 0, EO_OCCUPIED = 1, EO_MOCKED = 2 };
enum { EO_PURPOSE_REPAIR = 1, EO_PURPOSE_NEIGHBOUR = 2, EO_PURPOSE_NORMAL = 3 };

/* ---- RNG and elementary functions ---- */
void eo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float eo_u01(uint32_t x);
float eo_expf(float x);
float eo_logf(float x);
void eo_sincos2pi(float u, float *s, float *c);
void eo_normals5(uint32_t seed, uint32_t epoch, uint32_t sim, uint32_t step,
                 uint32_t site, float z[5]);
float eo_uniform(uint32_t seed, uint32_t epoch, uint32_t sim, uint32_t step,
                 uint32_t site, uint32_t purpose);

/* ---- cell rules (host restatement) ---- */
int eo_update_state(eo_cell *c, const float sub[3], const eo_params *p, int vacant);
int eo_progress_clock(eo_cell *c, float time_step);
void eo_metabolise(const eo_cell *c, float sub[3], const eo_params *p);
void eo_irradiate(eo_cell *c, float dose, const eo_params *p);
float eo_delay_time(float R, const eo_params *p);
int eo_try_repair(eo_cell *c, int cycle_changed, const eo_params *p, float u);
int eo_choose_neighbour(int r, int c, float u, int *cr, int *cc);
uint32_t eo_arbitration_key(uint32_t i);

/* ---- region of interest (reference float formulation) ---- */
/* roi[5] = {r0, c0, h, w, dist}; returns 0 when no site is occupied. */
int eo_find_roi(int H, int W, const uint8_t *st, int32_t roi[5]);
int eo_mask(int hm, int wm, int dist, int r, int c);

/* ---- diffusion ---- */
/* FP64 restatement of the reference formula; sub_in/out: 3 planes of H*W. */
void eo_diffuse_f64(const eo_params *p, int H, int W, const double *sub_in,
                    const int32_t roi[5], double *sub_out);
/* FP32 diffusion with the build's fixed operation order. */
void eo_diffuse_f32(const eo_params *p, int H, int W, float *sub, const int32_t roi[5]);

/* ---- whole simulator (FP32 port) ----
 * Arrays are laid out per simulation, B simulations back to back:
 *   sub   : B * 3 * H*W floats (planes cho, ox, gi)
 *   st    : B * H*W bytes
 *   cells : B * H*W records
 *   list  : B * cap uint16 site indices, nlist: B
 *   roi   : B * 5
 *   step  : B (current step counter of each simulation)
 *   doses : per simulation protocol: prot_index[b] selects a row of
 *           prot_data (n_slots floats); a dose fires at step s when
 *           s < prot_length && s % prot_res == 0.
 * Returns 0 on success, -1 if a list overflowed its capacity.
 */
int eo_run(const eo_params *p, int B, float *sub, uint8_t *st, eo_cell *cells,
           uint16_t *list, int32_t *nlist, int32_t cap, int32_t *roi, uint32_t *step,
           const int32_t *sim_id, const int32_t *prot_index, const float *prot_data,
           int32_t n_slots, uint32_t prot_res, uint32_t prot_length,
           uint32_t seed, uint32_t epoch, uint32_t nsteps, int nthreads);

/* eo_run with the instrumentation counters of common/instr.h:5-21 per
 * simulation: instr (may be NULL) holds B records of EO_INSTR_N uint64:
 * repair decisions, repair kills, irradiation events, divisions, the sum of
 * death probabilities in 32.32 fixed point, then the 300-bin histograms of
 * irradiation and of time in repair (0.1 per bin) at each repair decision. */
enum {
  EO_INSTR_REPAIR_DECISIONS = 0,
  EO_INSTR_REPAIR_KILLS = 1,
  EO_INSTR_IRRADIATION_EVENTS = 2,
  EO_INSTR_DIVISIONS = 3,
  EO_INSTR_DEATH_PROB_FX = 4,
  EO_INSTR_HIST_N = 300,
  EO_INSTR_IRR_HIST = 5,
  EO_INSTR_TIR_HIST = 5 + 300,
  EO_INSTR_N = 5 + 600
};
int eo_run_instr(const eo_params *p, int B, float *sub, uint8_t *st, eo_cell *cells,
                 uint16_t *list, int32_t *nlist, int32_t cap, int32_t *roi, uint32_t *step,
                 const int32_t *sim_id, const int32_t *prot_index, const float *prot_data,
                 int32_t n_slots, uint32_t prot_res, uint32_t prot_length,
                 uint32_t seed, uint32_t epoch, uint32_t nsteps, int nthreads,
                 uint64_t *instr);

/* eo_run_instr with two of the deterministic step rules switched to the
 * reference's literal behaviour (bit mask; 0 = the rules the product
 * implements).  For A/B statistics of the rule choice only. */
enum { EO_RULE_Q3_EVERY_ENTRY = 1, EO_RULE_Q4_LIST_ORDER = 2 };
int eo_run_rules(const eo_params *p, int B, float *sub, uint8_t *st, eo_cell *cells,
                 uint16_t *list, int32_t *nlist, int32_t cap, int32_t *roi, uint32_t *step,
                 const int32_t *sim_id, const int32_t *prot_index, const float *prot_data,
                 int32_t n_slots, uint32_t prot_res, uint32_t prot_length,
                 uint32_t seed, uint32_t epoch, uint32_t nsteps, int nthreads,
                 uint64_t *instr, int rules);

uint32_t eo_count_living(int H, int W, const uint8_t *st);

#ifdef __cplusplus
}
#endif
#endif  /* EMT6RO_ORACLE_H_ */
