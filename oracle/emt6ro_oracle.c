/*
 * EMT6/Ro CPU oracle -- TEST INFRASTRUCTURE ONLY (see emt6ro_oracle.h).
 *
 * Built with -ffp-contract=off: every float operation below is rounded on its
 * own, in the order written; fmaf() appears only where the product's FP32
 * stencil uses a fused multiply-add.  Reference file:line citations are
 * relative to the reference repository root.
 */
#include "emt6ro_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Philox-4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11).  Round function and Weyl key schedule as         */
/* published; constants are the paper's multipliers and golden-ratio bumps.  */
/* ------------------------------------------------------------------------ */
#define EO_PH_M0 0xD2511F53u
#define EO_PH_M1 0xCD9E8D57u
#define EO_PH_W0 0x9E3779B9u
#define EO_PH_W1 0xBB67AE85u

void eo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    uint64_t p0 = (uint64_t)EO_PH_M0 * c0;
    uint64_t p1 = (uint64_t)EO_PH_M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += EO_PH_W0;
    k1 += EO_PH_W1;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Uniform on (0, 1]: the top 24 bits plus one, times 2^-24 (exact). The
 * reference's generator also yields (0, 1] (random-engine.cu:34). */
float eo_u01(uint32_t x) {
  return (float)((x >> 8) + 1u) * 5.9604644775390625e-8f;
}

static float eo_bits_float(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static uint32_t eo_float_bits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

/* expf after Cephes (S. L. Moshier, Cephes Math Library, expf.c): Cody-Waite
 * reduction by ln2 split into 0.693359375 - 2.12194440e-4, degree-5
 * polynomial, scale by 2^n.  Arguments here stay well inside (-87, 88). */
float eo_expf(float x) {
  if (x > 88.0f) x = 88.0f;
  if (x < -87.0f) x = -87.0f;
  float z = floorf(1.44269504088896341f * x + 0.5f);
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
  return p * eo_bits_float((uint32_t)(n + 127) << 23);
}

/* logf after Cephes (logf.c): split x = m 2^e with m in [sqrt(1/2), sqrt(2)),
 * degree-9 polynomial in m-1, recombine with ln2 split as in eo_expf.
 * Valid for positive normal x (the RNG's uniforms are >= 2^-24). */
float eo_logf(float x) {
  uint32_t b = eo_float_bits(x);
  int e = (int)((b >> 23) & 0xffu) - 126;
  float m = eo_bits_float((b & 0x807fffffu) | (126u << 23));  /* [0.5, 1) */
  if (m < 0.707106781186547524f) {
    e -= 1;
    m = m + m;
    m = m - 1.0f;
  } else {
    m = m - 1.0f;
  }
  float z = m * m;
  float y = 7.0376836292e-2f;
  y = y * m - 1.1514610310e-1f;
  y = y * m + 1.1676998740e-1f;
  y = y * m - 1.2420140846e-1f;
  y = y * m + 1.4249322787e-1f;
  y = y * m - 1.6668057665e-1f;
  y = y * m + 2.0000714765e-1f;
  y = y * m - 2.4999993993e-1f;
  y = y * m + 3.3333331174e-1f;
  y = y * m;
  y = y * z;
  float fe = (float)e;
  y = y + -2.12194440e-4f * fe;
  y = y + -0.5f * z;
  float res = m + y;
  res = res + 0.693359375f * fe;
  return res;
}

/* sin and cos of 2*pi*u for u in (0, 1]: exact quadrant reduction on u, then
 * the Cephes sinf/cosf polynomials on [-pi/4, pi/4]. */
void eo_sincos2pi(float u, float *s_out, float *c_out) {
  float q = floorf(u * 4.0f + 0.5f);        /* exact: u has <= 24 bits */
  float r = u - q * 0.25f;                  /* exact, |r| <= 1/8 */
  int quad = ((int)q) & 3;
  float x = r * 6.28318530717958648f;
  float z = x * x;
  float s = -1.9515295891e-4f;
  s = s * z + 8.3321608736e-3f;
  s = s * z - 1.6666654611e-1f;
  s = s * z;
  s = s * x;
  s = s + x;
  float c = 2.443315711809948e-5f;
  c = c * z - 1.388731625493765e-3f;
  c = c * z + 4.166664568298827e-2f;
  c = c * z;
  c = c * z;
  c = c - 0.5f * z;
  c = c + 1.0f;
  switch (quad) {
    case 0: *s_out = s;  *c_out = c;  break;
    case 1: *s_out = c;  *c_out = -s; break;
    case 2: *s_out = -s; *c_out = -c; break;
    default: *s_out = -c; *c_out = s; break;
  }
}

static void eo_draw(uint32_t seed, uint32_t epoch, uint32_t sim, uint32_t step,
                    uint32_t site, uint32_t purpose, uint32_t index, uint32_t out[4]) {
  uint32_t ctr[4] = {step, site | (purpose << 24), sim, index};
  uint32_t key[2] = {seed, epoch};
  eo_philox4x32_10(ctr, key, out);
}

float eo_uniform(uint32_t seed, uint32_t epoch, uint32_t sim, uint32_t step,
                 uint32_t site, uint32_t purpose) {
  uint32_t o[4];
  eo_draw(seed, epoch, sim, step, site, purpose, 0, o);
  return eo_u01(o[0]);
}

/* Five standard normals by the Box-Muller transform (Box & Muller, 1958)
 * from six uniforms (two Philox blocks). */
void eo_normals5(uint32_t seed, uint32_t epoch, uint32_t sim, uint32_t step,
                 uint32_t site, float z[5]) {
  uint32_t a[4], b[4];
  eo_draw(seed, epoch, sim, step, site, EO_PURPOSE_NORMAL, 0, a);
  eo_draw(seed, epoch, sim, step, site, EO_PURPOSE_NORMAL, 1, b);
  float u[6] = {eo_u01(a[0]), eo_u01(a[1]), eo_u01(a[2]), eo_u01(a[3]),
                eo_u01(b[0]), eo_u01(b[1])};
  float zz[6];
  for (int k = 0; k < 3; ++k) {
    float rad = sqrtf(-2.0f * eo_logf(u[2 * k]));
    float s, c;
    eo_sincos2pi(u[2 * k + 1], &s, &c);
    zz[2 * k] = rad * c;
    zz[2 * k + 1] = rad * s;
  }
  for (int k = 0; k < 5; ++k) z[k] = zz[k];
}

/* ------------------------------------------------------------------------ */
/* Cell rules                                                               */
/* ------------------------------------------------------------------------ */

/* cell/cell.cu:33-82 (updateState, enterG1SStopping, tryProliferating,
 * tryQuiescence).  Returns 1 if the cell stays alive; sets mode. */
int eo_update_state(eo_cell *c, const float sub[3], const eo_params *p, int vacant) {
  if (c->prolif >= c->times[4] - p->time_step / 3600.f) return 0;
  int stopping = c->prolif > (c->times[0] - 2.f * p->time_step / 3600.f) &&
                 c->phase == 0 && vacant <= 1;
  if (!stopping) {
    if (sub[0] >= p->metab[0][0] && sub[1] >= p->metab[0][1] && sub[2] < p->quiescence_gi) {
      c->mode = 0;
      return 1;
    }
    if (sub[0] >= p->metab[1][0] && sub[1] >= p->metab[1][1] && sub[2] < p->quiescence_gi) {
      c->mode = 1;
      return 1;
    }
  }
  if (c->phase != 0 && c->phase != 2) return 0;
  if (sub[0] >= p->metab[2][0] && sub[1] >= p->metab[2][1] && sub[2] < p->death_gi) {
    c->mode = 2;
    return 1;
  }
  if (sub[0] >= p->metab[3][0] && sub[1] >= p->metab[3][1] && sub[2] < p->death_gi) {
    c->mode = 3;
    return 1;
  }
  return 0;
}

/* cell/cell.cu:9-12,21-32.  Phase index is clamped at D (4) for the lookup;
 * the reference never reads past D because a cell beyond D dies in the next
 * updateState before its clock moves again. */
int eo_progress_clock(eo_cell *c, float ts) {
  if ((c->mode == 0 || c->mode == 1) && c->t_rep == 0) {
    c->prolif = c->prolif + ts / 3600.f;
    uint8_t cur = c->phase;
    int idx = cur > 4 ? 4 : cur;
    c->phase = (uint8_t)(cur + (c->prolif > c->times[idx] ? 1 : 0));
    return c->phase != cur && c->phase != 4;
  }
  return 0;
}

/* cell/cell.cu:16-19 */
void eo_metabolise(const eo_cell *c, float sub[3], const eo_params *p) {
  sub[0] = sub[0] - p->metab[c->mode][0];
  sub[1] = sub[1] - p->metab[c->mode][1];
  sub[2] = sub[2] - p->metab[c->mode][2];
}

/* cell/cell.cu:93-100 (calcDelayTime) */
float eo_delay_time(float R, const eo_params *p) {
  return p->delay_coeff * eo_expf(p->delay_exp * R);
}

/* cell/cell.cu:84-91 (irradiate) */
void eo_irradiate(eo_cell *c, float dose, const eo_params *p) {
  c->R = c->R / (1.f + c->t_rep / p->repair_half_time) + dose;
  c->t_rep = 0.f;
  c->delay = eo_delay_time(c->R, p);
}

/* Histogram bin of the instrumentation (common/instr.h:42-44, cell.h:115-120):
 * value / 0.1f truncated, clamped to [0, 299]. */
static int eo_instr_bin(float v) {
  int b = (int)(v / 0.1f);
  if (b < 0) b = 0;
  if (b >= EO_INSTR_HIST_N) b = EO_INSTR_HIST_N - 1;
  return b;
}

/* cell/cell.h:102-136 with the uniform supplied by the caller; `instr` (may
 * be NULL) receives the counters of cell.h:111-126. */
static int eo_try_repair_instr(eo_cell *c, int cycle_changed, const eo_params *p, float u,
                               uint64_t *instr) {
  if (c->t_rep > 0 || (c->R > 0 && cycle_changed)) {
    c->t_rep = c->t_rep + p->time_step / 3600.f;
    if (c->t_rep >= c->delay) {
      float death_prob = 1.f - p->surv_coeff * eo_expf(p->surv_exp * c->R);
      if (instr) {
        instr[EO_INSTR_REPAIR_DECISIONS] += 1;
        instr[EO_INSTR_DEATH_PROB_FX] += (uint64_t)((double)death_prob * 4294967296.0);
        instr[EO_INSTR_IRR_HIST + eo_instr_bin(c->R)] += 1;
        instr[EO_INSTR_TIR_HIST + eo_instr_bin(c->t_rep)] += 1;
      }
      if (u < death_prob) {
        if (instr) instr[EO_INSTR_REPAIR_KILLS] += 1;
        return 0;
      }
      c->R = 0.f;
      c->t_rep = 0.f;
      c->delay = 0.f;
    }
  }
  return 1;
}

int eo_try_repair(eo_cell *c, int cycle_changed, const eo_params *p, float u) {
  return eo_try_repair_instr(c, cycle_changed, p, u, NULL);
}

/* simulation/cell-division.h:44-57 (chooseNeighbour): diagonal neighbours
 * get weight 1 each, orthogonal ones sqrt(2) each.  Returns the direction
 * number (0-3 diagonal, 4-7 orthogonal) and the child coordinates. */
int eo_choose_neighbour(int r, int c, float u, int *cr, int *cc) {
  const float diag = 4.f;
  const float ortho = (float)(4. * 1.41421356237309504880);
  float score = u * (diag + ortho);
  if (score < diag) {
    unsigned num = (unsigned)(uint8_t)score;
    int vert = (int)(num & 1u) * 2 - 1;
    int hor = (int)((num & 2u) >> 1) * 2 - 1;
    *cr = r + vert;
    *cc = c + hor;
    return (int)num;
  }
  unsigned num = (unsigned)(uint8_t)((double)(score - diag) / 1.41421356237309504880);
  int which = (int)(num & 1u);
  int diff = (int)((num & 2u) >> 1) * 2 - 1;
  if (which) { *cr = r + diff; *cc = c; } else { *cr = r; *cc = c + diff; }
  return 4 + (int)num;
}

/* ------------------------------------------------------------------------ */
/* Region of interest and mask (diffusion/diffusion.cu:16-100), in the      */
/* reference's float formulation.                                           */
/* ------------------------------------------------------------------------ */
int eo_find_roi(int H, int W, const uint8_t *st, int32_t roi[5]) {
  int minr = H, minc = W, maxr = 0, maxc = 0;
  for (int r = 0; r < H; ++r)
    for (int c = 0; c < W; ++c)
      if (st[r * W + c] == EO_OCCUPIED) {
        if (r < minr) minr = r;
        if (c < minc) minc = c;
        if (r > maxr) maxr = r;
        if (c > maxc) maxc = c;
      }
  if (maxr == 0 && maxc == 0) {
    roi[0] = roi[1] = roi[2] = roi[3] = roi[4] = 0;
    return 0;
  }
  int midr = minr + (maxr - minr + 1) / 2;
  int midc = minc + (maxc - minc + 1) / 2;
  int dist = 0;
  for (int r = minr; r <= maxr; ++r)
    for (int c = minc; c <= maxc; ++c)
      if (st[r * W + c] == EO_OCCUPIED) {
        float d2 = (float)((midr - r) * (midr - r) + (midc - c) * (midc - c));
        int d = (int)ceilf(sqrtf(d2));
        if (d > dist) dist = d;
      }
  int sr = (midr <= dist) ? 1 : midr - dist;
  int sc = (midc <= dist) ? 1 : midc - dist;
  int sw = (midc + dist >= W - 1) ? W - 1 - sc : midc + dist - sc + 1;
  int sh = (midr + dist >= H - 1) ? H - 1 - sr : midr + dist - sr + 1;
  roi[0] = sr; roi[1] = sc; roi[2] = sh; roi[3] = sw; roi[4] = dist;
  return 1;
}

/* fillBorderMask (diffusion.cu:16-30): mask dims (hm, wm) = ROI + 2. */
int eo_mask(int hm, int wm, int dist, int r, int c) {
  float midr = hm / 2.f, midc = wm / 2.f;
  float fr = (float)(r + 1), fc = (float)(c + 1);
  float d = sqrtf((fr - midr) * (fr - midr) + (fc - midc) * (fc - midc));
  return d >= (float)dist;
}

/* ------------------------------------------------------------------------ */
/* Diffusion                                                                */
/* ------------------------------------------------------------------------ */
static int eo_substeps(const eo_params *p) {
  return (int)(int16_t)(p->time_step / p->diff_time_step);
}

/* FP64 restatement of diffusionKernel (diffusion.cu:119-191): the temporary
 * grid is the ROI plus a two-site ring, the outer ring is zero, masked sites
 * hold the external levels, S-1 Jacobi substeps update the unmasked sites
 * only, a final substep updates every ROI site, and the ROI is written back.
 * The stencil keeps the reference's summation order, evaluated in double. */
void eo_diffuse_f64(const eo_params *p, int H, int W, const double *sub_in,
                    const int32_t roi[5], double *sub_out) {
  const int HW = H * W;
  memcpy(sub_out, sub_in, sizeof(double) * 3 * HW);
  const int r0 = roi[0], c0 = roi[1], h = roi[2], w = roi[3], dist = roi[4];
  if (h == 0 || w == 0) return;
  const int th = h + 4, tw = w + 4, tv = th * tw;
  double *tmp = (double *)malloc(sizeof(double) * 3 * tv);
  double *nxt = (double *)malloc(sizeof(double) * 3 * tv);
  uint8_t *act = (uint8_t *)malloc(tv);
  for (int dr = 0; dr < th; ++dr)
    for (int dc = 0; dc < tw; ++dc) {
      int i = dr * tw + dc;
      act[i] = 0;
      for (int s = 0; s < 3; ++s) {
        if (dr == 0 || dr == th - 1 || dc == 0 || dc == tw - 1) {
          tmp[s * tv + i] = 0.0;
        } else if (eo_mask(h + 2, w + 2, dist, dr - 1, dc - 1)) {
          tmp[s * tv + i] = (double)p->ext[s];
        } else {
          int r = r0 - 2 + dr, c = c0 - 2 + dc;
          tmp[s * tv + i] = sub_in[s * HW + r * W + c];
          act[i] = 1;
        }
      }
    }
  const double hs = 2.0 * 1.41421356237309504880;
  const double f = 4.0 + hs;
  double k[3];
  for (int s = 0; s < 3; ++s)
    k[s] = (double)p->diff_coeffs[s] * (double)p->diff_time_step * hs / f;
  /* S-1 Jacobi substeps then one final substep; with S <= 1 only the final
   * substep runs (diffusion.cu:165-184). */
  const int S = eo_substeps(p);
  const int total = S > 1 ? S : 1;
  for (int step = 0; step < total; ++step) {
    int final_sub = (step == total - 1);
    for (int dr = 1; dr < th - 1; ++dr)
      for (int dc = 1; dc < tw - 1; ++dc) {
        int i = dr * tw + dc;
        int upd = final_sub ? (dr >= 2 && dr <= h + 1 && dc >= 2 && dc <= w + 1) : act[i];
        for (int s = 0; s < 3; ++s) {
          const double *t = tmp + s * tv;
          if (!upd) { nxt[s * tv + i] = t[i]; continue; }
          double o = t[i - tw] + t[i - 1] + t[i + tw] + t[i + 1];
          double d = t[i - tw - 1] + t[i - tw + 1] + t[i + tw - 1] + t[i + tw + 1];
          double res = o + d * 0.70710678118654752440;
          res -= t[i] * f;
          res *= k[s];
          nxt[s * tv + i] = t[i] + res;
        }
      }
    for (int dr = 1; dr < th - 1; ++dr)
      for (int dc = 1; dc < tw - 1; ++dc)
        for (int s = 0; s < 3; ++s) tmp[s * tv + dr * tw + dc] = nxt[s * tv + dr * tw + dc];
  }
  for (int dr = 2; dr <= h + 1; ++dr)
    for (int dc = 2; dc <= w + 1; ++dc)
      for (int s = 0; s < 3; ++s)
        sub_out[s * HW + (r0 - 2 + dr) * W + (c0 - 2 + dc)] = tmp[s * tv + dr * tw + dc];
  free(tmp);
  free(nxt);
  free(act);
}

/* FP32 stencil in the build's fixed order, with H(r) = left(r) + right(r)
 * the sum of a row's two horizontal neighbours of the centre column and
 * G(r) = fma(H(r), 1/sqrt2, C(r)) its centre plus its diagonal share:
 *   L = (G(up row) + G(down row)) + H(mid);  L = fma(-F, C, L);
 *   C' = fma(k, L, C)
 * (a column walking down the rows computes each H and G once; 7 operations
 * per site instead of the 8 of (up + down) + H(mid) + (H(up) + H(down))/sqrt2)
 * with F = (float)(4 + (float)(2 sqrt2)) and k = ((D dt) HS) / F in float, the
 * coefficient expression of diffusion.cu:163. */
static float eo_stencil_f32(const float *t, int i, int tw, float k) {
  const float R2 = (float)0.70710678118654752440;
  const float HS = (float)(2.0 * 1.41421356237309504880);
  const float F = (float)(4.0 + (double)HS);
  float hu = t[i - tw - 1] + t[i - tw + 1];
  float hm = t[i - 1] + t[i + 1];
  float hd = t[i + tw - 1] + t[i + tw + 1];
  float gu = fmaf(hu, R2, t[i - tw]);
  float gd = fmaf(hd, R2, t[i + tw]);
  float L = (gu + gd) + hm;
  L = fmaf(-F, t[i], L);
  return fmaf(k, L, t[i]);
}

void eo_diffuse_f32(const eo_params *p, int H, int W, float *sub, const int32_t roi[5]) {
  const int HW = H * W;
  const int r0 = roi[0], c0 = roi[1], h = roi[2], w = roi[3], dist = roi[4];
  if (h == 0 || w == 0) return;
  const int th = h + 4, tw = w + 4, tv = th * tw;
  float *tmp = (float *)malloc(sizeof(float) * 3 * tv);
  float *nxt = (float *)malloc(sizeof(float) * 3 * tv);
  uint8_t *act = (uint8_t *)malloc(tv);
  for (int dr = 0; dr < th; ++dr)
    for (int dc = 0; dc < tw; ++dc) {
      int i = dr * tw + dc;
      act[i] = 0;
      for (int s = 0; s < 3; ++s) {
        if (dr == 0 || dr == th - 1 || dc == 0 || dc == tw - 1) {
          tmp[s * tv + i] = 0.f;
        } else if (eo_mask(h + 2, w + 2, dist, dr - 1, dc - 1)) {
          tmp[s * tv + i] = p->ext[s];
        } else {
          int r = r0 - 2 + dr, c = c0 - 2 + dc;
          tmp[s * tv + i] = sub[s * HW + r * W + c];
          act[i] = 1;
        }
      }
    }
  const float HS = (float)(2.0 * 1.41421356237309504880);
  const float F = (float)(4.0 + (double)HS);
  float k[3];
  for (int s = 0; s < 3; ++s) k[s] = ((p->diff_coeffs[s] * p->diff_time_step) * HS) / F;
  /* S-1 Jacobi substeps then one final substep; with S <= 1 only the final
   * substep runs (diffusion.cu:165-184). */
  const int S = eo_substeps(p);
  const int total = S > 1 ? S : 1;
  for (int step = 0; step < total; ++step) {
    int final_sub = (step == total - 1);
    for (int dr = 1; dr < th - 1; ++dr)
      for (int dc = 1; dc < tw - 1; ++dc) {
        int i = dr * tw + dc;
        int upd = final_sub ? (dr >= 2 && dr <= h + 1 && dc >= 2 && dc <= w + 1) : act[i];
        for (int s = 0; s < 3; ++s)
          nxt[s * tv + i] = upd ? eo_stencil_f32(tmp + s * tv, i, tw, k[s]) : tmp[s * tv + i];
      }
    for (int dr = 1; dr < th - 1; ++dr)
      for (int dc = 1; dc < tw - 1; ++dc)
        for (int s = 0; s < 3; ++s) tmp[s * tv + dr * tw + dc] = nxt[s * tv + dr * tw + dc];
  }
  for (int dr = 2; dr <= h + 1; ++dr)
    for (int dc = 2; dc <= w + 1; ++dc)
      for (int s = 0; s < 3; ++s)
        sub[s * HW + (r0 - 2 + dr) * W + (c0 - 2 + dc)] = tmp[s * tv + dr * tw + dc];
  free(tmp);
  free(nxt);
  free(act);
}

/* ------------------------------------------------------------------------ */
/* Whole simulator                                                          */
/* ------------------------------------------------------------------------ */
uint32_t eo_count_living(int H, int W, const uint8_t *st) {
  uint32_t n = 0;
  for (int i = 0; i < H * W; ++i) n += (st[i] == EO_OCCUPIED);
  return n;
}

/* findOccupied (simulation.cu:41-64): 32 scanners, scanner g takes columns
 * g, g+32, ... row by row; their lists are concatenated in scanner order. */
static int eo_rebuild_list(int H, int W, const uint8_t *st, uint16_t *list, int cap,
                           int32_t *owner) {
  int n = 0;
  for (int i = 0; i < H * W; ++i) owner[i] = -1;
  for (int g = 0; g < 32; ++g)
    for (int r = 0; r < H - 1; ++r)
      for (int c = g; c < W - 1; c += 32)
        if (st[r * W + c] == EO_OCCUPIED) {
          if (n >= cap) return -1;
          list[n] = (uint16_t)(r * W + c);
          owner[r * W + c] = n;
          ++n;
        }
  return n;
}

static int eo_vacant_neighbours(int W, const uint8_t *st, int x) {
  return (st[x - W - 1] == EO_VACANT) + (st[x - W + 1] == EO_VACANT) +
         (st[x + W - 1] == EO_VACANT) + (st[x + W + 1] == EO_VACANT) +
         (st[x - 1] == EO_VACANT) + (st[x + 1] == EO_VACANT) +
         (st[x - W] == EO_VACANT) + (st[x + W] == EO_VACANT);
}

static uint32_t eo_bitrev(uint32_t v, int bits) {
  uint32_t r = 0;
  for (int b = 0; b < bits; ++b) r |= ((v >> b) & 1u) << (bits - 1 - b);
  return r;
}

/* Division arbitration (simulation.cu:86-105, cuda-utils.h:23-49).  The
 * reference steps list entry i on thread i mod 512, where a later entry of the
 * thread overwrites its candidate, and reduces the per-thread candidates with
 * the first-non-zero operator a | (b * !a) through __shfl_down trees: first
 * within each 32-lane warp, then over the 16 warp results held by lanes 0-15
 * of warp 0.  A shuffle-down tree with offsets 16, 8, 4, 2, 1 ranks lane l by
 * bitrev5(l) (lane 0 first, then 16, 8, 24, 4, ...), so the winner is the
 * candidate whose thread has the smallest (bitrev4(warp), bitrev5(lane)),
 * ties (same thread) going to the largest i.  Smaller key wins. */
uint32_t eo_arbitration_key(uint32_t i) {
  uint32_t t = i & 511u;
  uint32_t prio = (eo_bitrev(t >> 5, 4) << 5) | eo_bitrev(t & 31u, 5);
  return (prio << 16) | (0xffffu - ((i >> 9) & 0xffffu));
}

/* One step of the cell automaton (simulation.cu:66-110, site.h:62-82) under
 * the deterministic rules: every occupied site is stepped once, with the
 * vacant-neighbour count from the state before the step; divisions look at
 * the state after every death of the step; at most one division per
 * lattice, won by the candidate whose list index i has the smallest i mod 512,
 * ties to the largest i (the reference's per-thread last-wins, lowest-thread
 * reduction); the child is appended to the list.
 *
 * `rules` switches two of these to the reference's literal behaviour, for
 * A/B statistics only (the product implements rules = 0):
 *   EO_RULE_Q3_EVERY_ENTRY  every list entry whose site is occupied steps it,
 *                           so a site re-occupied since the last rebuild is
 *                           stepped once per entry (simulation.cu:86-88 over
 *                           the pushed child, :105);
 *   EO_RULE_Q4_LIST_ORDER   a ready cell's target is tested for vacancy right
 *                           after its own step, in list order. */
static int eo_cells_step(const eo_params *p, int H, int W, float *sub, uint8_t *st,
                         eo_cell *cells, uint16_t *list, int32_t *pn, int cap,
                         int32_t *owner, uint8_t *flags, int32_t *seen, uint32_t t,
                         uint32_t sim, float dose, uint32_t seed, uint32_t epoch,
                         uint64_t *instr, int rules) {
  const int HW = H * W;
  const int n = *pn;
  const int q3 = (rules & EO_RULE_Q3_EVERY_ENTRY) != 0;
  const int q4 = (rules & EO_RULE_Q4_LIST_ORDER) != 0;
  int best = -1;
  uint32_t best_key = 0xffffffffu;
  int best_child = -1;
  for (int i = 0; i < n; ++i) {
    int x = list[i];
    int live = (st[x] == EO_OCCUPIED && (q3 || owner[x] == i));
    flags[i] = live ? (uint8_t)(0x10 | eo_vacant_neighbours(W, st, x)) : 0;
    seen[x] = 0;
  }
  for (int i = 0; i < n; ++i) {
    if (!(flags[i] & 0x10)) continue;
    int x = list[i];
    /* Q3 (every entry): a second entry of a site re-occupied since the last
     * rebuild steps the cell again, if it is still alive (site.h:65), with
     * draws of its own (Philox index = how many entries of x came before). */
    if (q3 && st[x] != EO_OCCUPIED) { flags[i] = 0; continue; }
    const uint32_t occ = (uint32_t)seen[x]++;
    eo_cell *c = &cells[x];
    float s3[3] = {sub[x], sub[HW + x], sub[2 * HW + x]};
    int ready = 0;
    int alive = eo_update_state(c, s3, p, flags[i] & 0x0f);
    if (alive) {
      if (dose > 0) {
        eo_irradiate(c, dose, p);
        if (instr) instr[EO_INSTR_IRRADIATION_EVENTS] += 1;
      }
      eo_metabolise(c, s3, p);
      int changed = eo_progress_clock(c, p->time_step);
      int need_u = (c->t_rep > 0 || (c->R > 0 && changed)) &&
                   (c->t_rep + p->time_step / 3600.f >= c->delay);
      float u = 1.f;
      if (need_u) {
        uint32_t o[4];
        eo_draw(seed, epoch, sim, t, (uint32_t)x, EO_PURPOSE_REPAIR, occ, o);
        u = eo_u01(o[0]);
      }
      alive = eo_try_repair_instr(c, changed, p, u, instr);
      if (alive) ready = (c->phase == 4);
    }
    sub[x] = s3[0];
    sub[HW + x] = s3[1];
    sub[2 * HW + x] = s3[2];
    if (!alive) st[x] = EO_VACANT;
    flags[i] = ready ? 0x20 : 0;
    if (ready && q4) {
      /* Q4 (list order): the target's vacancy is read right after this
       * entry's step, as the reference's thread reads it
       * (simulation.cu:88-93): deaths of earlier entries are seen, later
       * ones are not. */
      uint32_t o[4];
      eo_draw(seed, epoch, sim, t, (uint32_t)x, EO_PURPOSE_NEIGHBOUR, occ, o);
      int cr, cc;
      eo_choose_neighbour(x / W, x % W, eo_u01(o[0]), &cr, &cc);
      int child = cr * W + cc;
      uint32_t key = eo_arbitration_key((uint32_t)i);
      if (st[child] == EO_VACANT && key < best_key) {
        best_key = key; best = i; best_child = child;
      }
    }
  }
  for (int i = 0; !q4 && i < n; ++i) {
    if (!(flags[i] & 0x20)) continue;
    int x = list[i];
    float u = eo_uniform(seed, epoch, sim, t, (uint32_t)x, EO_PURPOSE_NEIGHBOUR);
    int cr, cc;
    eo_choose_neighbour(x / W, x % W, u, &cr, &cc);
    int child = cr * W + cc;
    if (st[child] != EO_VACANT) continue;
    uint32_t key = eo_arbitration_key((uint32_t)i);
    if (key < best_key) { best_key = key; best = i; best_child = child; }
  }
  if (best >= 0) {
    int parent = list[best];
    int child = best_child;
    st[child] = EO_OCCUPIED;
    eo_cell *pc = &cells[parent];
    pc->phase = 0;
    pc->prolif = 0.f;
    float z[5];
    eo_normals5(seed, epoch, sim, t, (uint32_t)child, z);
    eo_cell nc;
    memset(&nc, 0, sizeof(nc));
    float acc = 0.f;
    for (int k = 0; k < 5; ++k) {
      float v = fmaf(p->cycle_sd[k], z[k], p->cycle_mean[k]);
      acc = (k == 0) ? v : acc + v;
      nc.times[k] = acc;
    }
    nc.R = pc->R;
    if (nc.R > 0) nc.delay = eo_delay_time(nc.R, p);
    cells[child] = nc;
    if (n >= cap) return -1;
    list[n] = (uint16_t)child;
    owner[child] = n;
    *pn = n + 1;
    if (instr) instr[EO_INSTR_DIVISIONS] += 1;
  }
  return 0;
}

static int eo_run_one(const eo_params *p, float *sub, uint8_t *st, eo_cell *cells,
                      uint16_t *list, int32_t *pn, int32_t cap, int32_t *roi,
                      uint32_t *pstep, uint32_t sim, const float *dose_row,
                      int32_t n_slots, uint32_t res, uint32_t length, uint32_t seed,
                      uint32_t epoch, uint32_t nsteps, uint64_t *instr, int rules) {
  const int H = p->H, W = p->W, HW = H * W;
  int32_t *owner = (int32_t *)malloc(sizeof(int32_t) * HW);
  int32_t *seen = (int32_t *)calloc((size_t)HW, sizeof(int32_t));
  uint8_t *flags = (uint8_t *)malloc((size_t)cap + 1);
  int rc = 0;
  for (int i = 0; i < HW; ++i) owner[i] = -1;
  for (int i = 0; i < *pn; ++i) owner[list[i]] = i;
  for (uint32_t k = 0; k < nsteps; ++k) {
    uint32_t t = *pstep;
    if (t % 128u == 0) {
      int n = eo_rebuild_list(H, W, st, list, cap, owner);
      if (n < 0) { rc = -1; break; }
      *pn = n;
    }
    if (t % 32u == 0) eo_find_roi(H, W, st, roi);
    eo_diffuse_f32(p, H, W, sub, roi);
    float dose = 0.f;
    if (t < length && t % res == 0 && (int32_t)(t / res) < n_slots) dose = dose_row[t / res];
    if (eo_cells_step(p, H, W, sub, st, cells, list, pn, cap, owner, flags, seen, t, sim,
                      dose, seed, epoch, instr, rules) != 0) {
      rc = -1;
      break;
    }
    *pstep = t + 1;
  }
  free(owner);
  free(seen);
  free(flags);
  return rc;
}

int eo_run(const eo_params *p, int B, float *sub, uint8_t *st, eo_cell *cells,
           uint16_t *list, int32_t *nlist, int32_t cap, int32_t *roi, uint32_t *step,
           const int32_t *sim_id, const int32_t *prot_index, const float *prot_data,
           int32_t n_slots, uint32_t prot_res, uint32_t prot_length,
           uint32_t seed, uint32_t epoch, uint32_t nsteps, int nthreads) {
  return eo_run_instr(p, B, sub, st, cells, list, nlist, cap, roi, step, sim_id, prot_index,
                      prot_data, n_slots, prot_res, prot_length, seed, epoch, nsteps, nthreads,
                      NULL);
}

int eo_run_instr(const eo_params *p, int B, float *sub, uint8_t *st, eo_cell *cells,
                 uint16_t *list, int32_t *nlist, int32_t cap, int32_t *roi, uint32_t *step,
                 const int32_t *sim_id, const int32_t *prot_index, const float *prot_data,
                 int32_t n_slots, uint32_t prot_res, uint32_t prot_length,
                 uint32_t seed, uint32_t epoch, uint32_t nsteps, int nthreads,
                 uint64_t *instr) {
  return eo_run_rules(p, B, sub, st, cells, list, nlist, cap, roi, step, sim_id, prot_index,
                      prot_data, n_slots, prot_res, prot_length, seed, epoch, nsteps, nthreads,
                      instr, 0);
}

int eo_run_rules(const eo_params *p, int B, float *sub, uint8_t *st, eo_cell *cells,
                 uint16_t *list, int32_t *nlist, int32_t cap, int32_t *roi, uint32_t *step,
                 const int32_t *sim_id, const int32_t *prot_index, const float *prot_data,
                 int32_t n_slots, uint32_t prot_res, uint32_t prot_length,
                 uint32_t seed, uint32_t epoch, uint32_t nsteps, int nthreads,
                 uint64_t *instr, int rules) {
  const size_t HW = (size_t)p->H * p->W;
  int err = 0;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(| : err)
#endif
  for (int b = 0; b < B; ++b) {
    int rc = eo_run_one(p, sub + (size_t)b * 3 * HW, st + (size_t)b * HW, cells + (size_t)b * HW,
                        list + (size_t)b * cap, nlist + b, cap, roi + 5 * b, step + b,
                        (uint32_t)sim_id[b], prot_data + (size_t)prot_index[b] * n_slots,
                        n_slots, prot_res, prot_length, seed, epoch, nsteps,
                        instr ? instr + (size_t)b * EO_INSTR_N : NULL, rules);
    if (rc != 0) err |= 1;
  }
  (void)nthreads;
  return err ? -1 : 0;
}
