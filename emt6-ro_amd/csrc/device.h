// Device-side building blocks for the EMT6/Ro kernels: counter-based RNG,
// FP32 elementary functions with a fixed operation order, and the cell rules.
//
// Compiled with -ffp-contract=off: every float expression is rounded
// operation by operation in the order written (fmaf() is the only fused
// operation), so the results are the same bits on any IEEE-754 machine that
// evaluates the same sequence.
#ifndef EMT6RO_AMD_DEVICE_H_
#define EMT6RO_AMD_DEVICE_H_

#include <hip/hip_runtime.h>
#include <cstdint>
#include "model.h"

namespace emt6ro {
namespace dev {

// ---------------------------------------------------------------------------
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): counter (c0..c3), key (k0, k1).
// ---------------------------------------------------------------------------
struct U4 { uint32_t x, y, z, w; };

__device__ __forceinline__ U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                     uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
    const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return U4{c0, c1, c2, c3};
}

// (0, 1]: top 24 bits plus one, times 2^-24 (exact).  The reference's
// uniforms are also on (0, 1] (random-engine.cu:34).
__device__ __forceinline__ float u01(uint32_t x) {
  return (float)((x >> 8) + 1u) * 5.9604644775390625e-8f;
}

struct RngKey {
  uint32_t seed, epoch, sim;
};

__device__ __forceinline__ float uniform(const RngKey &k, uint32_t step, uint32_t site,
                                         uint32_t purpose) {
  U4 o = philox(step, site | (purpose << 24), k.sim, 0u, k.seed, k.epoch);
  return u01(o.x);
}

// expf: Cody-Waite reduction by ln2 = 0.693359375 - 2.12194440e-4 and a
// degree-5 polynomial, the method and coefficients of the Cephes library's
// expf (S. L. Moshier).  Arguments are clamped to [-87, 88].
__device__ __forceinline__ float expf_fixed(float x) {
  x = x > 88.0f ? 88.0f : x;
  x = x < -87.0f ? -87.0f : x;
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
  return p * __uint_as_float((uint32_t)(n + 127) << 23);
}

// logf for positive normal arguments: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
// degree-9 polynomial in m - 1 (Cephes logf).
__device__ __forceinline__ float logf_fixed(float x) {
  uint32_t b = __float_as_uint(x);
  int e = (int)((b >> 23) & 0xffu) - 126;
  float m = __uint_as_float((b & 0x807fffffu) | (126u << 23));
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

// sin(2 pi u), cos(2 pi u) for u in (0, 1]: exact quadrant split of u, then
// the Cephes sinf/cosf polynomials on [-pi/4, pi/4].
__device__ __forceinline__ void sincos2pi(float u, float &s_out, float &c_out) {
  float q = floorf(u * 4.0f + 0.5f);
  float r = u - q * 0.25f;
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
  if (quad == 0) { s_out = s; c_out = c; }
  else if (quad == 1) { s_out = c; c_out = -s; }
  else if (quad == 2) { s_out = -s; c_out = -c; }
  else { s_out = -c; c_out = s; }
}

// Five standard normals by the Box-Muller transform (Box & Muller, 1958)
// from six uniforms of two Philox blocks keyed by (step, site, sim).
__device__ __forceinline__ void normals5(const RngKey &k, uint32_t step, uint32_t site,
                                         float z[5]) {
  const uint32_t c1 = site | (kPurposeNormal << 24);
  U4 a = philox(step, c1, k.sim, 0u, k.seed, k.epoch);
  U4 b = philox(step, c1, k.sim, 1u, k.seed, k.epoch);
  const float u[6] = {u01(a.x), u01(a.y), u01(a.z), u01(a.w), u01(b.x), u01(b.y)};
  float zz[6];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float rad = sqrtf(-2.0f * logf_fixed(u[2 * i]));
    float s, c;
    sincos2pi(u[2 * i + 1], s, c);
    zz[2 * i] = rad * c;
    zz[2 * i + 1] = rad * s;
  }
#pragma unroll
  for (int i = 0; i < 5; ++i) z[i] = zz[i];
}

// ---------------------------------------------------------------------------
// Cell rules (cell/cell.cu:9-100, cell/cell.h:102-136, site/site.h:62-82).
// ---------------------------------------------------------------------------
// A cell as the step sees it: t0 and t4 are the G1 and D phase ends, tcur
// the end of the current phase (cycle_times.times[phase], clamped at D).
struct CellL {
  float t_rep, R, delay, prolif;
  float t0, t4, tcur;
  uint32_t mode, phase;
};

// calcDelayTime (cell.cu:93-100)
__device__ __forceinline__ float delay_time(float R, const DevParams &p) {
  return p.delay_coeff * expf_fixed(p.delay_exp * R);
}

// updateState, enterG1SStopping, tryProliferating, tryQuiescence
// (cell.cu:33-82); returns alive and sets mode.
__device__ __forceinline__ bool update_state(CellL &c, float cho, float ox, float gi,
                                             const DevParams &p, int vacant) {
  if (c.prolif >= c.t4 - p.death_margin) return false;
  const bool stopping = c.prolif > (c.t0 - p.g1s_margin) && c.phase == 0u && vacant <= 1;
  if (!stopping) {
    if (cho >= p.metab[0][0] && ox >= p.metab[0][1] && gi < p.quiescence_gi) {
      c.mode = 0u;
      return true;
    }
    if (cho >= p.metab[1][0] && ox >= p.metab[1][1] && gi < p.quiescence_gi) {
      c.mode = 1u;
      return true;
    }
  }
  if (c.phase != 0u && c.phase != 2u) return false;
  if (cho >= p.metab[2][0] && ox >= p.metab[2][1] && gi < p.death_gi) {
    c.mode = 2u;
    return true;
  }
  if (cho >= p.metab[3][0] && ox >= p.metab[3][1] && gi < p.death_gi) {
    c.mode = 3u;
    return true;
  }
  return false;
}

// progressClock (cell.cu:21-32); returns the "major cycle change" flag.
// The phase index is clamped at D for the lookup: a cell past D dies in the
// next updateState before its clock moves again.
__device__ __forceinline__ bool progress_clock(CellL &c, const DevParams &p) {
  if ((c.mode == 0u || c.mode == 1u) && c.t_rep == 0.f) {
    c.prolif = c.prolif + p.hour_step;
    const uint32_t cur = c.phase;
    c.phase = cur + (c.prolif > c.tcur ? 1u : 0u);
    return c.phase != cur && c.phase != 4u;
  }
  return false;
}

// irradiate (cell.cu:84-91)
__device__ __forceinline__ void irradiate(CellL &c, float dose, const DevParams &p) {
  c.R = c.R / (1.f + c.t_rep / p.repair_half_time) + dose;
  c.t_rep = 0.f;
  c.delay = delay_time(c.R, p);
}

// chooseNeighbour (cell-division.h:44-54): diagonal neighbours weigh 1 each,
// orthogonal ones sqrt(2) each; u in (0, 1].  Returns the neighbour's offset.
__device__ __forceinline__ void choose_neighbour(float u, int &dr, int &dc) {
  const float diag = 4.f;
  const float ortho = (float)(4. * 1.41421356237309504880);
  const float score = u * (diag + ortho);
  if (score < diag) {
    const uint32_t num = (uint32_t)(uint8_t)score;
    dr = (int)(num & 1u) * 2 - 1;
    dc = (int)((num & 2u) >> 1) * 2 - 1;
  } else {
    const uint32_t num =
        (uint32_t)(uint8_t)((double)(score - diag) / 1.41421356237309504880);
    const int diff = (int)((num & 2u) >> 1) * 2 - 1;
    if (num & 1u) { dr = diff; dc = 0; } else { dr = 0; dc = diff; }
  }
}

// Division arbitration (simulation.cu:86-105, cuda-utils.h:23-49).  The
// reference steps list entry i on thread i mod 512 (a later entry of a thread
// overwrites its candidate) and reduces the per-thread candidates with the
// first-non-zero operator through shuffle-down trees over 32-lane warps and
// then over the 16 warp results.  A shuffle-down tree visits lanes in
// bit-reversed order, so the winner is the candidate whose thread has the
// smallest (bitrev4(warp), bitrev5(lane)), ties broken by the largest i.
// Smaller key wins.
__device__ __forceinline__ uint32_t arbitration_key(uint32_t i) {
  const uint32_t t = i & 511u;
  const uint32_t prio = (__brev(t >> 5) >> 28) << 5 | (__brev(t & 31u) >> 27);
  return (prio << 16) | (0xffffu - ((i >> 9) & 0xffffu));
}

}  // namespace dev
}  // namespace emt6ro

#endif  // EMT6RO_AMD_DEVICE_H_
