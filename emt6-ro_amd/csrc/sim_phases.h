// The phases of one EMT6/Ro step as workgroup-wide device functions: occupied
// list rebuild, region of interest, substrate diffusion and the cell update.
// The persistent kernel calls them with the lattice resident in LDS; the
// per-step kernels wrap each phase with its own loads and stores.
//
// One workgroup of NT threads owns one simulation.  Every function here is
// entered by all NT threads with uniform arguments and contains its own
// barriers.
#ifndef EMT6RO_AMD_SIM_PHASES_H_
#define EMT6RO_AMD_SIM_PHASES_H_

#include "device.h"

namespace emt6ro {
namespace dev {

// Ready-to-divide entries kept in LDS per step; more spill to global memory.
constexpr int kReadyLds = 56;
constexpr int kRedInts = 128;
constexpr uint32_t kNoWinner = 0xffffffffu;

// Workgroup scalars in LDS.
struct alignas(16) Shared {
  uint64_t ready[kReadyLds];  // slot | child << 16 | R bits << 32
  int32_t n, cur, ready_cnt, sim;
  int32_t roi[5];
  uint32_t winner, divisions, status;
  int32_t red[kRedInts];      // reductions and diffusion planning scratch
};

// Kernel-wide arguments shared by all kernels of a run.
struct RunArgs {
  DevParams p;
  const SimDesc *desc;
  SimHdr *hdr;
  float *sub_pool;
  uint8_t *st_pool;
  uint32_t *slot_pool;
  const float *prot_data;     // n_prot rows of n_slots doses
  int32_t n_slots;
  uint32_t prot_res, prot_length;
  uint32_t seed, epoch;
  uint32_t nsteps;            // steps to advance each simulation (capped by horizon)
  const int32_t *order;       // simulation processing order (persistent kernel)
  int32_t nsims;
  int32_t *queue;             // work counter (persistent kernel)
  int32_t *scratch;           // per-workgroup prefix table, scratch_stride ints
  int64_t scratch_stride;
  uint64_t *ovf;              // per-workgroup ready-entry spill, ovf_stride entries
  int64_t ovf_stride;
  int32_t pv_max;             // LDS floats per substrate plane
  unsigned long long *instr;  // instrumentation counters (null: off), kInstr* layout
  const uint32_t *target;     // per-simulation step to reach (null: advance by nsteps)
  const int32_t *nsims_dev;   // queue length in device memory (window kernel)
};

// Instrumentation record (common/instr.h:5-21): repair decisions, repair
// kills, irradiation events, divisions, the sum of death probabilities in
// 32.32 fixed point (an integer sum, so the total does not depend on the
// order of the atomic adds), then 300-bin histograms of the irradiation and
// of the time in repair at each repair decision (0.1 per bin).
enum : int {
  kInstrRepairDecisions = 0,
  kInstrRepairKills = 1,
  kInstrIrradiationEvents = 2,
  kInstrDivisions = 3,
  kInstrDeathProbFx = 4,
  kInstrHistN = 300,
  kInstrIrrHist = 5,
  kInstrTirHist = 5 + 300,
  kInstrN = 5 + 600,
};

// Histogram bin of common/instr.h:42-44 (cell.h:115-120).
__device__ __forceinline__ int instr_bin(float v) {
  int b = (int)(v / 0.1f);
  return b < 0 ? 0 : (b >= kInstrHistN ? kInstrHistN - 1 : b);
}

// Protocol::getDose (protocol.h:9-14)
__device__ __forceinline__ float dose_at(const RunArgs &a, int prot, uint32_t t) {
  if (t < a.prot_length && t % a.prot_res == 0u) {
    const uint32_t slot = t / a.prot_res;
    if ((int32_t)slot < a.n_slots) return a.prot_data[(int64_t)prot * a.n_slots + slot];
  }
  return 0.f;
}

// ---------------------------------------------------------------------------
// Workgroup reductions (64-wide wavefronts).
// ---------------------------------------------------------------------------
template <int NT>
__device__ __forceinline__ int wg_reduce_min(int v, int32_t *red) {
  for (int m = 32; m >= 1; m >>= 1) { int y = __shfl_xor(v, m, 64); v = y < v ? y : v; }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  int r = red[0];
  for (int w = 1; w < NT / 64; ++w) r = red[w] < r ? red[w] : r;
  __syncthreads();
  return r;
}

template <int NT>
__device__ __forceinline__ int wg_reduce_max(int v, int32_t *red) {
  for (int m = 32; m >= 1; m >>= 1) { int y = __shfl_xor(v, m, 64); v = y > v ? y : v; }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  int r = red[0];
  for (int w = 1; w < NT / 64; ++w) r = red[w] > r ? red[w] : r;
  __syncthreads();
  return r;
}

// Four minima at once: (a, b, -c, -d) reduced together.
template <int NT>
__device__ __forceinline__ void wg_reduce_min4(int v[4], int32_t *red) {
  for (int m = 32; m >= 1; m >>= 1)
    for (int k = 0; k < 4; ++k) { int y = __shfl_xor(v[k], m, 64); v[k] = y < v[k] ? y : v[k]; }
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 4; ++k) red[(threadIdx.x >> 6) * 4 + k] = v[k];
  __syncthreads();
  for (int k = 0; k < 4; ++k) {
    int r = red[k];
    for (int w = 1; w < NT / 64; ++w) r = red[w * 4 + k] < r ? red[w * 4 + k] : r;
    v[k] = r;
  }
  __syncthreads();
}

template <int NT>
__device__ __forceinline__ uint64_t wg_reduce_sum_u64(uint64_t v, int32_t *red) {
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  uint64_t *r64 = reinterpret_cast<uint64_t *>(red);
  if ((threadIdx.x & 63) == 0) r64[threadIdx.x >> 6] = v;
  __syncthreads();
  uint64_t r = 0;
  for (int w = 0; w < NT / 64; ++w) r += r64[w];
  __syncthreads();
  return r;
}

// Exclusive prefix sum over the workgroup; returns the offset, sets total.
template <int NT>
__device__ __forceinline__ int wg_exclusive_scan(int v, int32_t *red, int &total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int inc = v;
  for (int d = 1; d < 64; d <<= 1) {
    int y = __shfl_up(inc, d, 64);
    if (lane >= d) inc += y;
  }
  if (lane == 63) red[wid] = inc;
  __syncthreads();
  int off = 0, tot = 0;
  for (int w = 0; w < NT / 64; ++w) {
    int s = red[w];
    off += (w < wid) ? s : 0;
    tot += s;
  }
  total = tot;
  __syncthreads();
  return off + inc - v;
}

// ---------------------------------------------------------------------------
// Occupied-list rebuild (findOccupied, simulation.cu:41-64).  The reference
// scans with 32 threads: thread g takes columns g, g+32, g+64, ... of rows
// 0..H-2, and the list is the concatenation of the threads' lists.  The new
// list order is written as a gather of the live slots into the other slot
// buffer; `pfx` is a per-workgroup table of 32*(H-1) ints in global memory.
// ---------------------------------------------------------------------------
template <int NT>
__device__ void rebuild_list(const uint8_t *st, int H, int W, uint32_t *slots, int cap,
                             int32_t *pfx, Shared &sm) {
  const int tid = threadIdx.x;
  const int rows = H - 1;
  const int npairs = 32 * rows;
  const int per = (npairs + NT - 1) / NT;
  const int q0 = tid * per;
  const int q1 = min(q0 + per, npairs);
  int local = 0;
  for (int q = q0; q < q1; ++q) {
    const int g = q / rows, r = q - g * rows;
    for (int c = g; c < W - 1; c += 32) local += (st[r * W + c] & 3) == kOccupied;
  }
  int total;
  int run = wg_exclusive_scan<NT>(local, sm.red, total);
  for (int q = q0; q < q1; ++q) {
    const int g = q / rows, r = q - g * rows;
    pfx[q] = run;
    for (int c = g; c < W - 1; c += 32) run += (st[r * W + c] & 3) == kOccupied;
  }
  __syncthreads();
  const int n_old = sm.n;
  const int cur = sm.cur;
  const uint32_t *src = slots + (int64_t)cur * kSlotFields * cap;
  uint32_t *dst = slots + (int64_t)(cur ^ 1) * kSlotFields * cap;
  for (int i = tid; i < n_old; i += NT) {
    const uint32_t w = src[(int64_t)kPacked * cap + i];
    if (!slot_live(w)) continue;
    const int x = (int)slot_site(w);
    const int r = x / W, c = x - r * W;
    const int g = c & 31;
    int pos = pfx[g * rows + r];
    for (int cc = g; cc < c; cc += 32) pos += (st[r * W + cc] & 3) == kOccupied;
#pragma unroll
    for (int f = 0; f < kSlotFields; ++f)
      dst[(int64_t)f * cap + pos] = src[(int64_t)f * cap + i];
  }
  __syncthreads();
  if (tid == 0) {
    sm.n = total;
    sm.cur = cur ^ 1;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Region of interest (findROI, diffusion.cu:65-100): bounding box of the
// occupied sites, its centre, the largest ceil(Euclidean distance) from the
// centre to an occupied site, and the square of that half-width clipped to
// the interior.  ceil(sqrtf(n)) of the reference equals the integer ceiling
// square root for every n a lattice can produce.
// ---------------------------------------------------------------------------
template <int NT>
__device__ void find_roi(const uint8_t *st, int H, int W, Shared &sm) {
  const int tid = threadIdx.x;
  int v[4] = {H, W, 0, 0};  // minr, minc, -maxr, -maxc
  for (int x = tid; x < H * W; x += NT) {
    if ((st[x] & 3) == kOccupied) {
      const int r = x / W, c = x - r * W;
      v[0] = min(v[0], r);
      v[1] = min(v[1], c);
      v[2] = min(v[2], -r);
      v[3] = min(v[3], -c);
    }
  }
  wg_reduce_min4<NT>(v, sm.red);
  const int minr = v[0], minc = v[1], maxr = -v[2], maxc = -v[3];
  if (maxr == 0 && maxc == 0) {
    if (tid == 0) for (int k = 0; k < 5; ++k) sm.roi[k] = 0;
    __syncthreads();
    return;
  }
  const int midr = minr + (maxr - minr + 1) / 2;
  const int midc = minc + (maxc - minc + 1) / 2;
  int d2 = 0;
  for (int x = tid; x < H * W; x += NT) {
    if ((st[x] & 3) == kOccupied) {
      const int r = x / W, c = x - r * W;
      const int e = (midr - r) * (midr - r) + (midc - c) * (midc - c);
      d2 = max(d2, e);
    }
  }
  d2 = wg_reduce_max<NT>(d2, sm.red);
  int dist = (int)sqrtf((float)d2);
  while (dist * dist < d2) ++dist;
  while (dist > 0 && (dist - 1) * (dist - 1) >= d2) --dist;
  if (tid == 0) {
    const int sr = (midr <= dist) ? 1 : midr - dist;
    const int sc = (midc <= dist) ? 1 : midc - dist;
    const int sw = (midc + dist >= W - 1) ? W - 1 - sc : midc + dist - sc + 1;
    const int sh = (midr + dist >= H - 1) ? H - 1 - sr : midr + dist - sr + 1;
    sm.roi[0] = sr; sm.roi[1] = sc; sm.roi[2] = sh; sm.roi[3] = sw; sm.roi[4] = dist;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Diffusion (diffusionKernel, diffusion.cu:119-191), in place on a padded
// lattice in LDS: plane s of substrate s holds lattice site (r, c) at
// (r+1)*PW + (c+1), with PW = W+2 and a zero ring around the lattice.
//
// The reference copies the ROI plus a two-site ring into a temporary grid:
// outer ring 0, masked sites the external levels, other sites the lattice
// values; it runs S-1 Jacobi substeps on the unmasked sites of the ROI plus
// its inner ring ("mask box"), then one substep on every ROI site, and writes
// the ROI back.  Here the two rings are saved to registers, set to their
// temporary values, and restored afterwards, which gives the same ROI values
// without a second grid.
//
// Mask: site (i, j) of the (h+2) x (w+2) mask box is masked when
// sqrtf((i+1-(h+2)/2)^2 + (j+1-(w+2)/2)^2) >= dist (fillBorderMask,
// diffusion.cu:16-30); all terms are multiples of 1/4, so this is the integer
// test a^2 + b^2 >= 4 dist^2 with a = 2(i+1)-(h+2), b = 2(j+1)-(w+2).
//
// Work split: the mask box is cut into bands of L rows.  An item is one
// column of one band; a band's items cover the columns where the band has
// unmasked sites (the union of the disc's chords over its rows), so lanes of
// a wave take consecutive columns and read consecutive LDS words.  Items are
// dealt to the waves in equal contiguous blocks.  The final substep has its
// own items over the ROI.  Items are planned when the ROI changes and kept in
// registers: bits 0-7 column, 8-15 first row (mask-box coordinates), 16-31
// the mask of rows the item updates.
//
// The LDS row stride RS is a compile-time constant, so the rows an item
// walks are immediate offsets from one address.  An item of the last band may
// read up to L-1 rows below the temporary grid; those rows only feed results
// that are not stored, and the LDS layout keeps them inside the allocation.
//
// Stencil, per substrate, in this order (no other fused operations), with
// H(row) = left + right, the row's neighbours of the centre column:
//   o = (up + down) + H(mid),  d = H(up row) + H(down row)
//   L = fma(d, 1/sqrt2, o);  L = fma(-F, x, L);  x' = fma(k, L, x)
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool masked(int i, int j, int hm, int wm, int d4) {
  const int a = 2 * (i + 1) - hm, b = 2 * (j + 1) - wm;
  return a * a + b * b >= d4;
}

// Item index for slot k of this thread: items fill the waves in order
// (item q on thread q mod NT, slot q / NT), so waves left without items
// leave their SIMDs to other workgroups; -1 if the slot is empty.
template <int NT>
__device__ __forceinline__ int dealt_item(int k, int n) {
  const int q = (int)threadIdx.x + k * NT;
  return q < n ? q : -1;
}

__device__ __forceinline__ uint32_t pack_item(int j, int i0, uint32_t mask) {
  return (uint32_t)j | ((uint32_t)i0 << 8) | (mask << 16);
}

// Plans this thread's Jacobi (unmasked mask-box sites) and final-substep
// (ROI sites) items for `roi`.  Uses sm.red as scratch; ends with a barrier.
template <int NT, int L, int KMAX>
__device__ void plan_diffusion(const int roi[5], uint32_t jac[KMAX], uint32_t fin[KMAX],
                               Shared &sm) {
  const int h = roi[2], w = roi[3], dist = roi[4];
#pragma unroll
  for (int k = 0; k < KMAX; ++k) { jac[k] = 0u; fin[k] = 0u; }
  if (h == 0 || w == 0) return;
  const int hm = h + 2, wm = w + 2, d4 = 4 * dist * dist;
  const int nb = (hm + L - 1) / L;
  int32_t *lo = sm.red, *cnt = sm.red + kRedInts / 2;
  const int tid = threadIdx.x;
  for (int b = tid; b < nb; b += NT) {
    int a = wm, e = 0;
    for (int l = 0; l < L; ++l) {
      const int i = b * L + l;
      if (i >= hm) break;
      const int ai = 2 * (i + 1) - hm;
      const int rem = d4 - ai * ai;
      if (rem <= 0) continue;
      for (int j = 0; j < wm; ++j) {
        const int bj = 2 * (j + 1) - wm;
        if (bj * bj < rem) { a = min(a, j); e = max(e, j + 1); }
      }
    }
    lo[b] = a;
    cnt[b] = e > a ? e - a : 0;
  }
  __syncthreads();
  int n = 0;
  for (int b = 0; b < nb; ++b) n += cnt[b];
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    const int q = dealt_item<NT>(k, n);
    if (q < 0) continue;
    int b = 0, base = 0;
    while (base + cnt[b] <= q) { base += cnt[b]; ++b; }
    const int j = lo[b] + (q - base), i0 = b * L;
    uint32_t m = 0u;
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const int i = i0 + l;
      if (i < hm && !masked(i, j, hm, wm, d4)) m |= 1u << l;
    }
    jac[k] = pack_item(j, i0, m);
  }
  const int nbf = (h + L - 1) / L;
  const int nf = nbf * w;
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    const int q = dealt_item<NT>(k, nf);
    if (q < 0) continue;
    const int band = q / w, j = 1 + (q - band * w), i0 = 1 + band * L;
    uint32_t m = 0u;
#pragma unroll
    for (int l = 0; l < L; ++l)
      if (i0 + l <= h) m |= 1u << l;
    fin[k] = pack_item(j, i0, m);
  }
  __syncthreads();
}

// Horizontal neighbours through the wave: DPP wave_shr:1 gives lane i the
// value of lane i - 1, wave_shl:1 that of lane i + 1 (lanes 0 and 63 get 0).
__device__ __forceinline__ float from_left_lane(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_right_lane(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}

// Item word of the lane-run layout (LaneRunDiffuser): column + 1 (bits 0-7),
// first band row (8-15), mask of rows to write (16-29), valid (31).
constexpr uint32_t kItemValid = 1u << 31;
__device__ __forceinline__ uint32_t pack_lane_item(int j, int i0, uint32_t mask) {
  return (uint32_t)(j + 1) | ((uint32_t)i0 << 8) | (mask << 16) | kItemValid;
}

// The reference's temporary grid (diffusion.cu:150-162) made in place: the
// two rings around the ROI are saved to registers and set to their
// temporary values (outer ring 0, masked sites the external levels), masked
// ROI sites get the external levels, and restore() puts the rings back.
template <int NT, int RS, int PS>
struct TempGrid {
  // Masked ROI sites a thread keeps between ROI changes (16-bit plane
  // indices, two per word); a ROI with more on some thread uses the loop.
  static constexpr int kMasked = 4;
  static constexpr uint32_t kOuter = 1u << 16, kMask = 1u << 17, kNone = 0xffffffffu;
  float sv[PS][3];
  uint32_t ring[PS];           // plane index | kOuter | kMask, or kNone
  uint32_t mk[kMasked / 2];    // masked ROI sites, 0xffff: none
  bool mk_all;                 // every thread's masked sites are in mk

  // The ring sites and masked ROI sites of `roi`; they depend on the ROI
  // only, so callers that keep the TempGrid across steps call this when the
  // ROI changes and apply() every step.
  __device__ void prepare(const int roi[5], Shared &sm) {
    const int r0 = roi[0], c0 = roi[1], h = roi[2], w = roi[3], dist = roi[4];
    const int tid = threadIdx.x;
    const int hm = h + 2, wm = w + 2;
    const int d4 = 4 * dist * dist;
    // Padded coordinates of the temporary grid: rows R0..R1, cols C0..C1.
    const int R0 = r0 - 1, R1 = r0 + h + 2, C0 = c0 - 1, C1 = c0 + w + 2;
    const int nper = 4 * w + 4 * h + 16;
#pragma unroll
    for (int k = 0; k < PS; ++k) {
      int q = tid + k * NT;
      ring[k] = kNone;
      if (q < nper && h > 0 && w > 0) {
        int pr, pc;
        bool outer;
        if (q < w + 4) { pr = R0; pc = C0 + q; outer = true; }
        else if ((q -= w + 4) < w + 4) { pr = R1; pc = C0 + q; outer = true; }
        else if ((q -= w + 4) < h + 2) { pr = R0 + 1 + q; pc = C0; outer = true; }
        else if ((q -= h + 2) < h + 2) { pr = R0 + 1 + q; pc = C1; outer = true; }
        else if ((q -= h + 2) < w + 2) { pr = R0 + 1; pc = C0 + 1 + q; outer = false; }
        else if ((q -= w + 2) < w + 2) { pr = R1 - 1; pc = C0 + 1 + q; outer = false; }
        else if ((q -= w + 2) < h) { pr = R0 + 2 + q; pc = C0 + 1; outer = false; }
        else { q -= h; pr = R0 + 2 + q; pc = C1 - 1; outer = false; }
        ring[k] = (uint32_t)(pr * RS + pc) |
                  (outer ? kOuter : masked(pr - r0, pc - c0, hm, wm, d4) ? kMask : 0u);
      }
    }
#pragma unroll
    for (int k = 0; k < kMasked / 2; ++k) mk[k] = 0xffffffffu;
    int n = 0;
    for (int q = tid; q < h * w; q += NT) {
      const int i = 1 + q / w, j = 1 + q % w;
      if (masked(i, j, hm, wm, d4)) {
        const uint32_t idx = (uint32_t)((r0 + i) * RS + (c0 + j));
#pragma unroll
        for (int k = 0; k < kMasked; ++k)
          if (k == n) {
            const int sh = 16 * (k & 1);
            mk[k >> 1] = (mk[k >> 1] & ~(0xffffu << sh)) | (idx << sh);
          }
        ++n;
      }
    }
    // Workgroup AND of (n <= kMasked) in the last reduction word, which
    // the chord table of a plan (at most 88 rows) leaves alone.
    int32_t &all = sm.red[kRedInts - 1];
    if (tid == 0) all = 1;
    __syncthreads();
    if (n > kMasked) all = 0;
    __syncthreads();
    mk_all = all != 0;
    __syncthreads();
  }

  // The reference's temporary grid (diffusion.cu:150-162) made in place:
  // ring values saved, outer ring 0, masked ring and ROI sites the external
  // levels.
  __device__ void apply(const DevParams &p, float *sub, int PV, const int roi[5]) {
#pragma unroll
    for (int k = 0; k < PS; ++k) {
      if (ring[k] != kNone) {
        const int idx = (int)(ring[k] & 0xffffu);
#pragma unroll
        for (int s = 0; s < 3; ++s) sv[k][s] = sub[s * PV + idx];
        if (ring[k] & kOuter) {
#pragma unroll
          for (int s = 0; s < 3; ++s) sub[s * PV + idx] = 0.f;
        } else if (ring[k] & kMask) {
#pragma unroll
          for (int s = 0; s < 3; ++s) sub[s * PV + idx] = p.ext[s];
        }
      }
    }
    if (mk_all) {
#pragma unroll
      for (int k = 0; k < kMasked; ++k) {
        const uint32_t idx = (mk[k >> 1] >> (16 * (k & 1))) & 0xffffu;
        if (idx != 0xffffu) {
#pragma unroll
          for (int s = 0; s < 3; ++s) sub[s * PV + idx] = p.ext[s];
        }
      }
    } else {
      masked_loop(p, sub, PV, roi);
    }
  }

  __device__ static void masked_loop(const DevParams &p, float *sub, int PV, const int roi[5]) {
    const int r0 = roi[0], c0 = roi[1], h = roi[2], w = roi[3], dist = roi[4];
    const int hm = h + 2, wm = w + 2, d4 = 4 * dist * dist;
    for (int q = threadIdx.x; q < h * w; q += NT) {
      const int i = 1 + q / w, j = 1 + q % w;
      if (masked(i, j, hm, wm, d4)) {
        const int idx = (r0 + i) * RS + (c0 + j);
#pragma unroll
        for (int s = 0; s < 3; ++s) sub[s * PV + idx] = p.ext[s];
      }
    }
  }

  // Per-step form for callers that do not keep the TempGrid: prepare()'s
  // ring decoding, then the masked ROI sites by the loop.
  __device__ void setup(const DevParams &p, float *sub, int PV, const int roi[5]) {
    const int r0 = roi[0], c0 = roi[1], h = roi[2], w = roi[3], dist = roi[4];
    const int tid = threadIdx.x;
    const int hm = h + 2, wm = w + 2;
    const int d4 = 4 * dist * dist;
    const int R0 = r0 - 1, R1 = r0 + h + 2, C0 = c0 - 1, C1 = c0 + w + 2;
    const int nper = 4 * w + 4 * h + 16;
#pragma unroll
    for (int k = 0; k < PS; ++k) {
      int q = tid + k * NT;
      ring[k] = kNone;
      if (q < nper) {
        int pr, pc;
        bool outer;
        if (q < w + 4) { pr = R0; pc = C0 + q; outer = true; }
        else if ((q -= w + 4) < w + 4) { pr = R1; pc = C0 + q; outer = true; }
        else if ((q -= w + 4) < h + 2) { pr = R0 + 1 + q; pc = C0; outer = true; }
        else if ((q -= h + 2) < h + 2) { pr = R0 + 1 + q; pc = C1; outer = true; }
        else if ((q -= h + 2) < w + 2) { pr = R0 + 1; pc = C0 + 1 + q; outer = false; }
        else if ((q -= w + 2) < w + 2) { pr = R1 - 1; pc = C0 + 1 + q; outer = false; }
        else if ((q -= w + 2) < h) { pr = R0 + 2 + q; pc = C0 + 1; outer = false; }
        else { q -= h; pr = R0 + 2 + q; pc = C1 - 1; outer = false; }
        const int idx = pr * RS + pc;
        ring[k] = (uint32_t)idx;
#pragma unroll
        for (int s = 0; s < 3; ++s) sv[k][s] = sub[s * PV + idx];
        if (outer) {
#pragma unroll
          for (int s = 0; s < 3; ++s) sub[s * PV + idx] = 0.f;
        } else if (masked(pr - r0, pc - c0, hm, wm, d4)) {
#pragma unroll
          for (int s = 0; s < 3; ++s) sub[s * PV + idx] = p.ext[s];
        }
      }
    }
    masked_loop(p, sub, PV, roi);
  }

  __device__ void restore(float *sub, int PV) const {
#pragma unroll
    for (int k = 0; k < PS; ++k) {
      if (ring[k] != kNone) {
        const int idx = (int)(ring[k] & 0xffffu);
#pragma unroll
        for (int s = 0; s < 3; ++s) sub[s * PV + idx] = sv[k][s];
      }
    }
  }
};

// One diffusion step on the padded lattice in LDS (planes PV floats apart,
// row stride RS).  `jac` and `fin` come from the plan for the same ROI.
// LANES = false: items are columns of bands and read their horizontal
// neighbours from LDS (plan_diffusion / MultiBand layout).  LANES = true:
// items are laid out in runs of consecutive lanes holding consecutive
// columns of one band, with a read-only item at each end of a run (mask 0),
// so every written site finds its left and right neighbours in the adjacent
// lanes and reads only its own column from LDS.
// `kept`: a TempGrid the caller prepared for this ROI and keeps between
// steps; without it one is set up for this step.
// UNROLL = 2: the substeps before the final one run in pairs, two copies of
// the substep body per loop pass (the compiler does not unroll a loop with
// barriers by itself); results are the same.
template <int NT, int L, int KMAX, int RS, int PS, bool LANES = false, int UNROLL = 1>
__device__ void diffuse(const DevParams &p, float *sub, int PV, const int roi[5],
                        const uint32_t jac[KMAX], const uint32_t fin[KMAX], uint32_t &updates,
                        TempGrid<NT, RS, PS> *kept = nullptr) {
  const int r0 = roi[0], c0 = roi[1], h = roi[2], w = roi[3];
  if (h == 0 || w == 0) return;
  TempGrid<NT, RS, PS> local;
  TempGrid<NT, RS, PS> &tg = kept ? *kept : local;
  if (kept) tg.apply(p, sub, PV, roi); else tg.setup(p, sub, PV, roi);
  const int S = p.substeps;
  const int total = S > 1 ? S : 1;
  {
    uint32_t na = 0, nf = 0;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      na += __popc((jac[k] >> 16) & 0x3fffu);
      nf += __popc((fin[k] >> 16) & 0x3fffu);
    }
    updates += na * (uint32_t)(total - 1) + nf;
  }
  const float R2 = (float)0.70710678118654752440;
  const float F = (float)(4.0 + (double)(float)(2.0 * 1.41421356237309504880));
  // Mask-box (0, 0) is padded (r0, c0).
  const int origin = r0 * RS + c0;
  __syncthreads();
  auto substep = [&](const bool final_sub) {
    float res[KMAX][L][3];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const uint32_t it = final_sub ? fin[k] : jac[k];
      if (LANES ? (it & kItemValid) == 0u : (it >> 16) == 0u) continue;
      const int j = LANES ? (int)(it & 0xffu) - 1 : (int)(it & 0xffu);
      const int i0 = (int)((it >> 8) & 0xffu);
      const int a0 = origin + (i0 - 1) * RS + j;
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const float *P = sub + s * PV + a0;
        const float kk = p.k[s];
        // Rolling window down the column: G(r) = fma(H(r), 1/sqrt2, C(r))
        // of the rows above (gu) and at (gm) the output row, its centre mc
        // and horizontal neighbour sum hm; each row's H and G once.
        float gu, mc, hm, gm;
        {
          const float uc = P[0];
          mc = P[RS];
          const float hu = LANES ? from_left_lane(uc) + from_right_lane(uc) : P[-1] + P[1];
          hm = LANES ? from_left_lane(mc) + from_right_lane(mc) : P[RS - 1] + P[RS + 1];
          gu = fmaf(hu, R2, uc);
          gm = fmaf(hm, R2, mc);
        }
#pragma unroll
        for (int l = 0; l < L; ++l) {
          const float *dn = P + (l + 2) * RS;
          const float dc = dn[0];
          const float hd = LANES ? from_left_lane(dc) + from_right_lane(dc) : dn[-1] + dn[1];
          const float gd = fmaf(hd, R2, dc);
          float Lp = (gu + gd) + hm;
          Lp = fmaf(-F, mc, Lp);
          res[k][l][s] = fmaf(kk, Lp, mc);
          gu = gm;
          mc = dc; hm = hd; gm = gd;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const uint32_t it = final_sub ? fin[k] : jac[k];
      const uint32_t m = (it >> 16) & 0x3fffu;
      if (m == 0u) continue;
      const int j = LANES ? (int)(it & 0xffu) - 1 : (int)(it & 0xffu);
      const int i0 = (int)((it >> 8) & 0xffu);
      float *P = sub + origin + i0 * RS + j;
#pragma unroll
      for (int l = 0; l < L; ++l) {
        if (m & (1u << l)) {
#pragma unroll
          for (int s = 0; s < 3; ++s) P[s * PV + l * RS] = res[k][l][s];
        }
      }
    }
    __syncthreads();
  };
  int step = 0;
  if (UNROLL == 2)
    for (; step + 2 < total; step += 2) {
      substep(false);
      substep(false);
    }
  for (; step < total; ++step) substep(step == total - 1);
  tg.restore(sub, PV);
  __syncthreads();
}

// Banded in-place diffusion whose band height is chosen per ROI.  For each
// candidate height the plan counts the band items and picks the height with
// the fewest wave instructions, estimated as (wave rounds of 64 items) x
// (21 L + 22) VALU operations per item; items are packed into waves in
// order (item q on thread q mod NT, round q / NT) so unused waves leave their
// SIMDs to other workgroups.
template <int NT, int RS>
struct MultiBand {
  static constexpr int kMaxItems(int L) { return ((RS - 2 + L - 1) / L) * (RS - 2); }
  static constexpr int kSlots(int L) { return (kMaxItems(L) + NT - 1) / NT; }
  static constexpr int kSlotsMax = kSlots(3);
};

// Band heights tried per ROI: 3, 4 and 6 rows.  (The hardware places the
// waves of co-resident workgroups on different SIMDs, so packing the items
// into the first waves of each workgroup does not load one SIMD.)
template <int NT, int RS, int PS>
struct MultiBandDiffuser {
  static constexpr int kPlanes = 3;
  static constexpr int KM = MultiBand<NT, RS>::kSlotsMax;
  uint32_t jac[KM], fin[KM];
  int L = 4;

  __device__ void set_sim(int) {}

  // Item index of this thread in slot k.
  __device__ int thread_q(int k) const { return (int)threadIdx.x + k * NT; }

  // Items of band height Lc for the chords in LDS; with item_q >= 0 also
  // returns that item's band and first column.
  __device__ static int count_items(const int32_t *chord, int hm, int wm, int Lc, int item_q,
                                    int &band, int &col) {
    int n = 0;
    for (int b = 0; b * Lc < hm; ++b) {
      int a = wm, e = 0;
      for (int l = 0; l < Lc && b * Lc + l < hm; ++l) {
        const int c = chord[b * Lc + l];
        if (c) { a = min(a, c & 0xffff); e = max(e, c >> 16); }
      }
      const int width = e > a ? e - a : 0;
      if (item_q >= n && item_q < n + width) { band = b; col = a + (item_q - n); }
      n += width;
    }
    return n;
  }

  __device__ void plan(const int roi[5], Shared &sm) {
    const int h = roi[2], w = roi[3], dist = roi[4];
#pragma unroll
    for (int k = 0; k < KM; ++k) { jac[k] = 0u; fin[k] = 0u; }
    if (h == 0 || w == 0) return;
    const int hm = h + 2, wm = w + 2, d4 = 4 * dist * dist;
    int32_t *chord = sm.red;  // per mask-box row: first | (past-last << 16)
    for (int i = threadIdx.x; i < hm; i += NT) {
      const int ai = 2 * (i + 1) - hm;
      const int rem = d4 - ai * ai;
      int a = wm, e = 0;
      if (rem > 0)
        for (int j = 0; j < wm; ++j) {
          const int bj = 2 * (j + 1) - wm;
          if (bj * bj < rem) { a = min(a, j); e = max(e, j + 1); }
        }
      chord[i] = e > a ? (a | (e << 16)) : 0;
    }
    __syncthreads();
    int best = 6, best_cost = -1, dummy0, dummy1;
    {
      const int lcs[3] = {3, 4, 6};
      const int slots[3] = {MultiBand<NT, RS>::kSlots(3), MultiBand<NT, RS>::kSlots(4),
                            MultiBand<NT, RS>::kSlots(6)};
#pragma unroll
      for (int ci = 0; ci < 3; ++ci) {
        const int Lc = lcs[ci];
        const int n = count_items(chord, hm, wm, Lc, -1, dummy0, dummy1);
        const int nf = ((h + Lc - 1) / Lc) * w;
        if (n > slots[ci] * NT || nf > slots[ci] * NT) continue;
        const int cost = ((n + 63) / 64) * (21 * Lc + 22);
        if (best_cost < 0 || cost < best_cost) { best = Lc; best_cost = cost; }
      }
    }
    L = best;
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      const int q = thread_q(k);
      int b = -1, j = 0;
      count_items(chord, hm, wm, L, q, b, j);
      if (b < 0) continue;
      const int i0 = b * L;
      uint32_t m = 0u;
      for (int l = 0; l < L; ++l) {
        const int i = i0 + l;
        if (i < hm && !masked(i, j, hm, wm, d4)) m |= 1u << l;
      }
      jac[k] = pack_item(j, i0, m);
    }
    const int nf = ((h + L - 1) / L) * w;
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      const int q = thread_q(k);
      if (q >= nf) continue;
      const int band = q / w, j = 1 + (q - band * w), i0 = 1 + band * L;
      uint32_t m = 0u;
      for (int l = 0; l < L; ++l)
        if (i0 + l <= h) m |= 1u << l;
      fin[k] = pack_item(j, i0, m);
    }
    __syncthreads();
  }

  template <int LL>
  __device__ void run_l(const DevParams &p, float *sub, int PV, const int roi[5], uint32_t &upd) {
    constexpr int K = MultiBand<NT, RS>::kSlots(LL);
    uint32_t j2[K], f2[K];
#pragma unroll
    for (int k = 0; k < K; ++k) { j2[k] = jac[k]; f2[k] = fin[k]; }
    diffuse<NT, LL, K, RS, PS>(p, sub, PV, roi, j2, f2, upd);
  }

  __device__ void run(const DevParams &p, float *sub, int PV, const int roi[5], uint32_t &upd) {
    if (L == 3) run_l<3>(p, sub, PV, roi, upd);
    else if (L == 4) run_l<4>(p, sub, PV, roi, upd);
    else run_l<6>(p, sub, PV, roi, upd);
  }
};

// Multi-band diffusion in the lane-run layout (diffuse<..., LANES = true>).
// The band height (3, 4 or 6 rows) is chosen per ROI like MultiBandDiffuser.
// A band's columns that hold unmasked sites form a run; the run, with one
// read-only column on each side, occupies consecutive lanes of one wave.  A
// run that does not fit the rest of a wave continues in the next one, the
// two columns at the cut being repeated so that both parts keep a read-only
// column at each end; a wave with fewer than 3 lanes left is not used.
template <int NT, int RS, int KCAP = 8>
struct LaneRuns {
  // Lanes a height-L layout can need: every band one run of at most RS
  // columns plus 2 read-only ones, 2 repeated columns per wave cut, up to 2
  // lanes lost per wave.
  static constexpr int kMaxLanes(int L) {
    return ((RS - 2 + L - 1) / L) * (RS + 2) + 4 * (((RS - 2 + L - 1) / L) * (RS + 2) / 60 + 2);
  }
  // Lane slots per thread for height L, at most KCAP (the results of every
  // slot stay in registers across a barrier); a plan uses a height only if
  // its layout fits.  Height 6 always fits.
  static constexpr int kSlots(int L) {
    return (kMaxLanes(L) + NT - 1) / NT < KCAP ? (kMaxLanes(L) + NT - 1) / NT : KCAP;
  }
  static constexpr int kSlotsUpTo(int L) { return L <= 0 ? 0 : (kSlots(L) > kSlotsUpTo(L - 1) ? kSlots(L) : kSlotsUpTo(L - 1)); }

  struct Lane {
    int run, col;
    bool owner;
  };

  // Places runs r = 0..nr-1 (columns [lo, hi) from runfn); returns the
  // lanes used and, for lane q >= 0, what it holds (run -1: nothing).
  template <typename RunFn>
  __device__ static int layout(int nr, const RunFn &runfn, int q, Lane &out) {
    out.run = -1;
    int pos = 0;
    for (int r = 0; r < nr; ++r) {
      int lo, hi;
      runfn(r, lo, hi);
      if (hi <= lo) continue;
      int c = lo - 1;
      for (;;) {
        const int room = 64 - (pos & 63);
        if (room < 3) {
          pos += room;
          continue;
        }
        const int take = min(room, hi + 1 - c);
        if (q >= pos && q < pos + take) {
          out.run = r;
          out.col = c + (q - pos);
          out.owner = q > pos && q < pos + take - 1;
        }
        pos += take;
        if (c + take - 1 >= hi) break;
        c += take - 2;
      }
    }
    return pos;
  }
};

// Band heights H1 < H2 < H3 are tried per ROI; layouts of height H3 always
// fit the KCAP lane slots.
template <int NT, int RS, int PS, int KCAP, int H1 = 3, int H2 = 4, int H3 = 6,
          int POLICY = 0, int UNROLL = 1>
struct LaneRunDiffuser {
  using LR = LaneRuns<NT, RS, KCAP>;
  static_assert(KCAP * NT >= LR::kMaxLanes(H3), "the tallest bands must fit");
  static_assert(RS < kRedInts, "the chord table and TempGrid's flag share sm.red");
  static constexpr int kPlanes = 3;
  static constexpr int KM = LR::kSlotsUpTo(H3);
  // Tallest mask box (rows) whose height-H3 layout has no more bands than
  // the lane slots were sized for (RS - 2 rows).
  static constexpr int kMaxBoxRows = H3 * ((RS - 2 + H3 - 1) / H3);
  static constexpr int kRingSites = PS * NT;  // TempGrid: 4 (h + w) + 16 at most
  uint32_t jac[KM], fin[KM];
  int L = 4;
  TempGrid<NT, RS, PS> tg;   // kept between ROI changes

  __device__ void set_sim(int) {}

  __device__ void plan(const int roi[5], Shared &sm) {
    const int h = roi[2], w = roi[3], dist = roi[4];
    tg.prepare(roi, sm);
#pragma unroll
    for (int k = 0; k < KM; ++k) { jac[k] = 0u; fin[k] = 0u; }
    if (h == 0 || w == 0) return;
    const int hm = h + 2, wm = w + 2, d4 = 4 * dist * dist;
    int32_t *chord = sm.red;  // per mask-box row: first | (past-last << 16)
    for (int i = threadIdx.x; i < hm; i += NT) {
      const int ai = 2 * (i + 1) - hm;
      const int rem = d4 - ai * ai;
      int a = wm, e = 0;
      if (rem > 0)
        for (int j = 0; j < wm; ++j) {
          const int bj = 2 * (j + 1) - wm;
          if (bj * bj < rem) { a = min(a, j); e = max(e, j + 1); }
        }
      chord[i] = e > a ? (a | (e << 16)) : 0;
    }
    __syncthreads();
    int Lc = H3;
    auto jac_runs = [&](int b, int &lo, int &hi) {
      int a = wm, e = 0;
      for (int l = 0; l < Lc && b * Lc + l < hm; ++l) {
        const int c = chord[b * Lc + l];
        if (c) { a = min(a, c & 0xffff); e = max(e, c >> 16); }
      }
      lo = a;
      hi = e > a ? e : a;
    };
    auto fin_runs = [&](int, int &lo, int &hi) { lo = 1; hi = w + 1; };
    typename LR::Lane dummy;
    int best = H3, best_cost = -1;
    {
      const int lcs[3] = {H1, H2, H3};
      const int slots[3] = {LR::kSlots(H1), LR::kSlots(H2), LR::kSlots(H3)};
#pragma unroll
      for (int ci = 0; ci < 3; ++ci) {
        Lc = lcs[ci];
        const int n = LR::layout((hm + Lc - 1) / Lc, jac_runs, -1, dummy);
        const int nf = LR::layout((h + Lc - 1) / Lc, fin_runs, -1, dummy);
        if (n > slots[ci] * NT || nf > slots[ci] * NT) continue;
        // POLICY 0: fewest wave instructions; 1: most waves (lowest bands);
        // 2: fewest wave instructions by the compiled stencil's count, 7 L +
        // 6 VALU per substrate and lane slot (3 x 34 + 4 at L = 4 in the ISA);
        // 3: as 2 plus four times the slots of the busiest wave (the
        // workgroup's first: lane slots fill thread by thread), which bounds
        // the time between the substep's barriers.
        const int waves = (n + 63) / 64, busiest = (n + NT - 1) / NT;
        const int cost = POLICY == 0 ? waves * (24 * Lc + 30)
                         : POLICY == 2 ? waves * (21 * Lc + 18)
                         : POLICY == 3 ? (4 * busiest + waves) * (21 * Lc + 18) : Lc;
        if (best_cost < 0 || cost < best_cost) { best = Lc; best_cost = cost; }
      }
    }
    // No height fits only for a mask box taller or wider than RS - 2, which
    // the kernels' launch checks exclude; flag it rather than diffuse a part.
    if (best_cost < 0 && threadIdx.x == 0) sm.status |= 2u;
    L = Lc = best;
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      typename LR::Lane ln;
      LR::layout((hm + L - 1) / L, jac_runs, (int)threadIdx.x + k * NT, ln);
      if (ln.run >= 0) {
        const int i0 = ln.run * L;
        uint32_t m = 0u;
        if (ln.owner)
          for (int l = 0; l < L; ++l) {
            const int i = i0 + l;
            if (i < hm && !masked(i, ln.col, hm, wm, d4)) m |= 1u << l;
          }
        jac[k] = pack_lane_item(ln.col, i0, m);
      }
      LR::layout((h + L - 1) / L, fin_runs, (int)threadIdx.x + k * NT, ln);
      if (ln.run >= 0) {
        const int i0 = 1 + ln.run * L;
        uint32_t m = 0u;
        if (ln.owner)
          for (int l = 0; l < L; ++l)
            if (i0 + l <= h) m |= 1u << l;
        fin[k] = pack_lane_item(ln.col, i0, m);
      }
    }
    __syncthreads();
  }

  template <int LL>
  __device__ void run_l(const DevParams &p, float *sub, int PV, const int roi[5], uint32_t &upd) {
    constexpr int K = LR::kSlots(LL);
    uint32_t j2[K], f2[K];
#pragma unroll
    for (int k = 0; k < K; ++k) { j2[k] = jac[k]; f2[k] = fin[k]; }
    diffuse<NT, LL, K, RS, PS, true, UNROLL>(p, sub, PV, roi, j2, f2, upd, &tg);
  }

  __device__ void run(const DevParams &p, float *sub, int PV, const int roi[5], uint32_t &upd) {
    if (L == H1) run_l<H1>(p, sub, PV, roi, upd);
    else if (L == H2) run_l<H2>(p, sub, PV, roi, upd);
    else run_l<H3>(p, sub, PV, roi, upd);
  }
};

// ---------------------------------------------------------------------------
// Pair-run layout: each lane holds two adjacent columns (ja, ja + 1) of one
// band, ja + padded column even, so a row of both is one 8-byte LDS read.
// The inner neighbour of each column is the lane's other column; the outer
// one comes from the adjacent lane through DPP.  A band's run of columns
// with unmasked sites, widened by one column on each side, becomes the
// pairs covering it, placed on consecutive lanes of one wave; a run cut at
// a wave boundary repeats the pair at the cut.  A column is written only by
// a lane whose neighbouring lane on that column's outer side belongs to the
// same run piece (so the first lane of a piece never writes ja, the last
// never writes ja + 1).
// Item word: ja + 2 (bits 0-7), first band row (8-15, 0xff: no item), two
// bits per row l (16 + 2l: write ja, 17 + 2l: write ja + 1).
// ---------------------------------------------------------------------------
constexpr uint32_t kNoPair = 0xff00u;

template <int NT, int RS, int KCAP>
struct PairRuns {
  static constexpr int kBands(int L) { return (RS - 2 + L - 1) / L; }
  // Pairs a height-L layout can need: per band at most RS / 2 + 2, one
  // repeated pair per wave cut, one lane lost per wave.
  static constexpr int kMaxLanes(int L) {
    return kBands(L) * (RS / 2 + 2) + 2 * (kBands(L) * (RS / 2 + 2) / 62 + 2);
  }
  static constexpr int kSlots(int L) {
    return (kMaxLanes(L) + NT - 1) / NT < KCAP ? (kMaxLanes(L) + NT - 1) / NT : KCAP;
  }

  struct Lane {
    int run, ja;
    bool own_a, own_b;
  };

  // Runs r = 0..nr-1 have owner columns [lo, hi) (runfn); e = parity of the
  // padded column of mask-box column 0.  Returns the lanes used; for lane
  // q >= 0 fills `out` (run -1: no item).
  template <typename RunFn>
  __device__ static int layout(int nr, const RunFn &runfn, int e, int q, Lane &out) {
    out.run = -1;
    int pos = 0;
    for (int r = 0; r < nr; ++r) {
      int lo, hi;
      runfn(r, lo, hi);
      if (hi <= lo) continue;
      // pair p covers columns 2p - e and 2p + 1 - e
      const int p0 = (lo - 1 + e + 2) / 2 - 1, p1 = (hi + e) / 2;
      int c = p0;
      for (;;) {
        const int room = 64 - (pos & 63);
        if (room < 2) {
          pos += room;
          continue;
        }
        const int take = min(room, p1 + 1 - c);
        if (q >= pos && q < pos + take) {
          const int t = q - pos, ja = 2 * (c + t) - e;
          out.run = r;
          out.ja = ja;
          out.own_a = t > 0 && ja >= lo && ja < hi;
          out.own_b = t < take - 1 && ja + 1 >= lo && ja + 1 < hi;
        }
        pos += take;
        if (c + take - 1 >= p1) break;
        c += take - 1;
      }
    }
    return pos;
  }
};

template <int NT, int L, int KMAX, int RS, int PS>
__device__ void diffuse_pairs(const DevParams &p, float *sub, int PV, const int roi[5],
                              const uint32_t jac[KMAX], const uint32_t fin[KMAX],
                              uint32_t &updates) {
  const int r0 = roi[0], c0 = roi[1], h = roi[2], w = roi[3];
  if (h == 0 || w == 0) return;
  TempGrid<NT, RS, PS> tg;
  tg.setup(p, sub, PV, roi);
  const int S = p.substeps;
  const int total = S > 1 ? S : 1;
  {
    uint32_t na = 0, nf = 0;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      na += __popc(jac[k] >> 16);
      nf += __popc(fin[k] >> 16);
    }
    updates += na * (uint32_t)(total - 1) + nf;
  }
  const float R2 = (float)0.70710678118654752440;
  const float F = (float)(4.0 + (double)(float)(2.0 * 1.41421356237309504880));
  const int origin = r0 * RS + c0;
  __syncthreads();
  for (int step = 0; step < total; ++step) {
    const bool final_sub = (step == total - 1);
    float ra[KMAX][L][3], rb[KMAX][L][3];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const uint32_t it = final_sub ? fin[k] : jac[k];
      if ((it & 0xff00u) == kNoPair) continue;
      const int ja = (int)(it & 0xffu) - 2, i0 = (int)((it >> 8) & 0xffu);
      const int a0 = origin + (i0 - 1) * RS + ja;
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const float2 *P = reinterpret_cast<const float2 *>(sub + s * PV + a0);
        const float kk = p.k[s];
        const float2 u = P[0], m = P[RS / 2];
        // Centre values and horizontal sums of the rows above, at and below.
        float mca = m.x, mcb = m.y;
        float hma = from_left_lane(m.y) + m.y, hmb = m.x + from_right_lane(m.x);
        float gua = fmaf(from_left_lane(u.y) + u.y, R2, u.x);
        float gub = fmaf(u.x + from_right_lane(u.x), R2, u.y);
        float gma = fmaf(hma, R2, mca), gmb = fmaf(hmb, R2, mcb);
#pragma unroll
        for (int l = 0; l < L; ++l) {
          const float2 d = P[(l + 2) * (RS / 2)];
          const float hda = from_left_lane(d.y) + d.y, hdb = d.x + from_right_lane(d.x);
          const float gda = fmaf(hda, R2, d.x), gdb = fmaf(hdb, R2, d.y);
          float La = (gua + gda) + hma, Lb = (gub + gdb) + hmb;
          La = fmaf(-F, mca, La);
          Lb = fmaf(-F, mcb, Lb);
          ra[k][l][s] = fmaf(kk, La, mca);
          rb[k][l][s] = fmaf(kk, Lb, mcb);
          gua = gma; gub = gmb;
          mca = d.x; mcb = d.y; hma = hda; hmb = hdb; gma = gda; gmb = gdb;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const uint32_t it = final_sub ? fin[k] : jac[k];
      const uint32_t m = it >> 16;
      if (m == 0u) continue;
      const int ja = (int)(it & 0xffu) - 2, i0 = (int)((it >> 8) & 0xffu);
      float *P = sub + origin + i0 * RS + ja;
#pragma unroll
      for (int l = 0; l < L; ++l) {
        const uint32_t two = (m >> (2 * l)) & 3u;
        if (two == 3u) {
#pragma unroll
          for (int s = 0; s < 3; ++s)
            *reinterpret_cast<float2 *>(P + s * PV + l * RS) = make_float2(ra[k][l][s], rb[k][l][s]);
        } else if (two == 1u) {
#pragma unroll
          for (int s = 0; s < 3; ++s) P[s * PV + l * RS] = ra[k][l][s];
        } else if (two == 2u) {
#pragma unroll
          for (int s = 0; s < 3; ++s) P[s * PV + l * RS + 1] = rb[k][l][s];
        }
      }
    }
    __syncthreads();
  }
  tg.restore(sub, PV);
  __syncthreads();
}

// Multi-band diffusion in the pair-run layout; band heights H1 < H2 < H3
// tried per ROI, H3 always fitting KCAP slots per thread.
template <int NT, int RS, int PS, int KCAP, int H1 = 4, int H2 = 6, int H3 = 8,
          int POLICY = 0>
struct PairRunDiffuser {
  using PR = PairRuns<NT, RS, KCAP>;
  static_assert(KCAP * NT >= PR::kMaxLanes(H3), "the tallest bands must fit");
  static_assert(H3 <= 8, "two mask bits per row in 16 bits");
  static_assert(RS % 2 == 0, "pairs need an even row stride");
  static constexpr int kPlanes = 3;
  static constexpr int KM = KCAP;
  uint32_t jac[KM], fin[KM];
  int L = H3;

  __device__ void set_sim(int) {}

  __device__ void plan(const int roi[5], Shared &sm) {
    const int c0 = roi[1], h = roi[2], w = roi[3], dist = roi[4];
#pragma unroll
    for (int k = 0; k < KM; ++k) { jac[k] = kNoPair; fin[k] = kNoPair; }
    if (h == 0 || w == 0) return;
    const int hm = h + 2, wm = w + 2, d4 = 4 * dist * dist;
    const int e = c0 & 1;  // parity of the padded column of mask-box column 0
    int32_t *chord = sm.red;  // per mask-box row: first | (past-last << 16)
    for (int i = threadIdx.x; i < hm; i += NT) {
      const int ai = 2 * (i + 1) - hm;
      const int rem = d4 - ai * ai;
      int a = wm, en = 0;
      if (rem > 0)
        for (int j = 0; j < wm; ++j) {
          const int bj = 2 * (j + 1) - wm;
          if (bj * bj < rem) { a = min(a, j); en = max(en, j + 1); }
        }
      chord[i] = en > a ? (a | (en << 16)) : 0;
    }
    __syncthreads();
    int Lc = H3;
    auto jac_runs = [&](int b, int &lo, int &hi) {
      int a = wm, en = 0;
      for (int l = 0; l < Lc && b * Lc + l < hm; ++l) {
        const int c = chord[b * Lc + l];
        if (c) { a = min(a, c & 0xffff); en = max(en, c >> 16); }
      }
      lo = a;
      hi = en > a ? en : a;
    };
    auto fin_runs = [&](int, int &lo, int &hi) { lo = 1; hi = w + 1; };
    typename PR::Lane dummy;
    int best = H3, best_cost = -1;
    {
      const int lcs[3] = {H1, H2, H3};
#pragma unroll
      for (int ci = 0; ci < 3; ++ci) {
        Lc = lcs[ci];
        const int n = PR::layout((hm + Lc - 1) / Lc, jac_runs, e, -1, dummy);
        const int nf = PR::layout((h + Lc - 1) / Lc, fin_runs, e, -1, dummy);
        if (n > KCAP * NT || nf > KCAP * NT) continue;
        // POLICY 0: fewest wave instructions; 1: most waves (lowest bands).
        const int cost = POLICY == 0 ? ((n + 63) / 64) * (42 * Lc + 40) : Lc;
        if (best_cost < 0 || cost < best_cost) { best = Lc; best_cost = cost; }
      }
    }
    L = Lc = best;
#pragma unroll
    for (int k = 0; k < KM; ++k) {
      typename PR::Lane ln;
      PR::layout((hm + L - 1) / L, jac_runs, e, (int)threadIdx.x + k * NT, ln);
      if (ln.run >= 0) {
        const int i0 = ln.run * L;
        uint32_t m = 0u;
        for (int l = 0; l < L; ++l) {
          const int i = i0 + l;
          if (i >= hm) break;
          if (ln.own_a && !masked(i, ln.ja, hm, wm, d4)) m |= 1u << (2 * l);
          if (ln.own_b && !masked(i, ln.ja + 1, hm, wm, d4)) m |= 2u << (2 * l);
        }
        jac[k] = (uint32_t)(ln.ja + 2) | ((uint32_t)i0 << 8) | (m << 16);
      }
      PR::layout((h + L - 1) / L, fin_runs, e, (int)threadIdx.x + k * NT, ln);
      if (ln.run >= 0) {
        const int i0 = 1 + ln.run * L;
        uint32_t m = 0u;
        for (int l = 0; l < L; ++l) {
          if (i0 + l > h) break;
          if (ln.own_a) m |= 1u << (2 * l);
          if (ln.own_b) m |= 2u << (2 * l);
        }
        fin[k] = (uint32_t)(ln.ja + 2) | ((uint32_t)i0 << 8) | (m << 16);
      }
    }
    __syncthreads();
  }

  template <int LL>
  __device__ void run_l(const DevParams &p, float *sub, int PV, const int roi[5], uint32_t &upd) {
    diffuse_pairs<NT, LL, KM, RS, PS>(p, sub, PV, roi, jac, fin, upd);
  }

  __device__ void run(const DevParams &p, float *sub, int PV, const int roi[5], uint32_t &upd) {
    if (L == H1) run_l<H1>(p, sub, PV, roi, upd);
    else if (L == H2) run_l<H2>(p, sub, PV, roi, upd);
    else run_l<H3>(p, sub, PV, roi, upd);
  }
};

// Banded in-place diffusion as a policy for the persistent kernel.
template <int NT, int L, int KMAX, int RS, int PS>
struct BandedDiffuser {
  static constexpr int kPlanes = 3;  // substrate planes in LDS
  uint32_t jac[KMAX], fin[KMAX];
  __device__ void set_sim(int) {}
  __device__ void plan(const int roi[5], Shared &sm) { plan_diffusion<NT, L, KMAX>(roi, jac, fin, sm); }
  __device__ void run(const DevParams &p, float *sub, int PV, const int roi[5], uint32_t &upd) {
    diffuse<NT, L, KMAX, RS, PS>(p, sub, PV, roi, jac, fin, upd);
  }
};

// ---------------------------------------------------------------------------
// Site-parallel diffusion with two LDS buffers.  The same temporary-grid
// values as diffuse() are set up in buffer A (rings saved to registers), the
// temporary grid is copied to buffer B, and each substep reads one buffer and
// writes the other, one lane per updated site, so a substep needs a single
// barrier and every wave is packed with sites.  Sites not updated by a
// substep hold the same value in both buffers.  The result lands in A (a
// copy back covers an odd substep count) and the rings are restored there.
// Sites are planned when the ROI changes: Jacobi sites are the unmasked
// sites of the mask box in row-major order, final sites the ROI's; each is
// kept as its LDS offset in the plane (0xffffffff: none).
// ---------------------------------------------------------------------------
template <int NT, int KJ, int KF>
__device__ void plan_sites(const int roi[5], int RS, uint32_t jac[KJ], uint32_t fin[KF],
                           Shared &sm) {
  const int r0 = roi[0], c0 = roi[1], h = roi[2], w = roi[3], dist = roi[4];
#pragma unroll
  for (int k = 0; k < KJ; ++k) jac[k] = 0xffffffffu;
#pragma unroll
  for (int k = 0; k < KF; ++k) fin[k] = 0xffffffffu;
  if (h == 0 || w == 0) return;
  const int hm = h + 2, wm = w + 2, d4 = 4 * dist * dist;
  int32_t *lo = sm.red, *cnt = sm.red + kRedInts / 2;
  const int tid = threadIdx.x;
  // Unmasked columns of row i: the chord b^2 < d4 - a^2, b = 2(j+1) - wm.
  for (int i = tid; i < hm; i += NT) {
    const int ai = 2 * (i + 1) - hm;
    const int rem = d4 - ai * ai;
    int a = wm, e = 0;
    if (rem > 0) {
      for (int j = 0; j < wm; ++j) {
        const int bj = 2 * (j + 1) - wm;
        if (bj * bj < rem) { a = min(a, j); e = max(e, j + 1); }
      }
    }
    lo[i] = a;
    cnt[i] = e > a ? e - a : 0;
  }
  __syncthreads();
  int n = 0;
  for (int i = 0; i < hm; ++i) n += cnt[i];
  const int origin = r0 * RS + c0;
#pragma unroll
  for (int k = 0; k < KJ; ++k) {
    const int q = tid + k * NT;
    if (q >= n) continue;
    int i = 0, base = 0;
    while (base + cnt[i] <= q) { base += cnt[i]; ++i; }
    const int j = lo[i] + (q - base);
    jac[k] = (uint32_t)(origin + i * RS + j);
  }
#pragma unroll
  for (int k = 0; k < KF; ++k) {
    const int q = tid + k * NT;
    if (q >= h * w) continue;
    const int i = 1 + q / w, j = 1 + q % w;
    fin[k] = (uint32_t)(origin + i * RS + j);
  }
  __syncthreads();
}

// One substep on the sites `site[]`: src -> dst (both plane sets, PV apart).
template <int K, int RS>
__device__ __forceinline__ void pp_substep(const DevParams &p, const float *src, float *dst,
                                           int PV, const uint32_t site[K]) {
  const float R2 = (float)0.70710678118654752440;
  const float F = (float)(4.0 + (double)(float)(2.0 * 1.41421356237309504880));
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (site[k] == 0xffffffffu) continue;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      const float *P = src + s * PV + site[k];
      const float mc = P[0];
      const float hu = P[-RS - 1] + P[-RS + 1];
      const float hm = P[-1] + P[1];
      const float hd = P[RS - 1] + P[RS + 1];
      const float gu = fmaf(hu, R2, P[-RS]), gd = fmaf(hd, R2, P[RS]);
      float Lp = (gu + gd) + hm;
      Lp = fmaf(-F, mc, Lp);
      dst[s * PV + site[k]] = fmaf(p.k[s], Lp, mc);
    }
  }
}

template <int NT, int KJ, int KF, int RS, int PS>
struct PingPongDiffuser {
  static constexpr int kPlanes = 6;  // two buffers of three planes
  uint32_t jac[KJ], fin[KF];
  __device__ void set_sim(int) {}
  __device__ void plan(const int roi[5], Shared &sm) { plan_sites<NT, KJ, KF>(roi, RS, jac, fin, sm); }
  __device__ void run(const DevParams &p, float *A, int PV, const int roi[5], uint32_t &upd) {
    const int r0 = roi[0], c0 = roi[1], h = roi[2], w = roi[3], dist = roi[4];
    if (h == 0 || w == 0) return;
    float *B = A + 3 * PV;
    const int tid = threadIdx.x;
    const int hm = h + 2, wm = w + 2, d4 = 4 * dist * dist;
    const int R0 = r0 - 1, R1 = r0 + h + 2, C0 = c0 - 1, C1 = c0 + w + 2;
    const int nper = 4 * w + 4 * h + 16;
    float sv[PS][3];
    int sidx[PS];
#pragma unroll
    for (int k = 0; k < PS; ++k) {
      int q = tid + k * NT;
      sidx[k] = -1;
      if (q < nper) {
        int pr, pc;
        bool outer;
        if (q < w + 4) { pr = R0; pc = C0 + q; outer = true; }
        else if ((q -= w + 4) < w + 4) { pr = R1; pc = C0 + q; outer = true; }
        else if ((q -= w + 4) < h + 2) { pr = R0 + 1 + q; pc = C0; outer = true; }
        else if ((q -= h + 2) < h + 2) { pr = R0 + 1 + q; pc = C1; outer = true; }
        else if ((q -= h + 2) < w + 2) { pr = R0 + 1; pc = C0 + 1 + q; outer = false; }
        else if ((q -= w + 2) < w + 2) { pr = R1 - 1; pc = C0 + 1 + q; outer = false; }
        else if ((q -= w + 2) < h) { pr = R0 + 2 + q; pc = C0 + 1; outer = false; }
        else { q -= h; pr = R0 + 2 + q; pc = C1 - 1; outer = false; }
        const int idx = pr * RS + pc;
        sidx[k] = idx;
#pragma unroll
        for (int s = 0; s < 3; ++s) sv[k][s] = A[s * PV + idx];
        if (outer) {
#pragma unroll
          for (int s = 0; s < 3; ++s) A[s * PV + idx] = 0.f;
        } else if (masked(pr - r0, pc - c0, hm, wm, d4)) {
#pragma unroll
          for (int s = 0; s < 3; ++s) A[s * PV + idx] = p.ext[s];
        }
      }
    }
    for (int q = tid; q < h * w; q += NT) {
      const int i = 1 + q / w, j = 1 + q % w;
      if (masked(i, j, hm, wm, d4)) {
        const int idx = (r0 + i) * RS + (c0 + j);
#pragma unroll
        for (int s = 0; s < 3; ++s) A[s * PV + idx] = p.ext[s];
      }
    }
    __syncthreads();
    // Temporary grid A -> B.
    const int th = h + 4, tw = w + 4;
    for (int q = tid; q < th * tw; q += NT) {
      const int idx = (R0 + q / tw) * RS + (C0 + q % tw);
#pragma unroll
      for (int s = 0; s < 3; ++s) B[s * PV + idx] = A[s * PV + idx];
    }
    const int S = p.substeps;
    const int total = S > 1 ? S : 1;
    {
      uint32_t na = 0, nf = 0;
#pragma unroll
      for (int k = 0; k < KJ; ++k) na += jac[k] != 0xffffffffu;
#pragma unroll
      for (int k = 0; k < KF; ++k) nf += fin[k] != 0xffffffffu;
      upd += na * (uint32_t)(total - 1) + nf;
    }
    __syncthreads();
    for (int step = 0; step < total - 1; ++step) {
      const float *src = (step & 1) ? B : A;
      float *dst = (step & 1) ? A : B;
      pp_substep<KJ, RS>(p, src, dst, PV, jac);
      __syncthreads();
    }
    {
      const bool odd = ((total - 1) & 1) != 0;
      const float *src = odd ? B : A;
      float *dst = odd ? A : B;
      pp_substep<KF, RS>(p, src, dst, PV, fin);
      __syncthreads();
      if (!odd) {
        // The final substep wrote B: copy the ROI back.
        for (int q = tid; q < h * w; q += NT) {
          const int idx = (r0 + 1 + q / w) * RS + (c0 + 1 + q % w);
#pragma unroll
          for (int s = 0; s < 3; ++s) A[s * PV + idx] = B[s * PV + idx];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < PS; ++k) {
      if (sidx[k] >= 0) {
#pragma unroll
        for (int s = 0; s < 3; ++s) A[s * PV + sidx[k]] = sv[k][s];
      }
    }
    __syncthreads();
  }
};

// ---------------------------------------------------------------------------
// Cell update (cellSimulationKernel, simulation.cu:66-110, with Site::step,
// site.h:62-82) under deterministic rules:
//   * every live slot (occupied site, newest list entry) is stepped once;
//   * the vacant-neighbour count uses the site states before the step;
//   * a cell ready to divide draws a neighbour; it is a candidate when that
//     neighbour is vacant after every death of the step;
//   * at most one division per lattice per step, won by the smallest
//     arbitration_key of the candidates' list positions;
//   * the child is appended to the occupied list.
// Site substrates are reached through `SubRef` (LDS in the persistent kernel,
// global memory in the per-step kernel).  During the step a dying site keeps
// its OCCUPIED bits with 0x80 set, so neighbour counts still see it occupied;
// the marks are cleared before the division is applied.
// ---------------------------------------------------------------------------
template <int RS>
struct SubLds {
  float *sub;
  int PV;
  __device__ float *at(int s, int r, int c) const { return sub + s * PV + (r + 1) * RS + (c + 1); }
};
struct SubGlobal {
  float *sub;
  int W, HW;
  __device__ float *at(int s, int r, int c) const { return sub + s * HW + r * W + c; }
};

template <int NT, typename SubRef>
__device__ void cells_step(const DevParams &p, const RngKey &key, uint32_t t, float dose,
                           const SubRef &sr, uint8_t *st, int H, int W, uint32_t *slots,
                           int cap, uint64_t *ovf, Shared &sm,
                           unsigned long long *instr = nullptr) {
  const int tid = threadIdx.x;
  if (tid == 0) { sm.ready_cnt = 0; sm.winner = kNoWinner; }
  __syncthreads();
  const int n = sm.n;
  uint32_t *S = slots + (int64_t)sm.cur * kSlotFields * cap;
  for (int i = tid; i < n; i += NT) {
    const uint32_t w = S[(int64_t)kPacked * cap + i];
    if (!slot_live(w)) continue;
    const int x = (int)slot_site(w);
    const int r = x / W, c = x - r * W;
    const int vn = ((st[x - W - 1] & 3) == kVacant) + ((st[x - W + 1] & 3) == kVacant) +
                   ((st[x + W - 1] & 3) == kVacant) + ((st[x + W + 1] & 3) == kVacant) +
                   ((st[x - 1] & 3) == kVacant) + ((st[x + 1] & 3) == kVacant) +
                   ((st[x - W] & 3) == kVacant) + ((st[x + W] & 3) == kVacant);
    CellL cl;
    cl.mode = slot_mode(w);
    cl.phase = slot_phase(w);
    cl.prolif = __uint_as_float(S[(int64_t)kProlif * cap + i]);
    cl.t_rep = __uint_as_float(S[(int64_t)kTRep * cap + i]);
    cl.R = __uint_as_float(S[(int64_t)kR * cap + i]);
    cl.t0 = __uint_as_float(S[(int64_t)kT0 * cap + i]);
    cl.t4 = __uint_as_float(S[(int64_t)kT4 * cap + i]);
    const uint32_t ph = cl.phase > 4u ? 4u : cl.phase;
    cl.tcur = ph == 0u ? cl.t0
            : ph == 4u ? cl.t4 : __uint_as_float(S[(int64_t)(kT0 + ph) * cap + i]);
    bool delay_loaded = false;
    float *ps0 = sr.at(0, r, c), *ps1 = sr.at(1, r, c), *ps2 = sr.at(2, r, c);
    float cho = *ps0, ox = *ps1, gi = *ps2;
    const float t_rep0 = cl.t_rep, R0 = cl.R;
    const float prolif0 = cl.prolif;
    bool alive = update_state(cl, cho, ox, gi, p, vn);
    bool ready = false;
    float delay = 0.f;
    if (alive) {
      if (dose > 0.f) {
        irradiate(cl, dose, p);
        delay = cl.delay;
        delay_loaded = true;
        if (instr) atomicAdd(&instr[kInstrIrradiationEvents], 1ull);
      }
      cho = cho - p.metab[cl.mode][0];
      ox = ox - p.metab[cl.mode][1];
      gi = gi - p.metab[cl.mode][2];
      const bool changed = progress_clock(cl, p);
      if (cl.t_rep > 0.f || (cl.R > 0.f && changed)) {
        cl.t_rep = cl.t_rep + p.hour_step;
        if (!delay_loaded) {
          delay = __uint_as_float(S[(int64_t)kDelay * cap + i]);
          delay_loaded = true;
        }
        if (cl.t_rep >= delay) {
          const float death_prob = 1.f - p.surv_coeff * expf_fixed(p.surv_exp * cl.R);
          const float u = uniform(key, t, (uint32_t)x, kPurposeRepair);
          if (instr) {
            atomicAdd(&instr[kInstrRepairDecisions], 1ull);
            atomicAdd(&instr[kInstrDeathProbFx],
                      (unsigned long long)((double)death_prob * 4294967296.0));
            atomicAdd(&instr[kInstrIrrHist + instr_bin(cl.R)], 1ull);
            atomicAdd(&instr[kInstrTirHist + instr_bin(cl.t_rep)], 1ull);
            if (u < death_prob) atomicAdd(&instr[kInstrRepairKills], 1ull);
          }
          if (u < death_prob) {
            alive = false;
          } else {
            cl.R = 0.f;
            cl.t_rep = 0.f;
            delay = 0.f;
          }
        }
      }
      *ps0 = cho; *ps1 = ox; *ps2 = gi;
      if (alive) ready = (cl.phase == 4u);
    }
    if (!alive) {
      st[x] = (uint8_t)(kOccupied | 0x80);
      S[(int64_t)kPacked * cap + i] = w & ~kLiveBit;
    } else {
      const uint32_t nw = pack_slot((uint32_t)x, cl.mode, cl.phase, true);
      if (nw != w) S[(int64_t)kPacked * cap + i] = nw;
      if (cl.prolif != prolif0) S[(int64_t)kProlif * cap + i] = __float_as_uint(cl.prolif);
      if (cl.t_rep != t_rep0) S[(int64_t)kTRep * cap + i] = __float_as_uint(cl.t_rep);
      if (cl.R != R0) S[(int64_t)kR * cap + i] = __float_as_uint(cl.R);
      if (delay_loaded) S[(int64_t)kDelay * cap + i] = __float_as_uint(delay);
      if (ready) {
        const float u = uniform(key, t, (uint32_t)x, kPurposeNeighbour);
        int dr, dc;
        choose_neighbour(u, dr, dc);
        const uint32_t child = (uint32_t)((r + dr) * W + (c + dc));
        const uint64_t ent = (uint64_t)(uint32_t)i | ((uint64_t)child << 16) |
                             ((uint64_t)__float_as_uint(cl.R) << 32);
        const int e = atomicAdd(&sm.ready_cnt, 1);
        if (e < kReadyLds) sm.ready[e] = ent; else ovf[e - kReadyLds] = ent;
      }
    }
  }
  __syncthreads();
  // Candidates: the child must be vacant after every death of this step.
  const int cnt = sm.ready_cnt;
  uint32_t best = kNoWinner;
  uint64_t best_ent = 0;
  for (int e = tid; e < cnt; e += NT) {
    const uint64_t ent = e < kReadyLds ? sm.ready[e] : ovf[e - kReadyLds];
    const uint32_t slot = (uint32_t)(ent & 0xffffu);
    const uint32_t child = (uint32_t)((ent >> 16) & 0xffffu);
    const uint8_t v = st[child];
    if ((v & 3) == kVacant || (v & 0x80)) {
      const uint32_t kk = arbitration_key(slot);
      if (kk < best) { best = kk; best_ent = ent; }
    }
  }
  if (best != kNoWinner) atomicMin(&sm.winner, best);
  // Clear the death marks (a mark or VACANT both read as vacant above).
  uint32_t *st32 = reinterpret_cast<uint32_t *>(st);
  const int nw32 = (H * W + 3) >> 2;
  for (int k = tid; k < nw32; k += NT) {
    const uint32_t v = st32[k];
    if (v & 0x80808080u) {
      uint32_t out = v;
#pragma unroll
      for (int b = 0; b < 4; ++b)
        if (v & (0x80u << (8 * b))) out &= ~(0xffu << (8 * b));
      st32[k] = out;
    }
  }
  __syncthreads();
  if (best != kNoWinner && best == sm.winner) {
    // divideCell / createCell (cell-division.h:10-30), simulation.cu:99-105.
    const uint32_t slot = (uint32_t)(best_ent & 0xffffu);
    const uint32_t child = (uint32_t)((best_ent >> 16) & 0xffffu);
    const float R = __uint_as_float((uint32_t)(best_ent >> 32));
    const uint32_t pw = S[(int64_t)kPacked * cap + slot];
    S[(int64_t)kPacked * cap + slot] = pack_slot(slot_site(pw), slot_mode(pw), 0u, true);
    S[(int64_t)kProlif * cap + slot] = 0u;
    float z[5];
    normals5(key, t, child, z);
    float acc = 0.f;
    const int nn = sm.n;
    if (nn < cap) {
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const float v = fmaf(p.cycle_sd[k], z[k], p.cycle_mean[k]);
        acc = (k == 0) ? v : acc + v;
        S[(int64_t)(kT0 + k) * cap + nn] = __float_as_uint(acc);
      }
      S[(int64_t)kTRep * cap + nn] = 0u;
      S[(int64_t)kR * cap + nn] = __float_as_uint(R);
      S[(int64_t)kDelay * cap + nn] = __float_as_uint(R > 0.f ? delay_time(R, p) : 0.f);
      S[(int64_t)kProlif * cap + nn] = 0u;
      S[(int64_t)kPacked * cap + nn] = pack_slot(child, 0u, 0u, true);
      st[child] = kOccupied;
      sm.n = nn + 1;
      sm.divisions += 1u;
      if (instr) atomicAdd(&instr[kInstrDivisions], 1ull);
    } else {
      sm.status |= 1u;
    }
  }
  __syncthreads();
}

}  // namespace dev
}  // namespace emt6ro

#endif  // EMT6RO_AMD_SIM_PHASES_H_
