// EMT6/Ro simulation kernels for gfx950.
//
//   persistent_kernel : one workgroup per simulation at a time, pulled from a
//                       work queue; the padded lattice substrates and site
//                       states stay in LDS for every step of the run, the
//                       cell slots stay in L2-resident global memory.
//   roi_kernel, diffuse_kernel, cells_kernel : the per-step path with the
//                       reference's structure (simulation.cu:153-180): one
//                       launch per phase per step, state in HBM between them.
//   count_living_kernel : living cells per simulation (statistics.cu:9-38).
#include <hip/hip_runtime.h>
#include <type_traits>
#include "kernels.h"
#include "sim_phases.h"

namespace emt6ro {
namespace dev {

// Loads a simulation's substrates into the padded LDS planes (zero ring) and
// its site states into LDS.
template <int NT, int RS>
__device__ void load_lattice(const RunArgs &a, const SimDesc &d, float *sub, int PV, uint8_t *st) {
  const int H = d.H, W = d.W, HW = H * W;
  const float *g = a.sub_pool + d.sub_off;
  for (int q = threadIdx.x; q < (H + 2) * RS; q += NT) {
    const int pr = q / RS, pc = q - pr * RS;
    const bool in = pr >= 1 && pr <= H && pc >= 1 && pc <= W;
    const int x = (pr - 1) * W + (pc - 1);
#pragma unroll
    for (int s = 0; s < 3; ++s) sub[s * PV + q] = in ? g[s * HW + x] : 0.f;
  }
  const uint32_t *gs = reinterpret_cast<const uint32_t *>(a.st_pool + d.st_off);
  uint32_t *ls = reinterpret_cast<uint32_t *>(st);
  for (int k = threadIdx.x; k < (HW + 3) / 4; k += NT) ls[k] = gs[k];
}

template <int NT, int RS>
__device__ void store_lattice(const RunArgs &a, const SimDesc &d, const float *sub, int PV,
                              const uint8_t *st) {
  const int H = d.H, W = d.W, HW = H * W;
  float *g = a.sub_pool + d.sub_off;
  for (int x = threadIdx.x; x < HW; x += NT) {
    const int r = x / W, c = x - r * W;
    const int q = (r + 1) * RS + (c + 1);
#pragma unroll
    for (int s = 0; s < 3; ++s) g[s * HW + x] = sub[s * PV + q];
  }
  uint32_t *gs = reinterpret_cast<uint32_t *>(a.st_pool + d.st_off);
  const uint32_t *ls = reinterpret_cast<const uint32_t *>(st);
  for (int k = threadIdx.x; k < (HW + 3) / 4; k += NT) gs[k] = ls[k];
}

// LDS of the persistent kernel: Shared | Diff::kPlanes substrate planes of
// pv_max floats (row stride RS; buffer A first) | site states.  The
// site-state bytes also keep the banded scheme's over-reads of its last band
// in bounds (persistent_lds_bytes).
template <int NT, int RS, typename Diff, int WPE>
__global__ __launch_bounds__(NT, WPE) void persistent_kernel(RunArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  Shared &sm = *reinterpret_cast<Shared *>(smem);
  float *sub = reinterpret_cast<float *>(smem + sizeof(Shared));
  const int PV = a.pv_max;
  uint8_t *st = reinterpret_cast<uint8_t *>(sub + Diff::kPlanes * PV);
  int32_t *pfx = a.scratch + (int64_t)blockIdx.x * a.scratch_stride;
  uint64_t *ovf = a.ovf + (int64_t)blockIdx.x * a.ovf_stride;
  const int tid = threadIdx.x;
  Diff diff;
  for (;;) {
    if (tid == 0) sm.sim = atomicAdd(a.queue, 1);
    __syncthreads();
    const int qi = sm.sim;
    if (qi >= a.nsims) break;
    const int sim = a.order[qi];
    const SimDesc d = a.desc[sim];
    SimHdr *hd = a.hdr + sim;
    const uint32_t step0 = hd->step;
    const uint32_t left = d.horizon > step0 ? d.horizon - step0 : 0u;
    const uint32_t nsteps = a.target ? (a.target[sim] > step0 ? a.target[sim] - step0 : 0u)
                                     : (a.nsteps < left ? a.nsteps : left);
    if (nsteps == 0u) {
      __syncthreads();
      continue;
    }
    const int H = d.H, W = d.W;
    load_lattice<NT, RS>(a, d, sub, PV, st);
    if (tid == 0) {
      sm.n = hd->n;
      sm.cur = hd->cur;
      for (int k = 0; k < 5; ++k) sm.roi[k] = hd->roi[k];
      sm.divisions = 0u;
      sm.status = 0u;
    }
    __syncthreads();
    uint32_t *slots = a.slot_pool + d.slot_off;
    const RngKey key{a.seed, a.epoch, d.sim_id};
    uint32_t updates = 0u;
    uint64_t updates64 = 0u;
    int roi[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) roi[q] = sm.roi[q];
    diff.set_sim(qi);
    diff.plan(roi, sm);
    for (uint32_t k = 0; k < nsteps; ++k) {
      const uint32_t t = step0 + k;
      if (t % 128u == 0u) rebuild_list<NT>(st, H, W, slots, d.cap, pfx, sm);
      if (t % 32u == 0u) {
        find_roi<NT>(st, H, W, sm);
        // The plan depends on the ROI only; most refreshes find it unchanged.
        bool same = true;
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          same = same && roi[q] == sm.roi[q];
          roi[q] = sm.roi[q];
        }
        if (!same) diff.plan(roi, sm);
      }
      diff.run(a.p, sub, PV, roi, updates);
      const float dose = dose_at(a, d.prot, t);
      cells_step<NT>(a.p, key, t, dose, SubLds<RS>{sub, PV}, st, H, W, slots, d.cap, ovf, sm,
                     a.instr);
      if ((k & 1023u) == 1023u) { updates64 += updates; updates = 0u; }
    }
    updates64 += updates;
    store_lattice<NT, RS>(a, d, sub, PV, st);
    const uint64_t tot = wg_reduce_sum_u64<NT>(updates64, sm.red);
    if (tid == 0) {
      hd->step = step0 + nsteps;
      hd->n = sm.n;
      hd->cur = sm.cur;
      for (int q = 0; q < 5; ++q) hd->roi[q] = sm.roi[q];
      hd->divisions += sm.divisions;
      hd->status |= sm.status;
      hd->site_substeps += tot;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Window kernel: the persistent kernel for lattices larger than its row
// stride RS, with only a window of NR x RS padded sites of the substrates in
// LDS (the site states stay whole).  The window holds the temporary grid of
// the ROI (the ROI and two rings: padded rows r0-1 .. r0+h+2, columns
// c0-1 .. c0+w+2); every substrate a step touches lies there: diffusion
// stays in the temporary grid, and cells lie in the ROI of the last refresh
// or, born since, one site outside it.  Window site (i, j) is padded site
// (wr + i, wc + j).  The diffusion and cell phases take ROI coordinates
// shifted into the window, so their addressing is the full-lattice one.
// When a refreshed ROI leaves the window it moves; when it no longer fits,
// the simulation stops before that step and a full-lattice launch takes it
// to its target.
// ---------------------------------------------------------------------------
// A window takes an ROI whose temporary grid (h + 4 rows, w + 4 columns)
// fits the window and whose mask box (h + 2 rows) is at most BOX rows, the
// tallest the diffusion layout's lane slots hold (they are sized for a
// lattice of row stride RS, however many rows the window has).
template <int RS, int BOX>
__device__ __forceinline__ bool window_origin(const int roi[5], int NR, int &wr, int &wc) {
  const int h = roi[2], w = roi[3];
  if (h == 0 || w == 0) { wr = 0; wc = 0; return true; }
  if (h + 4 > NR || h + 2 > BOX || w + 4 > RS) return false;
  wr = roi[0] - 1 - (NR - (h + 4)) / 2;
  wc = roi[1] - 1 - (RS - (w + 4)) / 2;
  return true;
}

template <int RS>
__device__ __forceinline__ bool window_holds(const int roi[5], int NR, int wr, int wc) {
  if (roi[2] == 0 || roi[3] == 0) return true;
  return roi[0] - 1 >= wr && roi[0] + roi[2] + 2 < wr + NR && roi[1] - 1 >= wc &&
         roi[1] + roi[3] + 2 < wc + RS;
}

template <int NT, int RS>
__device__ void load_window(const RunArgs &a, const SimDesc &d, float *win, int PV, int wr, int wc) {
  const int H = d.H, W = d.W, HW = H * W;
  const float *g = a.sub_pool + d.sub_off;
  for (int q = threadIdx.x; q < PV; q += NT) {
    const int qr = q / RS, pr = wr + qr, pc = wc + (q - qr * RS);
    const bool in = pr >= 1 && pr <= H && pc >= 1 && pc <= W;
    const int x = (pr - 1) * W + (pc - 1);
#pragma unroll
    for (int s = 0; s < 3; ++s) win[s * PV + q] = in ? g[s * HW + x] : 0.f;
  }
}

template <int NT, int RS>
__device__ void store_window(const RunArgs &a, const SimDesc &d, const float *win, int PV, int wr,
                             int wc) {
  const int H = d.H, W = d.W, HW = H * W;
  float *g = a.sub_pool + d.sub_off;
  for (int q = threadIdx.x; q < PV; q += NT) {
    const int qr = q / RS, pr = wr + qr, pc = wc + (q - qr * RS);
    if (pr >= 1 && pr <= H && pc >= 1 && pc <= W) {
      const int x = (pr - 1) * W + (pc - 1);
#pragma unroll
      for (int s = 0; s < 3; ++s) g[s * HW + x] = win[s * PV + q];
    }
  }
}

// Substrates of lattice site (r, c) in the window.
template <int RS>
struct SubWin {
  float *win;
  int PV, wr, wc;
  __device__ float *at(int s, int r, int c) const {
    return win + s * PV + (r + 1 - wr) * RS + (c + 1 - wc);
  }
};

template <int NT, int RS, typename Diff, int WPE>
__global__ __launch_bounds__(NT, WPE) void window_kernel(RunArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  Shared &sm = *reinterpret_cast<Shared *>(smem);
  float *win = reinterpret_cast<float *>(smem + sizeof(Shared));
  const int PV = a.pv_max;
  const int NR = PV / RS;
  uint8_t *st = reinterpret_cast<uint8_t *>(win + Diff::kPlanes * PV);
  int32_t *pfx = a.scratch + (int64_t)blockIdx.x * a.scratch_stride;
  uint64_t *ovf = a.ovf + (int64_t)blockIdx.x * a.ovf_stride;
  const int tid = threadIdx.x;
  const int nq = *a.nsims_dev;
  Diff diff;
  for (;;) {
    if (tid == 0) sm.sim = atomicAdd(a.queue, 1);
    __syncthreads();
    const int qi = sm.sim;
    if (qi >= nq) break;
    const int sim = a.order[qi];
    const SimDesc d = a.desc[sim];
    SimHdr *hd = a.hdr + sim;
    const uint32_t step0 = hd->step;
    const uint32_t tgt = a.target[sim];
    const uint32_t nsteps = tgt > step0 ? tgt - step0 : 0u;
    int roi[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) roi[q] = hd->roi[q];
    int wr = 0, wc = 0;
    if (nsteps == 0u || !window_origin<RS, Diff::kMaxBoxRows>(roi, NR, wr, wc)) {
      __syncthreads();
      continue;
    }
    const int H = d.H, W = d.W, HW = H * W;
    load_window<NT, RS>(a, d, win, PV, wr, wc);
    {
      const uint32_t *gs = reinterpret_cast<const uint32_t *>(a.st_pool + d.st_off);
      uint32_t *ls = reinterpret_cast<uint32_t *>(st);
      for (int k = tid; k < (HW + 3) / 4; k += NT) ls[k] = gs[k];
    }
    if (tid == 0) {
      sm.n = hd->n;
      sm.cur = hd->cur;
      for (int k = 0; k < 5; ++k) sm.roi[k] = roi[k];
      sm.divisions = 0u;
      sm.status = 0u;
    }
    __syncthreads();
    uint32_t *slots = a.slot_pool + d.slot_off;
    const RngKey key{a.seed, a.epoch, d.sim_id};
    uint32_t updates = 0u;
    uint64_t updates64 = 0u;
    int roiw[5] = {roi[0] - wr, roi[1] - wc, roi[2], roi[3], roi[4]};
    diff.set_sim(qi);
    diff.plan(roiw, sm);
    uint32_t k = 0;
    for (; k < nsteps; ++k) {
      const uint32_t t = step0 + k;
      // The ROI refresh comes before the list rebuild here (the rebuild does
      // not change the site states, so the ROI is the same), so that a
      // simulation leaving the window stops before any work of step t.
      if (t % 32u == 0u) {
        find_roi<NT>(st, H, W, sm);
        int nr[5];
        bool same = true;
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          nr[q] = sm.roi[q];
          same = same && nr[q] == roi[q];
        }
        if (!same) {
          if (!window_holds<RS>(nr, NR, wr, wc)) {
            int nwr, nwc;
            if (!window_origin<RS, Diff::kMaxBoxRows>(nr, NR, nwr, nwc)) break;
            store_window<NT, RS>(a, d, win, PV, wr, wc);
            // The new window is loaded from what the old one just stored:
            // the stores must be complete before any thread's loads.
            __threadfence_block();
            __syncthreads();
            wr = nwr;
            wc = nwc;
            load_window<NT, RS>(a, d, win, PV, wr, wc);
            __syncthreads();
          }
#pragma unroll
          for (int q = 0; q < 5; ++q) roi[q] = nr[q];
          roiw[0] = roi[0] - wr; roiw[1] = roi[1] - wc;
          roiw[2] = roi[2]; roiw[3] = roi[3]; roiw[4] = roi[4];
          diff.plan(roiw, sm);
        }
      }
      if (t % 128u == 0u) rebuild_list<NT>(st, H, W, slots, d.cap, pfx, sm);
      diff.run(a.p, win, PV, roiw, updates);
      const float dose = dose_at(a, d.prot, t);
      cells_step<NT>(a.p, key, t, dose, SubWin<RS>{win, PV, wr, wc}, st, H, W, slots, d.cap, ovf,
                     sm, a.instr);
      if ((k & 1023u) == 1023u) { updates64 += updates; updates = 0u; }
    }
    updates64 += updates;
    __syncthreads();
    store_window<NT, RS>(a, d, win, PV, wr, wc);
    {
      uint32_t *gs = reinterpret_cast<uint32_t *>(a.st_pool + d.st_off);
      const uint32_t *ls = reinterpret_cast<const uint32_t *>(st);
      for (int q = tid; q < (HW + 3) / 4; q += NT) gs[q] = ls[q];
    }
    const uint64_t tot = wg_reduce_sum_u64<NT>(updates64, sm.red);
    if (tid == 0) {
      hd->step = step0 + k;
      hd->n = sm.n;
      hd->cur = sm.cur;
      for (int q = 0; q < 5; ++q) hd->roi[q] = sm.roi[q];
      hd->divisions += sm.divisions;
      hd->status |= sm.status;
      hd->site_substeps += tot;
    }
    __syncthreads();
  }
}

// Diffusion policy of a variant: kind 0 banded in place (A = band rows,
// B = items per thread), kind 1 ping-pong site-parallel (A = Jacobi sites,
// B = final sites per thread).
template <int KIND, int NT, int A, int B, int RS, int PS>
struct DiffSel {
  using T = BandedDiffuser<NT, A, B, RS, PS>;
};
template <int NT, int A, int B, int RS, int PS>
struct DiffSel<1, NT, A, B, RS, PS> {
  using T = PingPongDiffuser<NT, A, B, RS, PS>;
};
template <int NT, int A, int B, int RS, int PS>
struct DiffSel<2, NT, A, B, RS, PS> {
  using T = MultiBandDiffuser<NT, RS, PS>;
};
template <int NT, int A, int B, int RS, int PS>
struct DiffSel<4, NT, A, B, RS, PS> {
  using T = typename std::conditional<B == 0, PairRunDiffuser<NT, RS, PS, A>,
                                      PairRunDiffuser<NT, RS, PS, A, 2, 4, 8, 1>>::type;
};
// Lane-run band heights and policy by code B: 0 3/4/6, 1 4/6/8 (fewest
// instructions); 2 4/6/8, 3 2/3/4, 4 3/4/6 (lowest bands that fit); 5 3/4/5
// (fewest instructions by the stencil's measured count).
// Band heights, height policy and diffusion-substep unroll (1, or 2: the
// substeps run in pairs) of a lane-run variant.
template <int B> struct LaneRunShape;
template <> struct LaneRunShape<0> { static constexpr int h1 = 3, h2 = 4, h3 = 6, pol = 0, unroll = 1; };
template <> struct LaneRunShape<1> { static constexpr int h1 = 4, h2 = 6, h3 = 8, pol = 0, unroll = 1; };
template <> struct LaneRunShape<2> { static constexpr int h1 = 4, h2 = 6, h3 = 8, pol = 1, unroll = 1; };
template <> struct LaneRunShape<3> { static constexpr int h1 = 2, h2 = 3, h3 = 4, pol = 1, unroll = 1; };
template <> struct LaneRunShape<4> { static constexpr int h1 = 3, h2 = 4, h3 = 6, pol = 1, unroll = 1; };
template <> struct LaneRunShape<5> { static constexpr int h1 = 3, h2 = 4, h3 = 5, pol = 2, unroll = 1; };
template <> struct LaneRunShape<6> { static constexpr int h1 = 4, h2 = 6, h3 = 8, pol = 1, unroll = 2; };
template <> struct LaneRunShape<7> { static constexpr int h1 = 3, h2 = 4, h3 = 5, pol = 2, unroll = 2; };
template <> struct LaneRunShape<8> { static constexpr int h1 = 3, h2 = 4, h3 = 6, pol = 1, unroll = 2; };
template <> struct LaneRunShape<9> { static constexpr int h1 = 4, h2 = 6, h3 = 8, pol = 3, unroll = 1; };
template <int NT, int A, int B, int RS, int PS>
struct DiffSel<3, NT, A, B, RS, PS> {
  using S = LaneRunShape<B>;
  using T = LaneRunDiffuser<NT, RS, PS, A, S::h1, S::h2, S::h3, S::pol, S::unroll>;
};


// ---------------------------------------------------------------------------
// Per-step path.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool sim_active(const RunArgs &a, int sim) {
  return a.hdr[sim].step < a.desc[sim].horizon;
}

template <int NT>
__global__ __launch_bounds__(NT) void roi_kernel(RunArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  Shared &sm = *reinterpret_cast<Shared *>(smem);
  uint8_t *st = smem + sizeof(Shared);
  const int sim = blockIdx.x;
  if (!sim_active(a, sim) || a.hdr[sim].step % 32u != 0u) return;
  const SimDesc d = a.desc[sim];
  const int HW = d.H * d.W;
  const uint32_t *gs = reinterpret_cast<const uint32_t *>(a.st_pool + d.st_off);
  uint32_t *ls = reinterpret_cast<uint32_t *>(st);
  for (int k = threadIdx.x; k < (HW + 3) / 4; k += NT) ls[k] = gs[k];
  __syncthreads();
  find_roi<NT>(st, d.H, d.W, sm);
  if (threadIdx.x < 5) a.hdr[sim].roi[threadIdx.x] = sm.roi[threadIdx.x];
}

// LDS: Shared | 3 planes of pv_max floats (row stride RS) | L*RS floats of
// slack for the last band's over-reads.
template <int NT, int L, int KMAX, int RS, int PS>
__global__ __launch_bounds__(NT) void diffuse_kernel(RunArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  Shared &sm = *reinterpret_cast<Shared *>(smem);
  float *sub = reinterpret_cast<float *>(smem + sizeof(Shared));
  const int PV = a.pv_max;
  const int sim = blockIdx.x;
  if (!sim_active(a, sim)) return;
  const SimDesc d = a.desc[sim];
  int roi[5];
  for (int q = 0; q < 5; ++q) roi[q] = a.hdr[sim].roi[q];
  if (roi[2] == 0 || roi[3] == 0) return;
  const int H = d.H, W = d.W, HW = H * W;
  // Temporary grid: padded rows r0-1 .. r0+h+2, cols c0-1 .. c0+w+2.
  const int R0 = roi[0] - 1, C0 = roi[1] - 1, th = roi[2] + 4, tw = roi[3] + 4;
  const float *g = a.sub_pool + d.sub_off;
  for (int q = threadIdx.x; q < th * tw; q += NT) {
    const int pr = R0 + q / tw, pc = C0 + q % tw;
    const bool in = pr >= 1 && pr <= H && pc >= 1 && pc <= W;
    const int x = (pr - 1) * W + (pc - 1);
#pragma unroll
    for (int s = 0; s < 3; ++s) sub[s * PV + pr * RS + pc] = in ? g[s * HW + x] : 0.f;
  }
  uint32_t jac[KMAX], fin[KMAX];
  // Also orders the loads above before diffuse() rewrites rings in place.
  plan_diffusion<NT, L, KMAX>(roi, jac, fin, sm);
  __syncthreads();
  uint32_t updates = 0u;
  diffuse<NT, L, KMAX, RS, PS>(a.p, sub, PV, roi, jac, fin, updates);
  float *gw = a.sub_pool + d.sub_off;
  for (int q = threadIdx.x; q < roi[2] * roi[3]; q += NT) {
    const int r = roi[0] + q / roi[3], c = roi[1] + q % roi[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) gw[s * HW + r * W + c] = sub[s * PV + (r + 1) * RS + (c + 1)];
  }
  const uint64_t tot = wg_reduce_sum_u64<NT>(updates, sm.red);
  if (threadIdx.x == 0) atomicAdd((unsigned long long *)&a.hdr[sim].site_substeps,
                                  (unsigned long long)tot);
}

template <int NT>
__global__ __launch_bounds__(NT) void cells_kernel(RunArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  Shared &sm = *reinterpret_cast<Shared *>(smem);
  uint8_t *st = smem + sizeof(Shared);
  const int sim = blockIdx.x;
  if (!sim_active(a, sim)) return;
  const SimDesc d = a.desc[sim];
  SimHdr *hd = a.hdr + sim;
  const int H = d.H, W = d.W, HW = H * W;
  uint32_t *gs = reinterpret_cast<uint32_t *>(a.st_pool + d.st_off);
  uint32_t *ls = reinterpret_cast<uint32_t *>(st);
  for (int k = threadIdx.x; k < (HW + 3) / 4; k += NT) ls[k] = gs[k];
  const uint32_t t = hd->step;
  if (threadIdx.x == 0) {
    sm.n = hd->n;
    sm.cur = hd->cur;
    sm.divisions = 0u;
    sm.status = 0u;
  }
  __syncthreads();
  uint32_t *slots = a.slot_pool + d.slot_off;
  int32_t *pfx = a.scratch + (int64_t)blockIdx.x * a.scratch_stride;
  uint64_t *ovf = a.ovf + (int64_t)blockIdx.x * a.ovf_stride;
  if (t % 128u == 0u) rebuild_list<NT>(st, H, W, slots, d.cap, pfx, sm);
  const RngKey key{a.seed, a.epoch, d.sim_id};
  const float dose = dose_at(a, d.prot, t);
  cells_step<NT>(a.p, key, t, dose, SubGlobal{a.sub_pool + d.sub_off, W, HW}, st, H, W, slots,
                 d.cap, ovf, sm, a.instr);
  for (int k = threadIdx.x; k < (HW + 3) / 4; k += NT) gs[k] = ls[k];
  if (threadIdx.x == 0) {
    hd->step = t + 1u;
    hd->n = sm.n;
    hd->cur = sm.cur;
    hd->divisions += sm.divisions;
    hd->status |= sm.status;
  }
}

// countLiving (statistics.cu:9-38)
template <int NT>
__global__ __launch_bounds__(NT) void count_living_kernel(const SimDesc *desc,
                                                          const uint8_t *st_pool,
                                                          uint32_t *out, int nsims) {
  __shared__ int32_t red[NT / 64];
  const int sim = blockIdx.x;
  if (sim >= nsims) return;
  const SimDesc d = desc[sim];
  const uint8_t *st = st_pool + d.st_off;
  int cnt = 0;
  for (int x = threadIdx.x; x < d.H * d.W; x += NT) cnt += (st[x] & 3) == kOccupied;
  for (int m = 32; m >= 1; m >>= 1) cnt += __shfl_xor(cnt, m, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < NT / 64; ++w) s += red[w];
    out[sim] = (uint32_t)s;
  }
}

__global__ void fill_i32_kernel(int32_t *p, int32_t v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Longest-first queue order for the next persistent launch: the simulations
// of src[0..n) into dst[0..n) by decreasing ROI area (the cost of the next
// steps is close to that of the last ones; before a simulation's first step
// its occupied-list length stands in), finished simulations last.  A
// counting sort over kOrderBuckets buckets of area / 8 in one workgroup;
// the order within a bucket is arbitrary.  Each simulation's results do not
// depend on when it runs.
constexpr int kOrderBuckets = 1024;
constexpr int kOrderThreads = 1024;
static_assert(kOrderThreads == kOrderBuckets && kOrderThreads % 64 == 0,
              "one bucket per thread in the prefix pass");

__device__ __forceinline__ int order_bucket(const SimHdr &h, const SimDesc &d) {
  if (h.step >= d.horizon) return 0;
  const int area = h.roi[2] * h.roi[3];
  return 1 + min((area > 0 ? area : h.n) >> 3, kOrderBuckets - 2);
}

__global__ __launch_bounds__(kOrderThreads) void order_by_cost_kernel(
    const int32_t *src, int32_t *dst, int n, const SimHdr *hdr, const SimDesc *desc) {
  __shared__ int32_t cnt[kOrderBuckets];
  __shared__ int32_t wsum[kOrderThreads / 64];
  const int tid = threadIdx.x;
  cnt[tid] = 0;
  __syncthreads();
  for (int i = tid; i < n; i += kOrderThreads) {
    const int sim = src[i];
    atomicAdd(&cnt[order_bucket(hdr[sim], desc[sim])], 1);
  }
  __syncthreads();
  // Exclusive prefix over buckets from the highest down: thread t holds
  // bucket kOrderBuckets - 1 - t.
  const int b = kOrderBuckets - 1 - tid;
  const int c = cnt[b];
  int incl = c;
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(incl, o, 64);
    if ((tid & 63) >= o) incl += v;
  }
  if ((tid & 63) == 63) wsum[tid >> 6] = incl;
  __syncthreads();
  int base = 0;
  for (int w = 0; w < (tid >> 6); ++w) base += wsum[w];
  __syncthreads();
  cnt[b] = base + incl - c;
  __syncthreads();
  for (int i = tid; i < n; i += kOrderThreads) {
    const int sim = src[i];
    dst[atomicAdd(&cnt[order_bucket(hdr[sim], desc[sim])], 1)] = sim;
  }
}

// Per-simulation target of a launch: k steps further, capped by the horizon.
__global__ void set_target_kernel(const int32_t *sims, int n, const SimHdr *hdr,
                                  const SimDesc *desc, uint32_t *target, uint32_t k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int sim = sims[i];
  const uint32_t step = hdr[sim].step, hz = desc[sim].horizon;
  const uint32_t left = hz > step ? hz - step : 0u;
  target[sim] = step + (k < left ? k : left);
}

// The window kernel's queue: the simulations short of their target whose
// ROI's temporary grid fits an NR x RS window with a site to spare on each
// side, longest first; the count goes to *count.
__global__ __launch_bounds__(kOrderThreads) void order_window_kernel(
    const int32_t *src, int32_t *dst, int n, const SimHdr *hdr, const SimDesc *desc,
    const uint32_t *target, int nr, int rs, int box, int32_t *count) {
  __shared__ int32_t cnt[kOrderBuckets];
  __shared__ int32_t wsum[kOrderThreads / 64];
  const int tid = threadIdx.x;
  auto bucket = [&](int sim) -> int {
    const SimHdr &h = hdr[sim];
    if (h.step >= target[sim]) return -1;
    const int hh = h.roi[2], ww = h.roi[3];
    // One site of room on each side for the ROI to grow during the launch
    // (within the mask-box rows the window kernel's layout holds).
    if (hh != 0 && ww != 0 && (hh + 6 > nr || hh + 4 > box || ww + 6 > rs)) return -1;
    return order_bucket(h, desc[sim]);
  };
  cnt[tid] = 0;
  __syncthreads();
  for (int i = tid; i < n; i += kOrderThreads) {
    const int b = bucket(src[i]);
    if (b >= 0) atomicAdd(&cnt[b], 1);
  }
  __syncthreads();
  const int b = kOrderBuckets - 1 - tid;
  const int c = cnt[b];
  int incl = c;
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(incl, o, 64);
    if ((tid & 63) >= o) incl += v;
  }
  if ((tid & 63) == 63) wsum[tid >> 6] = incl;
  __syncthreads();
  int base = 0;
  for (int w = 0; w < (tid >> 6); ++w) base += wsum[w];
  int total = 0;
  for (int w = 0; w < kOrderThreads / 64; ++w) total += wsum[w];
  __syncthreads();
  cnt[b] = base + incl - c;
  if (tid == 0) *count = total;
  __syncthreads();
  for (int i = tid; i < n; i += kOrderThreads) {
    const int sim = src[i];
    const int bk = bucket(sim);
    if (bk >= 0) dst[atomicAdd(&cnt[bk], 1)] = sim;
  }
}

}  // namespace dev

// ---------------------------------------------------------------------------
// Host-side launchers.
// ---------------------------------------------------------------------------
namespace {

size_t shared_bytes() { return sizeof(dev::Shared); }

template <typename K>
hipError_t launch_lds(K k, int grid, int nt, size_t lds, hipStream_t s, const dev::RunArgs &a) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(grid), dim3(nt), lds, s, a);
  return hipGetLastError();
}

template <int NT>
hipError_t launch_small(bool cells, const dev::RunArgs &a, int nsims, size_t lds, hipStream_t s) {
  if (cells) hipLaunchKernelGGL(dev::cells_kernel<NT>, dim3(nsims), dim3(NT), lds, s, a);
  else hipLaunchKernelGGL(dev::roi_kernel<NT>, dim3(nsims), dim3(NT), lds, s, a);
  return hipGetLastError();
}

}  // namespace

// Kernel variants: id, kind (0 banded in place, 1 ping-pong site-parallel,
// 2 banded with the band height chosen per ROI, 3 as 2 in the lane-run
// layout with horizontal neighbours from adjacent lanes, 4 as 3 with two
// columns per lane),
// workgroup size, A, B (kind 0: band rows, items per thread; kind 1: Jacobi
// and final sites per thread; kind 3: A = lane slots per thread, B = band
// heights, policy and substep unroll (LaneRunShape; 45-48 are 32, 44, 43
// and 39 with the substeps in pairs, 49 is 32 with band policy 3); kind 4:
// A = lane slots per thread, B = heights 4/6/8 (0) or 2/4/8 preferring low
// bands (1)), ring-save
// slots per thread, LDS row stride (lattices up to stride-2 on a side),
// minimum waves per SIMD the compiler must allow (bounds the registers so
// that many workgroups fit a CU).
#define EMT6RO_VARIANTS(X)        \
  X(0, 0, 128, 4, 6, 4, 56, 1)    \
  X(1, 0, 256, 4, 3, 2, 56, 1)    \
  X(2, 0, 128, 8, 3, 4, 56, 1)    \
  X(3, 0, 256, 2, 6, 2, 56, 1)    \
  X(4, 1, 512, 6, 6, 1, 56, 4)    \
  X(5, 1, 256, 12, 11, 2, 56, 2)  \
  X(6, 0, 128, 4, 3, 3, 36, 1)    \
  X(7, 0, 64, 4, 5, 5, 36, 1)     \
  X(8, 1, 256, 5, 5, 2, 36, 4)    \
  X(9, 1, 512, 3, 2, 1, 36, 4)    \
  X(10, 0, 256, 8, 4, 3, 88, 1)   \
  X(11, 0, 512, 4, 4, 2, 88, 1)   \
  X(12, 0, 256, 4, 8, 3, 88, 1)   \
  X(13, 1, 512, 6, 6, 1, 56, 2)   \
  X(14, 2, 256, 0, 0, 2, 56, 4)   \
  X(15, 2, 128, 0, 0, 4, 56, 1)   \
  X(16, 2, 256, 0, 0, 2, 56, 1)   \
  X(17, 2, 256, 0, 0, 2, 36, 4)   \
  X(18, 2, 128, 0, 0, 3, 36, 1)   \
  X(19, 2, 256, 0, 0, 3, 88, 2)   \
  X(20, 3, 512, 2, 0, 1, 56, 4)   \
  X(21, 3, 512, 2, 0, 1, 56, 2)   \
  X(22, 3, 256, 3, 0, 2, 56, 2)   \
  X(23, 3, 256, 1, 0, 2, 36, 4)   \
  X(24, 3, 512, 3, 0, 2, 88, 2)   \
  X(25, 3, 256, 2, 1, 2, 56, 4)   \
  X(26, 3, 256, 2, 1, 2, 56, 3)   \
  X(27, 4, 256, 1, 0, 2, 56, 4)   \
  X(28, 4, 256, 1, 0, 2, 56, 3)   \
  X(29, 4, 128, 2, 0, 4, 56, 4)   \
  X(30, 4, 256, 1, 0, 2, 36, 4)   \
  X(31, 4, 256, 3, 0, 3, 88, 2)   \
  X(32, 3, 256, 2, 2, 2, 56, 4)   \
  X(33, 4, 256, 2, 1, 2, 56, 4)   \
  X(34, 3, 256, 2, 2, 2, 36, 4)   \
  X(35, 3, 256, 4, 3, 2, 56, 4)   \
  X(36, 3, 256, 3, 4, 2, 56, 4)   \
  X(37, 3, 512, 1, 2, 1, 56, 4)   \
  X(38, 3, 256, 2, 2, 2, 56, 3)   \
  X(39, 3, 512, 3, 4, 2, 88, 2)   \
  X(40, 3, 512, 3, 2, 2, 88, 2)   \
  X(41, 3, 256, 2, 4, 2, 36, 4)   \
  X(42, 3, 128, 4, 4, 3, 36, 4)   \
  X(43, 3, 128, 2, 2, 3, 36, 4)   \
  X(44, 3, 256, 3, 5, 2, 56, 4)   \
  X(45, 3, 256, 2, 6, 2, 56, 4)   \
  X(46, 3, 256, 3, 7, 2, 56, 4)   \
  X(47, 3, 128, 2, 6, 3, 36, 4)   \
  X(48, 3, 512, 3, 8, 2, 88, 2)   \
  X(49, 3, 256, 2, 9, 2, 56, 4)

const VariantInfo *variant_table(int *count) {
  static const VariantInfo table[] = {
#define X(id, kind, nt, A, B, ps, rs, wpe) \
  {id, kind, nt, kind == 0 ? A : kind == 1 ? 1 : 8,                                  \
   kind == 0 ? B : kind == 1 ? A : kind == 2 ? dev::MultiBand<nt, rs>::kSlotsMax : A,      \
   kind == 1 ? B : 0, ps, rs},
      EMT6RO_VARIANTS(X)
#undef X
  };
  *count = (int)(sizeof(table) / sizeof(table[0]));
  return table;
}

size_t persistent_lds_bytes(const VariantInfo &v, int hmax, int wmax) {
  const size_t planes = (v.kind == 1 ? 6 : 3) * (size_t)v.rs * (hmax + 2) * sizeof(float);
  const size_t stb = ((size_t)hmax * wmax + 15) & ~(size_t)15;
  const size_t slack = v.kind != 1 ? (size_t)v.l * v.rs * sizeof(float) + 16 : 0;
  return shared_bytes() + planes + (stb > slack ? stb : slack);
}

size_t diffuse_lds_bytes(const VariantInfo &v, int hmax) {
  const size_t planes = 3 * (size_t)v.rs * (hmax + 2) * sizeof(float);
  return shared_bytes() + planes + (size_t)(v.l + 1) * v.rs * sizeof(float);
}

size_t cells_lds_bytes(int hmax, int wmax) {
  return shared_bytes() + (((size_t)hmax * wmax + 15) & ~(size_t)15);
}

hipError_t launch_persistent_lds(int variant, const dev::RunArgs &a, int grid, size_t lds,
                                 hipStream_t s) {
  switch (variant) {
#define X(id, kind, nt, A, B, ps, rs, wpe)                                                      \
  case id:                                                                                \
    return launch_lds(                                                                    \
        dev::persistent_kernel<nt, rs, typename dev::DiffSel<kind, nt, A, B, rs, ps>::T, wpe>, \
        grid, nt, lds, s, a);
    EMT6RO_VARIANTS(X)
#undef X
    default: return hipErrorInvalidValue;
  }
}

int persistent_occupancy(int variant, size_t lds) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
  switch (variant) {
#define X(id, kind, nt, A, B, ps, rs, wpe)                                                          \
  case id:                                                                                    \
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(                                         \
        &nb,                                                                                  \
        reinterpret_cast<const void *>(                                                       \
            dev::persistent_kernel<nt, rs, typename dev::DiffSel<kind, nt, A, B, rs, ps>::T, wpe>), \
        nt, lds);                                                                             \
    break;
    EMT6RO_VARIANTS(X)
#undef X
    default: break;
  }
  return e == hipSuccess ? nb : 0;
}

hipError_t launch_step(int variant, const dev::RunArgs &a, int nsims, int hmax, int wmax,
                       hipStream_t s) {
  int cnt = 0;
  const VariantInfo *vt = variant_table(&cnt);
  if (variant < 0 || variant >= cnt) return hipErrorInvalidValue;
  const VariantInfo &v = vt[variant];
  const size_t small = cells_lds_bytes(hmax, wmax);
  auto small_launch = [&](bool cells) -> hipError_t {
    switch (v.nt) {
      case 64: return launch_small<64>(cells, a, nsims, small, s);
      case 128: return launch_small<128>(cells, a, nsims, small, s);
      case 512: return launch_small<512>(cells, a, nsims, small, s);
      default: return launch_small<256>(cells, a, nsims, small, s);
    }
  };
  // ROI (every 32 steps, decided per simulation inside the kernel).
  hipError_t e = small_launch(false);
  if (e != hipSuccess) return e;
  if (v.kind != 0) return hipErrorInvalidValue;
  const size_t lds = diffuse_lds_bytes(v, hmax);
  switch (variant) {
#define X(id, kind, nt, A, B, ps, rs, wpe)                                                         \
  case id:                                                                                  \
    if (kind == 0)                                                                          \
      e = launch_lds(dev::diffuse_kernel<nt, (kind == 0 ? A : 4), (kind == 0 ? B : 1), rs, ps>, \
                     nsims, nt, lds, s, a);                                                 \
    break;
    EMT6RO_VARIANTS(X)
#undef X
    default: e = hipErrorInvalidValue; break;
  }
  if (e != hipSuccess) return e;
  return small_launch(true);
}

hipError_t launch_count_living(const SimDesc *desc, const uint8_t *st_pool, uint32_t *out,
                               int nsims, hipStream_t s) {
  hipLaunchKernelGGL(dev::count_living_kernel<256>, dim3(nsims), dim3(256), 0, s, desc, st_pool,
                     out, nsims);
  return hipGetLastError();
}

hipError_t launch_order_by_cost(const int32_t *src, int32_t *dst, int n, const SimHdr *hdr,
                                const SimDesc *desc, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(dev::order_by_cost_kernel, dim3(1), dim3(dev::kOrderThreads), 0, s, src, dst,
                     n, hdr, desc);
  return hipGetLastError();
}

// The window kernels, lane-run diffusion, 4 waves per SIMD: row stride 36
// with 128 threads (eight workgroups per CU), band heights 3/4/5 chosen by
// the stencil's instruction count with the busiest wave's share weighted
// (policy 3), 3 lane slots (over the whole 53x53 workload, heights 3/4/5 by
// instruction count alone were 1.3% faster than heights 4/6/8, lowest first,
// in 2 slots, and policy 3 0.8% faster than that); row stride 56
// with 256 threads (four per CU, for lattices wider than 54 sites, whose
// whole-lattice kernel holds one workgroup per CU), heights 4/6/8.  (A
// 44-column window at six 128-thread workgroups per CU ran slower than the
// 53x53 whole-lattice kernel.)
#ifndef EMT6RO_WINDOW36_SHAPE
#define EMT6RO_WINDOW36_SHAPE 3, 3, 4, 5, 3
#endif
using WindowDiff36 = dev::LaneRunDiffuser<128, 36, 3, EMT6RO_WINDOW36_SHAPE>;
using WindowDiff56 = dev::LaneRunDiffuser<256, 56, 2, 2, 4, 6, 8, 1>;
// The window kernels take ROIs up to RS - 4 wide (window_origin) and at
// most kMaxBoxRows - 2 tall: their two rings must fit the ring slots.
static_assert(WindowDiff36::kRingSites >= 4 * (36 - 4 + WindowDiff36::kMaxBoxRows - 2) + 16,
              "ring slots of the 36-column window");
static_assert(WindowDiff56::kRingSites >= 4 * (56 - 4 + WindowDiff56::kMaxBoxRows - 2) + 16,
              "ring slots of the 56-column window");
#define EMT6RO_WINDOW_KERNEL_36 dev::window_kernel<128, 36, WindowDiff36, 4>
#define EMT6RO_WINDOW_KERNEL_56 dev::window_kernel<256, 56, WindowDiff56, 4>

size_t window_lds_bytes(int rs, int rows, int hmax, int wmax) {
  const size_t planes = 3 * (size_t)rs * rows * sizeof(float);
  const size_t stb = ((size_t)hmax * wmax + 15) & ~(size_t)15;
  const size_t slack = (size_t)8 * rs * sizeof(float) + 16;
  return shared_bytes() + planes + (stb > slack ? stb : slack);
}

int window_occupancy(int rs, size_t lds) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
  if (rs == 36)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &nb, reinterpret_cast<const void *>(EMT6RO_WINDOW_KERNEL_36), 128, lds);
  else if (rs == 56)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &nb, reinterpret_cast<const void *>(EMT6RO_WINDOW_KERNEL_56), 256, lds);
  return e == hipSuccess ? nb : 0;
}

hipError_t launch_window(int rs, const dev::RunArgs &a, int grid, size_t lds, hipStream_t s) {
  if (rs == 36) return launch_lds(EMT6RO_WINDOW_KERNEL_36, grid, 128, lds, s, a);
  if (rs == 56) return launch_lds(EMT6RO_WINDOW_KERNEL_56, grid, 256, lds, s, a);
  return hipErrorInvalidValue;
}

hipError_t launch_set_target(const int32_t *sims, int n, const SimHdr *hdr, const SimDesc *desc,
                             uint32_t *target, uint32_t k, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(dev::set_target_kernel, dim3((n + 255) / 256), dim3(256), 0, s, sims, n,
                     hdr, desc, target, k);
  return hipGetLastError();
}

hipError_t launch_order_window(const int32_t *src, int32_t *dst, int n, const SimHdr *hdr,
                               const SimDesc *desc, const uint32_t *target, int rows, int rs,
                               int32_t *count, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int box = rs == 36 ? WindowDiff36::kMaxBoxRows : WindowDiff56::kMaxBoxRows;
  hipLaunchKernelGGL(dev::order_window_kernel, dim3(1), dim3(dev::kOrderThreads), 0, s, src, dst,
                     n, hdr, desc, target, rows, rs, box, count);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t *p, int32_t v, int n, hipStream_t s) {
  hipLaunchKernelGGL(dev::fill_i32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

}  // namespace emt6ro
