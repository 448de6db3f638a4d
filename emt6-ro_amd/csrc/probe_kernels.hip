// LDS bandwidth probe: the peak against which the simulator's LDS traffic is
// reported.  Each workgroup fills 32 KB of LDS, then every lane reads it back
// in lane-contiguous order (no bank conflicts) with 4-byte or 16-byte loads,
// summing what it reads (v_add3_u32) so that no load can be dropped.  Thread 0 of
// each workgroup records the shader-clock and the constant-rate wall-clock
// ticks its loop took, so the probe reports both bytes per second and bytes
// per shader clock per CU.  No reference counterpart (measurement only).
// Also a grid-barrier probe (below).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "../../include/emt6ro_amd.h"
#include "host.h"

namespace emt6ro {
namespace {

constexpr int kProbeThreads = 512;
constexpr int kProbeWords = 8192;  // 32 KB of LDS per workgroup

template <int VEC>
__global__ __launch_bounds__(kProbeThreads) void lds_probe_kernel(uint32_t *sink, int iters,
                                                                  unsigned long long *ticks) {
  __shared__ uint32_t lds[kProbeWords];
  const int t = threadIdx.x;
  for (int i = t; i < kProbeWords; i += kProbeThreads) lds[i] = (uint32_t)(i * 2654435761u);
  __syncthreads();
  const unsigned long long c0 = clock64();
  const unsigned long long w0 = wall_clock64();
  uint32_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  // Each pass reads the whole 32 KB once: lane t takes words (or 16-byte
  // slots) t + 512 k, so the addresses are one base plus immediate offsets.
  // The empty asm with a memory clobber makes every pass load again.
  if (VEC == 4) {
    const uint4 *v = reinterpret_cast<const uint4 *>(lds) + t;
    for (int it = 0; it < iters; ++it) {
      asm volatile("" ::: "memory");
      const uint4 x0 = v[0], x1 = v[512], x2 = v[1024], x3 = v[1536];
      a0 += x0.x + x1.x + x2.x + x3.x;
      a1 += x0.y + x1.y + x2.y + x3.y;
      a2 += x0.z + x1.z + x2.z + x3.z;
      a3 += x0.w + x1.w + x2.w + x3.w;
    }
  } else {
    const uint32_t *v = lds + t;
    for (int it = 0; it < iters; ++it) {
      asm volatile("" ::: "memory");
      uint32_t x[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) x[k] = v[k * kProbeThreads];
      a0 += x[0] + x[1] + x[2] + x[3];
      a1 += x[4] + x[5] + x[6] + x[7];
      a2 += x[8] + x[9] + x[10] + x[11];
      a3 += x[12] + x[13] + x[14] + x[15];
    }
  }
  const unsigned long long c1 = clock64();
  const unsigned long long w1 = wall_clock64();
  sink[blockIdx.x * kProbeThreads + t] = a0 ^ a1 ^ a2 ^ a3;
  if (t == 0) {
    ticks[2 * blockIdx.x] = c1 - c0;
    ticks[2 * blockIdx.x + 1] = w1 - w0;
  }
}

// ---------------------------------------------------------------------------
// Grid-barrier probe: what one workgroup-wide barrier across the whole GPU
// costs inside a kernel, against ending one kernel and starting the next.
// One workgroup per CU (a cooperative launch, so all are resident) runs
// 2 * kPhases barriers; before the first barrier of a phase each workgroup
// writes `words` words of its own slice (dirty lines for the release fence
// to write back), after it reads the slice of workgroup (b + 37) % G and
// checks every word (visibility across XCDs).  Every wait is bounded by
// the wall clock (1 s); a workgroup that times out raises an error flag
// that makes all the others leave too.
// ---------------------------------------------------------------------------
constexpr int kLine = 32;        // words per 128-byte line
constexpr int kBarPhases = 100;
struct GridBar {
  unsigned central[kLine];
  unsigned top[kLine];
  unsigned grp[8 * kLine];       // per-group arrival counters, one line each
  unsigned rel[8 * kLine];       // per-group release words
  unsigned err[kLine];
  unsigned bad[kLine];
};

__device__ __forceinline__ unsigned bar_load(unsigned *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void bar_store(unsigned *p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned bar_add(unsigned *p) {
  return __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until *p >= v; false on the time limit or another workgroup's error.
__device__ __forceinline__ bool bar_wait(unsigned *p, unsigned v, GridBar *bar) {
  const uint64_t t0 = wall_clock64();
  while (bar_load(p) < v) {
    if (bar_load(bar->err) != 0u || wall_clock64() - t0 > 100000000ull) {
      bar_store(bar->err, 1u);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

// Barrier number k (1, 2, ...).  mode 0: one counter that every workgroup
// increments and polls; mode 1: eight group counters, the last arrival of a
// group increments a top counter, the last of those writes the eight group
// release words, and each workgroup polls its group's.  Both publish the
// workgroup's writes with an agent-scope release fence before arriving and
// take the others' with an acquire fence after leaving.
__device__ __forceinline__ bool grid_barrier(GridBar *bar, unsigned k, int mode) {
  const int G = (int)gridDim.x, g = (int)blockIdx.x & 7;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __syncthreads();
  __shared__ int ok;
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (mode == 0) {
      bar_add(bar->central);
      ok = bar_wait(bar->central, k * (unsigned)G, bar);
    } else {
      const unsigned gsize = (unsigned)((G - g + 7) / 8);
      if (bar_add(&bar->grp[g * kLine]) == k * gsize - 1u) {
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
        if (bar_add(bar->top) == k * 8u - 1u) {
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
          for (int j = 0; j < 8; ++j) bar_store(&bar->rel[j * kLine], k);
        }
      }
      ok = bar_wait(&bar->rel[g * kLine], k, bar);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  return ok != 0;
}

__global__ __launch_bounds__(256) void grid_barrier_kernel(GridBar *bar, uint32_t *buf, int words,
                                                           int mode) {
  const int G = (int)gridDim.x, b = (int)blockIdx.x;
  unsigned nbad = 0, k = 0;
  for (int p = 0; p < kBarPhases; ++p) {
    uint32_t *mine = buf + (size_t)b * words;
    for (int i = threadIdx.x; i < words; i += blockDim.x) mine[i] = (uint32_t)(p * G + b);
    if (!grid_barrier(bar, ++k, mode)) break;
    const int o = (b + 37) % G;
    const uint32_t *other = buf + (size_t)o * words;
    for (int i = threadIdx.x; i < words; i += blockDim.x)
      if (other[i] != (uint32_t)(p * G + o)) ++nbad;
    if (!grid_barrier(bar, ++k, mode)) break;
  }
  if (nbad) atomicAdd(bar->bad, nbad);
}

// The same writes, one phase per launch.
__global__ __launch_bounds__(256) void phase_write_kernel(uint32_t *buf, int words, int p) {
  uint32_t *mine = buf + (size_t)blockIdx.x * words;
  for (int i = threadIdx.x; i < words; i += blockDim.x) mine[i] = (uint32_t)(p * (int)gridDim.x + (int)blockIdx.x);
}

}  // namespace
}  // namespace emt6ro

extern "C" int emt6ro_probe_grid_barrier(int32_t device, int32_t mode, int32_t words,
                                         double out[2]) {
  using namespace emt6ro;
  constexpr int kMaxWords = 16384;  // 64 KB per workgroup
  if (!out || (mode != 0 && mode != 1) || words < 0 || words > kMaxWords) return EMT6RO_ERR_ARG;
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess) return EMT6RO_ERR_DEVICE;
  int cus = 0, coop = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device) != hipSuccess)
    return EMT6RO_ERR_DEVICE;
  if (!coop || cus < 1) return EMT6RO_ERR_DEVICE;
  const int G = cus;
  GridBar *bar = nullptr;
  uint32_t *buf = nullptr;
  if (hipMalloc(&bar, sizeof(GridBar)) != hipSuccess) return EMT6RO_ERR_DEVICE;
  if (hipMalloc(&buf, (size_t)G * kMaxWords * 4) != hipSuccess) {
    (void)hipFree(bar);
    return EMT6RO_ERR_DEVICE;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  float ms_bar = 0.f, ms_launch = 0.f;
  GridBar hb{};
  e = hipSuccess;
  for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {  // the first is a warm-up
    if ((e = hipMemset(bar, 0, sizeof(GridBar))) != hipSuccess) break;
    if ((e = hipDeviceSynchronize()) != hipSuccess) break;
    int w = words, m = mode;
    void *args[] = {&bar, &buf, &w, &m};
    (void)hipEventRecord(e0, nullptr);
    if ((e = hipLaunchCooperativeKernel((const void *)grid_barrier_kernel, dim3(G), dim3(256), args,
                                        0, nullptr)) != hipSuccess)
      break;
    (void)hipEventRecord(e1, nullptr);
    if ((e = hipEventSynchronize(e1)) != hipSuccess) break;
    e = hipEventElapsedTime(&ms_bar, e0, e1);
  }
  if (e == hipSuccess) e = hipMemcpy(&hb, bar, sizeof(GridBar), hipMemcpyDeviceToHost);
  for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
    (void)hipEventRecord(e0, nullptr);
    for (int p = 0; p < 2 * kBarPhases; ++p)
      hipLaunchKernelGGL(phase_write_kernel, dim3(G), dim3(256), 0, nullptr, buf, words, p);
    (void)hipEventRecord(e1, nullptr);
    if ((e = hipEventSynchronize(e1)) != hipSuccess) break;
    e = hipEventElapsedTime(&ms_launch, e0, e1);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(bar);
  (void)hipFree(buf);
  if (e != hipSuccess) return EMT6RO_ERR_DEVICE;
  if (hb.err[0] != 0u || hb.bad[0] != 0u) return EMT6RO_ERR_DEVICE;  // timed out or saw stale data
  out[0] = ms_bar * 1e3 / (2 * kBarPhases);     // microseconds per grid barrier
  out[1] = ms_launch * 1e3 / (2 * kBarPhases);  // microseconds per launch
  return EMT6RO_OK;
}

extern "C" int emt6ro_probe_lds(int32_t device, int32_t width_bytes, double out[4]) {
  using namespace emt6ro;
  if (!out || (width_bytes != 4 && width_bytes != 16)) return EMT6RO_ERR_ARG;
  auto fail = [](hipError_t) { return EMT6RO_ERR_DEVICE; };
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess) return fail(e);
  int cus = 0, wall_khz = 0;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device)) !=
      hipSuccess)
    return fail(e);
  if ((e = hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device)) !=
      hipSuccess)
    return fail(e);
  // Four resident workgroups per CU (16 waves each way of LDS and threads),
  // four rounds of them.
  const int blocks = cus * 16;
  const int iters = 6000;
  uint32_t *sink = nullptr;
  unsigned long long *ticks = nullptr;
  hipEvent_t e0, e1;
  if ((e = hipMalloc(&sink, sizeof(uint32_t) * blocks * kProbeThreads)) != hipSuccess)
    return fail(e);
  if ((e = hipMalloc(&ticks, sizeof(unsigned long long) * 2 * blocks)) != hipSuccess) {
    (void)hipFree(sink);
    return fail(e);
  }
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  auto launch = [&]() {
    if (width_bytes == 16)
      hipLaunchKernelGGL(lds_probe_kernel<4>, dim3(blocks), dim3(kProbeThreads), 0, nullptr,
                         sink, iters, ticks);
    else
      hipLaunchKernelGGL(lds_probe_kernel<1>, dim3(blocks), dim3(kProbeThreads), 0, nullptr,
                         sink, iters, ticks);
  };
  launch();  // warm-up
  (void)hipEventRecord(e0, nullptr);
  launch();
  (void)hipEventRecord(e1, nullptr);
  e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  std::vector<unsigned long long> h(2 * blocks);
  if (e == hipSuccess)
    e = hipMemcpy(h.data(), ticks, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(sink);
  (void)hipFree(ticks);
  if (e != hipSuccess) return fail(e);
  // Bytes read: per workgroup, iters x (16 loads of 4 B or 4 of 16 B) x 512 lanes.
  const double bytes = (double)blocks * iters * 64.0 * kProbeThreads;
  double clk = 0, wall = 0;
  for (int b = 0; b < blocks; ++b) { clk += (double)h[2 * b]; wall += (double)h[2 * b + 1]; }
  const double hz = wall > 0 ? clk / wall * (double)wall_khz * 1e3 : 0.0;  // shader clock
  const double sec = ms * 1e-3;
  out[0] = bytes / sec;                              // bytes per second, whole GPU
  out[1] = hz > 0 ? bytes / sec / hz / cus : 0.0;    // bytes per shader clock per CU
  out[2] = hz;                                       // shader clock under the probe
  out[3] = (double)cus;
  return EMT6RO_OK;
}
