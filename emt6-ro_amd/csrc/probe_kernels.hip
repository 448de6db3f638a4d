// LDS bandwidth probe: the peak against which the simulator's LDS traffic is
// reported.  Each workgroup fills 32 KB of LDS, then every lane reads it back
// in lane-contiguous order (no bank conflicts) with 4-byte or 16-byte loads,
// summing what it reads (v_add3_u32) so that no load can be dropped.  Thread 0 of
// each workgroup records the shader-clock and the constant-rate wall-clock
// ticks its loop took, so the probe reports both bytes per second and bytes
// per shader clock per CU.  No reference counterpart (measurement only).
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

}  // namespace
}  // namespace emt6ro

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
