// Host-visible entry points of the simulation kernels (sim_kernels.hip).
#ifndef EMT6RO_AMD_KERNELS_H_
#define EMT6RO_AMD_KERNELS_H_

#include <hip/hip_runtime.h>
#include <cstddef>
#include "model.h"

namespace emt6ro {
namespace dev { struct RunArgs; }

// A compiled shape of the simulation kernels: diffusion scheme `kind`
// (0 banded in place, 1 ping-pong site-parallel), workgroup size `nt`, band
// rows `l` (1 for kind 1), diffusion items per thread `kmax` (kind 1: Jacobi
// sites), final-substep sites per thread `kf` (kind 1), ring-save slots per
// thread `ps` and LDS row stride `rs`.  It fits lattices of bordered dims
// H x W when H + 2 <= rs, W + 2 <= rs, 4 (H + W) <= ps * nt and the
// diffusion items fit: kind 0 ceil(H / l) * W <= kmax * nt; kind 1
// H * W <= kmax * nt and (H - 2) * (W - 2) <= kf * nt.
struct VariantInfo {
  int id, kind, nt, l, kmax, kf, ps, rs;
};

const VariantInfo *variant_table(int *count);
inline bool variant_fits(const VariantInfo &v, int H, int W) {
  if (H + 2 > v.rs || W + 2 > v.rs || 4 * (H + W) > v.ps * v.nt) return false;
  if (v.kind == 0) return ((H + v.l - 1) / v.l) * W <= v.kmax * v.nt;
  if (v.kind >= 2) return true;  // items sized for any lattice up to rs - 2
  return H * W <= v.kmax * v.nt && (H - 2) * (W - 2) <= v.kf * v.nt;
}

size_t persistent_lds_bytes(const VariantInfo &v, int hmax, int wmax);
size_t diffuse_lds_bytes(const VariantInfo &v, int hmax);
size_t cells_lds_bytes(int hmax, int wmax);

hipError_t launch_persistent_lds(int variant, const dev::RunArgs &a, int grid, size_t lds,
                                 hipStream_t s);
int persistent_occupancy(int variant, size_t lds);
// Window kernels (row stride rs = 36 or 56, for lattices wider than that;
// only the ROI's temporary grid in LDS): LDS bytes for `rows` window rows,
// workgroups per CU, launch.
size_t window_lds_bytes(int rs, int rows, int hmax, int wmax);
int window_occupancy(int rs, size_t lds);
hipError_t launch_window(int rs, const dev::RunArgs &a, int grid, size_t lds, hipStream_t s);
hipError_t launch_set_target(const int32_t *sims, int n, const SimHdr *hdr, const SimDesc *desc,
                             uint32_t *target, uint32_t k, hipStream_t s);
hipError_t launch_order_window(const int32_t *src, int32_t *dst, int n, const SimHdr *hdr,
                               const SimDesc *desc, const uint32_t *target, int rows, int rs,
                               int32_t *count, hipStream_t s);
hipError_t launch_step(int variant, const dev::RunArgs &a, int nsims, int hmax, int wmax,
                       hipStream_t s);
hipError_t launch_count_living(const SimDesc *desc, const uint8_t *st_pool, uint32_t *out,
                               int nsims, hipStream_t s);
hipError_t launch_fill_i32(int32_t *p, int32_t v, int n, hipStream_t s);
// Longest-first (largest ROI area) copy of the queue order src[0..n) into dst.
hipError_t launch_order_by_cost(const int32_t *src, int32_t *dst, int n, const SimHdr *hdr,
                                const SimDesc *desc, hipStream_t s);

}  // namespace emt6ro

#endif  // EMT6RO_AMD_KERNELS_H_
