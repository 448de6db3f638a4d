// Host-visible entry points of the simulation kernels (sim_kernels.hip).
#ifndef EMT6RO_AMD_KERNELS_H_
#define EMT6RO_AMD_KERNELS_H_

#include <hip/hip_runtime.h>
#include <cstddef>
#include "model.h"

namespace emt6ro {
namespace dev { struct RunArgs; }

// A compiled shape of the simulation kernels: workgroup size `nt`, diffusion
// band rows `l`, diffusion items per thread `kmax`, ring-save slots per
// thread `ps` and LDS row stride `rs`.  A variant fits lattices of bordered
// dims H x W when H + 2 <= rs, W + 2 <= rs, ceil(H / l) * W <= kmax * nt and
// 4 (H + W) <= ps * nt.
struct VariantInfo {
  int id, nt, l, kmax, ps, rs;
};

const VariantInfo *variant_table(int *count);
inline bool variant_fits(const VariantInfo &v, int H, int W) {
  return H + 2 <= v.rs && W + 2 <= v.rs && ((H + v.l - 1) / v.l) * W <= v.kmax * v.nt &&
         4 * (H + W) <= v.ps * v.nt;
}

size_t persistent_lds_bytes(const VariantInfo &v, int hmax, int wmax);
size_t diffuse_lds_bytes(const VariantInfo &v, int hmax);
size_t cells_lds_bytes(int hmax, int wmax);

hipError_t launch_persistent_lds(int variant, const dev::RunArgs &a, int grid, size_t lds,
                                 hipStream_t s);
int persistent_occupancy(int variant, size_t lds);
hipError_t launch_step(int variant, const dev::RunArgs &a, int nsims, int hmax, int wmax,
                       hipStream_t s);
hipError_t launch_count_living(const SimDesc *desc, const uint8_t *st_pool, uint32_t *out,
                               int nsims, hipStream_t s);
hipError_t launch_fill_i32(int32_t *p, int32_t v, int n, hipStream_t s);

}  // namespace emt6ro

#endif  // EMT6RO_AMD_KERNELS_H_
