// EMT6/Ro model types shared by the host library and the HIP kernels.
//
// Canonical per-simulation state in HBM (one pool, offsets in SimDesc):
//   sub   : 3 planes (cho, ox, gi) of H*W floats, row-major, row stride W
//   st    : H*W site-state bytes (VACANT / OCCUPIED / MOCKED), padded to 16 B
//   slots : 2 buffers x kSlotFields planes x cap 32-bit words.  Slot i holds the
//           cell of entry i of the occupied list (simulation.cu:41-64, 105):
//           the list is rebuilt in the reference's column-group order every
//           128 steps and a newborn cell is appended.  A rebuild gathers the
//           live cells into the other buffer, so a step reads its cells with
//           coalesced loads in list order.
//   hdr   : SimHdr, the counters that persist across runs
#ifndef EMT6RO_AMD_MODEL_H_
#define EMT6RO_AMD_MODEL_H_

#include <cstdint>

// Functions usable from both the host library and the kernels.
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define EMT_HD __host__ __device__
#else
#define EMT_HD
#endif

namespace emt6ro {

enum : uint8_t { kVacant = 0, kOccupied = 1, kMocked = 2 };

// Slot planes (cell/cell.h:46-52).  kPacked = site index (bits 0-15),
// mode (bits 16-17), phase (bits 18-20), live flag (bit 24).
enum SlotField : int {
  kTRep = 0, kR = 1, kDelay = 2, kProlif = 3,
  kT0 = 4, kT1 = 5, kT2 = 6, kT3 = 7, kT4 = 8,
  kPacked = 9, kSlotFields = 10
};
constexpr uint32_t kLiveBit = 1u << 24;

EMT_HD inline uint32_t pack_slot(uint32_t site, uint32_t mode, uint32_t phase,
                                              bool live) {
  return (site & 0xffffu) | ((mode & 3u) << 16) | ((phase & 7u) << 18) |
         (live ? kLiveBit : 0u);
}
EMT_HD inline uint32_t slot_site(uint32_t w) { return w & 0xffffu; }
EMT_HD inline uint32_t slot_mode(uint32_t w) { return (w >> 16) & 3u; }
EMT_HD inline uint32_t slot_phase(uint32_t w) { return (w >> 18) & 7u; }
EMT_HD inline bool slot_live(uint32_t w) { return (w & kLiveBit) != 0u; }

// RNG purposes: the Philox counter is (step, site | purpose << 24, sim, index)
// and the key is (seed, epoch).
enum : uint32_t { kPurposeRepair = 1, kPurposeNeighbour = 2, kPurposeNormal = 3 };

// Model parameters in the form the kernels use (parameters/parameters.h:10-72).
struct DevParams {
  float time_step;          // seconds per step
  float hour_step;          // time_step / 3600 (clock and repair increments)
  float death_margin;       // time_step / 3600 (updateState, cell.cu:74)
  float g1s_margin;         // 2 * time_step / 3600 (enterG1SStopping, cell.cu:68)
  float ext[3];             // external substrate levels
  float k[3];               // ((D * dt_diff) * HS) / F per substrate (diffusion.cu:163)
  int32_t substeps;         // (int16)(time_step / dt_diff)
  float metab[4][3];        // per metabolic mode: cho, ox, gi consumed per step
  float death_gi, quiescence_gi;
  float cycle_mean[5], cycle_sd[5];
  float delay_coeff, delay_exp, surv_coeff, surv_exp, repair_half_time;
};

// Per-simulation descriptor (fixed for the life of a batch).
struct SimDesc {
  int64_t sub_off;          // float offset of the substrate planes
  int64_t st_off;           // byte offset of the site states
  int64_t slot_off;         // word offset of the two slot buffers
  int32_t H, W;             // bordered lattice dims
  int32_t cap;              // slot capacity: interior sites + 128
  int32_t prot;             // protocol row
  uint32_t sim_id;          // RNG stream id
  uint32_t horizon;         // a run stops this simulation at this step
};
static_assert(sizeof(SimDesc) == 48, "SimDesc is 48 bytes");

// Per-simulation counters that survive between runs.
struct SimHdr {
  uint32_t step;            // steps done since reset
  int32_t n;                // occupied-list length (slots in use)
  int32_t cur;              // slot buffer holding the list (0 or 1)
  int32_t roi[5];           // r0, c0, h, w, dist (h == 0: no diffusion)
  uint32_t divisions;       // divisions performed
  uint32_t status;          // 0 ok; bit 0: slot overflow; bit 1: no diffusion layout fit
  uint64_t site_substeps;   // diffusion site-updates performed (roofline count)
  uint32_t pad[4];
};
static_assert(sizeof(SimHdr) == 64, "SimHdr is 64 bytes");

// Bytes of site-state storage per simulation (16-byte aligned).
EMT_HD inline int64_t st_bytes(int32_t H, int32_t W) {
  return ((int64_t)H * W + 15) & ~(int64_t)15;
}
EMT_HD inline int32_t slot_cap(int32_t H, int32_t W) {
  return (H - 2) * (W - 2) + 128;
}

}  // namespace emt6ro

#endif  // EMT6RO_AMD_MODEL_H_
