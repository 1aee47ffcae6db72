// Device buffers of a resident batch and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include "layout.h"

namespace lch {

struct DeviceBatch {
  uint32_t n_sets = 0, n_items = 0;
  uint32_t max_set_lanes = 0;  // largest SetDev::special0 of a valid set
  uint32_t n_items_plain = 0;  // leading items with no beaming / lensing and no staged eclipse intervals
  uint32_t n_items_ecl = 0;    // then items with staged eclipse intervals (no beaming / lensing)
  const SetDev* sets = nullptr;
  const SegDev* segs = nullptr;
  const ItemDev* items = nullptr;
  const char* recs = nullptr;
  const double* lane_phi = nullptr;
  const float4* lane_e = nullptr;        // earth vector per lane
  const uint32_t* lane_meta = nullptr;
  const uint32_t* chunk_info = nullptr;  // use_mask | kind << 16, per chunk
  const double* moms = nullptr;          // MOM_STRIDE per patch
  const double* imoms = nullptr;         // MOM_STRIDE per item with PT_MOM patches (ItemDev::imom)
  const uint32_t* t_lane0 = nullptr;
  const int* t_ndiv = nullptr;
  const double* t_slfac = nullptr;
  const double* t_y = nullptr;
  const double* t_ferr = nullptr;
  const double* t_w = nullptr;
  const double* t_phase = nullptr;       // central phase per time
  // iscale work space (null when no set uses iscale): per-lane component
  // sums [lane][5], per-time columns [time][5], least-squares work [time][6].
  double* lane_comp = nullptr;
  double* comp = nullptr;
  double* lsw = nullptr;
  double* partials = nullptr;
  double* lane_tot = nullptr;
  double* flux = nullptr;
  SetOut* out = nullptr;
  // Limb lanes (none when n_limb_chunks == 0).
  uint32_t n_limb_chunks = 0;
  const LimbChunk* limb_chunks = nullptr;
  const LimbItem* limb_items = nullptr;
  const uint16_t* limb_idx = nullptr;    // 64 per limb item
  const uint32_t* limb_lanes = nullptr;  // global lane per chunk-list entry
  const uint64_t* lane_masks = nullptr;  // per pmom group of an item
};

// Work items, one workgroup (one wavefront) each; then the limb chunks,
// which add to the partial rows the items wrote. With a side stream, the
// plain items run on `stream` and the other item kernels on `side`, forked
// after and joined before the limb pass by the two events (also inside a
// graph capture of `stream`).
struct SideStream {
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};
hipError_t launch_flux_items(const DeviceBatch& d, hipStream_t stream, const SideStream* ss = nullptr);
// FP64 assembly: lane_kernel (per-lane sums, in parallel over sets and
// lane blocks), then assemble_kernel (per set: scale factors, smearing,
// autoscale and chi-squared).
hipError_t launch_assemble(const DeviceBatch& d, int scale, hipStream_t stream);
// Eclipse visibility of one segment at `nphase` phases (debug / parity).
hipError_t launch_visibility(const DeviceBatch& d, const SegDev& sg, const double* phases,
                             uint32_t nphase, uint8_t* mask, hipStream_t stream);

}  // namespace lch
