// Surface-element grids handed over by the caller as arrays (the reference's
// Point grids, crates/lcurve/src/types.rs:7-21, flattened as its GPU backend
// does in crates/lcurve-cuda/src/transfer.rs:134-159), kept resident in HBM,
// and comp_light (crates/lcurve/src/flux.rs:7-111) at the caller's phases.
#pragma once

#include <memory>
#include <vector>

#include "../../../include/lcurve_hip.h"
#include "engine.h"

namespace lch {

// Host grids and flux-path scalars of lch_model_upload's arguments
// (validated; throws std::invalid_argument on a malformed grid).
std::unique_ptr<lc::PreparedModel> prepared_from_grids(const lch_grid_soa* grids, const lch_flux_params& fp);

class GridModel {
 public:
  GridModel(Context& ctx, const lch_grid_soa* grids, const lch_flux_params& fp);
  ~GridModel();
  GridModel(const GridModel&) = delete;
  GridModel& operator=(const GridModel&) = delete;

  // out[i] = slfac[i] * comp_light(phase[i], expose[i], ndiv[i]) + third.
  // With `arena` the call's work tables are carved from that device memory
  // and nothing is allocated on the device; without, the tables are kept
  // for the next call at the same phases, exposures and sub-phase counts.
  void eval(const double* phase, const double* expose, const int* ndiv, const double* slfac, double third,
            double* out, int n, const BatchMemory* arena);
  // Device bytes the last eval in given memory used.
  size_t last_arena_bytes() const { return last_arena_; }

 private:
  std::unique_ptr<SetPack> lanes_pack(const double* phase, const double* expose, const int* ndiv,
                                      const double* slfac, double third, int n);

  Context& ctx_;
  std::unique_ptr<lc::PreparedModel> P_;
  std::unique_ptr<SetPack> rec_;   // plain records, segments, cones, moments
  std::vector<SegDev> segs_dev_;   // segments as on the device (linear-law rescaling)
  char* d_recs_ = nullptr;
  uint64_t rec_bytes_ = 0;
  std::unique_ptr<Batch> cache_;
  std::vector<double> key_phase_, key_expose_;
  std::vector<int> key_ndiv_;
  size_t last_arena_ = 0;
};

}  // namespace lch
