// Host side of the flux engine: device context, batch packing (CPU setup
// output -> HBM layout of layout.h), launch, graph replay and autotuning.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "batch_build.h"
#include "kernels.h"

namespace lch {

// Raised for HIP runtime failures; the C ABI maps it to LCH_ERR_RUNTIME.
struct DeviceError : public std::runtime_error {
  explicit DeviceError(const std::string& m) : std::runtime_error(m) {}
};

// Raised when a batch's tables do not fit the device memory it was given.
struct ArenaFull : public std::runtime_error {
  explicit ArenaFull(const std::string& m) : std::runtime_error(m) {}
};

// Launch configuration of the flux kernel; chosen by the autotuner.
struct Tuning {
  uint32_t chunk = 1024;        // elements per work item
  // Limb lanes: 0 off; otherwise a no-eclipse (patch, tile) pair is split
  // per lane when at most `limb` of its lanes see the patch in part.
  int limb = 48;
  // The host packer's element-level check of the lanes the split's cone
  // test leaves undecided (on), or none: they stay limb lanes (the device
  // pack's rule).
  bool limb_check = true;
  bool tuned = false;
};

class Context {
 public:
  explicit Context(int device);
  ~Context();
  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  int cu_count() const { return cu_count_; }
  // This device's static scratch (lch_flux_eval_static), or nullptr with
  // the reason in scratch_error().
  void* scratch() const { return scratch_; }
  const char* scratch_error() const { return scratch_err_; }
  Tuning tuning;

 private:
  int device_ = 0;
  int cu_count_ = 0;
  hipStream_t stream_ = nullptr;
  void* scratch_ = nullptr;
  const char* scratch_err_ = "";
};

// Device memory a batch may use instead of allocating its own.
//  * recs: element records already resident (a batch of one set whose
//    pack carries no records; its segment offsets are relative to recs).
//  * arena: a region every other device buffer of the batch is carved
//    from, so that building and running the batch calls no hipMalloc. Such
//    a batch creates no events, side stream or graph, and its work items
//    are built once (no autotune, no rebuild).
struct BatchMemory {
  const char* recs = nullptr;
  uint64_t rec_bytes = 0;
  char* arena = nullptr;
  size_t arena_bytes = 0;
};

class Batch {
 public:
  // Takes the packed sets (host staging is released while uploading).
  Batch(Context& ctx, std::vector<std::unique_ptr<SetPack>> packs, size_t n_times,
        const BatchMemory* mem = nullptr);
  ~Batch();
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  // Rebuild the work-item table for a chunk size (keeps resident data).
  void build_items(uint32_t chunk);
  // Enqueue flux + assembly on the context stream (graph replay if enabled).
  void launch(int scale);
  // Launch, wait and copy results back (any pointer may be null).
  void run(int scale, double* chisq, double* wnok, double* sfac, double* wdwarf, double* flux);
  // Time `reps` launches of the flux kernel alone and of the whole batch
  // with HIP events on the context stream. Returns mean milliseconds.
  void time(int scale, int reps, double* ms_flux, double* ms_total);
  // GPU eclipse visibility of one grid of one set at `nphase` phases,
  // written as mask[p * n + e] in the grid's original element order.
  void visibility(int set, int grid, const double* phases, int nphase, uint8_t* mask,
                  uint32_t* n_out);
  // Pick the chunk size / launch form with the fastest flux kernel.
  void autotune(int scale);
  // Replace observations, [n_sets][n_times] each (null keeps the current).
  void set_obs(const double* flux2d, const double* ferr2d, const double* weight2d);
  // Replace the per-time polynomial factors [n_sets][n_times] and each
  // set's third light (a one-set batch of uploaded grids).
  void set_curve_terms(const double* slfac, double third);
  // Device bytes taken from the arena so far (0 without one).
  size_t arena_used() const { return arena_used_; }
  // Host copies of the batch's tables, for parity checks of other packers:
  // element records, segments, patch moments, work codes as packed, and the
  // work items and item moments of the current item build.
  struct Tables {
    std::vector<char> recs;
    std::vector<SegDev> segs;
    std::vector<double> moms, imoms;
    std::vector<uint8_t> codes;
    std::vector<ItemDev> items;
    std::vector<uint64_t> lane_masks;
    std::vector<LimbChunk> limb_chunks;
    std::vector<LimbItem> limb_items;
    std::vector<uint16_t> limb_idx;
    std::vector<uint32_t> limb_lanes;
  };
  Tables tables() const;

  uint32_t n_sets() const { return static_cast<uint32_t>(sets_.size()); }
  // Per-set results of the last run() (empty before the first).
  const std::vector<SetOut>& last_results() const { return last_out_; }
  const SetHost& set(uint32_t i) const { return host_[i]; }
  uint64_t visits() const;
  // Resident element records plus patch moments.
  uint64_t record_bytes() const { return rec_bytes_ + moms_.size() * sizeof(double); }
  uint64_t pairs_evaluated() const { return pairs_eval_; }
  uint64_t pairs_by_moments() const { return pairs_mom_; }
  // Of pairs_evaluated(): (element, lane) pairs of limb lanes.
  uint64_t pairs_limb() const { return pairs_limb_; }
  uint64_t n_limb_chunks() const { return limb_chunks_.size(); }
  uint64_t n_limb_items() const { return limb_items_.size(); }
  // (element, lane) slots the limb pass computes, padding included.
  uint64_t limb_slots() const { return limb_slots_; }
  // Moment sums evaluated per launch, per lane (one per item with moment
  // patches, one per split patch).
  uint64_t moment_evals() const { return mom_evals_; }
  uint64_t n_items() const { return n_items_; }
  uint64_t n_lanes() const { return n_lanes_; }
  size_t n_times() const { return n_times_; }
  uint32_t chunk() const { return chunk_; }
  bool use_graph = true;

 private:
  void init(std::vector<std::unique_ptr<SetPack>> packs);
  // Frees every device resource held; safe on a partly built batch.
  void release_all();
  void upload();
  void release_graph();
  template <class T>
  T* dalloc(size_t n);
  void dfree_all();
  // Device memory of the work-item tables: from the arena when there is
  // one (released with the batch), else hipMalloc (freed by item_free).
  void* item_alloc(size_t bytes, const char* what);
  void item_free(void* p);

  Context& ctx_;
  size_t n_times_;
  BatchMemory mem_;
  size_t arena_used_ = 0;
  bool items_built_ = false;
  std::vector<SetHost> host_;
  // Host mirrors (kept small: records are uploaded set by set).
  std::vector<SetDev> sets_;
  std::vector<SegDev> segs_;
  std::vector<double> lane_phi_;
  std::vector<float4> lane_e_;
  std::vector<uint32_t> lane_meta_, tile_cp_;
  std::vector<uint32_t> t_lane0_;
  std::vector<int> t_ndiv_;
  std::vector<double> t_slfac_, t_y_, t_ferr_, t_w_, t_phase_;
  bool any_iscale_ = false;
  std::vector<SetOut> last_out_;
  std::vector<uint32_t> chunk_info_;
  // Per-set patch codes (ptcode of SetPack), kept for rebuilding items.
  std::vector<uint8_t> ptcode_;
  std::vector<double> moms_;
  std::vector<double> item_moms_;  // summed moments of each item's PT_MOM patches
  std::vector<uint64_t> pt_base_;
  std::vector<uint32_t> pt_tiles_;
  std::vector<std::vector<SetPack::LaneSplit>> lsplit_;  // per set
  // Limb lanes of the current item build.
  std::vector<LimbChunk> limb_chunks_;
  std::vector<LimbItem> limb_items_;
  std::vector<uint16_t> limb_idx_;
  std::vector<uint32_t> limb_lanes_;
  std::vector<uint64_t> lane_masks_;
  uint64_t rec_bytes_ = 0, n_items_ = 0, n_lanes_ = 0, n_prows_ = 0, pairs_eval_ = 0, pairs_mom_ = 0,
           n_items_plain_ = 0, n_items_ecl_ = 0, pairs_limb_ = 0, limb_slots_ = 0, mom_evals_ = 0;
  void* d_limb_ = nullptr;  // one allocation for the limb tables
  uint32_t chunk_ = 0;

  DeviceBatch d_;
  std::vector<void*> allocs_;
  ItemDev* d_items_ = nullptr;
  double* d_imoms_ = nullptr;
  double* d_partials_ = nullptr;
  hipGraphExec_t graph_ = nullptr;
  int graph_scale_ = -1;
  hipEvent_t ev_[3] = {nullptr, nullptr, nullptr};
  // Side stream (default priority) and fork / join events for the item
  // kernels that run beside the plain one.
  SideStream side_;
};

}  // namespace lch
