// A batch whose per-set pack runs on the GPU. The host keeps only what the
// pack does not compute: the CPU setup (the reference's grids, eclipses and
// brightness), each set's segment layout, lanes and per-time tables, and
// the raw element grids in struct-of-arrays form. Every step then builds,
// on the device, the element records in patch order, patch cones and
// moments, the per-(patch, tile) work codes, the work items and their
// moment sums (pack_kernels.hip), and runs the flux pass and the assembly
// on them. The tables equal the host pack's bit for bit for the same chunk
// size and limb-lane threshold (check()).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "batch_build.h"
#include "engine.h"
#include "kernels.h"
#include "pack_kernels.h"

namespace lch {

struct DevicePackStats {
  uint64_t n_sets = 0, n_valid = 0, n_segs = 0, n_grids = 0, n_elements = 0, n_pairs = 0, n_patches = 0;
  uint64_t n_tiles = 0, n_chunks = 0, n_codes = 0, n_cand = 0, n_prows = 0, n_lanes = 0, n_times = 0;
  uint64_t rec_bytes = 0;     // element records the pack writes
  uint64_t input_bytes = 0;   // raw grids and layout tables copied by upload()
  uint64_t work_bytes = 0;    // device tables the pack and the flux pass write, per slot
  uint64_t visits = 0;        // element-visits in the reference's loop order per pass
  uint64_t n_items = 0, n_items_plain = 0, n_items_ecl = 0;  // of the last pack
  uint64_t n_limb_chunks = 0, n_limb_items = 0;              // of the last pack
  // The per-lane split of the last run(): (patch, tile) pairs split and
  // their limb lanes (lanes that see the patch in part).
  uint64_t split_pairs = 0, split_limb_lanes = 0;
  // Work of the flux pass of the last run(): (element, lane) pairs
  // evaluated one by one, pairs summed from patch moments, moment sums
  // evaluated (per lane); of the first, the pairs of limb lanes.
  uint64_t pairs_evaluated = 0, pairs_by_moments = 0, moment_evals = 0, pairs_limb = 0;
  double setup_core_s = 0;    // CPU setup (prepare_model), summed over threads
  double layout_core_s = 0;   // host layout and grid flattening, summed over threads
  double build_wall_s = 0;    // both, on the setup threads
  double concat_s = 0;        // batch-wide tables from the per-set layouts
  double upload_s = 0;        // the last upload()
  uint32_t chunk = 0;
  uint32_t limb = 0;
  int threads = 0;
};

// The host layout of one set, before the batch-wide offsets.
struct SetLayout;

class DevicePackBatch {
 public:
  // CPU setup and host layout of every set on `nthreads` threads (0: all
  // cores or LCURVE_HIP_THREADS), batch-wide tables, device memory, then
  // upload(). `data` holds one data set shared by all models or one per
  // model. Work items cover `chunk` elements (a multiple of 64, at most
  // 4096). limb: 0 no limb lanes; otherwise a no-eclipse (patch, tile) pair
  // is split per lane when at most `limb` (<= 64) of its lanes see the
  // patch in part (Tuning::limb).
  DevicePackBatch(Context& ctx, const std::vector<lc::Model>& models, const std::vector<lc::Data>& data,
                  uint32_t chunk, int limb, int nthreads = 0);
  ~DevicePackBatch();
  DevicePackBatch(const DevicePackBatch&) = delete;
  DevicePackBatch& operator=(const DevicePackBatch&) = delete;

  // Copies the host-staged raw grids and layout tables to the device (one
  // copy from pinned memory) and waits.
  void upload();
  // Enqueues the pack on the context stream, waits for its item counts,
  // then enqueues the flux pass and the assembly (no copy back).
  void launch(int scale);
  // launch(), wait, copy back (any pointer may be null; flux [n_sets][n]).
  void run(int scale, double* chisq, double* wnok, double* sfac, double* wdwarf, double* flux);
  // Mean milliseconds per step over `reps` launches: of the pack alone and
  // of the flux pass alone (HIP events around their kernels), and of the
  // whole step.
  void time(int scale, int reps, double* ms_pack, double* ms_flux, double* ms_total);
  // Replace the observations of every set ([n_sets][n_times] each; a null
  // array keeps the current values), in the staging block and on the device.
  void set_obs(const double* flux2d, const double* ferr2d, const double* weight2d);
  // Packs set `set` on the host (the same CPU setup, pack_set, this
  // batch's chunk and limb threshold) and compares it with the tables of a
  // fresh device pack. diff[0] differing record bytes, [1] patch cones,
  // [2] patch moment values, [3] work codes, [4] segment flags, [5] work
  // items (a count mismatch counts every item), [6] item moment values and
  // lane-mask groups, [7] the host's item count, [8] limb chunks, limb
  // items, their lane indices and lane lists that differ (a count mismatch
  // counts every entry), [9] the host's limb item count.
  void check(int set, uint64_t diff[10]);

  // Work-table sets that launches alternate between (1 or 2; built with
  // 1). With 2, the pack of a launch runs beside the flux pass of the one
  // before, at twice the work memory.
  void set_slots(int n);
  int slots() const { return static_cast<int>(slots_.size()); }
  const DevicePackStats& stats() const { return st_; }
  uint32_t n_sets() const { return static_cast<uint32_t>(host_.size()); }
  const SetHost& set(uint32_t i) const { return host_[i]; }
  size_t n_times() const { return n_times_; }
  const std::vector<SetOut>& last_results() const { return last_out_; }

 private:
  // Offsets of the work tables in a slot's work block.
  struct WorkLayout {
    size_t keys, vals, keys_s, vals_s, recs, cones, moms, wmax, ten, tval, trng, tcp, codes, efl;
    size_t hsp, sal, ssm, gln, gl, sst, crev, csh, cnq, clc, clp, lch;
    size_t cand, cinf, cpos, items, counts, part, ltot, flux, out, lcomp, comp, lsw;
    size_t total;
    bool iscale;
  };
  // One set of the pack's and the flux pass's work tables with its own
  // streams: launches alternate between slots, so that the pack of one
  // launch can run beside the flux pass of the one before.
  struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;  // recorded after each launch; the context stream waits on it
    char* work = nullptr;
    void* scan_temp = nullptr;
    PackCounts* h_counts = nullptr;  // pinned
    PackCounts last_counts{};
    // Tables of the pack's second phase, each its own allocation so it can
    // grow: item moments, lane-mask groups, limb items, their lane indices,
    // limb lane lists, the moment entries' first patches and patch masks.
    void* grow[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    PackDev pd;
    DeviceBatch d;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    SideStream side;
  };

  void release();
  void release_slot(Slot& s);
  void build(std::vector<std::unique_ptr<SetLayout>> lay);
  void add_slot();
  // The pack's first phase, its counts read back, the second phase's tables
  // grown to them, then the second phase; leaves s.d ready for the flux
  // pass. All on s.stream.
  void pack(Slot& s);
  void ensure_capacity(Slot& s, const PackCounts& c);
  // The next launch's slot, after its previous launch (in its stream order).
  Slot& next_slot();
  // Waits for every slot.
  void sync_all();
  const lc::Data& data_of(size_t s) const { return data_.size() == 1 ? data_[0] : data_[s]; }

  Context& ctx_;
  std::vector<lc::Model> models_;
  std::vector<lc::Data> data_;
  size_t n_times_ = 0;
  uint32_t chunk_ = 0;
  uint32_t limb_ = 0;
  DevicePackStats st_;
  std::vector<SetHost> host_;
  std::vector<SetOut> last_out_;
  // Host mirrors used by check(): per set its first segment, patch, tile,
  // code and chunk; its PackSet.
  std::vector<uint32_t> seg0_, nseg_;
  std::vector<uint64_t> rec0_;  // first record byte of each set
  std::vector<PackSet> psets_;
  std::vector<SetDev> sets_;

  // One pinned host block and one device block of the same layout hold
  // every input; each slot's device block holds its work tables.
  char* h_in_ = nullptr;
  char* d_in_ = nullptr;
  size_t in_bytes_ = 0;
  WorkLayout wl_{};
  size_t scan_temp_bytes_ = 0;
  uint64_t n_cand_ = 0;
  // The input pointers and sizes every slot starts from.
  PackDev pd0_;
  DeviceBatch d0_;
  std::vector<std::unique_ptr<Slot>> slots_;
  size_t cur_ = 0;   // the slot of the next launch
  size_t last_ = 0;  // the slot of the last launch
  size_t o_ty_ = 0, o_tfe_ = 0, o_tw_ = 0;  // observation tables in the input block
  size_t segs_off_ = 0, segs_bytes_ = 0;     // the segment table in the input block
};

}  // namespace lch
