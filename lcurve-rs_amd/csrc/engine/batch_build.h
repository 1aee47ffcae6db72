// Packing of one parameter set's CPU-setup output into the device layout,
// and the threaded setup + pack of a whole batch.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "layout.h"
#include "pack_math.h"
#include "../setup/prepare.h"

namespace lch {

// Host-side description of one set of a batch.
struct SetHost {
  bool valid = false;
  std::string error;
  double logg1 = 0, logg2 = 0, rv1 = 0, rv2 = 0;
  uint32_t n_times = 0;
  uint64_t visits = 0;  // element-visits in the reference's loop order
  bool star1_alias = false, star2_alias = false;  // coarse grid = fine grid
  // Debug visibility map: per grid (0 star1f, 1 star1c, 2 star2f, 3 star2c,
  // 4 disc, 5 edge, 6 spot) its segments and each segment element's index
  // in the grid's original order.
  struct GridMap {
    std::vector<uint32_t> segs;
    std::vector<std::vector<uint32_t>> orig;
    uint32_t n = 0;
  } grids[7];
};

// A grid of a prepared set as the pack uses it: grid id (0..6 as
// lch_batch_visibility numbers them, 7..10 calibration copies), its
// elements, segment kind, the lanes that add it (cp_bit mask) and whether
// its records carry positions (beaming or lensing).
struct GridRef {
  uint32_t id;
  const std::vector<lc::Point>* pts;
  uint32_t kind;
  uint32_t use;
  bool posn;
};

// One set, packed on the host, offsets relative to the set.
struct SetPack {
  SetDev dev{};
  SetHost host;
  std::vector<uint8_t> recs;
  std::vector<SegDev> segs;
  std::vector<uint32_t> seg_grid;                 // grid id per segment (>= 7: calibration copy)
  std::vector<std::vector<uint32_t>> seg_orig;    // optional element maps
  // Work code per (patch, tile): ptcode[patch * n_tiles + tile], patches
  // numbered in segment order (SegDev::patch_off).
  std::vector<uint8_t> ptcode;
  uint32_t n_tiles = 0;
  std::vector<Cone> cones;   // normal cone per patch
  std::vector<double> moms;  // MOM_STRIDE per patch (zero where unused)
  // Per-lane split of PT_FREE (patch, tile) pairs of moment-capable
  // segments, over the lanes that add the segment: `all` lanes see every
  // element of the patch with mu above the cut-off (the moments give the
  // exact sum), `some` lanes see part of it (element by element); the
  // other lanes see none of it. Sorted by pt = patch * n_tiles + tile.
  struct LaneSplit {
    uint32_t pt;
    uint64_t all, some;
  };
  std::vector<LaneSplit> lsplit;
  std::vector<double> lane_phi;
  std::vector<float4> lane_e;  // earth vector per lane (x, y, z, 0), FP32
  std::vector<uint32_t> lane_meta;
  std::vector<uint32_t> t_lane0;
  std::vector<int> t_ndiv;
  std::vector<double> t_slfac, t_y, t_ferr, t_w, t_phase;
};

// Pack a prepared model; `data` supplies the observations of this set.
// pack_records then pack_lanes on the same pack. elem_check: the lane
// split decides the lanes its cone test leaves undecided from the patch's
// elements (the host default); without it they stay limb lanes (the
// device pack's rule).
std::unique_ptr<SetPack> pack_set(const lc::PreparedModel& P, const lc::Data& data, bool keep_maps,
                                  bool elem_check = true);
// The pieces of pack_records: the set's scalars (SetDev, SetHost), its
// grids in segment order, and its segments without records (one per grid
// and eclipse-pair count: rec_off, padded n, use mask, kind, necl, flags,
// patch_off), with each segment's element count and grid (index in
// `grids`); returns the record bytes.
void set_scalars(SetPack& S, const lc::PreparedModel& P);
std::vector<GridRef> grid_refs(const lc::PreparedModel& P);
uint64_t lay_out_segments(SetPack& S, const std::vector<GridRef>& grids, std::vector<uint32_t>* n_real,
                          std::vector<uint32_t>* seg_ref);
// The part of a pack that does not depend on the phases: flux-path scalars,
// segments, element records in patch order, patch cones and moments.
std::unique_ptr<SetPack> pack_records(const lc::PreparedModel& P, bool keep_maps);
// The part that does: lanes for P's phases / exposures / sub-phase counts,
// the per-(patch, tile) work codes and lane splits made from the plain
// (not yet rescaled) records and cones of R, visit counts, per-time tables
// and the observations of `data`. S carries R's segments (S may be R).
void pack_lanes(SetPack& S, const SetPack& R, const lc::PreparedModel& P, const lc::Data& data,
                bool elem_check = true);
// The two halves of pack_lanes: lane and per-time tables (lanes, earth
// vectors, visit counts, observations), then the work codes and lane
// splits made from them and R's records.
void pack_lane_tables(SetPack& S, const lc::PreparedModel& P, const lc::Data& data);
void pack_codes(SetPack& S, const SetPack& R, const lc::PreparedModel& P, bool elem_check = true);
// Rescales the records of the set's linear-law star segments for the device
// (SEG_LINW); call after every host table has been made from the records.
void scale_linear_records(SetPack& S);
// The work-code rule of one segment of P: an element adds nothing unless
// mu > thr, and is seen whole when mu > thr_front for the whole patch;
// cull / front_ok say whether the skip / seen-whole tests apply, mom_ok
// whether the segment's contribution is a polynomial in mu of degree <= 3
// (patch moments), K the constants of the cone tests.
struct SegCodeRule {
  bool cull = false, front_ok = false, mom_ok = false;
  double thr = 0.0, thr_front = 0.0;
  CodeConsts K{};
};
SegCodeRule seg_code_rule(const SegDev& sg, const SetDev& d, const lc::PreparedModel& P);
CodeConsts code_consts(bool cull, double thr, bool front_ok, double thr_front);
// Degree of a segment's contribution as a polynomial in mu for elements
// that are visible with mu above the cut-off, or 0 where it is none
// (beaming, lensing, Claret or cubic / quartic Poly limb darkening). The
// FP32 coefficients are the ones the kernel uses.
int moment_degree(const SegDev& sg, const SetDev& d);
// Whether a star segment takes the kernel's linear form (SEG_LINW): Poly
// law, a1 = a2 = a3 = 0 and mucrit <= 0 in the FP32 values the kernel
// reads, no beaming or lensing.
bool linear_law_segment(const SegDev& sg, const SetDev& d);
// A set whose setup failed: contributes NaN results, no work.
std::unique_ptr<SetPack> invalid_set(const std::string& err, const lc::Data& data);

// CPU setup + packing for every set on `nthreads` threads (0 = all cores).
// `data` holds either one data set shared by all models or one per model.
std::vector<std::unique_ptr<SetPack>> build_packs(const std::vector<lc::Model>& models,
                                                  const std::vector<lc::Data>& data, int nthreads,
                                                  bool keep_maps, bool elem_check = true);

}  // namespace lch
