// The per-set pack on the GPU: from the sets' raw element grids resident in
// HBM to the tables the flux kernels read (element records in patch order,
// patch cones and moments, per-(patch, tile) work codes, work items and
// their summed moments). The arithmetic is pack_math.h's, so the tables
// equal the host pack's (batch_build.cpp) bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "layout.h"
#include "pack_math.h"

namespace lch {

// One segment of a batch (a set's grid elements with one eclipse-pair count).
struct PackSeg {
  uint64_t rec_off;   // first record, bytes into the batch's records
  uint64_t e0;        // first raw element of the segment's grid
  uint32_t pos0;      // first raw position of that grid (grids with positions)
  uint32_t n_real;    // elements
  uint32_t n_pad;     // elements padded to patches
  uint32_t necl;      // eclipse pairs per element
  uint32_t posn;      // 1: records carry positions
  uint32_t patch0;    // first patch (batch-wide numbering)
  uint32_t sorted0;   // first position of its elements in sorted order
  uint32_t set;
  uint32_t flags;     // PSEG_* below
  uint32_t base_flags;  // SegDev::flags before the linear-law rescaling
  CodeConsts K;
  // The per-lane split's element-level cut-offs: a lane sees the patch
  // whole when every element's FP32 mu exceeds `whole` (= the seen-whole
  // threshold + 1e-6) and none of it when every mu is below `none` (= the
  // cut-off - 1e-6).
  float whole, none;
};
constexpr uint32_t PSEG_CULL = 1u, PSEG_FRONT = 2u, PSEG_MOM = 4u, PSEG_LINW = 8u;

// One (set, grid) block of raw elements, contiguous in the raw arrays.
struct PackGrid {
  uint64_t e0;
  uint32_t n;
  uint32_t seg_of[MAX_ECL + 1];  // segment of each eclipse-pair count (0xFFFFFFFF: none)
};

struct PackSet {
  uint32_t lane_off, n_tiles;  // first lane (batch-wide, multiple of TILE), lane tiles
  uint32_t tile0;              // first tile (batch-wide)
  uint32_t patch0, n_patches;  // first patch, patches
  uint32_t n_chunks;           // work chunks (partial-row columns)
  uint32_t chunk0;             // first chunk (batch-wide)
  uint64_t code0;              // first code: code0 + tile * n_patches + local patch (tile-major)
  uint64_t prow_base;          // first partial row
};

// A work chunk: up to MAX_ITEM_PATCHES patches of one segment.
struct PackChunk {
  uint32_t set, seg, e0, ne, col;  // col: chunk index within the set
};

// A run of up to KEY_TILE elements of one grid (the key and record
// kernels' workgroups).
constexpr uint32_t KEY_TILE = 4096;
struct KeyTile {
  uint32_t grid, i0, n;
};

// A block of up to 256 patches of one set (the code kernel's workgroups).
struct PatchBlock {
  uint32_t set, patch0, n;
};

// Per work-item candidate: items per kernel class (plain, staged
// intervals, beaming / lensing), lane-mask groups and item-moment entries.
struct ItemCounts {
  uint32_t a, b, c, masks, moms;
};

// Per chunk: limb chunks (0 or 1), limb items and limb lanes.
struct LimbCounts {
  uint32_t chunks, items, lanes;
};

// What the pack's first phase counted (totals of the scans), read by the
// host before the second phase writes the tables these size.
struct PackCounts {
  ItemCounts items;
  LimbCounts limb;
};

struct PackDev {
  // inputs
  const float4* raw_df = nullptr;    // dirn.xyz, flux per raw element
  const float4* raw_pos = nullptr;   // posn.xyz per raw element of grids with positions
  const uint32_t* raw_eoff = nullptr;  // [n_raw + 1] offsets into raw_ecl
  const double2* raw_ecl = nullptr;  // (ingress, egress) per eclipse pair
  uint64_t n_raw = 0;
  const PackGrid* grids = nullptr;
  uint32_t n_grids = 0;
  const uint32_t* grid_order = nullptr;  // grids by size, largest first (the sort's workgroups)
  const KeyTile* ktiles = nullptr;
  uint32_t n_ktiles = 0;
  const PackSeg* psegs = nullptr;
  SegDev* segs = nullptr;            // the flux kernels' segment table (flags set here)
  uint32_t n_segs = 0;
  const PackSet* sets = nullptr;
  uint32_t n_sets = 0;
  const uint32_t* patch_seg = nullptr;  // segment of each patch
  uint32_t n_patches = 0;
  const PatchBlock* pblocks = nullptr;
  uint32_t n_pblocks = 0;
  const uint32_t* ecl_patches = nullptr;  // patches of segments with eclipse pairs
  uint32_t n_ecl_patches = 0;
  const PackChunk* chunks = nullptr;
  uint32_t n_chunks = 0;
  uint32_t max_tiles = 0;
  uint32_t n_tiles = 0;              // all sets' tiles
  const uint32_t* tile_set = nullptr;  // set of each tile
  const double* lane_phi = nullptr;
  const float4* lane_e = nullptr;
  const uint32_t* lane_meta = nullptr;
  const SetDev* setdev = nullptr;      // the flux kernels' set table (limb-darkening laws)
  // Limb lanes: 0 off; otherwise a PT_FREE (patch, tile) pair is split per
  // lane when at most `limb` of its lanes see the patch in part.
  uint32_t limb = 0;
  // work and outputs
  uint32_t* wmax = nullptr;          // per segment: largest weight (FP32 bits), then a bad-weight flag
  // Sort keys and raw element indices, [n_raw] each. The sort's first pass
  // goes from keys / vals to keys_s / vals_s, its second back: the kernels
  // after it read the sorted order through vals_s, which launch_pack_count
  // points at vals for them.
  uint32_t* keys = nullptr;
  uint32_t* vals = nullptr;
  uint32_t* keys_s = nullptr;
  uint32_t* vals_s = nullptr;
  char* recs = nullptr;
  Cone* cones = nullptr;
  double* moms = nullptr;
  float4* tile_en = nullptr;         // [n_tiles * TILE] normalised earth vectors, rounded to FP32
  uint64_t* tile_valid = nullptr;
  double2* tile_range = nullptr;     // (lo, hi) of phi - floor(phi) over valid lanes
  uint32_t* tile_cp = nullptr;       // cp_bit mask of the valid lanes
  uint8_t* codes = nullptr;          // tile-major, see PackSet::code0
  // Per (patch, tile), patch-major (code0 + local patch * n_tiles + tile):
  // bit 0 no lane eclipsed, bit 1 every lane eclipsed.
  uint8_t* eflags = nullptr;
  // Per-lane split of each PT_FREE pair (tile-major, as the codes):
  // has_split 1 where the pair was split, then the lanes of the tile that
  // see the patch whole (split_all) and in part (split_some).
  uint8_t* has_split = nullptr;
  uint64_t* split_all = nullptr;
  uint64_t* split_some = nullptr;
  // The split's work: pairs split and their limb lanes.
  unsigned long long* split_stats = nullptr;

  // Per chunk: 1 when its limb lanes exceed LIMB_UNION_MAX (its split
  // patches go back to the element loops), the fixed-point scale, the
  // limb lanes of each of its patches ([n_chunks][MAX_ITEM_PATCHES]), and
  // its limb counts with their exclusive scan.
  uint8_t* chunk_rev = nullptr;
  int32_t* chunk_shift = nullptr;
  uint32_t* chunk_nq = nullptr;
  LimbCounts* chunk_lcnt = nullptr;
  LimbCounts* chunk_lpos = nullptr;
  ItemDev* cand = nullptr;           // [n_chunks * max_tiles]
  // class | ngrp << 8 | (item has PT_MOM patches) << 16
  uint32_t* cand_info = nullptr;
  // Candidates with split patches (their lane-mask groups are counted and
  // written a wavefront each): the list's length, then the list.
  uint32_t* glist_n = nullptr;
  uint32_t* glist = nullptr;
  ItemCounts* cand_pos = nullptr;
  ItemDev* items = nullptr;
  // Tables the second phase writes, sized by the counts of the first.
  double* imoms = nullptr;           // item-moment entries
  // Per entry: the moments of the item's first patch, and the mask of the
  // item's patches it sums.
  uint32_t* ment_base = nullptr;
  uint64_t* ment_mask = nullptr;
  uint64_t cap_imoms = 0;
  uint64_t* lane_masks = nullptr;    // lane-mask groups of the items' split patches
  uint64_t cap_masks = 0;
  LimbChunk* lchunks = nullptr;      // capacity n_chunks
  LimbItem* litems = nullptr;
  uint16_t* lidx = nullptr;          // TILE per limb item
  uint64_t cap_litems = 0;
  uint32_t* llanes = nullptr;
  uint64_t cap_llanes = 0;
  PackCounts* counts = nullptr;
  // Partial rows of the flux pass: the row of a (chunk, tile) pair without
  // an item is zeroed here; the rows with one are written by the item.
  double* partials = nullptr;
};

size_t pack_scan_temp_bytes(uint64_t n_cand, uint64_t n_chunks);
// The pack in two phases on `st`. The first makes the records, cones,
// moments, codes, lane splits and item candidates and leaves the counts of
// what the second writes in d.counts; the host reads them and sizes the
// second phase's tables. The second writes the work items, their lane-mask
// groups and moment sums, and the limb tables.
hipError_t launch_pack_count(const PackDev& d, void* scan_temp, size_t scan_temp_bytes, hipStream_t st);
hipError_t launch_pack_fill(const PackDev& d, uint64_t n_entries, hipStream_t st);

}  // namespace lch
