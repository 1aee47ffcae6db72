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

// Item counts per kernel class (plain, staged intervals, beaming / lensing).
struct ItemCounts {
  uint32_t a, b, c;
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
  // work and outputs
  uint32_t* wmax = nullptr;          // per segment: largest weight (FP32 bits), then a bad-weight flag
  uint32_t* keys = nullptr;          // [n_raw] x 2 (sort in / out)
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
  ItemDev* cand = nullptr;           // [n_chunks * max_tiles]
  uint8_t* cand_class = nullptr;
  ItemCounts* cand_pos = nullptr;
  ItemDev* items = nullptr;
  double* imoms = nullptr;
  ItemCounts* counts = nullptr;      // one: items per class
  // Partial rows of the flux pass: the row of a (chunk, tile) pair without
  // an item is zeroed here; the rows with one are written by the item.
  double* partials = nullptr;
};

// Bytes of the radix sort's temporary storage for n keys. The keys are
// (segment << 16 | direction key): at most 65536 segments.
size_t pack_sort_temp_bytes(uint64_t n);
size_t pack_scan_temp_bytes(uint64_t n);
// Every step of the pack on `st`; the item counts land in d.counts.
hipError_t launch_pack(const PackDev& d, void* sort_temp, size_t sort_temp_bytes, void* scan_temp,
                       size_t scan_temp_bytes, uint32_t key_bits, hipStream_t st);

}  // namespace lch
