// Device-side data layout of a batch of light curves, shared by the host
// packer (engine.cpp) and the kernels (flux_kernels.hip).
//
// A batch holds one or more parameter sets. Each set owns:
//  * lanes   — one per exposure sub-phase (plus up to nine calibration /
//              white-dwarf lanes), padded to tiles of 64 = one wavefront;
//  * segments— its surface-element grids, split by eclipse-pair count so a
//              segment's elements share one record layout;
//  * chunks  — fixed-size slices of segments;
//  * partial rows — one row of 64 FP64 partial sums per (tile, chunk).
// A work item is one wavefront computing one (tile, chunk) partial row.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lch {

// Segment kinds. Star 1 and star 2 use their own limb darkening; disc and
// edge use the disc's quadratic law and test mu > 0; the spot has none.
enum SegKind : uint32_t { KIND_STAR1 = 0, KIND_STAR2 = 1, KIND_DISC = 2, KIND_EDGE = 3, KIND_SPOT = 4 };

// Lane classes. NORMAL lanes are exposure sub-phases (flux.rs:7-111
// conventions); CAL1 / CAL2 lanes evaluate comp_star1 / comp_star2 at the
// Ginterp calibration phases (orchestration.rs:93-131); WD evaluates
// comp_star1 at phase 0.5 (orchestration.rs:256-259).
enum LaneClass : uint32_t { CLS_NORMAL = 0, CLS_CAL1 = 1, CLS_CAL2 = 2, CLS_WD = 3 };

// lane_meta bits: [0,2) class, [2,4) grid type 1..3, bit 4 valid.
constexpr uint32_t META_VALID = 1u << 4;
__host__ __device__ inline uint32_t meta_class(uint32_t m) { return m & 3u; }
__host__ __device__ inline uint32_t meta_ptype(uint32_t m) { return (m >> 2) & 3u; }
// Bit of a (class, grid type) pair inside a segment use mask.
__host__ __device__ inline uint32_t cp_bit(uint32_t cls, uint32_t ptype) { return 1u << (cls * 4u + ptype); }

constexpr int TILE = 64;     // lanes per tile = wavefront width on CDNA
constexpr int MAX_ECL = 4;   // eclipse pairs per element supported by the kernels

// Element record, 16-byte aligned, stored back to back for a segment:
//   float4 (dirn.x, dirn.y, dirn.z, flux)
//   float4 (posn.x, posn.y, posn.z, 0)        only if SEG_POSN is set
//   NECL x { double ingress, egress, egress - 1 }
constexpr uint32_t SEG_POSN = 1u;
// Star segments with the linear law (Poly with only the first coefficient
// set), no mu cut-off and no beaming or lensing may carry SEG_LINW: their
// device records hold (s d, s) in place of (d, w), s = sqrt(w) / 2^k, with k
// in flags bits [8, 16) biased by 128. With v = clamp(s d.e, 0, 1) = s mu
// (mu in [0, 1]), w mu (c0 + a0 mu) = 4^k (c0 v s + a0 v^2): the weight
// multiply folds into the dot product. Host tables (moments, patch codes,
// limb lanes) are made from the plain records before the rescaling.
constexpr uint32_t SEG_LINW = 2u;
__host__ __device__ inline int seg_linw_exp(uint32_t flags) {
  return static_cast<int>((flags >> 8) & 0xffu) - 128;
}

// Segments with 1..LDS_ECL_MAX eclipse intervals per element have their
// element-loop patches staged in LDS by a kernel form of their own.
constexpr uint32_t LDS_ECL_MAX = 2;

__host__ __device__ constexpr uint32_t rec_stride(uint32_t necl, bool posn) {
  return 16u + (posn ? 16u : 0u) + 24u * necl;
}

// Elements of a segment are stored in patches of PATCH (padded with
// zero-flux records), in an order that keeps each patch's normals close
// together. For every (patch, lane tile) pair the host decides, exactly,
// whether the patch must be evaluated with the eclipse test (PT_FULL),
// needs no eclipse test because no lane of the tile is eclipsed by any
// element of it (PT_FREE), adds exactly zero to every lane of the tile
// (PT_SKIP: all normals face away from all the tile's lines of sight, or
// every element is eclipsed at every lane), or is seen whole by every lane
// (PT_MOM: every element visible with mu above the cut-off), when a
// contribution polynomial in mu of degree <= 3 is a polynomial in the earth
// vector whose coefficients are the patch's direction moments.
constexpr uint32_t PATCH = 64;
constexpr uint8_t PT_FULL = 0, PT_FREE = 1, PT_SKIP = 2, PT_MOM = 3;

// Direction moments of a patch, FP64, MOM_STRIDE doubles per patch, from the
// FP32 normals d and fluxes w the element records hold:
//   [0..2]  sum w d                       (x, y, z)
//   [3..8]  sum w (e.d)^2 coefficients of  x2 y2 z2 xy xz yz
//   [9..18] sum w (e.d)^3 coefficients of  x3 y3 z3 x2y x2z y2x y2z z2x z2y xyz
//   [19]    sum |w|                        (a bound for fixed-point sums)
// so that sum_w mu^k = the k-th block dotted with the same monomials of e.
constexpr uint32_t MOM_STRIDE = 20;
// At most 64 patches per work item: the item's patch codes fit 64-bit masks.
constexpr uint32_t MAX_ITEM_PATCHES = 64;

struct SegDev {
  uint64_t rec_off;   // byte offset of the first record
  uint32_t n;         // number of elements, padded to a multiple of PATCH
  uint32_t set;       // owning set
  uint32_t use_mask;  // cp_bit()s of the lanes that add this segment
  uint32_t kind;      // SegKind
  uint32_t necl;      // eclipse pairs per element
  uint32_t flags;     // SEG_POSN
  uint32_t patch_off; // first patch, numbered within the set (host only)
  uint32_t mom_off;   // first patch's moments (global patch index)
};

struct ItemDev {
  uint32_t seg;   // segment index (global)
  uint32_t e0;    // first element (multiple of PATCH)
  uint32_t ne;    // element count (multiple of PATCH)
  uint32_t lane0; // first lane of the item's tile (global index, multiple of TILE)
  uint64_t prow;  // partial row index
  uint64_t todo;  // bit q: patch e0 / PATCH + q is evaluated element by element
  uint64_t free;  // bit q: that patch needs no eclipse test (PT_FREE)
  uint64_t mom;   // bit q: that patch is evaluated from its moments (PT_MOM)
  // bit q: that patch (PT_FREE) is evaluated from its moments on the lanes
  // that see all of it only; its limb lanes, which see part of it, are
  // summed by a limb chunk, and its other lanes see none of it. The pmom
  // patches fall into ngrp groups by lane mask: group g adds on the lanes of
  // lane_masks[lm_off + g].
  uint64_t pmom;
  uint32_t lm_off;
  // First entry of the item-moment table (MOM_STRIDE doubles each): with
  // mom != 0, the sum in patch order of the moments of the patches in mom;
  // then one entry per pmom group, the sum in patch order of its patches'.
  uint32_t imom;
  uint32_t ngrp;  // pmom groups (distinct lane masks), 0 without pmom
  uint32_t pad;
};

// Limb lanes. For the PT_FREE patches of one segment chunk, the lanes (of
// any tile) that see only part of a patch are evaluated element by element
// by one workgroup per chunk: its wavefronts take limb items (one patch and
// up to 64 such lanes) and add each lane's patch sum to the lane's LDS slot
// as a 64-bit fixed-point integer (2^shift units; integer addition makes the
// total independent of the order of the adds). After the work items have
// written the chunk's partial rows, the totals are added to them. A chunk's
// limb lanes are listed once (`lanes`, global lane indices); a limb item
// names them by index into that list (16-bit, 0xFFFF = unused).
constexpr uint32_t LIMB_UNION_MAX = 2048;  // limb lanes per chunk (LDS)
constexpr uint32_t LIMB_WAVES = 8;         // wavefronts per limb workgroup
struct LimbChunk {
  uint32_t seg;         // segment (global index)
  uint32_t chunk;       // chunk index within the set (partial row column)
  uint32_t item0, n_items;  // limb items
  uint32_t lane_off, n_lanes;  // entries of the chunk's lane list
  int32_t shift;        // fixed-point scale 2^shift: |lane total| * 2^shift < 2^62
  uint32_t pad;
};
struct LimbItem {
  uint32_t e0;     // first element of the patch within the segment
  uint32_t split;  // log2 of the lane groups sharing the patch's elements: 0, 1 or 2
};

// Set flags.
constexpr uint32_t SET_VALID = 1u, SET_CAL1 = 2u, SET_CAL2 = 4u, SET_BEAM1 = 8u, SET_BEAM2 = 16u,
                   SET_LENS = 32u;
// Per-component scale factors (iscale): the set's time lanes follow
// comp_star1 / comp_star2 / comp_disc / comp_edge / comp_spot (flux.rs:115-279)
// and the assembly fits one factor per component (orchestration.rs:181-297).
// SET_NCOL5: star 2 is a component (t2 > 0).
constexpr uint32_t SET_ISCALE = 64u, SET_NCOL5 = 128u;

struct SetDev {
  uint32_t lane_off;   // first lane (global index, multiple of TILE)
  uint32_t n_lanes;    // lanes incl. padding (multiple of TILE)
  uint32_t special0;   // first calibration / white-dwarf lane, relative to lane_off
  uint32_t flags;
  uint32_t time_off;   // first time (global index)
  uint32_t n_times;
  uint32_t chunk_off;  // first entry of this set in chunk_info
  uint32_t n_chunks;
  uint64_t prow_base;  // first partial row; row = prow_base + tile * n_chunks + chunk
  uint32_t ltype1, ltype2;  // 0 Poly, 1 Claret
  float ldc1[4], ldc2[4];
  float mucrit1, mucrit2, lin_disc, quad_disc;
  float beam1, beam2, vfac, xcofm, spin1, spin2, rlens1, pad0;
  double cosi_lc, sini_lc;    // cos/sin(TWOPI*iangle/360): comp_light lanes
  double cosi_rad, sini_rad;  // iangle.to_radians(): comp_star1/2 lanes
  double phase1, phase2;      // Ginterp switch phases
  double third;
  // Ginterp factors s11, s12, s21, s22 (types.rs:129-202) the assembly
  // starts from: 1 when the set's calibration lanes find them
  // (SET_CAL1 / SET_CAL2, orchestration.rs:85-131), or the caller's own
  // factors for a set uploaded as grids with no calibration lanes.
  double gscl[4];
};

// Per-set results written by the assembly kernel.
struct SetOut {
  double chisq, wnok, scale, wdwarf;
  double sfac[5];  // per-component factors (all equal to scale without iscale)
  double gint[4];  // Ginterp factors s11, s12, s21, s22 the assembly used
};

}  // namespace lch
