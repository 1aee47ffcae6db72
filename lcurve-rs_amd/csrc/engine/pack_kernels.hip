// The per-set pack on the GPU (pack_kernels.h). Everything here is exact
// FP64 / integer work in a fixed order: built without FMA contraction, the
// tables equal the host pack's bit for bit.
#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

#include <rocprim/block/block_radix_rank.hpp>
#include <rocprim/block/block_scan.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include "pack_kernels.h"

namespace lch {

namespace {

constexpr uint32_t kBlock = 256;
// Workgroups of the kernels that stride over a list whose length is on the
// device.
constexpr uint32_t kListBlocks = 2048;

__device__ inline uint32_t stride_of(const PackSeg& S) { return rec_stride(S.necl, S.posn != 0); }

// Stores 8-byte words (records are 8-byte aligned; strides of 40 or 64
// bytes leave 16-byte alignment to every other record).
__device__ inline void store_f4(char* r, float4 v) {
  float2* o = reinterpret_cast<float2*>(r);
  o[0] = make_float2(v.x, v.y);
  o[1] = make_float2(v.z, v.w);
}
__device__ inline float4 load_f4(const char* r) {
  const float2* o = reinterpret_cast<const float2*>(r);
  const float2 a = o[0], b = o[1];
  return make_float4(a.x, a.y, b.x, b.y);
}

// Sort key of every raw element: its eclipse-pair count (the segment
// within its grid: a grid's segments are numbered by that count, in order)
// above the spatial key. With
// it, per segment, the largest weight and whether a weight is negative or
// not finite, for the linear-law rescaling. One workgroup per key tile.
__global__ __launch_bounds__(kBlock) void keys_kernel(PackDev d) {
  constexpr int NK = MAX_ECL + 1;
  __shared__ float red_w[kBlock / TILE][NK];
  __shared__ uint32_t red_b[kBlock / TILE];
  const KeyTile kt = d.ktiles[blockIdx.x];
  const PackGrid G = d.grids[kt.grid];
  float mx[NK];
#pragma unroll
  for (int c = 0; c < NK; ++c) mx[c] = 0.0f;
  uint32_t bad = 0;  // bit k: a weight of eclipse-pair count k is negative or not finite
  for (uint32_t i = threadIdx.x; i < kt.n; i += kBlock) {
    const uint64_t e = G.e0 + kt.i0 + i;
    uint32_t k = d.raw_eoff[e + 1] - d.raw_eoff[e];
    k = k < static_cast<uint32_t>(MAX_ECL) ? k : static_cast<uint32_t>(MAX_ECL);
    const float4 v = d.raw_df[e];
    d.keys[e] = (k << DIR_KEY_BITS) | direction_key(v.x, v.y, v.z);
    d.vals[e] = static_cast<uint32_t>(e);
    const float w = v.w;
    if (!(w >= 0.0f) || !isfinite(w)) {
      bad |= 1u << k;
    } else {
#pragma unroll
      for (int c = 0; c < NK; ++c) mx[c] = (static_cast<uint32_t>(c) == k && w > mx[c]) ? w : mx[c];
    }
  }
#pragma unroll
  for (int c = 0; c < NK; ++c)
    for (int off = 32; off > 0; off >>= 1) {
      const float o = __shfl_xor(mx[c], off);
      mx[c] = o > mx[c] ? o : mx[c];
    }
  for (int off = 32; off > 0; off >>= 1) bad |= __shfl_xor(bad, off);
  const uint32_t wv = threadIdx.x / TILE;
  if (threadIdx.x % TILE == 0) {
#pragma unroll
    for (int c = 0; c < NK; ++c) red_w[wv][c] = mx[c];
    red_b[wv] = bad;
  }
  __syncthreads();
  if (threadIdx.x < NK) {
    const uint32_t c = threadIdx.x;
    const uint32_t seg = G.seg_of[c];
    if (seg != 0xFFFFFFFFu) {
      float wm = 0.0f;
      uint32_t b = 0;
      for (uint32_t i = 0; i < kBlock / TILE; ++i) {
        wm = red_w[i][c] > wm ? red_w[i][c] : wm;
        b |= red_b[i];
      }
      // Non-negative FP32 values order as their bit patterns.
      if (wm > 0.0f) atomicMax(&d.wmax[2u * seg], __float_as_uint(wm));
      if ((b >> c) & 1u) atomicOr(&d.wmax[2u * seg + 1u], 1u);
    }
  }
}

// The sort: each grid's elements in place in their own range (a grid's
// raw elements are also its sorted positions: its segments are numbered in
// eclipse-pair order, after the previous grid's), by (eclipse-pair count,
// direction key), a 15-bit key, stable. A workgroup per grid, largest grids
// first, two 8-bit LSD passes: a histogram of the grid's digits in LDS,
// its exclusive scan, then the grid in tiles of SS_TILE elements in order,
// each ranked by rocPRIM's block radix rank, placed in LDS in its ranked
// order and written out from there, consecutive threads to consecutive
// positions of each digit's run (from registers, a wavefront's 64 stores
// went to up to 64 runs and left partly written cache lines). The rank's
// match form keeps its own LDS small (per-wavefront digit counters) and is
// stable in warp-striped order: wavefront w holds elements
// w * 64 * SS_ITEMS + 64 i + lane of the tile, which is also the coalesced
// way to load them.
constexpr uint32_t SS_BLOCK = 512, SS_ITEMS = 16, SS_TILE = SS_BLOCK * SS_ITEMS;
using SsRank = rocprim::block_radix_rank<SS_BLOCK, 8, rocprim::block_radix_rank_algorithm::match>;
using SsScan = rocprim::block_scan<uint32_t, SS_BLOCK>;
static_assert(SsRank::digits_per_thread == 1 && SS_BLOCK >= 256, "digit t on thread t");
static_assert(MAX_ECL < 8 && DIR_KEY_BITS + 3 <= 16, "keys of at most 16 bits: two 8-bit passes");

__global__ __launch_bounds__(SS_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void segsort_kernel(PackDev d) {
  __shared__ typename SsRank::storage_type rank_storage;
  __shared__ typename SsScan::storage_type scan_storage;
  // hist[0]: the first pass's digit counts (a loop over the grid); hist[1]:
  // the second's, counted while the first pass scatters.
  __shared__ uint32_t hist[2][256], offs[256], pref[256];
  // The tile in its ranked order: keys, then values.
  __shared__ uint32_t xk[SS_TILE], xv[SS_TILE];
  const PackGrid G = d.grids[d.grid_order[blockIdx.x]];
  const uint64_t e0 = G.e0;
  const uint32_t n = G.n, t = threadIdx.x;
  if (t < 256) {
    hist[0][t] = 0u;
    hist[1][t] = 0u;
  }
  __syncthreads();
  for (uint32_t i = t; i < n; i += SS_BLOCK) atomicAdd(&hist[0][d.keys[e0 + i] & 0xffu], 1u);
  for (uint32_t pass = 0; pass < 2; ++pass) {
    const uint32_t* kin = (pass == 0 ? d.keys : d.keys_s) + e0;
    const uint32_t* vin = (pass == 0 ? d.vals : d.vals_s) + e0;
    uint32_t* kout = (pass == 0 ? d.keys_s : d.keys) + e0;
    uint32_t* vout = (pass == 0 ? d.vals_s : d.vals) + e0;
    const uint32_t shift = 8u * pass;
    __syncthreads();  // the pass's counts are complete; the previous pass's writes are done
    uint32_t start = 0u;
    SsScan().exclusive_scan(t < 256 ? hist[pass][t] : 0u, start, 0u, scan_storage);
    if (t < 256) offs[t] = start;
    for (uint32_t base = 0; base < n; base += SS_TILE) {
      uint32_t key[SS_ITEMS], val[SS_ITEMS], rank[SS_ITEMS];
      const uint32_t first = base + (t / TILE) * (TILE * SS_ITEMS) + t % TILE;  // warp-striped
#pragma unroll
      for (uint32_t i = 0; i < SS_ITEMS; ++i) {
        const uint32_t idx = first + i * TILE;
        // Past the grid's end: digit 255, ranked after every element of
        // the tile (they come last in its order), never written.
        key[i] = idx < n ? kin[idx] : 0xffffffffu;
        val[i] = idx < n ? vin[idx] : 0u;
      }
      uint32_t prefix[1], counts[1];
      SsRank().rank_keys(
          key, rank, rank_storage, [shift](const uint32_t& k) { return (k >> shift) & 0xffu; }, prefix, counts);
      if (t < 256) pref[t] = prefix[0];
#pragma unroll
      for (uint32_t i = 0; i < SS_ITEMS; ++i) {
        xk[rank[i]] = key[i];
        xv[rank[i]] = val[i];
      }
      __syncthreads();  // the ranked tile, pref, and offs of the previous tile
      // The tile's m elements of the grid hold ranks 0..m-1: the ones past
      // the grid's end have the last digit and come last in the tile's order.
      const uint32_t m = n - base < SS_TILE ? n - base : SS_TILE;
      for (uint32_t j = t; j < m; j += SS_BLOCK) {
        const uint32_t k = xk[j];
        const uint32_t dg = (k >> shift) & 0xffu;
        const uint32_t pos = offs[dg] + j - pref[dg];
        kout[pos] = k;
        vout[pos] = xv[j];
        if (pass == 0) atomicAdd(&hist[1][(k >> 8) & 0xffu], 1u);
      }
      __syncthreads();  // offs, pref and the ranked tile are read
      if (t < 256) offs[t] += counts[0];
    }
  }
}

// Flags of each segment for the flux kernels: the base flags, plus the
// linear-law rescaling with its exponent k (sqrt(w max) < 2^k) where the
// segment allows it and all its weights are finite and >= 0, the largest
// > 0.
__global__ __launch_bounds__(kBlock) void seg_flags_kernel(PackDev d) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= d.n_segs) return;
  const PackSeg S = d.psegs[s];
  uint32_t flags = S.base_flags;
  const float wm = __uint_as_float(d.wmax[2u * s]);
  if ((S.flags & PSEG_LINW) && d.wmax[2u * s + 1u] == 0u && wm != 0.0f) {
    int k = 0;
    (void)frexp(sqrt(static_cast<double>(wm)), &k);
    if (k >= -60 && k <= 60) flags |= SEG_LINW | (static_cast<uint32_t>(k + 128) << 8);
  }
  d.segs[s].flags = flags;
}

// What writing a segment's element records needs.
struct RecSeg {
  uint64_t rec_off;
  uint32_t stride, eoff, necl, posn, pos0, flags;
  uint64_t e0;
};
__device__ inline RecSeg rec_seg(const PackDev& d, uint32_t seg) {
  const PackSeg& S = d.psegs[seg];
  RecSeg R;
  R.rec_off = S.rec_off;
  R.stride = stride_of(S);
  R.eoff = S.posn ? 32u : 16u;
  R.necl = S.necl;
  R.posn = S.posn;
  R.pos0 = S.pos0;
  R.e0 = S.e0;
  R.flags = d.segs[seg].flags;
  return R;
}

// The record of the element of rank `rank` in its segment's sorted order,
// raw element e with plain values f. Linear-law segments (SEG_LINW) take
// (s d, s), s = sqrt(w) / 2^k, in place of (d, w).
__device__ inline void write_record(const PackDev& d, const RecSeg& R, uint32_t rank, uint32_t e, float4 f) {
  char* r = d.recs + R.rec_off + static_cast<uint64_t>(rank) * R.stride;
  if (R.flags & SEG_LINW) {
    const double inv = ldexp(1.0, -seg_linw_exp(R.flags));
    const double sc = sqrt(static_cast<double>(f.w)) * inv;
    f = make_float4(static_cast<float>(static_cast<double>(f.x) * sc), static_cast<float>(static_cast<double>(f.y) * sc),
                    static_cast<float>(static_cast<double>(f.z) * sc), static_cast<float>(sc));
  }
  store_f4(r, f);
  if (R.posn) {
    float4 q = d.raw_pos[R.pos0 + (e - R.e0)];
    q.w = 0.0f;
    store_f4(r + 16, q);
  }
  const uint32_t i0 = d.raw_eoff[e];
  for (uint32_t j = 0; j < R.necl; ++j) {
    const double2 pr = d.raw_ecl[i0 + j];
    double* o = reinterpret_cast<double*>(r + R.eoff + 24u * j);
    o[0] = pr.x;
    o[1] = pr.y;
    o[2] = pr.y - 1.0;
  }
}

// The plain (dirn, flux) of slot j of patch q of segment S (zeros for a
// padding slot), from the raw grid through the sorted order.
__device__ inline float4 plain_slot(const PackDev& d, const PackSeg& S, uint32_t q, uint32_t j) {
  const uint32_t slot = q * PATCH + j;
  if (slot >= S.n_real) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  return d.raw_df[d.vals_s[S.sorted0 + slot]];
}

// Padding records of each segment: zero normal and weight; intervals that
// hold no phase of [0, 1).
__global__ __launch_bounds__(PATCH) void pad_kernel(PackDev d) {
  const PackSeg S = d.psegs[blockIdx.x];
  const uint32_t slot = S.n_real + threadIdx.x;
  if (slot >= S.n_pad) return;
  const uint32_t stride = stride_of(S);
  const uint32_t eoff = S.posn ? 32u : 16u;
  char* r = d.recs + S.rec_off + static_cast<uint64_t>(slot) * stride;
  for (uint32_t b = 0; b < eoff; b += 8) *reinterpret_cast<double*>(r + b) = 0.0;
  for (uint32_t j = 0; j < S.necl; ++j) {
    double* o = reinterpret_cast<double*>(r + eoff + 24u * j);
    o[0] = 2.0;
    o[1] = -1.0;
    o[2] = -2.0;
  }
}

__device__ inline double wave_min(double v) {
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  return v;
}
__device__ inline double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off);
    v = o > v ? o : v;
  }
  return v;
}
__device__ inline uint32_t wave_or(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off);
  return v;
}

// Patches per wavefront of the patch kernel (consecutive patch numbers).
constexpr uint32_t PATCH_RUN = 32;

// Lane x of a wavefront's value at lane x + 32, for lanes below 32 (lanes
// 32-63 get their own value back): v_permlane32_swap on each 32-bit half.
__device__ inline double upper_half(double v) {
  const uint32_t lo = __double2loint(v), hi = __double2hiint(v);
  const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  return __hiloint2double(static_cast<int>(h[1]), static_cast<int>(l[1]));
}

// Cone and (for moment-capable segments) direction moments of each patch,
// from the plain values: a wavefront per run of PATCH_RUN patches, a lane
// per slot. Each sum is tree_sum64's: its first level (lane i + lane i +
// 32) by a lane swap, the rest by one lane per sum over the 32 partial sums
// in LDS. The same pass writes the patch's element records (sorted order,
// each segment's elements by spatial key, ties in raw order) from the values
// it gathered. The gather is pipelined: while a patch is summed, the plain
// values of the next one and the sorted positions of the one after it are
// in flight. The segment's fields are re-read only when a run crosses into
// the next segment.
__global__ __launch_bounds__(TILE) void patches_kernel(PackDev d) {
  constexpr uint32_t NC = MOM_STRIDE + 3;  // cone sums x, y, z, then the moments
  // Row stride of 23 doubles: conflict-free row writes and column reads.
  __shared__ double part[PATCH / 2][NC];
  const uint32_t j = threadIdx.x;
  const uint32_t p0 = blockIdx.x * PATCH_RUN;
  if (p0 >= d.n_patches) return;
  const uint32_t pend = d.n_patches - p0 < PATCH_RUN ? d.n_patches : p0 + PATCH_RUN;
  // The segment of the patch whose positions are read next.
  uint32_t s_seg = 0, s_patch0 = 0, s_end = 0, s_nreal = 0, s_sorted0 = 0, s_mom = 0;
  // Segment, first slot, real slots, moment flag and this lane's raw
  // element of a patch.
  struct Ahead {
    uint32_t seg, first, nr, mom, e;
  };
  auto positions = [&](uint32_t pp) {
    if (pp >= s_end) {
      s_seg = d.patch_seg[pp];
      const PackSeg& S = d.psegs[s_seg];
      s_patch0 = S.patch0;
      s_end = S.patch0 + S.n_pad / PATCH;
      s_nreal = S.n_real;
      s_sorted0 = S.sorted0;
      s_mom = (S.flags & PSEG_MOM) != 0u;
    }
    const uint32_t first = (pp - s_patch0) * PATCH;
    Ahead a;
    a.seg = s_seg;
    a.first = first;
    a.nr = s_nreal > first ? (s_nreal - first < PATCH ? s_nreal - first : PATCH) : 0u;
    a.mom = s_mom;
    a.e = j < a.nr ? d.vals_s[s_sorted0 + first + j] : 0xFFFFFFFFu;
    return a;
  };
  auto plain = [&](const Ahead& a) {
    return a.e != 0xFFFFFFFFu ? d.raw_df[a.e] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  };
  Ahead cur = positions(p0);
  float4 f = plain(cur);
  // The segment whose records are written.
  uint32_t r_seg = 0xFFFFFFFFu;
  RecSeg R{};
  Ahead nxt = cur;
  if (p0 + 1 < pend) nxt = positions(p0 + 1);
  for (uint32_t p = p0; p < pend; ++p) {
    Ahead after = nxt;
    float4 fn = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (p + 1 < pend) fn = plain(nxt);
    if (p + 2 < pend) after = positions(p + 2);

    const uint32_t nr = cur.nr;
    const bool real = j < nr;
    const bool mom = cur.mom != 0u;
    if (cur.seg != r_seg) {
      r_seg = cur.seg;
      R = rec_seg(d, r_seg);
    }
    if (real) write_record(d, R, cur.first + j, cur.e, f);
    auto level1 = [&](double v, uint32_t c) {
      const double o = upper_half(v);
      if (j < PATCH / 2) part[j][c] = v + o;
    };
    level1(real ? static_cast<double>(f.x) : 0.0, 0);
    level1(real ? static_cast<double>(f.y) : 0.0, 1);
    level1(real ? static_cast<double>(f.z) : 0.0, 2);
    if (mom) {
      double T[MOM_STRIDE];
      moment_terms(T, f.x, f.y, f.z, f.w);
#pragma unroll
      for (uint32_t c = 0; c < MOM_STRIDE; ++c) level1(T[c], 3 + c);
    }
    __syncthreads();
    double sum = 0.0;
    if (j < (mom ? NC : 3u)) {
      double a[PATCH / 2];
#pragma unroll
      for (uint32_t i = 0; i < PATCH / 2; ++i) a[i] = part[i][j];
#pragma unroll
      for (uint32_t off = PATCH / 4; off > 0; off >>= 1)
#pragma unroll
        for (uint32_t i = 0; i < off; ++i) a[i] = a[i] + a[i + off];
      sum = a[0];
    }
    __syncthreads();  // the partial sums are read before the next patch writes them
    ConeBuilder cb;
    cb.set_sum(__shfl(sum, 0), __shfl(sum, 1), __shfl(sum, 2), nr > 0);
    cb.axis();
    cb.mn = wave_min(real ? cb.cosine(f.x, f.y, f.z) : 1.0);
    const Cone c = cb.finish();
    if (j == 0) d.cones[p] = c;
    if (j >= 3 && j < NC) d.moms[static_cast<uint64_t>(p) * MOM_STRIDE + (j - 3)] = mom ? sum : 0.0;

    cur = nxt;
    f = fn;
    nxt = after;
  }
}

// Per lane tile: normalised FP32 earth vectors, valid lanes, the range of
// their phases in [0, 1) and the (class, grid type) bits of its lanes.
__global__ __launch_bounds__(TILE) void tiles_kernel(PackDev d) {
  const uint32_t tg = blockIdx.x;
  const PackSet set = d.sets[d.tile_set[tg]];
  const uint32_t l = set.lane_off + (tg - set.tile0) * TILE + threadIdx.x;
  const uint32_t m = d.lane_meta[l];
  const bool valid = (m & META_VALID) != 0;
  const float4 e = d.lane_e[l];
  double x = e.x, y = e.y, z = e.z;
  const double len = sqrt(x * x + y * y + z * z);
  if (len > 0.0) {
    const double inv = 1.0 / len;
    x = inv * x;
    y = inv * y;
    z = inv * z;
  }
  d.tile_en[static_cast<uint64_t>(tg) * TILE + threadIdx.x] =
      make_float4(static_cast<float>(x), static_cast<float>(y), static_cast<float>(z), 0.0f);
  const double phi = d.lane_phi[l];
  const double phif = phi - floor(phi);
  const double lo = wave_min(valid ? phif : 4.0), hi = wave_max(valid ? phif : -4.0);
  const uint64_t vm = __ballot(valid);
  const uint32_t cp = wave_or(valid ? cp_bit(meta_class(m), meta_ptype(m)) : 0u);
  if (threadIdx.x == 0) {
    d.tile_valid[tg] = vm;
    d.tile_range[tg] = make_double2(lo, hi);
    d.tile_cp[tg] = cp;
  }
}

// Lane k's value (k wave-uniform), by v_readlane on each 32-bit half.
__device__ inline uint64_t readlane_u64(uint64_t v, uint32_t k) {
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v)), k));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v >> 32)), k));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ inline double readlane_f64(double v, uint32_t k) {
  return __longlong_as_double(static_cast<long long>(readlane_u64(static_cast<uint64_t>(__double_as_longlong(v)), k)));
}

// Eclipse interval test of every patch of a segment with eclipse pairs
// against every tile of its set: one wavefront per patch, a lane per
// element; zero-weight records count as both clear and eclipsed.
__global__ __launch_bounds__(TILE) void eclflags_kernel(PackDev d) {
  const uint32_t p = d.ecl_patches[blockIdx.x];
  const uint32_t lane = threadIdx.x;
  const PackSeg S = d.psegs[d.patch_seg[p]];
  const PackSet set = d.sets[S.set];
  const uint32_t q = p - S.patch0;
  const uint32_t stride = stride_of(S);
  const uint32_t eoff = S.posn ? 32u : 16u;
  const char* r = d.recs + S.rec_off + (static_cast<uint64_t>(q) * PATCH + lane) * stride;
  const float w = plain_slot(d, S, q, lane).w;
  double e3[MAX_ECL][3];
  for (uint32_t j = 0; j < S.necl; ++j) {
    const double* s = reinterpret_cast<const double*>(r + eoff + 24u * j);
    e3[j][0] = s[0];
    e3[j][1] = s[1];
    e3[j][2] = s[2];
  }
  const uint64_t cbase = set.code0 + static_cast<uint64_t>(p - set.patch0) * set.n_tiles;
  // Per run of 64 tiles: lane k loads tile t0 + k's valid mask and phase
  // range, the loop reads them with readlane (no scalar load per tile), and
  // lane k keeps tile t0 + k's flags, written at once.
  for (uint32_t t0 = 0; t0 < set.n_tiles; t0 += TILE) {
    const uint32_t nt = set.n_tiles - t0 < TILE ? set.n_tiles - t0 : TILE;
    const bool in = lane < nt;
    const uint64_t vl = in ? d.tile_valid[set.tile0 + t0 + lane] : 0ull;
    const double2 rl = in ? d.tile_range[set.tile0 + t0 + lane] : make_double2(0.0, 0.0);
    uint8_t mine = 0;
    for (uint32_t k = 0; k < nt; ++k) {
      uint8_t f = 0;
      if (readlane_u64(vl, k) != 0ull) {  // else coded skip without the test
        const double lo = readlane_f64(rl.x, k), hi = readlane_f64(rl.y, k);
        bool vis = true, ecl = false;
        for (uint32_t j = 0; j < S.necl; ++j) {
          bool misses, holds;
          interval_vs_range(e3[j], lo, hi, &misses, &holds);
          vis = vis && misses;
          ecl = ecl || holds;
        }
        if (w == 0.0f) vis = ecl = true;
        const uint64_t bv = __ballot(vis), be = __ballot(ecl);
        f = static_cast<uint8_t>((bv == ~0ull ? 1u : 0u) | (be == ~0ull ? 2u : 0u));
      }
      if (lane == k) mine = f;
    }
    if (in) d.eflags[cbase + t0 + lane] = mine;
  }
}

// Work code of every (patch, tile) pair: a workgroup per block of 256
// patches of one set, a thread per patch, looping over the set's tiles with
// each tile's lanes in LDS. The patch's axis and cone bounds (FP64) are
// made once, before the tile loop.
// With limb lanes, each PT_FREE pair of a segment that allows moments is
// also split per lane (the host packer's split_lanes, batch_build.cpp, with
// its element-level check off): a lane of the tile whose FP32 cosine with
// the patch axis (the one the code decision takes) is above cos_front sees
// it whole, one below cos_skip sees none of it, and one between sees it in
// part (a limb lane).
template <bool SPLIT>
__global__ __launch_bounds__(kBlock) void codes_kernel(PackDev d) {
  __shared__ float ex[TILE], ey[TILE], ez[TILE];
  __shared__ uint32_t lcp[TILE];
  __shared__ unsigned long long cnt[2];
  const PatchBlock pb = d.pblocks[blockIdx.x];
  const PackSet set = d.sets[pb.set];
  const bool live = threadIdx.x < pb.n;
  const uint32_t p = pb.patch0 + (live ? threadIdx.x : 0u);
  const uint32_t seg = d.patch_seg[p];
  const PackSeg& S = d.psegs[seg];
  const uint32_t lp = p - set.patch0;
  const uint32_t sflags = S.flags;
  const bool cull = (sflags & PSEG_CULL) != 0, front_ok = (sflags & PSEG_FRONT) != 0;
  const uint32_t necl = S.necl;
  // Per patch: the axis in FP32 and the cone's bounds.
  const Cone c = d.cones[p];
  double px, py, pz, plen;
  cone_axis(c, &px, &py, &pz, &plen);
  const float fx = static_cast<float>(px), fy = static_cast<float>(py), fz = static_cast<float>(pz);
  const bool can_all = c.cosw > S.K.c_front, can_none = cull && c.cosw > S.K.c_skip;
  const double cf = can_all ? cos_front(S.K, c.cosw) : 2.0;
  const double cg = can_none ? cos_skip(S.K, c.cosw) : -2.0;
  // Skip when every lane is beyond cos_skip; moments when every lane is
  // inside cos_front (both need a cone that bounds something).
  const bool may_skip = can_none && plen > 0.0, may_front = front_ok && can_all && plen > 0.0;
  // The split's per-lane bounds (split_lanes): cos_front / cos_skip of the
  // cone where it allows them.
  const bool split = SPLIT && front_ok && plen > 0.0;
  const uint32_t use = split ? d.segs[seg].use_mask : 0u;
  if (SPLIT && threadIdx.x < 2) cnt[threadIdx.x] = 0ull;
  for (uint32_t t = 0; t < set.n_tiles; ++t) {
    const uint32_t tg = set.tile0 + t;
    const uint64_t vm = d.tile_valid[tg];
    __syncthreads();  // the previous tile's lanes are read
    if (threadIdx.x < TILE) {
      // A lane outside the tile's valid lanes takes the first valid lane's
      // line of sight: the extremes over the tile are the same, and the
      // lane loop below needs no test per lane. (A tile without valid
      // lanes is coded skip without the loop.)
      const uint32_t src = ((vm >> threadIdx.x) & 1ull) || vm == 0ull
                               ? threadIdx.x
                               : static_cast<uint32_t>(__builtin_ctzll(vm));
      const float4 v = d.tile_en[static_cast<uint64_t>(tg) * TILE + src];
      ex[threadIdx.x] = v.x;
      ey[threadIdx.x] = v.y;
      ez[threadIdx.x] = v.z;
      if (SPLIT) {
        // The lane's (class, grid type) bit: the lanes a segment's split covers.
        const uint32_t m = d.lane_meta[set.lane_off + t * TILE + threadIdx.x];
        lcp[threadIdx.x] = (m & META_VALID) ? cp_bit(meta_class(m), meta_ptype(m)) : 0u;
      }
    }
    __syncthreads();
    const uint64_t co = set.code0 + static_cast<uint64_t>(t) * set.n_patches + lp;  // codes, tile-major
    uint8_t code = PT_SKIP;
    bool has = false;
    uint64_t all = 0, some = 0;
    if (live && vm != 0ull) {
      float emin = 1.0f, emax = -1.0f;
      for (uint32_t j = 0; j < TILE; ++j) {
        const float v = los_cosine(ex[j], ey[j], ez[j], fx, fy, fz);
        emin = fminf(v, emin);  // v is finite: the same as v < emin ? v : emin
        emax = fmaxf(v, emax);
        if (SPLIT && (use & lcp[j])) {
          const double cv = v;
          if (cv > cf) all |= 1ull << j;
          else if (!(cv < cg)) some |= 1ull << j;
        }
      }
      if (may_skip && emax < cg) {
        code = PT_SKIP;
      } else {
        const bool front = may_front && emin > cf;
        if (necl == 0) {
          code = front ? PT_MOM : PT_FREE;
        } else {
          const uint8_t ef = d.eflags[set.code0 + static_cast<uint64_t>(lp) * set.n_tiles + t];  // patch-major
          code = (ef & 2u) ? PT_SKIP : ((ef & 1u) ? (front ? PT_MOM : PT_FREE) : PT_FULL);
        }
      }
      has = split && code == PT_FREE;
    }
    if (live) d.codes[co] = code;
    if (SPLIT) {
      if (!has) all = some = 0;
      if (live) {
        d.has_split[co] = has ? 1u : 0u;
        d.split_all[co] = all;
        d.split_some[co] = some;
      }
      // Counts per wavefront, then one LDS add per wavefront.
      const uint32_t nh = static_cast<uint32_t>(__popcll(__ballot(has)));
      uint32_t nl = static_cast<uint32_t>(__popcll(some));
      for (int off = 32; off > 0; off >>= 1) nl += __shfl_xor(nl, off);
      if (threadIdx.x % TILE == 0 && nh) {
        atomicAdd(&cnt[0], static_cast<unsigned long long>(nh));
        atomicAdd(&cnt[1], static_cast<unsigned long long>(nl));
      }
    }
  }
  if (!SPLIT) return;  // uniform
  __syncthreads();
  if (threadIdx.x < 2 && cnt[threadIdx.x]) atomicAdd(&d.split_stats[threadIdx.x], cnt[threadIdx.x]);
}

// Whether (patch, tile) pair `ci` is taken as a split patch: PT_FREE, split,
// with at most `limb` lanes that see it in part.
__device__ inline bool take_split(const PackDev& d, uint64_t ci) {
  return d.codes[ci] == PT_FREE && d.has_split[ci] && static_cast<uint32_t>(__popcll(d.split_some[ci])) <= d.limb;
}

__device__ inline uint64_t wave_or64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off);
  return v;
}
__device__ inline uint32_t wave_sum(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
// Sum of v over the lanes below this one.
__device__ inline uint32_t wave_excl(uint32_t v, uint32_t lane) {
  uint32_t x = v;
  for (int off = 1; off < TILE; off <<= 1) {
    const uint32_t o = __shfl_up(x, off);
    if (lane >= static_cast<uint32_t>(off)) x += o;
  }
  return x - v;
}

// Limb items of a patch with n limb lanes: 64 lanes each, the last one
// holding the rest; and the lane-group split of item k (the host packer's
// rule: a remainder of at most 16 (32) lanes is shared by 4 (2) groups).
__device__ inline uint32_t limb_items_of(uint32_t n) { return n / TILE + (n % TILE ? 1u : 0u); }
__device__ inline uint32_t limb_split_of(uint32_t n, uint32_t k) {
  const uint32_t rem = n - k * TILE;
  return rem <= TILE / 4 ? 2u : (rem <= TILE / 2 ? 1u : 0u);
}

// Per chunk (a wavefront, a lane per patch): the limb lanes of each patch
// (lanes of its split pairs that see it in part), the union of limb lanes
// over the chunk's tiles, whether the union exceeds LIMB_UNION_MAX (then
// the chunk's split patches go back to the element loops), the number of
// limb items and the fixed-point scale of its sums (engine.cpp's
// build_items, in the same order).
__global__ __launch_bounds__(kBlock) void limb_chunk_kernel(PackDev d) {
  const uint32_t ch = blockIdx.x * (kBlock / TILE) + threadIdx.x / TILE;
  if (ch >= d.n_chunks) return;  // whole wavefront
  const uint32_t q = threadIdx.x % TILE;
  const PackChunk C = d.chunks[ch];
  const PackSet set = d.sets[C.set];
  const PackSeg& S = d.psegs[C.seg];
  const SegDev sg = d.segs[C.seg];
  const uint32_t np = C.ne / PATCH;
  const uint32_t lp = (S.patch0 - set.patch0) + C.e0 / PATCH + q;
  uint32_t nq = 0, nu = 0;
  for (uint32_t t = 0; t < set.n_tiles; ++t) {
    if (!(d.tile_cp[set.tile0 + t] & sg.use_mask)) continue;  // uniform
    uint64_t s = 0;
    if (q < np) {
      const uint64_t ci = set.code0 + static_cast<uint64_t>(t) * set.n_patches + lp;
      if (take_split(d, ci)) s = d.split_some[ci];
    }
    nq += static_cast<uint32_t>(__popcll(s));
    nu += static_cast<uint32_t>(__popcll(wave_or64(s)));
  }
  const bool rev = nu > LIMB_UNION_MAX;
  const bool has = !rev && nu > 0;
  if (!has) nq = 0;
  const uint32_t n_items = wave_sum(limb_items_of(nq));
  // A lane's total over the chunk's limb patches is at most
  // sum |w| * max |I(m)|, m in [0, 1]; sum |w| in patch order.
  const double mw = nq ? d.moms[static_cast<uint64_t>(S.patch0 + C.e0 / PATCH + q) * MOM_STRIDE + 19] : 0.0;
  double wsum = 0.0;
  for (uint32_t k = 0; k < np; ++k) {
    const double v = __shfl(mw, static_cast<int>(k));
    if (__shfl(nq, static_cast<int>(k))) wsum += v;
  }
  if (q == 0) {
    double imax = 1.0;
    const SetDev& sd = d.setdev[C.set];
    if (sg.kind <= KIND_STAR2) {
      const float* a = sg.kind == KIND_STAR1 ? sd.ldc1 : sd.ldc2;
      imax += fabsf(a[0]) + fabsf(a[1]) + fabsf(a[2]) + fabsf(a[3]);
    } else if (sg.kind != KIND_SPOT) {
      imax += fabsf(sd.lin_disc) + fabsf(sd.quad_disc);
    }
    const double bound = 1.01 * imax * wsum;
    int ex = 0;
    if (bound > 0.0) (void)frexp(bound, &ex);  // bound < 2^ex
    d.chunk_shift[ch] = bound > 0.0 ? 61 - ex : 0;
    d.chunk_rev[ch] = rev ? 1u : 0u;
    d.chunk_lcnt[ch] = LimbCounts{has ? 1u : 0u, has ? n_items : 0u, has ? nu : 0u};
  }
  d.chunk_nq[static_cast<uint64_t>(ch) * MAX_ITEM_PATCHES + q] = nq;
}

// The patch masks of candidate (chunk, tile) in a wavefront, a lane per
// patch, and per lane of a split patch: its lane mask (split_all) and its
// group, the first split patch with the same mask.
struct CandMasks {
  bool live = false;  // the tile uses the segment
  uint64_t todo = 0, free = 0, mom = 0, pmom = 0;
  uint64_t mask = 0;  // this lane's split mask (pmom lanes)
  uint64_t lead = 0;  // lanes whose mask no lower pmom lane has
  uint32_t ngrp = 0;
  uint32_t grp = 0;   // this lane's group: its leader's rank among the leaders
};
__device__ inline CandMasks cand_masks(const PackDev& d, const PackChunk& C, const PackSet& set, uint32_t ch,
                                       uint32_t t, uint32_t q) {
  CandMasks r;
  const PackSeg& S = d.psegs[C.seg];
  const uint32_t use = d.segs[C.seg].use_mask;
  if (!(d.tile_cp[set.tile0 + t] & use)) return r;
  r.live = true;
  const uint32_t np = C.ne / PATCH;
  const uint32_t lp = (S.patch0 - set.patch0) + C.e0 / PATCH + q;
  const uint64_t ci = set.code0 + static_cast<uint64_t>(t) * set.n_patches + lp;
  const uint8_t code = q < np ? d.codes[ci] : PT_SKIP;
  const bool split = q < np && d.limb && !d.chunk_rev[ch] && take_split(d, ci);
  r.todo = __ballot(!split && (code == PT_FULL || code == PT_FREE));
  r.free = __ballot(!split && code == PT_FREE);
  r.mom = __ballot(code == PT_MOM);
  r.pmom = __ballot(split);
  if (r.pmom) {
    r.mask = split ? d.split_all[ci] : 0ull;
    // A pmom lane leads its group unless a lower pmom lane has its mask;
    // its leader is the lowest pmom lane with its mask.
    bool lead = split;
    uint32_t first = q;
    for (uint64_t b = r.pmom; b; b &= b - 1ull) {
      const uint32_t k = static_cast<uint32_t>(__builtin_ctzll(b));
      const uint64_t mk = __shfl(r.mask, static_cast<int>(k));
      if (k < q && mk == r.mask) {
        lead = false;
        first = k < first ? k : first;
      }
    }
    r.lead = __ballot(lead);
    r.ngrp = static_cast<uint32_t>(__popcll(r.lead));
    r.grp = split ? static_cast<uint32_t>(__popcll(r.lead & ((1ull << first) - 1ull))) : 0u;
  }
  return r;
}

// Candidate work item of every (chunk, tile), a thread each: its patch
// masks and kernel class (0: adds nothing, not launched; 1 plain; 2 staged
// intervals; 3 beaming / lensing). Candidates with split patches are listed
// for groups_kernel, which counts their lane-mask groups. The partial row
// of a candidate without an item is zeroed here.
__global__ __launch_bounds__(kBlock) void items_kernel(PackDev d) {
  const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  const uint64_t n = static_cast<uint64_t>(d.n_chunks) * d.max_tiles;
  uint32_t cls = 0;
  bool listed = false;
  if (idx < n) {
    const uint32_t ch = static_cast<uint32_t>(idx / d.max_tiles), t = static_cast<uint32_t>(idx % d.max_tiles);
    const PackChunk C = d.chunks[ch];
    const PackSet set = d.sets[C.set];
    uint64_t row = ~0ull;
    uint64_t mom = 0;
    if (t < set.n_tiles) {
      row = set.prow_base + static_cast<uint64_t>(t) * set.n_chunks + C.col;
      const PackSeg& S = d.psegs[C.seg];
      const SegDev sg = d.segs[C.seg];
      if (d.tile_cp[set.tile0 + t] & sg.use_mask) {
        ItemDev it{};
        it.seg = C.seg;
        it.e0 = C.e0;
        it.ne = C.ne;
        it.lane0 = set.lane_off + t * TILE;
        it.prow = row;
        const uint32_t np = C.ne / PATCH;
        const uint64_t co0 = set.code0 + static_cast<uint64_t>(t) * set.n_patches + (S.patch0 - set.patch0) + C.e0 / PATCH;
        const bool can_split = d.limb && !d.chunk_rev[ch];
        for (uint32_t q = 0; q < np; ++q) {
          const uint8_t code = d.codes[co0 + q];
          if (code == PT_FREE && can_split && take_split(d, co0 + q)) {
            it.pmom |= 1ull << q;
            continue;
          }
          if (code == PT_FULL || code == PT_FREE) it.todo |= 1ull << q;
          if (code == PT_FREE) it.free |= 1ull << q;
          if (code == PT_MOM) it.mom |= 1ull << q;
        }
        if (it.todo || it.mom || it.pmom) {
          const bool fx = sg.kind <= KIND_STAR2 && S.posn;
          const bool staged = (it.todo & ~it.free) != 0ull && S.necl >= 1 && S.necl <= LDS_ECL_MAX;
          cls = fx ? 3u : (staged ? 2u : 1u);
          mom = it.mom;
          listed = it.pmom != 0ull;
          d.cand[idx] = it;
        }
      }
    }
    d.cand_info[idx] = cls | ((mom ? 1u : 0u) << 16);
    if (cls == 0 && row != ~0ull) {
      double2* o = reinterpret_cast<double2*>(d.partials + row * TILE);
      for (int j = 0; j < TILE / 2; ++j) o[j] = make_double2(0.0, 0.0);
    }
  }
  // The candidates with split patches, appended by wavefront.
  if (d.limb) {
    const uint64_t b = __ballot(listed);
    if (b) {
      uint32_t base = 0;
      const uint32_t lane = threadIdx.x % TILE;
      if (lane == static_cast<uint32_t>(__builtin_ctzll(b))) base = atomicAdd(d.glist_n, static_cast<uint32_t>(__popcll(b)));
      base = static_cast<uint32_t>(__shfl(static_cast<int>(base), __builtin_ctzll(b)));
      if (listed) d.glist[base + static_cast<uint32_t>(__popcll(b & ((1ull << lane) - 1ull)))] = static_cast<uint32_t>(idx);
    }
  }
}

// Lane-mask groups of each listed candidate, a wavefront each (over the
// list): their number goes into the candidate's item and info.
__global__ __launch_bounds__(kBlock) void groups_kernel(PackDev d) {
  const uint32_t n = *d.glist_n;
  const uint32_t lane = threadIdx.x % TILE;
  const uint32_t nw = gridDim.x * (kBlock / TILE);
  for (uint32_t i = blockIdx.x * (kBlock / TILE) + threadIdx.x / TILE; i < n; i += nw) {
    const uint32_t idx = d.glist[i];
    const uint32_t ch = idx / d.max_tiles, t = idx % d.max_tiles;
    const PackChunk C = d.chunks[ch];
    const CandMasks r = cand_masks(d, C, d.sets[C.set], ch, t, lane);
    if (lane == 0) {
      d.cand[idx].ngrp = r.ngrp;
      d.cand_info[idx] |= r.ngrp << 8;
    }
  }
}

struct CandCount {
  __device__ ItemCounts operator()(uint32_t v) const {
    const uint32_t c = v & 0xffu, g = (v >> 8) & 0xffu, hm = (v >> 16) & 1u;
    return ItemCounts{c == 1 ? 1u : 0u, c == 2 ? 1u : 0u, c == 3 ? 1u : 0u, g, hm + g};
  }
};
struct AddCounts {
  __device__ ItemCounts operator()(const ItemCounts& a, const ItemCounts& b) const {
    return ItemCounts{a.a + b.a, a.b + b.b, a.c + b.c, a.masks + b.masks, a.moms + b.moms};
  }
};
struct AddLimb {
  __device__ LimbCounts operator()(const LimbCounts& a, const LimbCounts& b) const {
    return LimbCounts{a.chunks + b.chunks, a.items + b.items, a.lanes + b.lanes};
  }
};

// Totals of the two scans: the last exclusive prefix plus the last entry.
__global__ void totals_kernel(PackDev d) {
  PackCounts c{};
  const uint64_t n = static_cast<uint64_t>(d.n_chunks) * d.max_tiles;
  if (n) {
    const ItemCounts last = CandCount()(d.cand_info[n - 1]);
    c.items = AddCounts()(d.cand_pos[n - 1], last);
  }
  if (d.n_chunks && d.limb) c.limb = AddLimb()(d.chunk_lpos[d.n_chunks - 1], d.chunk_lcnt[d.n_chunks - 1]);
  *d.counts = c;
}

// The launch position of a candidate's item: plain, then staged intervals,
// then beaming / lensing, each class in candidate order.
__device__ inline uint32_t item_slot(uint32_t cls, const ItemCounts& pos, const ItemCounts& tot) {
  return cls == 1 ? pos.a : (cls == 2 ? tot.a + pos.b : tot.a + tot.b + pos.c);
}

// Items in launch order, a thread per candidate, with their first moment
// entry and lane-mask group; the entry of its PT_MOM patches (moment sums
// are made per entry: the patch mask and the item's first patch).
__global__ __launch_bounds__(kBlock) void scatter_kernel(PackDev d) {
  const uint64_t n = static_cast<uint64_t>(d.n_chunks) * d.max_tiles;
  const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= n) return;
  const uint32_t cls = d.cand_info[idx] & 0xffu;
  if (!cls) return;
  const ItemCounts tot = d.counts->items;
  const ItemCounts pos = d.cand_pos[idx];
  ItemDev it = d.cand[idx];
  it.lm_off = pos.masks;
  it.imom = pos.moms;
  d.items[item_slot(cls, pos, tot)] = it;
  if (it.mom && pos.moms < d.cap_imoms) {
    d.ment_base[pos.moms] = d.segs[it.seg].mom_off + it.e0 / PATCH;
    d.ment_mask[pos.moms] = it.mom;
  }
}

// The lane-mask groups of each listed candidate's item, a wavefront each:
// the group masks (split_all of each group's first split patch, in order)
// and one moment entry per group (its split patches).
__global__ __launch_bounds__(kBlock) void scatter_groups_kernel(PackDev d) {
  const uint32_t n = *d.glist_n;
  const uint32_t lane = threadIdx.x % TILE;
  const uint32_t nw = gridDim.x * (kBlock / TILE);
  for (uint32_t i = blockIdx.x * (kBlock / TILE) + threadIdx.x / TILE; i < n; i += nw) {
    const uint32_t idx = d.glist[i];
    const uint32_t ch = idx / d.max_tiles, t = idx % d.max_tiles;
    const PackChunk C = d.chunks[ch];
    const CandMasks r = cand_masks(d, C, d.sets[C.set], ch, t, lane);
    const ItemCounts pos = d.cand_pos[idx];
    const bool split = (r.pmom >> lane) & 1ull;
    if ((r.lead >> lane) & 1ull) {
      const uint64_t g = pos.masks + static_cast<uint64_t>(__popcll(r.lead & ((1ull << lane) - 1ull)));
      if (g < d.cap_masks) d.lane_masks[g] = r.mask;
    }
    const uint32_t base = d.segs[C.seg].mom_off + C.e0 / PATCH;
    const uint64_t e0 = static_cast<uint64_t>(pos.moms) + ((d.cand_info[idx] >> 16) & 1u);
    for (uint32_t k = 0; k < r.ngrp; ++k) {
      const uint64_t gm = __ballot(split && r.grp == k);
      if (lane == 0 && e0 + k < d.cap_imoms) {
        d.ment_base[e0 + k] = base;
        d.ment_mask[e0 + k] = gm;
      }
    }
  }
}

// Moment entries: the sum, in patch order, of the moments of the patches
// in each entry's mask. Ten lanes per entry, one pair of the 20 components
// each, six entries per wavefront (lanes 60-63 idle): an entry's patch rows
// are read, and its sums written, as whole 160-byte runs.
constexpr uint32_t IM_LANES = MOM_STRIDE / 2, IM_PER_WAVE = TILE / IM_LANES;
constexpr uint32_t IM_PER_BLOCK = IM_PER_WAVE * (kBlock / TILE);
static_assert(MOM_STRIDE % 2 == 0 && IM_PER_WAVE >= 1, "component pairs per lane");
__global__ __launch_bounds__(kBlock) void imoms_kernel(PackDev d) {
  const uint32_t wave = threadIdx.x / TILE, lane = threadIdx.x % TILE;
  const uint32_t sub = lane / IM_LANES, c = lane % IM_LANES;
  if (sub >= IM_PER_WAVE) return;
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * IM_PER_BLOCK + wave * IM_PER_WAVE + sub;
  const uint64_t n = d.counts->items.moms;
  if (e >= n || e >= d.cap_imoms) return;
  const double* m0 = d.moms + static_cast<uint64_t>(d.ment_base[e]) * MOM_STRIDE + 2u * c;
  double sx = 0.0, sy = 0.0;
  for (uint64_t mm = d.ment_mask[e]; mm; mm &= mm - 1ull) {
    const double2 v = *reinterpret_cast<const double2*>(m0 + static_cast<uint64_t>(__builtin_ctzll(mm)) * MOM_STRIDE);
    sx = sx + v.x;
    sy = sy + v.y;
  }
  reinterpret_cast<double2*>(d.imoms + e * MOM_STRIDE)[c] = make_double2(sx, sy);
}

// The limb tables of each chunk with limb lanes, a wavefront (a workgroup)
// per chunk and a lane per patch: its LimbChunk, its lane list (the union
// of its limb lanes, by tile and then lane), and per patch its limb items
// (64 lanes each, the last one split into lane groups) with each item's 64
// union indices (0xFFFF unused), in the host packer's order. A chunk's
// union indices are assembled in LDS when they fit and then copied out
// whole; a larger chunk writes them in place.
constexpr uint32_t LIMB_FILL_LDS = 16384;  // union indices (256 limb items)
__global__ __launch_bounds__(TILE) void limb_fill_kernel(PackDev d) {
  __shared__ uint64_t lbuf8[LIMB_FILL_LDS / 4];
  uint16_t* lbuf = reinterpret_cast<uint16_t*>(lbuf8);
  const uint32_t ch = blockIdx.x;
  if (ch >= d.n_chunks) return;  // whole workgroup
  const LimbCounts cnt = d.chunk_lcnt[ch];
  if (!cnt.chunks) return;
  const LimbCounts pos = d.chunk_lpos[ch];
  if (pos.chunks >= d.n_chunks || static_cast<uint64_t>(pos.items) + cnt.items > d.cap_litems ||
      static_cast<uint64_t>(pos.lanes) + cnt.lanes > d.cap_llanes)
    return;
  const uint32_t q = threadIdx.x;
  const PackChunk C = d.chunks[ch];
  const PackSet set = d.sets[C.set];
  const PackSeg& S = d.psegs[C.seg];
  const uint32_t use = d.segs[C.seg].use_mask;
  const uint32_t np = C.ne / PATCH;
  const uint32_t nq = d.chunk_nq[static_cast<uint64_t>(ch) * MAX_ITEM_PATCHES + q];
  const uint32_t ni = limb_items_of(nq);
  const uint32_t ib = wave_excl(ni, q);  // this patch's first item, relative to the chunk's
  const bool staged = static_cast<uint64_t>(cnt.items) * TILE <= LIMB_FILL_LDS;
  uint16_t* out = staged ? lbuf : d.lidx + static_cast<uint64_t>(pos.items) * TILE;
  if (q == 0) {
    LimbChunk lc{};
    lc.seg = C.seg;
    lc.chunk = C.col;
    lc.item0 = pos.items;
    lc.n_items = cnt.items;
    lc.lane_off = pos.lanes;
    lc.n_lanes = cnt.lanes;
    lc.shift = d.chunk_shift[ch];
    d.lchunks[pos.chunks] = lc;
  }
  for (uint32_t k = 0; k < ni; ++k) d.litems[pos.items + ib + k] = LimbItem{C.e0 + q * PATCH, limb_split_of(nq, k)};
  // Unused slots: those past the last item's lanes, in each of its groups.
  if (ni) {
    const uint32_t k = ni - 1, sp = limb_split_of(nq, k), gl = TILE >> sp, rem = nq - k * TILE;
    for (uint32_t g = 0; g < (1u << sp); ++g)
      for (uint32_t slot = rem; slot < gl; ++slot) out[(ib + k) * TILE + g * gl + slot] = 0xFFFFu;
  }
  const uint32_t lp = (S.patch0 - set.patch0) + C.e0 / PATCH + q;
  uint32_t run = 0, posq = 0;
  for (uint32_t t = 0; t < set.n_tiles; ++t) {
    if (!(d.tile_cp[set.tile0 + t] & use)) continue;  // uniform
    uint64_t sm = 0;
    if (q < np) {
      const uint64_t ci = set.code0 + static_cast<uint64_t>(t) * set.n_patches + lp;
      if (take_split(d, ci)) sm = d.split_some[ci];
    }
    const uint64_t u = wave_or64(sm);
    // Lane q lists lane q of the tile if it is a limb lane.
    if ((u >> q) & 1ull)
      d.llanes[pos.lanes + run + static_cast<uint32_t>(__popcll(u & ((1ull << q) - 1ull)))] = set.lane_off + t * TILE + q;
    for (uint64_t b = sm; b; b &= b - 1ull, ++posq) {
      const uint32_t jb = static_cast<uint32_t>(__builtin_ctzll(b));
      const uint16_t ui = static_cast<uint16_t>(run + static_cast<uint32_t>(__popcll(u & ((1ull << jb) - 1ull))));
      const uint32_t k = posq / TILE, slot = posq % TILE;
      const uint32_t sp = limb_split_of(nq, k), gl = TILE >> sp;
      for (uint32_t g = 0; g < (1u << sp); ++g) out[(ib + k) * TILE + g * gl + slot] = ui;
    }
    run += static_cast<uint32_t>(__popcll(u));
  }
  if (staged) {
    __syncthreads();
    // Copied out 8 bytes per lane and step (the chunk's indices start at a
    // multiple of 64 entries).
    const uint64_t* src = lbuf8;
    uint64_t* dst = reinterpret_cast<uint64_t*>(d.lidx + static_cast<uint64_t>(pos.items) * TILE);
    const uint32_t n8 = cnt.items * TILE / 4;
    for (uint32_t i = q; i < n8; i += TILE) dst[i] = src[i];
  }
}

uint32_t blocks(uint64_t n, uint32_t b = kBlock) { return static_cast<uint32_t>((n + b - 1) / b); }

}  // namespace


size_t pack_scan_temp_bytes(uint64_t n_cand, uint64_t n_chunks) {
  size_t a = 0, b = 0;
  auto in = rocprim::make_transform_iterator(static_cast<const uint32_t*>(nullptr), CandCount());
  (void)rocprim::exclusive_scan(nullptr, a, in, static_cast<ItemCounts*>(nullptr), ItemCounts{0, 0, 0, 0, 0},
                                static_cast<size_t>(n_cand), AddCounts());
  (void)rocprim::exclusive_scan(nullptr, b, static_cast<const LimbCounts*>(nullptr), static_cast<LimbCounts*>(nullptr),
                                LimbCounts{0, 0, 0}, static_cast<size_t>(n_chunks), AddLimb());
  return a > b ? a : b;
}

hipError_t launch_pack_count(const PackDev& d0, void* scan_temp, size_t scan_temp_bytes, hipStream_t st) {
  // After the sort's two passes the sorted order is in keys / vals.
  PackDev d = d0;
  d.keys_s = d0.keys;
  d.vals_s = d0.vals;
  hipError_t e = hipSuccess;
  if (d.limb) {
    e = hipMemsetAsync(d.split_stats, 0, 2 * sizeof(unsigned long long), st);
    if (e != hipSuccess) return e;
  }
  if (d.n_segs) {
    e = hipMemsetAsync(d.wmax, 0, static_cast<size_t>(d.n_segs) * 2u * sizeof(uint32_t), st);
    if (e != hipSuccess) return e;
  }
  if (d.n_raw) {
    hipLaunchKernelGGL(keys_kernel, dim3(d0.n_ktiles), dim3(kBlock), 0, st, d0);
    if (d0.n_grids) hipLaunchKernelGGL(segsort_kernel, dim3(d0.n_grids), dim3(SS_BLOCK), 0, st, d0);
  }
  if (d.n_segs) hipLaunchKernelGGL(seg_flags_kernel, dim3(blocks(d.n_segs)), dim3(kBlock), 0, st, d);
  if (d.n_segs) hipLaunchKernelGGL(pad_kernel, dim3(d.n_segs), dim3(PATCH), 0, st, d);
  if (d.n_patches) hipLaunchKernelGGL(patches_kernel, dim3(blocks(d.n_patches, PATCH_RUN)), dim3(TILE), 0, st, d);
  if (d.n_tiles) hipLaunchKernelGGL(tiles_kernel, dim3(d.n_tiles), dim3(TILE), 0, st, d);
  if (d.n_ecl_patches) hipLaunchKernelGGL(eclflags_kernel, dim3(d.n_ecl_patches), dim3(TILE), 0, st, d);
  if (d.n_pblocks && d.max_tiles) {
    if (d.limb) hipLaunchKernelGGL(codes_kernel<true>, dim3(d.n_pblocks), dim3(kBlock), 0, st, d);
    else hipLaunchKernelGGL(codes_kernel<false>, dim3(d.n_pblocks), dim3(kBlock), 0, st, d);
  }
  const uint32_t wpb = kBlock / TILE;  // wavefronts per workgroup
  if (d.limb && d.n_chunks) {
    hipLaunchKernelGGL(limb_chunk_kernel, dim3(blocks(d.n_chunks, wpb)), dim3(kBlock), 0, st, d);
    size_t tb = scan_temp_bytes;
    e = rocprim::exclusive_scan(scan_temp, tb, d.chunk_lcnt, d.chunk_lpos, LimbCounts{0, 0, 0},
                                static_cast<size_t>(d.n_chunks), AddLimb(), st);
    if (e != hipSuccess) return e;
  }
  const uint64_t nc = static_cast<uint64_t>(d.n_chunks) * d.max_tiles;
  if (nc) {
    if (d.limb) {
      e = hipMemsetAsync(d.glist_n, 0, sizeof(uint32_t), st);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(items_kernel, dim3(blocks(nc)), dim3(kBlock), 0, st, d);
    // The list's length stays on the device: a fixed grid strides over it.
    if (d.limb) hipLaunchKernelGGL(groups_kernel, dim3(kListBlocks), dim3(kBlock), 0, st, d);
    auto in = rocprim::make_transform_iterator(static_cast<const uint32_t*>(d.cand_info), CandCount());
    size_t tb = scan_temp_bytes;
    e = rocprim::exclusive_scan(scan_temp, tb, in, d.cand_pos, ItemCounts{0, 0, 0, 0, 0}, static_cast<size_t>(nc),
                                AddCounts(), st);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(totals_kernel, dim3(1), dim3(1), 0, st, d);
  return hipGetLastError();
}

hipError_t launch_pack_fill(const PackDev& d, uint64_t n_entries, hipStream_t st) {
  const uint64_t nc = static_cast<uint64_t>(d.n_chunks) * d.max_tiles;
  if (nc) {
    hipLaunchKernelGGL(scatter_kernel, dim3(blocks(nc)), dim3(kBlock), 0, st, d);
    if (d.limb) hipLaunchKernelGGL(scatter_groups_kernel, dim3(kListBlocks), dim3(kBlock), 0, st, d);
  }
  if (n_entries) hipLaunchKernelGGL(imoms_kernel, dim3(blocks(n_entries, IM_PER_BLOCK)), dim3(kBlock), 0, st, d);
  if (d.limb && d.n_chunks) hipLaunchKernelGGL(limb_fill_kernel, dim3(d.n_chunks), dim3(TILE), 0, st, d);
  return hipGetLastError();
}

}  // namespace lch
