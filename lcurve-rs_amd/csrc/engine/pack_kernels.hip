// The per-set pack on the GPU (pack_kernels.h). Everything here is exact
// FP64 / integer work in a fixed order: built without FMA contraction, the
// tables equal the host pack's bit for bit.
#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include "pack_kernels.h"

namespace lch {

namespace {

constexpr uint32_t kBlock = 256;

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

// Sort key of every raw element: its segment above the spatial key. With
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
    d.keys[e] = (G.seg_of[k] << 16) | direction_key16(v.x, v.y, v.z);
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

// The record of sorted position p, an element of segment seg.
__device__ inline void record_at(const PackDev& d, uint64_t p, uint32_t seg) {
  const uint32_t e = d.vals_s[p];
  const PackSeg S = d.psegs[seg];
  const uint32_t flags = d.segs[seg].flags;
  const uint32_t rank = static_cast<uint32_t>(p) - S.sorted0;
  const uint32_t stride = stride_of(S);
  const uint32_t eoff = S.posn ? 32u : 16u;
  char* r = d.recs + S.rec_off + static_cast<uint64_t>(rank) * stride;
  float4 f = d.raw_df[e];
  if (flags & SEG_LINW) {
    const double inv = ldexp(1.0, -seg_linw_exp(flags));
    const double sc = sqrt(static_cast<double>(f.w)) * inv;
    f = make_float4(static_cast<float>(static_cast<double>(f.x) * sc), static_cast<float>(static_cast<double>(f.y) * sc),
                    static_cast<float>(static_cast<double>(f.z) * sc), static_cast<float>(sc));
  }
  store_f4(r, f);
  if (S.posn) {
    float4 q = d.raw_pos[S.pos0 + (e - S.e0)];
    q.w = 0.0f;
    store_f4(r + 16, q);
  }
  const uint32_t i0 = d.raw_eoff[e];
  for (uint32_t j = 0; j < S.necl; ++j) {
    const double2 pr = d.raw_ecl[i0 + j];
    double* o = reinterpret_cast<double*>(r + eoff + 24u * j);
    o[0] = pr.x;
    o[1] = pr.y;
    o[2] = pr.y - 1.0;
  }
}

// Element records in sorted order: each segment's elements by spatial key,
// ties in raw (grid) order. Linear-law segments (SEG_LINW) take (s d, s),
// s = sqrt(w) / 2^k, in place of (d, w).
__global__ __launch_bounds__(kBlock) void records_kernel(PackDev d) {
  const KeyTile kt = d.ktiles[blockIdx.x];
  const PackGrid G = d.grids[kt.grid];
  for (uint32_t i = threadIdx.x; i < kt.n; i += kBlock) {
    const uint64_t p = G.e0 + kt.i0 + i;
    record_at(d, p, d.keys_s[p] >> 16);
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

// Cone and (for moment-capable segments) direction moments of each patch,
// from the plain values: one wavefront per patch, a lane per slot. Each sum
// is tree_sum64's: its first level (lane i + lane i + 32) by a shuffle, the
// rest by one lane per sum over the 32 partial sums in LDS.
__global__ __launch_bounds__(TILE) void patches_kernel(PackDev d) {
  constexpr uint32_t NC = MOM_STRIDE + 3;  // cone sums x, y, z, then the moments
  // Row stride of 23 doubles: conflict-free row writes and column reads.
  __shared__ double part[PATCH / 2][NC];
  const uint32_t p = blockIdx.x;
  const uint32_t j = threadIdx.x;
  const PackSeg S = d.psegs[d.patch_seg[p]];
  const uint32_t q = p - S.patch0;
  const uint32_t first = q * PATCH;
  const uint32_t nr = S.n_real > first ? (S.n_real - first < PATCH ? S.n_real - first : PATCH) : 0u;
  const bool real = j < nr;
  const float4 f = plain_slot(d, S, q, j);
  const bool mom = (S.flags & PSEG_MOM) != 0;
  auto level1 = [&](double v, uint32_t c) {
    const double o = __shfl_xor(v, 32);
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
  ConeBuilder cb;
  cb.set_sum(__shfl(sum, 0), __shfl(sum, 1), __shfl(sum, 2), nr > 0);
  cb.axis();
  cb.mn = wave_min(real ? cb.cosine(f.x, f.y, f.z) : 1.0);
  const Cone c = cb.finish();
  if (j == 0) d.cones[p] = c;
  if (j >= 3 && j < NC) d.moms[static_cast<uint64_t>(p) * MOM_STRIDE + (j - 3)] = mom ? sum : 0.0;
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
  // Lane t % 64 keeps tile t's flags; each run of 64 tiles is written at once.
  uint8_t mine = 0;
  for (uint32_t t = 0; t < set.n_tiles; ++t) {
    const uint32_t tg = set.tile0 + t;
    uint8_t f = 0;
    if (d.tile_valid[tg] != 0ull) {  // else coded skip without the test
      const double2 rg = d.tile_range[tg];
      bool vis = true, ecl = false;
      for (uint32_t j = 0; j < S.necl; ++j) {
        bool misses, holds;
        interval_vs_range(e3[j], rg.x, rg.y, &misses, &holds);
        vis = vis && misses;
        ecl = ecl || holds;
      }
      if (w == 0.0f) vis = ecl = true;
      const uint64_t bv = __ballot(vis), be = __ballot(ecl);
      f = static_cast<uint8_t>((bv == ~0ull ? 1u : 0u) | (be == ~0ull ? 2u : 0u));
    }
    if (lane == t % TILE) mine = f;
    if (t % TILE == TILE - 1 || t + 1 == set.n_tiles) {
      const uint32_t t0 = t - t % TILE;
      if (lane <= t - t0) d.eflags[cbase + t0 + lane] = mine;
    }
  }
}

// Work code of every (patch, tile) pair: a workgroup per (block of 256
// patches of one set, tile), the tile's lanes in LDS, a thread per patch.
__global__ __launch_bounds__(kBlock) void codes_kernel(PackDev d) {
  __shared__ float ex[TILE], ey[TILE], ez[TILE];
  const PatchBlock pb = d.pblocks[blockIdx.x];
  const PackSet set = d.sets[pb.set];
  const uint32_t t = blockIdx.y;
  if (t >= set.n_tiles) return;  // uniform over the workgroup, before the barrier
  const uint32_t tg = set.tile0 + t;
  if (threadIdx.x < TILE) {
    const float4 v = d.tile_en[static_cast<uint64_t>(tg) * TILE + threadIdx.x];
    ex[threadIdx.x] = v.x;
    ey[threadIdx.x] = v.y;
    ez[threadIdx.x] = v.z;
  }
  __syncthreads();
  if (threadIdx.x >= pb.n) return;
  const uint32_t p = pb.patch0 + threadIdx.x;
  const PackSeg S = d.psegs[d.patch_seg[p]];
  const uint32_t lp = p - set.patch0;
  const uint64_t ci = set.code0 + static_cast<uint64_t>(lp) * set.n_tiles + t;  // eclipse flags, patch-major
  const uint64_t vm = d.tile_valid[tg];
  uint8_t code = PT_SKIP;
  if (vm != 0ull) {
    const Cone c = d.cones[p];
    double px, py, pz, plen;
    cone_axis(c, &px, &py, &pz, &plen);
    const float fx = static_cast<float>(px), fy = static_cast<float>(py), fz = static_cast<float>(pz);
    float emin = 1.0f, emax = -1.0f;
    for (uint32_t j = 0; j < TILE; ++j) {
      if (!((vm >> j) & 1ull)) continue;
      const float v = los_cosine(ex[j], ey[j], ez[j], fx, fy, fz);
      emin = v < emin ? v : emin;
      emax = v > emax ? v : emax;
    }
    const bool cull = (S.flags & PSEG_CULL) != 0, front_ok = (S.flags & PSEG_FRONT) != 0;
    if (cull && plen > 0.0 && c.cosw > S.K.c_skip && emax < cos_skip(S.K, c.cosw)) {
      code = PT_SKIP;
    } else {
      const bool front = front_ok && plen > 0.0 && c.cosw > S.K.c_front && emin > cos_front(S.K, c.cosw);
      if (S.necl == 0) {
        code = front ? PT_MOM : PT_FREE;
      } else {
        const uint8_t ef = d.eflags[ci];
        code = (ef & 2u) ? PT_SKIP : ((ef & 1u) ? (front ? PT_MOM : PT_FREE) : PT_FULL);
      }
    }
  }
  d.codes[set.code0 + static_cast<uint64_t>(t) * set.n_patches + lp] = code;
}

// Candidate work item of every (chunk, tile): its patch masks and kernel
// class (0: adds nothing, not launched; 1 plain; 2 staged intervals; 3
// beaming / lensing).
__global__ __launch_bounds__(kBlock) void items_kernel(PackDev d) {
  const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= static_cast<uint64_t>(d.n_chunks) * d.max_tiles) return;
  const uint32_t ch = static_cast<uint32_t>(idx / d.max_tiles), t = static_cast<uint32_t>(idx % d.max_tiles);
  const PackChunk C = d.chunks[ch];
  const PackSet set = d.sets[C.set];
  uint8_t cls = 0;
  uint64_t row = ~0ull;
  if (t < set.n_tiles) {
    row = set.prow_base + static_cast<uint64_t>(t) * set.n_chunks + C.col;
    const PackSeg S = d.psegs[C.seg];
    const SegDev sg = d.segs[C.seg];
    if (d.tile_cp[set.tile0 + t] & sg.use_mask) {
      ItemDev it{};
      it.seg = C.seg;
      it.e0 = C.e0;
      it.ne = C.ne;
      it.lane0 = set.lane_off + t * TILE;
      it.prow = row;
      const uint32_t q0 = (S.patch0 - set.patch0) + C.e0 / PATCH;
      const uint32_t np = C.ne / PATCH;
      const uint8_t* codes = d.codes + set.code0 + static_cast<uint64_t>(t) * set.n_patches + q0;
      for (uint32_t q = 0; q < np; ++q) {
        const uint8_t code = codes[q];
        if (code == PT_FULL || code == PT_FREE) it.todo |= 1ull << q;
        if (code == PT_FREE) it.free |= 1ull << q;
        if (code == PT_MOM) it.mom |= 1ull << q;
      }
      if (it.todo || it.mom) {
        const bool fx = sg.kind <= KIND_STAR2 && S.posn;
        const bool staged = (it.todo & ~it.free) != 0ull && S.necl >= 1 && S.necl <= LDS_ECL_MAX;
        cls = fx ? 3 : (staged ? 2 : 1);
        d.cand[idx] = it;
      }
    }
  }
  d.cand_class[idx] = cls;
  if (cls == 0 && row != ~0ull) {
    double2* o = reinterpret_cast<double2*>(d.partials + row * TILE);
    for (int j = 0; j < TILE / 2; ++j) o[j] = make_double2(0.0, 0.0);
  }
}

struct ClassCount {
  __device__ ItemCounts operator()(uint8_t c) const {
    return ItemCounts{c == 1 ? 1u : 0u, c == 2 ? 1u : 0u, c == 3 ? 1u : 0u};
  }
};
struct AddCounts {
  __device__ ItemCounts operator()(const ItemCounts& a, const ItemCounts& b) const {
    return ItemCounts{a.a + b.a, a.b + b.b, a.c + b.c};
  }
};

// Items in launch order (plain, then staged intervals, then beaming /
// lensing), each class in candidate order; each item's moment entry is its
// own index.
__global__ __launch_bounds__(kBlock) void scatter_kernel(PackDev d) {
  const uint64_t n = static_cast<uint64_t>(d.n_chunks) * d.max_tiles;
  const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (idx >= n) return;
  const uint8_t lc = d.cand_class[n - 1];
  ItemCounts tot = d.cand_pos[n - 1];
  tot.a += lc == 1 ? 1u : 0u;
  tot.b += lc == 2 ? 1u : 0u;
  tot.c += lc == 3 ? 1u : 0u;
  if (idx == 0) *d.counts = tot;
  const uint8_t cls = d.cand_class[idx];
  if (!cls) return;
  const ItemCounts pos = d.cand_pos[idx];
  const uint32_t out = cls == 1 ? pos.a : (cls == 2 ? tot.a + pos.b : tot.a + tot.b + pos.c);
  ItemDev it = d.cand[idx];
  it.imom = out;
  d.items[out] = it;
}

// Summed moments of each item's PT_MOM patches, in patch order: a thread
// per item.
__global__ __launch_bounds__(kBlock) void imoms_kernel(PackDev d, uint64_t capacity) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= capacity) return;
  const ItemCounts n = *d.counts;
  if (i >= static_cast<uint64_t>(n.a) + n.b + n.c) return;
  const ItemDev it = d.items[i];
  if (!it.mom) return;
  const double* m0 = d.moms + (static_cast<uint64_t>(d.segs[it.seg].mom_off) + it.e0 / PATCH) * MOM_STRIDE;
  double s[MOM_STRIDE];
#pragma unroll
  for (uint32_t c = 0; c < MOM_STRIDE; ++c) s[c] = 0.0;
  for (uint64_t mm = it.mom; mm; mm &= mm - 1ull) {
    const double2* a = reinterpret_cast<const double2*>(m0 + static_cast<uint64_t>(__builtin_ctzll(mm)) * MOM_STRIDE);
#pragma unroll
    for (uint32_t c = 0; c < MOM_STRIDE / 2; ++c) {
      const double2 v = a[c];
      s[2 * c] = s[2 * c] + v.x;
      s[2 * c + 1] = s[2 * c + 1] + v.y;
    }
  }
  double2* o = reinterpret_cast<double2*>(d.imoms + i * MOM_STRIDE);
#pragma unroll
  for (uint32_t c = 0; c < MOM_STRIDE / 2; ++c) o[c] = make_double2(s[2 * c], s[2 * c + 1]);
}

uint32_t blocks(uint64_t n, uint32_t b = kBlock) { return static_cast<uint32_t>((n + b - 1) / b); }

}  // namespace

size_t pack_sort_temp_bytes(uint64_t n) {
  size_t bytes = 0;
  (void)rocprim::radix_sort_pairs(nullptr, bytes, static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                  static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                  static_cast<size_t>(n), 0, 32);
  return bytes;
}

size_t pack_scan_temp_bytes(uint64_t n) {
  size_t bytes = 0;
  auto in = rocprim::make_transform_iterator(static_cast<const uint8_t*>(nullptr), ClassCount());
  (void)rocprim::exclusive_scan(nullptr, bytes, in, static_cast<ItemCounts*>(nullptr), ItemCounts{0, 0, 0},
                                static_cast<size_t>(n), AddCounts());
  return bytes;
}

hipError_t launch_pack(const PackDev& d, void* sort_temp, size_t sort_temp_bytes, void* scan_temp,
                       size_t scan_temp_bytes, uint32_t key_bits, hipStream_t st) {
  hipError_t e = hipSuccess;
  if (d.n_segs) {
    e = hipMemsetAsync(d.wmax, 0, static_cast<size_t>(d.n_segs) * 2u * sizeof(uint32_t), st);
    if (e != hipSuccess) return e;
  }
  if (d.n_raw) {
    hipLaunchKernelGGL(keys_kernel, dim3(d.n_ktiles), dim3(kBlock), 0, st, d);
    size_t tb = sort_temp_bytes;
    e = rocprim::radix_sort_pairs(sort_temp, tb, d.keys, d.keys_s, d.vals, d.vals_s, static_cast<size_t>(d.n_raw),
                                  0, key_bits, st);
    if (e != hipSuccess) return e;
  }
  if (d.n_segs) hipLaunchKernelGGL(seg_flags_kernel, dim3(blocks(d.n_segs)), dim3(kBlock), 0, st, d);
  if (d.n_raw) hipLaunchKernelGGL(records_kernel, dim3(d.n_ktiles), dim3(kBlock), 0, st, d);
  if (d.n_segs) hipLaunchKernelGGL(pad_kernel, dim3(d.n_segs), dim3(PATCH), 0, st, d);
  if (d.n_patches) hipLaunchKernelGGL(patches_kernel, dim3(d.n_patches), dim3(TILE), 0, st, d);
  if (d.n_tiles) hipLaunchKernelGGL(tiles_kernel, dim3(d.n_tiles), dim3(TILE), 0, st, d);
  if (d.n_ecl_patches) hipLaunchKernelGGL(eclflags_kernel, dim3(d.n_ecl_patches), dim3(TILE), 0, st, d);
  if (d.n_pblocks && d.max_tiles)
    hipLaunchKernelGGL(codes_kernel, dim3(d.n_pblocks, d.max_tiles), dim3(kBlock), 0, st, d);
  const uint64_t nc = static_cast<uint64_t>(d.n_chunks) * d.max_tiles;
  if (nc) {
    hipLaunchKernelGGL(items_kernel, dim3(blocks(nc)), dim3(kBlock), 0, st, d);
    auto in = rocprim::make_transform_iterator(static_cast<const uint8_t*>(d.cand_class), ClassCount());
    size_t tb = scan_temp_bytes;
    e = rocprim::exclusive_scan(scan_temp, tb, in, d.cand_pos, ItemCounts{0, 0, 0}, static_cast<size_t>(nc),
                                AddCounts(), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(scatter_kernel, dim3(blocks(nc)), dim3(kBlock), 0, st, d);
    hipLaunchKernelGGL(imoms_kernel, dim3(blocks(nc)), dim3(kBlock), 0, st, d, nc);
  } else {
    e = hipMemsetAsync(d.counts, 0, sizeof(ItemCounts), st);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

}  // namespace lch
