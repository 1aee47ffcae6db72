// HIP kernels of the flux engine for CDNA4 (gfx950).
//
// flux_items_kernel — the hot path. One wavefront = one work item =
//   64 sub-phases (one per lane) x one chunk of surface elements, stored as
//   patches of 64 elements. Each lane owns one sub-phase and loops over the
//   item's patches; the element record address is wave-uniform, so records
//   arrive through scalar loads and feed the VALU as scalar operands. There
//   is no cross-lane reduction in the inner loop. The item carries two
//   64-bit masks worked out exactly on the host: patches that add zero to
//   every lane of the tile (back-facing or fully eclipsed) are not visited,
//   and patches that no lane of the tile sees eclipsed skip the eclipse
//   test. Per element and lane: FP64 eclipse-interval test (bit-exact with
//   the reference's Point::visible, types.rs:50-58), FP32 mu, limb
//   darkening, beaming and lensing. Summation is cascaded: four FP32
//   accumulators over a patch (16 terms each), FP64 across patches; the
//   lane's item sum is written as an FP64 partial.
// assemble_kernel — one workgroup per set, all FP64: sums the partials per
//   sub-phase and component, derives the Ginterp scale factors from the
//   calibration lanes, applies the scale factors, smears exposures, adds the
//   polynomial factor and third light, autoscales and forms chi-squared.
// visibility_kernel — debug: eclipse visibility of every element of one
//   segment at a list of phases, for the bit-exact mask check.
//
// Reference: crates/lcurve/src/flux.rs:7-279, orchestration.rs:85-131,
// 181-310, 361-401; crates/lcurve-cuda/kernels/flux_kernel.cu:384-485 is
// the kernel this replaces.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace lch {

namespace {


// Limb-darkening forms: LD_POLY / LD_CLARET are LDC::imu (types.rs:101-119);
// LD_LIN is the Poly law with only the linear coefficient set; LD_LINW the
// same law on a segment whose records are rescaled by sqrt(w) (SEG_LINW).
enum : int { LD_POLY = 0, LD_CLARET = 1, LD_LIN = 2, LD_LINW = 3 };

// ------------------------------------------------------------ helpers
struct LaneCtx {
  double phif;          // sub-phase minus its floor: the visibility phase
  float ex, ey, ez;     // earth vector
  float kx, ky, k0;     // beaming: vx = kx*py, vy = ky*px + k0
};

struct StarPar {
  float a0, a1, a2, a3;
  float c0;      // LD_LIN(W): 1 - a0; LD_CLARET: 1 - (a0 + a1 + a2 + a3)
  float mucrit, beam, rlens;
  float w2;      // LD_LINW: 4^k, the records' weight scale
};

// Point::visible for NECL intervals (types.rs:50-58), in FP64 so the mask is
// identical to the reference's. egm1 is egress - 1 rounded on the host, the
// same value the reference computes inline.
template <int NECL>
__device__ __forceinline__ bool eclipse_free(const double* __restrict__ e, double phif) {
  bool ecl = false;
#pragma unroll
  for (int k = 0; k < NECL; ++k) {
    const double ing = e[3 * k], eg = e[3 * k + 1], egm1 = e[3 * k + 2];
    ecl = ecl | ((phif >= ing) & (phif <= eg)) | (phif <= egm1);
  }
  return !ecl;
}

__device__ __forceinline__ float clamp01(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f); }

// sqrt(m) for the Claret law, m in [0, 1]: the hardware square root
// (v_sqrt_f32, within 1 ulp). The correctly rounded sqrtf adds range
// scaling and a Newton correction, about 13 instructions per element-lane
// pair; the FP32 term does not need it (its error budget is 1e-6 of the
// flux).
__device__ __forceinline__ float claret_sqrt(float m) { return __builtin_amdgcn_sqrtf(m); }

// LDC::imu (types.rs:101-119) in FP32, including its mu <= 0 test.
template <int LD>
__device__ __forceinline__ float imu(float mu, const StarPar& P) {
  float m = fminf(mu, 1.0f);
  float im;
  if (LD == LD_CLARET) {
    float s = claret_sqrt(m);  // NaN for m < 0, dropped by the mu > 0 select below
    im = P.c0 + s * (P.a0 + s * (P.a1 + s * (P.a2 + s * P.a3)));
  } else {
    float o = 1.0f - m;
    im = 1.0f - o * (P.a0 + o * (P.a1 + o * (P.a2 + o * P.a3)));
  }
  return mu > 0.0f ? im : 0.0f;
}

// LDC::imu for m = clamp(mu, 0, 1), for use where m == 0 already zeroes the
// term (the element's mu factor is m too).
template <int LD>
__device__ __forceinline__ float imu_clamped(float m, const StarPar& P) {
  if (LD == LD_LIN) return fmaf(P.a0, m, P.c0);
  if (LD == LD_CLARET) {
    float s = claret_sqrt(m);
    return P.c0 + s * (P.a0 + s * (P.a1 + s * (P.a2 + s * P.a3)));
  }
  float o = 1.0f - m;
  return 1.0f - o * (P.a0 + o * (P.a1 + o * (P.a2 + o * P.a3)));
}

// lensing_mag (flux.rs:390-406) in FP32.
__device__ __forceinline__ float lensing_mag(float px, float py, float pz, const LaneCtx& L,
                                             float rlens) {
  float d = -(px * L.ex + py * L.ey + pz * L.ez);
  if (!(d > 0.0f)) return 1.0f;
  float qx = px + d * L.ex, qy = py + d * L.ey, qz = pz + d * L.ez;
  float p = sqrtf(qx * qx + qy * qy + qz * qz);
  float ph = p * 0.5f;
  float phsq = ph * ph;
  float rd = rlens * d;
  float pd = phsq > 25.0f * rd ? p + rd / p : ph + sqrtf(phsq + rd);
  return pd * pd / (pd - ph) / ph * 0.25f;
}

__device__ __forceinline__ float mu_of(const float4& d, const LaneCtx& L) {
  return fmaf(L.ez, d.z, fmaf(L.ey, d.y, L.ex * d.x));
}

// ---------------------------------------------------- one patch of 64
// Star elements without beaming or lensing (flux.rs:36-80 / 115-200):
// mu * flux * imu(mu) where mu > mucrit and the element is visible. With
// m = clamp(mu, 0, 1) the term m * flux * imu(m) equals the reference's for
// 0 < mu <= 1 and is 0 for mu <= 0, so the mucrit test is only needed for
// mucrit > 0 (MCRIT) and the select only with it or the eclipse test.
// Four consecutive 16-byte records (segments without eclipse intervals):
// one 64-byte scalar load instead of four.
struct Rec4 {
  float4 r[4];
};

// Record i + j of a patch: from global memory (wave-uniform address, scalar
// loads) or, with LDSR, from a copy of the patch staged in LDS (a broadcast
// read; segments without eclipse intervals only).
template <int NECL, bool LDSR = false>
__device__ __forceinline__ float4 rec_dirn(const char* __restrict__ p, uint32_t i, uint32_t j,
                                           const float4* lp = nullptr) {
  if (LDSR) return lp[i + j];
  if (NECL == 0) return reinterpret_cast<const Rec4*>(p + i * 16u)->r[j];
  return *reinterpret_cast<const float4*>(p + (i + j) * rec_stride(NECL, false));
}

// Eclipse intervals of record k: from the record in global memory, or with
// LDSR from the patch's interval block staged in LDS (3 NECL doubles per
// record).
template <int NECL, bool LDSR>
__device__ __forceinline__ const double* rec_ecl(const char* __restrict__ r, uint32_t k, const double* le) {
  if (LDSR) return le + k * 3u * NECL;
  return reinterpret_cast<const double*>(r + 16);
}

template <int NECL, bool FREE, int LD, bool MCRIT, bool LDSR = false>
__device__ __forceinline__ float star_patch(const char* __restrict__ p, const LaneCtx& L,
                                            const StarPar& P, const float4* lp = nullptr,
                                            const double* le = nullptr, uint32_t n = PATCH) {
  constexpr uint32_t STRIDE = rec_stride(NECL, false);
  constexpr bool ECL = NECL > 0 && !FREE;
  if constexpr (LD == LD_LINW) {
    // Records (s d, s): v = clamp(s d.e) = s m, and w m (c0 + a0 m) =
    // 4^k (c0 v s + a0 v^2). No cut-off on these segments.
    static_assert(!MCRIT, "SEG_LINW segments have no mu cut-off");
    float s1[2] = {0.0f, 0.0f}, s2[2] = {0.0f, 0.0f};
#pragma unroll(LDSR ? 1 : 2)
    for (uint32_t i = 0; i < n; i += 4) {
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) {
        const char* r = p + (i + j) * STRIDE;
        const float4 d = rec_dirn<NECL, LDSR>(p, i, j, lp);
        float v = clamp01(mu_of(d, L));
        if (ECL) v = eclipse_free<NECL>(rec_ecl<NECL, LDSR>(r, i + j, le), L.phif) ? v : 0.0f;
        s1[j & 1] = fmaf(v, d.w, s1[j & 1]);
        s2[j & 1] = fmaf(v, v, s2[j & 1]);
      }
    }
    return P.w2 * fmaf(P.c0, s1[0] + s1[1], P.a0 * (s2[0] + s2[1]));
  }
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll(LDSR ? 1 : 2)
  for (uint32_t i = 0; i < n; i += 4) {
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
      const char* r = p + (i + j) * STRIDE;
      const float4 d = rec_dirn<NECL, LDSR>(p, i, j, lp);
      const float mu = mu_of(d, L);
      const float m = clamp01(mu);
      float wm = m * d.w;
      if (ECL || MCRIT) {
        bool ok = true;
        if (ECL) ok = eclipse_free<NECL>(rec_ecl<NECL, LDSR>(r, i + j, le), L.phif);
        if (MCRIT) ok = ok & (mu > P.mucrit);
        wm = ok ? wm : 0.0f;
      }
      acc[j] = fmaf(wm, imu_clamped<LD>(m, P), acc[j]);
    }
  }
  return (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

// Star elements with beaming and / or lensing (FX bit 0 beaming, bit 1
// lensing): the reference's expression as written, imu with its own test.
template <int NECL, bool FREE, int LD, int FX>
__device__ __forceinline__ float star_patch_fx(const char* __restrict__ p, const LaneCtx& L,
                                               const StarPar& P) {
  constexpr bool BEAM = (FX & 1) != 0, LENS = (FX & 2) != 0;
  constexpr uint32_t STRIDE = rec_stride(NECL, true);
  constexpr bool ECL = NECL > 0 && !FREE;
  float acc[2] = {0.0f, 0.0f};
  for (uint32_t i = 0; i < PATCH; i += 2) {
#pragma unroll
    for (uint32_t j = 0; j < 2; ++j) {
      const char* r = p + (i + j) * STRIDE;
      const float4 d = *reinterpret_cast<const float4*>(r);
      const float4 q = *reinterpret_cast<const float4*>(r + 16);
      const float mu = mu_of(d, L);
      const float mag = LENS ? lensing_mag(q.x, q.y, q.z, L, P.rlens) : 1.0f;
      float term;
      if (BEAM) {
        const float vx = L.kx * q.y;
        const float vy = fmaf(L.ky, q.x, L.k0);
        const float vr = -(L.ex * vx + L.ey * vy);
        const float vn = d.x * vx + d.y * vy;
        const float mud = mu - mu * vr - vn;
        term = mu * mag * d.w * (1.0f - P.beam * vr) * imu<LD>(mud, P);
      } else {
        term = mu * mag * d.w * imu<LD>(mu, P);
      }
      bool ok = mu > P.mucrit;
      if (ECL) ok = ok & eclipse_free<NECL>(reinterpret_cast<const double*>(r + 32), L.phif);
      acc[j] += ok ? term : 0.0f;
    }
  }
  return acc[0] + acc[1];
}

// Disc and disc-edge elements (flux.rs:83-100): mu > 0 and visible.
template <int NECL, bool FREE, bool LDSR = false>
__device__ __forceinline__ float disc_patch(const char* __restrict__ p, const LaneCtx& L, float lin,
                                            float quad, const float4* lp = nullptr, const double* le = nullptr,
                                            uint32_t n = PATCH) {
  constexpr uint32_t STRIDE = rec_stride(NECL, false);
  constexpr bool ECL = NECL > 0 && !FREE;
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 2
  for (uint32_t i = 0; i < n; i += 4) {
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
      const char* r = p + (i + j) * STRIDE;
      const float4 d = rec_dirn<NECL, LDSR>(p, i, j, lp);
      const float m = clamp01(mu_of(d, L));
      float wm = m * d.w;
      if (ECL) wm = eclipse_free<NECL>(rec_ecl<NECL, LDSR>(r, i + j, le), L.phif) ? wm : 0.0f;
      const float o = 1.0f - m;
      acc[j] = fmaf(wm, 1.0f - o * (lin + quad * o), acc[j]);
    }
  }
  return (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

// Bright-spot elements (flux.rs:102-106): mu * flux where mu > 0 and visible.
template <int NECL, bool FREE, bool LDSR = false>
__device__ __forceinline__ float spot_patch(const char* __restrict__ p, const LaneCtx& L,
                                            const float4* lp = nullptr, const double* le = nullptr,
                                            uint32_t n = PATCH) {
  constexpr uint32_t STRIDE = rec_stride(NECL, false);
  constexpr bool ECL = NECL > 0 && !FREE;
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 2
  for (uint32_t i = 0; i < n; i += 4) {
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
      const char* r = p + (i + j) * STRIDE;
      const float4 d = rec_dirn<NECL, LDSR>(p, i, j, lp);
      float w = d.w;
      if (ECL) w = eclipse_free<NECL>(rec_ecl<NECL, LDSR>(r, i + j, le), L.phif) ? w : 0.0f;
      acc[j] = fmaf(clamp01(mu_of(d, L)), w, acc[j]);
    }
  }
  return (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

// ------------------------------------- linear-law star patches on MFMA
// The cosines of a rescaled (LD_LINW) patch against a tile: a 64 x 64
// product of the patch's directions (64 x 3) and the lanes' earth vectors
// (3 x 64), as 16 v_mfma_f32_16x16x4_f32 (K = 4, the fourth term zero).
// Operand layouts of one 16 x 16 block: lane l holds A[l % 16][l / 16],
// B[l / 16][l % 16], and D[4 (l / 16) + i][l % 16] in result element i.
// Lane l then owns column (time lane) 16 cb + l % 16 of column block cb
// for the 4 rows 4 (l / 16) + i of each row block, and sums v s and v^2
// over them (v = clamp(mu), s the record's weight scale); the four lanes
// of a column meet in FP64 once per item.
typedef float f32x4 __attribute__((ext_vector_type(4)));

// The B operands (earth vectors of the tile's lanes) of the four column
// blocks: every lane puts its earth vector in `eb` (64 float4 of LDS), then
// reads its four operands.
__device__ __forceinline__ void linw_mfma_b(const LaneCtx& L, float4* eb, float b[4]) {
  const uint32_t lane = __lane_id(), r = lane % 16u, k = lane / 16u;
  eb[lane] = make_float4(L.ex, L.ey, L.ez, 0.0f);
  __builtin_amdgcn_wave_barrier();  // every lane's vector is in LDS
  const float* ef = reinterpret_cast<const float*>(eb);
#pragma unroll
  for (uint32_t cb = 0; cb < 4; ++cb) b[cb] = ef[(cb * 16u + r) * 4u + k];  // k = 3: the zero row
  __builtin_amdgcn_wave_barrier();  // reads done before the LDS is reused
}

// One staged patch: adds w2 (c0 S1 + a0 S2) over this lane's rows of each
// column block to acc[cb] (S1 = sum v s, S2 = sum v^2 in FP32 per patch).
__device__ __forceinline__ void linw_mfma_patch(const float4* lp, const float b[4], const StarPar& P,
                                                double acc[4]) {
  const uint32_t lane = __lane_id(), r = lane % 16u, k = lane / 16u;
  const float* lpf = reinterpret_cast<const float*>(lp);
  float s1[4] = {0.0f, 0.0f, 0.0f, 0.0f}, s2[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (uint32_t rb = 0; rb < 4; ++rb) {
    const float a = lpf[(rb * 16u + r) * 4u + k];  // k = 3: the weight scale, times B's zero row
    float sw[4];
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) sw[i] = lp[rb * 16u + 4u * k + i].w;
    // The four blocks' products first, so they overlap in the matrix unit.
    f32x4 mu[4];
#pragma unroll
    for (uint32_t cb = 0; cb < 4; ++cb)
      mu[cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b[cb], f32x4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
#pragma unroll
    for (uint32_t cb = 0; cb < 4; ++cb) {
#pragma unroll
      for (uint32_t i = 0; i < 4; ++i) {
        const float v = clamp01(mu[cb][i]);
        s1[cb] = fmaf(v, sw[i], s1[cb]);
        s2[cb] = fmaf(v, v, s2[cb]);
      }
    }
  }
#pragma unroll
  for (uint32_t cb = 0; cb < 4; ++cb) acc[cb] += static_cast<double>(P.w2 * fmaf(P.c0, s1[cb], P.a0 * s2[cb]));
}

// The item's total for this lane's own time lane: lane l owns column
// 16 cb + l % 16 of block cb for the rows 4 (l / 16) + i, so the four lanes
// l % 16 + 16 g (g = 0..3) of a column meet through LDS (`red`, 256
// doubles), in the order g = 0, 1, 2, 3, and lane l takes block l / 16.
__device__ __forceinline__ double linw_mfma_total(const double acc[4], double* red) {
  const uint32_t lane = __lane_id(), c = lane % 16u, cb = lane / 16u;
#pragma unroll
  for (uint32_t j = 0; j < 4; ++j) red[lane * 4u + j] = acc[j];
  __builtin_amdgcn_wave_barrier();  // every lane's partials are in LDS
  double s = red[c * 4u + cb] + red[(16u + c) * 4u + cb];
  s += red[(32u + c) * 4u + cb];
  s += red[(48u + c) * 4u + cb];
  __builtin_amdgcn_wave_barrier();  // reads done before the LDS is reused
  return s;
}

// ----------------------------------------------------- one work item
// Visits the patches in `todo` (wave-uniform), choosing the patch body by
// the `fr` bit; each patch's FP32 sum is added in FP64.
enum : int { PK_STAR = 0, PK_STARFX = 1, PK_DISC = 2, PK_SPOT = 3 };

// Patches of segments with up to LDS_ECL_MAX eclipse intervals per element
// are staged in LDS too: directions (PATCH float4) then intervals
// (PATCH * 3 * LDS_ECL_MAX doubles), 4 KB per wavefront.
constexpr int PBUF_F4 = PATCH + PATCH * 3 * LDS_ECL_MAX / 2;

template <int PK, int NECL, int LD, int X, bool ECLS = false>
__device__ __forceinline__ double item_sum(const char* __restrict__ p, uint64_t todo, uint64_t fr,
                                           const LaneCtx& L, const StarPar& P, float lin,
                                           float quad, float4* lp) {
  constexpr uint32_t PB = PATCH * rec_stride(NECL, PK == PK_STARFX);
  double acc = 0.0;
  if (PK != PK_STARFX && lp && (NECL == 0 || (todo & ~fr) == 0)) {
    // Records staged in LDS: lane j fetches the direction and weight of
    // record j of the next patch while the current patch is summed from LDS.
    // Also for segments with eclipse intervals when every patch of the item
    // needs no eclipse test (the intervals are then not read).
    constexpr uint32_t STRIDE = rec_stride(NECL, false);
    const uint32_t lane = __lane_id();
    auto rec = [&](uint64_t t) {
      return *reinterpret_cast<const float4*>(p + (static_cast<size_t>(__builtin_ctzll(t)) * PATCH + lane) * STRIDE);
    };
    float4 nxt = rec(todo);
    if constexpr (PK == PK_STAR && LD == LD_LINW) {
      // Linear-law star patches: the cosines on MFMA (linw_mfma_patch),
      // for items of two patches or more (the set-up and the final sum
      // cost about what one patch saves). lp holds 2 PATCH float4 here:
      // the earth vectors go to its second half before the first patch,
      // and the final sum uses all of it after the last.
      if (__builtin_popcountll(todo) >= 2) {
        float b[4];
        linw_mfma_b(L, lp + PATCH, b);
        double part[4] = {0.0, 0.0, 0.0, 0.0};
        while (todo) {
          todo &= todo - 1ull;
          lp[lane] = nxt;
          __builtin_amdgcn_wave_barrier();  // the patch is in LDS for every lane
          if (todo) nxt = rec(todo);
          linw_mfma_patch(lp, b, P, part);
          __builtin_amdgcn_wave_barrier();  // reads done before the next write
        }
        return linw_mfma_total(part, reinterpret_cast<double*>(lp));
      }
    }
    while (todo) {
      todo &= todo - 1ull;
      lp[lane] = nxt;
      __builtin_amdgcn_wave_barrier();  // the patch is in LDS for every lane
      if (todo) nxt = rec(todo);
      float s;
      if (PK == PK_STAR) s = star_patch<0, true, LD, X != 0, true>(nullptr, L, P, lp);
      else if (PK == PK_DISC) s = disc_patch<0, true, true>(nullptr, L, lin, quad, lp);
      else s = spot_patch<0, true, true>(nullptr, L, lp);
      __builtin_amdgcn_wave_barrier();  // reads done before the next write
      acc += static_cast<double>(s);
    }
    return acc;
  }
  if (ECLS && NECL > 0 && NECL <= LDS_ECL_MAX && PK != PK_STARFX && lp) {
    // The same with eclipse intervals: lane j fetches record j's direction
    // and its 3 NECL interval bounds (for eclipse-tested and free patches
    // alike, so the staged values never cross a branch); the patch's
    // directions go to lp[0..63], its intervals to the block after them.
    constexpr uint32_t STRIDE = rec_stride(NECL, false);
    constexpr int NE = 3 * (NECL > 0 ? NECL : 1);
    const uint32_t lane = __lane_id();
    double* le = reinterpret_cast<double*>(lp + PATCH);
    uint32_t q = static_cast<uint32_t>(__builtin_ctzll(todo));
    const char* rq = p + static_cast<size_t>(q) * PB + lane * STRIDE;
    float4 nd = *reinterpret_cast<const float4*>(rq);
    double ne[NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) ne[k] = reinterpret_cast<const double*>(rq + 16)[k];
    while (todo) {
      const bool free = (fr >> q) & 1ull;
      todo &= todo - 1ull;
      lp[lane] = nd;
#pragma unroll
      for (int k = 0; k < NE; ++k) le[lane * NE + k] = ne[k];
      __builtin_amdgcn_wave_barrier();  // the patch is in LDS for every lane
      if (todo) {
        q = static_cast<uint32_t>(__builtin_ctzll(todo));
        rq = p + static_cast<size_t>(q) * PB + lane * STRIDE;
        nd = *reinterpret_cast<const float4*>(rq);
#pragma unroll
        for (int k = 0; k < NE; ++k) ne[k] = reinterpret_cast<const double*>(rq + 16)[k];
      }
      float s;
      if (PK == PK_STAR) {
        s = free ? star_patch<NECL, true, LD, X != 0, true>(nullptr, L, P, lp, le)
                 : star_patch<NECL, false, LD, X != 0, true>(nullptr, L, P, lp, le);
      } else if (PK == PK_DISC) {
        s = free ? disc_patch<NECL, true, true>(nullptr, L, lin, quad, lp, le)
                 : disc_patch<NECL, false, true>(nullptr, L, lin, quad, lp, le);
      } else {
        s = free ? spot_patch<NECL, true, true>(nullptr, L, lp, le) : spot_patch<NECL, false, true>(nullptr, L, lp, le);
      }
      __builtin_amdgcn_wave_barrier();  // reads done before the next write
      acc += static_cast<double>(s);
    }
    return acc;
  }
  while (todo) {
    const uint32_t q = static_cast<uint32_t>(__builtin_ctzll(todo));
    todo &= todo - 1ull;
    const char* pp = p + static_cast<size_t>(q) * PB;
    const bool free = NECL > 0 && ((fr >> q) & 1ull);
    float s;
    if (PK == PK_STAR) {
      s = free ? star_patch<NECL, true, LD, X != 0>(pp, L, P) : star_patch<NECL, false, LD, X != 0>(pp, L, P);
    } else if (PK == PK_STARFX) {
      s = free ? star_patch_fx<NECL, true, LD, X>(pp, L, P) : star_patch_fx<NECL, false, LD, X>(pp, L, P);
    } else if (PK == PK_DISC) {
      s = free ? disc_patch<NECL, true>(pp, L, lin, quad) : disc_patch<NECL, false>(pp, L, lin, quad);
    } else {
      s = free ? spot_patch<NECL, true>(pp, L) : spot_patch<NECL, false>(pp, L);
    }
    acc += static_cast<double>(s);
  }
  return acc;
}

// FX: whether the beaming / lensing star paths are compiled in. The kernel
// without them needs fewer registers (more waves per SIMD); the host sends
// it no item of a beamed or lensed star segment.
template <int NECL, bool FX, bool ECLS = false>
__device__ __forceinline__ double item_dispatch(uint32_t kind, int ld, uint32_t fx, bool mcrit,
                                                const char* p, uint64_t todo, uint64_t fr,
                                                const LaneCtx& L, const StarPar& P, float lin,
                                                float quad, float4* lp = nullptr) {
  if (kind <= KIND_STAR2) {
    if (!FX && fx != 0) return __builtin_nan("");  // never sent here: fail loudly
    if (!FX || fx == 0) {
      if (ld == LD_LINW) return item_sum<PK_STAR, NECL, LD_LINW, 0, ECLS>(p, todo, fr, L, P, lin, quad, lp);
      if (ld == LD_LIN)
        return mcrit ? item_sum<PK_STAR, NECL, LD_LIN, 1, ECLS>(p, todo, fr, L, P, lin, quad, lp)
                     : item_sum<PK_STAR, NECL, LD_LIN, 0, ECLS>(p, todo, fr, L, P, lin, quad, lp);
      if (ld == LD_CLARET)
        return mcrit ? item_sum<PK_STAR, NECL, LD_CLARET, 1, ECLS>(p, todo, fr, L, P, lin, quad, lp)
                     : item_sum<PK_STAR, NECL, LD_CLARET, 0, ECLS>(p, todo, fr, L, P, lin, quad, lp);
      return mcrit ? item_sum<PK_STAR, NECL, LD_POLY, 1, ECLS>(p, todo, fr, L, P, lin, quad, lp)
                   : item_sum<PK_STAR, NECL, LD_POLY, 0, ECLS>(p, todo, fr, L, P, lin, quad, lp);
    }
    const bool cl = ld == LD_CLARET;
    switch (fx) {
      case 1:
        return cl ? item_sum<PK_STARFX, NECL, LD_CLARET, 1, ECLS>(p, todo, fr, L, P, lin, quad, lp)
                  : item_sum<PK_STARFX, NECL, LD_POLY, 1, ECLS>(p, todo, fr, L, P, lin, quad, lp);
      case 2:
        return cl ? item_sum<PK_STARFX, NECL, LD_CLARET, 2, ECLS>(p, todo, fr, L, P, lin, quad, lp)
                  : item_sum<PK_STARFX, NECL, LD_POLY, 2, ECLS>(p, todo, fr, L, P, lin, quad, lp);
      default:
        return cl ? item_sum<PK_STARFX, NECL, LD_CLARET, 3, ECLS>(p, todo, fr, L, P, lin, quad, lp)
                  : item_sum<PK_STARFX, NECL, LD_POLY, 3, ECLS>(p, todo, fr, L, P, lin, quad, lp);
    }
  }
  if (kind == KIND_SPOT) return item_sum<PK_SPOT, NECL, 0, 0, ECLS>(p, todo, fr, L, P, lin, quad, lp);
  return item_sum<PK_DISC, NECL, 0, 0, ECLS>(p, todo, fr, L, P, lin, quad, lp);
}

// One patch staged in LDS (segments without eclipse intervals, no beaming
// or lensing, no eclipse test), by segment kind and limb-darkening law.
__device__ __forceinline__ float patch_sum_lds(uint32_t kind, int ld, bool mcrit, const float4* lp,
                                               const LaneCtx& L, const StarPar& P, float lin,
                                               float quad, uint32_t n = PATCH) {
  if (kind <= KIND_STAR2) {
    if (ld == LD_LINW) return star_patch<0, true, LD_LINW, false, true>(nullptr, L, P, lp, nullptr, n);
    if (ld == LD_LIN)
      return mcrit ? star_patch<0, true, LD_LIN, true, true>(nullptr, L, P, lp, nullptr, n)
                   : star_patch<0, true, LD_LIN, false, true>(nullptr, L, P, lp, nullptr, n);
    if (ld == LD_CLARET)
      return mcrit ? star_patch<0, true, LD_CLARET, true, true>(nullptr, L, P, lp, nullptr, n)
                   : star_patch<0, true, LD_CLARET, false, true>(nullptr, L, P, lp, nullptr, n);
    return mcrit ? star_patch<0, true, LD_POLY, true, true>(nullptr, L, P, lp, nullptr, n)
                 : star_patch<0, true, LD_POLY, false, true>(nullptr, L, P, lp, nullptr, n);
  }
  if (kind == KIND_SPOT) return spot_patch<0, true, true>(nullptr, L, lp, nullptr, n);
  return disc_patch<0, true, true>(nullptr, L, lin, quad, lp, nullptr, n);
}

// An item's PT_MOM patches from their summed moments (one table entry per
// item): w mu (c1 + c2 mu + c3 mu^2) summed over the patches, in FP64.
__device__ __forceinline__ double moment_item(const double* __restrict__ M, const LaneCtx& L, double c1,
                                              double c2, double c3) {
  const double x = L.ex, y = L.ey, z = L.ez;
  const double xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
  const double s1 = M[0] * x + M[1] * y + M[2] * z;
  const double s2 = M[3] * xx + M[4] * yy + M[5] * zz + M[6] * xy + M[7] * xz + M[8] * yz;
  double v = c1 * s1 + c2 * s2;
  if (c3 != 0.0) {
    const double s3 = M[9] * (xx * x) + M[10] * (yy * y) + M[11] * (zz * z) + M[12] * (xx * y) +
                      M[13] * (xx * z) + M[14] * (yy * x) + M[15] * (yy * z) + M[16] * (zz * x) +
                      M[17] * (zz * y) + M[18] * (xy * z);
    v += c3 * s3;
  }
  return v;
}

// Patches split per lane (pmom), summed by lane mask: group g's moments
// (entry M + g * MOM_STRIDE) add only on the lanes of masks[g].
__device__ __forceinline__ double moment_groups(const double* __restrict__ M, uint32_t ngrp,
                                                const uint64_t* __restrict__ masks, uint32_t lane,
                                                const LaneCtx& L, double c1, double c2, double c3) {
  const double x = L.ex, y = L.ey, z = L.ez;
  const double xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
  const double t0 = xx * x, t1 = yy * y, t2 = zz * z, t3 = xx * y, t4 = xx * z, t5 = yy * x,
               t6 = yy * z, t7 = zz * x, t8 = zz * y, t9 = xy * z;
  double acc = 0.0;
  for (uint32_t g = 0; g < ngrp; ++g, M += MOM_STRIDE) {
    const uint64_t mk = masks[g];
    const double s1 = M[0] * x + M[1] * y + M[2] * z;
    const double s2 = M[3] * xx + M[4] * yy + M[5] * zz + M[6] * xy + M[7] * xz + M[8] * yz;
    double v = c1 * s1 + c2 * s2;
    if (c3 != 0.0) {
      const double s3 = M[9] * t0 + M[10] * t1 + M[11] * t2 + M[12] * t3 + M[13] * t4 +
                        M[14] * t5 + M[15] * t6 + M[16] * t7 + M[17] * t8 + M[18] * t9;
      v += c3 * s3;
    }
    acc += ((mk >> lane) & 1ull) ? v : 0.0;
  }
  return acc;
}

// Star limb darkening and cut-off of a segment, for the paths without
// beaming or lensing.
struct SegLaw {
  StarPar P;
  int ld;
  bool mcrit;
};
__device__ __forceinline__ SegLaw seg_law(const SegDev& sg, const SetDev& sp) {
  SegLaw w;
  const bool star1 = sg.kind == KIND_STAR1;
  w.P.a0 = star1 ? sp.ldc1[0] : sp.ldc2[0];
  w.P.a1 = star1 ? sp.ldc1[1] : sp.ldc2[1];
  w.P.a2 = star1 ? sp.ldc1[2] : sp.ldc2[2];
  w.P.a3 = star1 ? sp.ldc1[3] : sp.ldc2[3];
  w.P.mucrit = star1 ? sp.mucrit1 : sp.mucrit2;
  w.P.beam = star1 ? sp.beam1 : sp.beam2;
  w.P.rlens = sp.rlens1;
  w.ld = (star1 ? sp.ltype1 : sp.ltype2) == 1u ? LD_CLARET : LD_POLY;
  if (w.ld == LD_POLY && w.P.a1 == 0.0f && w.P.a2 == 0.0f && w.P.a3 == 0.0f) w.ld = LD_LIN;
  w.P.c0 = w.ld == LD_LIN ? 1.0f - w.P.a0 : 1.0f - (w.P.a0 + w.P.a1 + w.P.a2 + w.P.a3);
  w.mcrit = w.P.mucrit > 0.0f;
  w.P.w2 = 1.0f;
  // The host flags only linear-law segments without cut-off.
  if (sg.flags & SEG_LINW) {
    w.ld = LD_LINW;
    w.P.w2 = ldexpf(1.0f, 2 * seg_linw_exp(sg.flags));
  }
  return w;
}

// Workgroup b of a launch lands on XCD b % 8. Give every XCD a contiguous
// run of items so that the tiles of one chunk share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_map(uint32_t b, uint32_t n) {
  uint32_t q = n / 8u, r = n % 8u;
  uint32_t x = b % 8u, i = b / 8u;
  return x * q + (x < r ? x : r) + i;
}

template <bool FX, bool ECLS = false>
__device__ __forceinline__ void process_item(uint32_t id, const ItemDev* __restrict__ items,
                             const SegDev* __restrict__ segs, const SetDev* __restrict__ sets,
                             const char* __restrict__ recs, const double* __restrict__ moms,
                             const double* __restrict__ imoms,
                             const double* __restrict__ lane_phi, const float4* __restrict__ lane_e,
                             const uint32_t* __restrict__ lane_meta, const uint64_t* __restrict__ lane_masks,
                             double* __restrict__ partials, float4* lp) {
  const ItemDev it = items[id];
  const uint32_t lane = threadIdx.x;
  const uint32_t gl = it.lane0 + lane;
  // Lane data first: it depends on the item only.
  const double ph = lane_phi[gl];
  const float4 e = lane_e[gl];
  const uint32_t meta = lane_meta[gl];
  const SegDev sg = segs[it.seg];
  const SetDev& sp = sets[sg.set];
  const uint32_t cls = meta_class(meta);
  const bool active = (meta & META_VALID) && (sg.use_mask & cp_bit(cls, meta_ptype(meta)));

  LaneCtx L;
  L.phif = ph - floor(ph);
  // comp_light lanes; the time lanes of iscale sets follow comp_star1/2.
  const bool lc_lane = cls == CLS_NORMAL && !(sp.flags & SET_ISCALE);
  // Earth vector (set_earth, flux.rs:7-22 conventions) made on the host in
  // FP64 and rounded to FP32: comp_light lanes use the TWOPI*iangle/360
  // inclination, the single-star lanes iangle.to_radians().
  L.ex = e.x;
  L.ey = e.y;
  L.ez = e.z;

  const bool star1 = sg.kind == KIND_STAR1;
  const SegLaw law = seg_law(sg, sp);
  const StarPar& P = law.P;
  const int ld = law.ld;
  uint32_t fx = 0;
  if (sg.kind <= KIND_STAR2) {
    const bool beam = (sp.flags & (star1 ? SET_BEAM1 : SET_BEAM2)) != 0;
    const bool lens = !star1 && (sp.flags & SET_LENS);
    fx = (beam ? 1u : 0u) | (lens ? 2u : 0u);
    // Velocity of an element: comp_light lanes include the spin and the
    // star's centre of mass (flux.rs:41-42, 68-69); comp_star1/2 lanes use
    // the spin-free form of flux.rs:138-139 / 183-184.
    const float spin = star1 ? sp.spin1 : sp.spin2;
    if (lc_lane) {
      L.kx = -sp.vfac * spin;
      L.ky = sp.vfac * spin;
      L.k0 = star1 ? sp.vfac * (-sp.xcofm) : sp.vfac * (1.0f - sp.xcofm - spin);
    } else {
      L.kx = -sp.vfac;
      L.ky = sp.vfac;
      L.k0 = -sp.vfac * sp.xcofm;
    }
  } else {
    L.kx = L.ky = L.k0 = 0.0f;
  }
  const bool mcrit = law.mcrit;

  const uint32_t stride = rec_stride(sg.necl, (sg.flags & SEG_POSN) != 0);
  const char* p = recs + sg.rec_off + static_cast<uint64_t>(it.e0) * stride;
  const float lin = sp.lin_disc, quad = sp.quad_disc;
  double sum = 0.0;
  if (it.mom | it.pmom) {
    // w mu (1 - A o - B o^2) with o = 1 - mu, as c1 mu + c2 mu^2 + c3 mu^3:
    // stars (A, B) = (a0, a1) (the host sends only Poly laws with a2 = a3 = 0
    // and no beaming or lensing here), disc and edge (lin, quad), spot (0, 0).
    const bool star = sg.kind <= KIND_STAR2;
    const double A = star ? P.a0 : (sg.kind == KIND_SPOT ? 0.0 : lin);
    const double B = star ? P.a1 : (sg.kind == KIND_SPOT ? 0.0 : quad);
    const double* im = imoms + static_cast<size_t>(it.imom) * MOM_STRIDE;
    if (it.mom) {
      sum = moment_item(im, L, 1.0 - A - B, A + 2.0 * B, -B);
      im += MOM_STRIDE;
    }
    if (it.pmom) sum += moment_groups(im, it.ngrp, lane_masks + it.lm_off, lane, L, 1.0 - A - B, A + 2.0 * B, -B);
  }
  if (it.todo) switch (sg.necl) {
    case 0: sum += item_dispatch<0, FX, ECLS>(sg.kind, ld, fx, mcrit, p, it.todo, it.free, L, P, lin, quad, lp); break;
    case 1: sum += item_dispatch<1, FX, ECLS>(sg.kind, ld, fx, mcrit, p, it.todo, it.free, L, P, lin, quad, lp); break;
    case 2: sum += item_dispatch<2, FX, ECLS>(sg.kind, ld, fx, mcrit, p, it.todo, it.free, L, P, lin, quad, lp); break;
    case 3: sum += item_dispatch<3, FX, ECLS>(sg.kind, ld, fx, mcrit, p, it.todo, it.free, L, P, lin, quad, lp); break;
    default: sum += item_dispatch<4, FX, ECLS>(sg.kind, ld, fx, mcrit, p, it.todo, it.free, L, P, lin, quad, lp); break;
  }
  partials[it.prow * TILE + lane] = active ? sum : 0.0;
}

}  // namespace

// One workgroup = one wavefront = one work item. ECLS: patches of segments
// with 1..LDS_ECL_MAX eclipse intervals are staged in LDS (more registers;
// the host launches this form only for items that have such patches).
// At least 5 waves per SIMD: the ECLS form then takes 78 VGPRs without
// spilling (113 unbounded, 4 waves); measured, the time is the same.
template <bool FX, bool ECLS>
__global__ __launch_bounds__(TILE, 5) void flux_items_kernel(
    const ItemDev* __restrict__ items, uint32_t n_items, const SegDev* __restrict__ segs,
    const SetDev* __restrict__ sets, const char* __restrict__ recs,
    const double* __restrict__ moms, const double* __restrict__ imoms, const double* __restrict__ lane_phi,
    const float4* __restrict__ lane_e, const uint32_t* __restrict__ lane_meta,
    const uint64_t* __restrict__ lane_masks, double* __restrict__ partials) {
  // One patch of records (one wavefront per workgroup); the linear-law
  // MFMA path also uses a second PATCH float4 (ECLS: within PBUF_F4).
  __shared__ float4 pbuf[ECLS ? PBUF_F4 : 2 * PATCH];
  const uint32_t id = xcd_map(blockIdx.x, n_items);
  if (id >= n_items) return;
  process_item<FX, ECLS>(id, items, segs, sets, recs, moms, imoms, lane_phi, lane_e, lane_meta, lane_masks, partials,
                   pbuf);
}

// Limb lanes: one workgroup of LIMB_WAVES wavefronts per limb chunk. The
// chunk's limb lanes' earth vectors are staged in LDS (48 KB with the sums
// and the patch buffers: three workgroups per CU). Each wavefront takes
// the chunk's limb items w, w + LIMB_WAVES, ...: every lane sums one patch
// element by element (no eclipse test: the patches are PT_FREE) for its
// listed lane and adds the sum, as a 64-bit fixed-point integer, to that
// lane's LDS slot. Integer addition is exact and associative, so the total
// does not depend on the order of the adds. The totals are then added to
// the partial rows the work items wrote earlier in the stream.
__global__ __launch_bounds__(LIMB_WAVES * TILE) void limb_kernel(const DeviceBatch d) {
  __shared__ unsigned long long acc[LIMB_UNION_MAX];
  __shared__ float elx[LIMB_UNION_MAX], ely[LIMB_UNION_MAX], elz[LIMB_UNION_MAX];
  __shared__ float4 lrec[LIMB_WAVES][PATCH];
  const uint32_t ci = xcd_map(blockIdx.x, d.n_limb_chunks);
  if (ci >= d.n_limb_chunks) return;  // whole workgroup, before any barrier
  const LimbChunk c = d.limb_chunks[ci];
  for (uint32_t u = threadIdx.x; u < c.n_lanes; u += blockDim.x) {
    acc[u] = 0ull;
    const float4 e = d.lane_e[d.limb_lanes[c.lane_off + u]];
    elx[u] = e.x;
    ely[u] = e.y;
    elz[u] = e.z;
  }
  __syncthreads();
  // The wavefront index is uniform across the wave; saying so keeps item
  // and record addresses in SGPRs (records then arrive by scalar loads).
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x / TILE), lane = threadIdx.x % TILE;
  const SegDev sg = d.segs[c.seg];
  const SetDev& sp = d.sets[sg.set];
  const SegLaw law = seg_law(sg, sp);
  const float lin = sp.lin_disc, quad = sp.quad_disc;
  const uint32_t stride = rec_stride(sg.necl, false);
  const double scale = ldexp(1.0, c.shift);
  const uint32_t iend = c.item0 + c.n_items;
  uint32_t i = c.item0 + w;
  // Each item's lane slots are fetched one item ahead (the index is clamped
  // to the chunk's last item: no branch around the load).
  const uint32_t ilast = iend - 1u;
  uint32_t u_next = d.limb_idx[static_cast<size_t>(i < ilast ? i : ilast) * TILE + lane];
  // Lane j fetches the direction and weight of record j of the next item's
  // patch while the current patch, staged in LDS, is summed from broadcast
  // reads. Limb lanes come from no-eclipse (patch, tile) pairs only, so no
  // eclipse test is made here and a segment's intervals are never read.
  const char* rbase = d.recs + sg.rec_off;
  float4 rec_next;
  {
    const uint32_t e = d.limb_items[i < ilast ? i : ilast].e0;
    rec_next = *reinterpret_cast<const float4*>(rbase + static_cast<uint64_t>(e + lane) * stride);
  }
  for (; i < iend; i += LIMB_WAVES) {
    const uint32_t u = u_next;
    const uint32_t inext = i + LIMB_WAVES;
    u_next = d.limb_idx[static_cast<size_t>(inext < ilast ? inext : ilast) * TILE + lane];
    const LimbItem li = d.limb_items[i];
    // A split item (li.split = log2 of 2 or 4) gives each group of
    // 64 >> split lanes (the same limb lanes in every group) its own share
    // of the patch's elements; the groups' parts meet in the integer sums.
    const uint32_t nel = PATCH >> li.split;
    const uint32_t base = (lane / (TILE >> li.split)) * nel;
    const bool valid = u != 0xFFFFu;
    const uint32_t uu = valid ? u : 0u;
    LaneCtx L;
    L.phif = 0.0;
    L.ex = elx[uu];
    L.ey = ely[uu];
    L.ez = elz[uu];
    L.kx = L.ky = L.k0 = 0.0f;
    lrec[w][lane] = rec_next;
    __builtin_amdgcn_wave_barrier();  // the patch is in LDS for every lane
    {
      const uint32_t e = d.limb_items[inext < ilast ? inext : ilast].e0;
      rec_next = *reinterpret_cast<const float4*>(rbase + static_cast<uint64_t>(e + lane) * stride);
    }
    const double s = patch_sum_lds(sg.kind, law.ld, law.mcrit, &lrec[w][base], L, law.P, lin, quad, nel);
    __builtin_amdgcn_wave_barrier();  // reads done before the next write
    if (valid) atomicAdd(&acc[u], static_cast<unsigned long long>(__builtin_llrint(s * scale)));
  }
  __syncthreads();
  const double inv = ldexp(1.0, -c.shift);
  for (uint32_t u = threadIdx.x; u < c.n_lanes; u += blockDim.x) {
    const double v = static_cast<double>(static_cast<long long>(acc[u])) * inv;
    const uint32_t sl = d.limb_lanes[c.lane_off + u] - sp.lane_off;
    const uint64_t row = sp.prow_base + static_cast<uint64_t>(sl / TILE) * sp.n_chunks + c.chunk;
    d.partials[row * TILE + sl % TILE] += v;
  }
}

// ------------------------------------------------------------ assembly
// FP64 from here on without FMA contraction (the reference's operation
// order), helpers included: the file is built with
// -ffp-contract=fast-honor-pragmas, which keeps contraction in the FP32
// flux loops above and honours this pragma for every function below.
#pragma clang fp contract(off)
namespace {

// Ginterp::scale1 / scale2 (types.rs:161-184) in FP64.
__device__ __forceinline__ double gscale1(double phase, double p1, double s11, double s12) {
  double pnorm = phase - floor(phase);
  if (pnorm <= p1 || pnorm >= 1.0 - p1) return 1.0;
  return (s11 * (1.0 - p1 - pnorm) + s12 * (pnorm - p1)) / (1.0 - 2.0 * p1);
}
__device__ __forceinline__ double gscale2(double phase, double p2, double s21, double s22) {
  double pnorm = phase - floor(phase);
  if (pnorm >= p2 && pnorm <= 1.0 - p2) return 1.0;
  if (pnorm < 0.5) return (s22 * (p2 - pnorm) + s21 * (pnorm + p2)) / (2.0 * p2);
  return (s21 * (1.0 + p2 - pnorm) + s22 * (pnorm - 1.0 + p2)) / (2.0 * p2);
}

// Sum of a lane's partial rows per segment kind: S[KIND_STAR1 .. KIND_SPOT].
__device__ __forceinline__ void lane_sums(const SetDev& sp, uint32_t l, const DeviceBatch& d,
                                          double S[5]) {
  const uint32_t t = l / TILE, j = l % TILE;
  S[0] = S[1] = S[2] = S[3] = S[4] = 0.0;
  const uint64_t row0 = sp.prow_base + static_cast<uint64_t>(t) * sp.n_chunks;
  // Every row is read: a chunk whose segment the tile does not use has no
  // item there, and its row stays zero, as do the slots of lanes an item
  // does not use. Without a branch around them the loads overlap. Each row
  // goes to its segment kind's sum, in chunk order, through selects (no
  // indexed register array).
#pragma unroll 4
  for (uint32_t c = 0; c < sp.n_chunks; ++c) {
    const uint32_t info = d.chunk_info[sp.chunk_off + c];
    const double v = d.partials[(row0 + c) * TILE + j];
    const uint32_t kind = info >> 16;
    S[0] += kind == 0u ? v : 0.0;
    S[1] += kind == 1u ? v : 0.0;
    S[2] += kind == 2u ? v : 0.0;
    S[3] += kind == 3u ? v : 0.0;
    S[4] += kind >= 4u ? v : 0.0;
  }
}

constexpr int ABLOCK = 256;
constexpr int NWAVE = ABLOCK / 64;

__device__ double block_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double r = 0.0;
  for (int i = 0; i < NWAVE; ++i) r += red[i];
  return r;
}

// Block sums of n <= 6 values at once; red holds NWAVE * 6 doubles.
__device__ void block_sum_n(double* v, int n, double* red) {
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  for (int k = 0; k < n; ++k) {
    double x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    v[k] = x;
  }
  __syncthreads();
  if (l == 0)
    for (int k = 0; k < n; ++k) red[w * 6 + k] = v[k];
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    double r = 0.0;
    for (int i = 0; i < NWAVE; ++i) r += red[i * 6 + k];
    v[k] = r;
  }
}

// Minimum-norm solution of R x ~ c for an upper-triangular n x n R (n <= 5)
// through a one-sided Jacobi SVD of R, singular values <= eps treated as
// zero. R's singular values are those of the matrix it factors, so this is
// the reference's svd.solve(&b, 1e-12) (orchestration.rs:431-432).
__device__ void jacobi_solve(double U[5][5], const double* c, int n, double eps, double* x) {
  double V[5][5];
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) V[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < n; ++p) {
      for (int q = p + 1; q < n; ++q) {
        double alpha = 0.0, beta = 0.0, gamma = 0.0;
        for (int i = 0; i < n; ++i) {
          alpha += U[i][p] * U[i][p];
          beta += U[i][q] * U[i][q];
          gamma += U[i][p] * U[i][q];
        }
        if (gamma == 0.0 || fabs(gamma) <= 1e-15 * sqrt(alpha * beta)) continue;
        rotated = true;
        const double zeta = (beta - alpha) / (2.0 * gamma);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
        for (int i = 0; i < n; ++i) {
          const double up = U[i][p], uq = U[i][q];
          U[i][p] = cs * up - sn * uq;
          U[i][q] = sn * up + cs * uq;
          const double vp = V[i][p], vq = V[i][q];
          V[i][p] = cs * vp - sn * vq;
          V[i][q] = sn * vp + cs * vq;
        }
      }
    }
    if (!rotated) break;
  }
  for (int k = 0; k < n; ++k) x[k] = 0.0;
  for (int j = 0; j < n; ++j) {
    double s2 = 0.0, uc = 0.0;
    for (int i = 0; i < n; ++i) {
      s2 += U[i][j] * U[i][j];
      uc += U[i][j] * c[i];
    }
    if (!(sqrt(s2) > eps)) continue;
    const double f = uc / s2;
    for (int k = 0; k < n; ++k) x[k] += V[k][j] * f;
  }
}

// Per-component scale factors of one iscale set (orchestration.rs:405-436):
// least squares over the rows with weight > 0, each scaled by
// sqrt(weight) / ferr, by modified Gram-Schmidt on [A | b] (stable for least
// squares) and an SVD of the small triangular factor.
__device__ void iscale_fit(const SetDev& sp, const DeviceBatch& d, int ncol, double* red,
                           double* sf_sh) {
  const uint32_t tid = threadIdx.x;
  __shared__ double R[5][6];
  double cnt[1] = {0.0};
  for (uint32_t t = tid; t < sp.n_times; t += ABLOCK) {
    const uint32_t g = sp.time_off + t;
    const bool ok = d.t_w[g] > 0.0;
    cnt[0] += ok ? 1.0 : 0.0;
    const double wt = ok ? sqrt(d.t_w[g]) / d.t_ferr[g] : 0.0;
    for (int j = 0; j < 5; ++j) d.lsw[g * 6 + j] = j < ncol ? d.comp[g * 5 + j] * wt : 0.0;
    d.lsw[g * 6 + 5] = d.t_y[g] * wt;
  }
  block_sum_n(cnt, 1, red);
  if (cnt[0] < ncol || cnt[0] == 0.0) {
    if (tid == 0)
      for (int j = 0; j < 5; ++j) sf_sh[j] = j < ncol ? 1.0 : 0.0;
    __syncthreads();
    return;
  }
  for (int j = 0; j < ncol; ++j) {
    double nrm[1] = {0.0};
    for (uint32_t t = tid; t < sp.n_times; t += ABLOCK) {
      const double a = d.lsw[(sp.time_off + t) * 6 + j];
      nrm[0] += a * a;
    }
    block_sum_n(nrm, 1, red);
    const double rjj = sqrt(nrm[0]);
    const double inv = rjj > 0.0 ? 1.0 / rjj : 0.0;
    // q_j and the projections of the later columns and b on it.
    double dots[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t t = tid; t < sp.n_times; t += ABLOCK) {
      double* row = d.lsw + (sp.time_off + t) * 6;
      const double q = row[j] * inv;
      row[j] = q;
      for (int k = j + 1; k < 6; ++k) dots[k] += q * row[k];
    }
    block_sum_n(dots, 6, red);
    for (uint32_t t = tid; t < sp.n_times; t += ABLOCK) {
      double* row = d.lsw + (sp.time_off + t) * 6;
      const double q = row[j];
      for (int k = j + 1; k < 6; ++k) row[k] -= dots[k] * q;
    }
    if (tid == 0) {
      for (int k = 0; k < 6; ++k) R[j][k] = k < j ? 0.0 : (k == j ? rjj : dots[k]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    double U[5][5], c[5], x[5];
    for (int i = 0; i < 5; ++i) {
      c[i] = i < ncol ? R[i][5] : 0.0;
      for (int j = 0; j < 5; ++j) U[i][j] = (i < ncol && j < ncol) ? R[i][j] : 0.0;
    }
    jacobi_solve(U, c, ncol, 1e-12, x);
    for (int j = 0; j < 5; ++j) sf_sh[j] = j < ncol ? x[j] : 0.0;
  }
  __syncthreads();
}

}  // namespace

// Calibration and white-dwarf lanes of a set: their raw component sums in
// special[] and the Ginterp factors s11, s12, s21, s22 in scl[]. Every
// thread of the workgroup calls it (it synchronises the workgroup).
__device__ void set_scales(const SetDev& sp, const DeviceBatch& d, double* special, double* scl) {
#pragma clang fp contract(off)
  const uint32_t tid = threadIdx.x;
  const double* lane_phi = d.lane_phi;
  const bool cal1 = (sp.flags & SET_CAL1) != 0, cal2 = (sp.flags & SET_CAL2) != 0;
  const uint32_t n_special = (cal1 ? 4u : 0u) + (cal2 ? 4u : 0u) + 1u;
  if (tid < n_special) {
    double S[5];
    lane_sums(sp, sp.special0 + tid, d, S);
    const bool star2_lane = cal2 && tid >= (cal1 ? 4u : 0u) && tid < n_special - 1u;
    special[tid] = star2_lane ? S[1] : S[0];
  }
  __syncthreads();
  if (tid == 0) {
    // The reference updates one Ginterp in place, so each ratio sees the
    // factors found before it: fc12 is weighted by scale1 with scale11 set,
    // fc22 by scale2 with scale21 set (orchestration.rs:93-131,
    // comp_star1/2 at flux.rs:151 / 197).
    const uint32_t b = sp.lane_off + sp.special0;
    double s11 = sp.gscl[0], s12 = sp.gscl[1], s21 = sp.gscl[2], s22 = sp.gscl[3];
    uint32_t k = 0;
    if (cal1) {
      double ff = (1.0 * gscale1(lane_phi[b + 0], sp.phase1, s11, s12)) * special[0] / 1.0;
      double fc = (1.0 * gscale1(lane_phi[b + 1], sp.phase1, s11, s12)) * special[1] / 1.0;
      s11 = ff / fc;
      ff = (1.0 * gscale1(lane_phi[b + 2], sp.phase1, s11, s12)) * special[2] / 1.0;
      fc = (1.0 * gscale1(lane_phi[b + 3], sp.phase1, s11, s12)) * special[3] / 1.0;
      s12 = ff / fc;
      k = 4;
    }
    if (cal2) {
      double ff = (1.0 * gscale2(lane_phi[b + k + 0], sp.phase2, s21, s22)) * special[k + 0] / 1.0;
      double fc = (1.0 * gscale2(lane_phi[b + k + 1], sp.phase2, s21, s22)) * special[k + 1] / 1.0;
      s21 = ff / fc;
      ff = (1.0 * gscale2(lane_phi[b + k + 2], sp.phase2, s21, s22)) * special[k + 2] / 1.0;
      fc = (1.0 * gscale2(lane_phi[b + k + 3], sp.phase2, s21, s22)) * special[k + 3] / 1.0;
      s22 = ff / fc;
    }
    scl[0] = s11; scl[1] = s12; scl[2] = s21; scl[3] = s22;
  }
  __syncthreads();
}

// Per-lane sums of the partial rows, in parallel over (set, block of
// ABLOCK lanes): a sub-phase's total with the Ginterp factors applied
// (flux.rs:74-107), or for iscale sets its five raw component sums. Each
// workgroup recomputes its set's factors from the special lanes. The
// per-set assembly that follows reads these.
__global__ __launch_bounds__(ABLOCK) void lane_kernel(const DeviceBatch d) {
#pragma clang fp contract(off)
  __shared__ double special[9];
  __shared__ double scl[4];
  const SetDev sp = d.sets[blockIdx.x];
  // Both exits are uniform over the workgroup (before any barrier).
  if (!(sp.flags & SET_VALID) || blockIdx.y * ABLOCK >= sp.special0) return;
  const uint32_t l = blockIdx.y * ABLOCK + threadIdx.x;
  if (sp.flags & SET_ISCALE) {
    if (l >= sp.special0) return;
    double S[5];
    lane_sums(sp, l, d, S);
    double* o = d.lane_comp + static_cast<uint64_t>(sp.lane_off + l) * 5;
    for (int k = 0; k < 5; ++k) o[k] = S[k];
    return;
  }
  set_scales(sp, d, special, scl);
  if (l >= sp.special0) return;
  const double s11 = scl[0], s12 = scl[1], s21 = scl[2], s22 = scl[3];
  double S[5];
  lane_sums(sp, l, d, S);
  const double phi = d.lane_phi[sp.lane_off + l];
  const double ssum = S[0] * gscale1(phi, sp.phase1, s11, s12) +
                      gscale2(phi, sp.phase2, s21, s22) * S[1] + (S[2] + S[3] + S[4]);
  d.lane_tot[sp.lane_off + l] = ssum;
}

__global__ __launch_bounds__(ABLOCK) void assemble_kernel(const DeviceBatch d, int scale) {
  // FP64 assembly without contraction, in the reference's operation order.
#pragma clang fp contract(off)
  __shared__ double special[9];
  __shared__ double scl[4];
  __shared__ double red[NWAVE * 6];
  __shared__ double sf_sh[5];
  const SetDev sp = d.sets[blockIdx.x];
  const uint32_t tid = threadIdx.x;
  const double* lane_phi = d.lane_phi;
  const uint32_t* t_lane0 = d.t_lane0;
  const int* t_ndiv = d.t_ndiv;
  const double* t_slfac = d.t_slfac;
  const double* t_y = d.t_y;
  const double* t_ferr = d.t_ferr;
  const double* t_w = d.t_w;
  double* flux_out = d.flux;
  if (!(sp.flags & SET_VALID)) {
    const double nan = __builtin_nan("");
    for (uint32_t t = tid; t < sp.n_times; t += ABLOCK) flux_out[sp.time_off + t] = nan;
    if (tid == 0) d.out[blockIdx.x] = SetOut{nan, nan, nan, nan, {nan, nan, nan, nan, nan}, {nan, nan, nan, nan}};
    return;
  }
  const bool cal1 = (sp.flags & SET_CAL1) != 0, cal2 = (sp.flags & SET_CAL2) != 0;
  const uint32_t n_special = (cal1 ? 4u : 0u) + (cal2 ? 4u : 0u) + 1u;

  // 1. Calibration and white-dwarf lanes: the Ginterp factors. The other
  //    lanes' sums come from lane_kernel, launched before in the stream.
  set_scales(sp, d, special, scl);
  const double s11 = scl[0], s12 = scl[1], s21 = scl[2], s22 = scl[3];
  double wdwarf = (1.0 * gscale1(0.5, sp.phase1, s11, s12)) * special[n_special - 1u] / 1.0;

  if (sp.flags & SET_ISCALE) {
    // Per-component columns (orchestration.rs:181-227): comp_star1 with
    // scale1(phase) (flux.rs:151), disc, edge, spot, comp_star2 with
    // scale2(phi) (flux.rs:197), each times slfac, no third light.
    const int ncol = (sp.flags & SET_NCOL5) ? 5 : 4;
    for (uint32_t t = tid; t < sp.n_times; t += ABLOCK) {
      const uint32_t g = sp.time_off + t;
      const uint32_t l0 = t_lane0[g];
      const int nd = t_ndiv[g];
      const double s1 = gscale1(d.t_phase[g], sp.phase1, s11, s12);
      double c[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      for (int k = 0; k < nd; ++k) {
        const double wgt = nd == 1 ? 1.0 : ((k == 0 || k == nd - 1) ? 0.5 : 1.0);
        const double* S = d.lane_comp + static_cast<uint64_t>(sp.lane_off + l0 + k) * 5;
        c[0] += wgt * s1 * S[KIND_STAR1];
        c[1] += wgt * S[KIND_DISC];
        c[2] += wgt * S[KIND_EDGE];
        c[3] += wgt * S[KIND_SPOT];
        c[4] += wgt * gscale2(lane_phi[sp.lane_off + l0 + k], sp.phase2, s21, s22) * S[KIND_STAR2];
      }
      const double denom = static_cast<double>(nd - 1 > 1 ? nd - 1 : 1);
      for (int j = 0; j < 5; ++j) d.comp[static_cast<uint64_t>(g) * 5 + j] = j < ncol ? t_slfac[g] * (c[j] / denom) : 0.0;
    }
    __syncthreads();
    if (scale) {
      iscale_fit(sp, d, ncol, red, sf_sh);
    } else {
      if (tid == 0)
        for (int j = 0; j < 5; ++j) sf_sh[j] = 1.0;
      __syncthreads();
    }
    const double sf[5] = {sf_sh[0], sf_sh[1], sf_sh[2], sf_sh[3], sf_sh[4]};
    double cs = 0.0, wn = 0.0;
    for (uint32_t t = tid; t < sp.n_times; t += ABLOCK) {
      const uint32_t g = sp.time_off + t;
      const double* c = d.comp + static_cast<uint64_t>(g) * 5;
      double v = sf[0] * c[0] + sf[1] * c[1] + sf[2] * c[2] + sf[3] * c[3];
      if (ncol == 5) v += sf[4] * c[4];
      flux_out[g] = v;
      if (t_w[g] > 0.0) {
        wn += t_w[g];
        const double r = (t_y[g] - v) / t_ferr[g];
        cs += t_w[g] * (r * r);
      }
    }
    const double chisq = block_sum(cs, red);
    const double wnok = block_sum(wn, red);
    if (tid == 0)
      d.out[blockIdx.x] =
          SetOut{chisq, wnok, sf[0], wdwarf * sf[0], {sf[0], sf[1], sf[2], sf[3], sf[4]}, {s11, s12, s21, s22}};
    return;
  }

  // 2. Sub-phase totals: lane_tot, from lane_kernel.
  // 3. Exposure smearing and the per-time factors (flux.rs:24-111,
  //    orchestration.rs:243-251).
  for (uint32_t t = tid; t < sp.n_times; t += ABLOCK) {
    const uint32_t l0 = t_lane0[sp.time_off + t];
    const int nd = t_ndiv[sp.time_off + t];
    double sum = 0.0;
    for (int k = 0; k < nd; ++k) {
      const double wgt = nd == 1 ? 1.0 : ((k == 0 || k == nd - 1) ? 0.5 : 1.0);
      sum += wgt * d.lane_tot[sp.lane_off + l0 + k];
    }
    const double denom = static_cast<double>(nd - 1 > 1 ? nd - 1 : 1);
    flux_out[sp.time_off + t] = t_slfac[sp.time_off + t] * (sum / denom) + sp.third;
  }
  __syncthreads();

  // 4. Autoscale (re_scale, orchestration.rs:361-388) or plain
  //    chi-squared (compute_chisq, orchestration.rs:391-401).
  double sfac = 1.0, chisq = 0.0, wnok = 0.0;
  if (scale) {
    double sdy = 0.0, syy = 0.0, wn = 0.0;
    for (uint32_t t = tid; t < sp.n_times; t += ABLOCK) {
      const uint32_t g = sp.time_off + t;
      if (t_w[g] > 0.0) {
        const double f = flux_out[g];
        const double wgt = t_w[g] / (t_ferr[g] * t_ferr[g]);
        sdy += wgt * t_y[g] * f;
        syy += wgt * f * f;
        wn += t_w[g];
      }
    }
    sdy = block_sum(sdy, red);
    syy = block_sum(syy, red);
    wn = block_sum(wn, red);
    if (wn > 0.0 && syy > 0.0) {
      sfac = sdy / syy;
      double cs = 0.0;
      for (uint32_t t = tid; t < sp.n_times; t += ABLOCK) {
        const uint32_t g = sp.time_off + t;
        const double f = flux_out[g] * sfac;
        flux_out[g] = f;
        if (t_w[g] > 0.0) {
          const double r = (t_y[g] - f) / t_ferr[g];
          cs += t_w[g] * (r * r);
        }
      }
      chisq = block_sum(cs, red);
      wnok = wn;
      wdwarf *= sfac;
    }
  } else {
    double cs = 0.0, wn = 0.0;
    for (uint32_t t = tid; t < sp.n_times; t += ABLOCK) {
      const uint32_t g = sp.time_off + t;
      if (t_w[g] > 0.0) {
        wn += t_w[g];
        const double r = (t_y[g] - flux_out[g]) / t_ferr[g];
        cs += t_w[g] * (r * r);
      }
    }
    chisq = block_sum(cs, red);
    wnok = block_sum(wn, red);
  }
  if (tid == 0)
    d.out[blockIdx.x] = SetOut{chisq, wnok, sfac, wdwarf, {sfac, sfac, sfac, sfac, sfac}, {s11, s12, s21, s22}};
}

// ------------------------------------------------------------ visibility
__global__ void visibility_kernel(const char* __restrict__ recs, SegDev sg,
                                  const double* __restrict__ phases, uint32_t nphase,
                                  uint8_t* __restrict__ mask) {
  const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t total = static_cast<uint64_t>(sg.n) * nphase;
  if (idx >= total) return;
  const uint32_t p = static_cast<uint32_t>(idx / sg.n);
  const uint32_t e = static_cast<uint32_t>(idx % sg.n);
  const double ph = phases[p];
  const double phif = ph - floor(ph);
  const uint32_t stride = rec_stride(sg.necl, (sg.flags & SEG_POSN) != 0);
  const uint32_t eoff = (sg.flags & SEG_POSN) ? 32u : 16u;
  const double* ecl = reinterpret_cast<const double*>(recs + sg.rec_off + static_cast<uint64_t>(e) * stride + eoff);
  bool vis;
  switch (sg.necl) {
    case 0: vis = eclipse_free<0>(ecl, phif); break;
    case 1: vis = eclipse_free<1>(ecl, phif); break;
    case 2: vis = eclipse_free<2>(ecl, phif); break;
    case 3: vis = eclipse_free<3>(ecl, phif); break;
    default: vis = eclipse_free<4>(ecl, phif); break;
  }
  mask[static_cast<uint64_t>(p) * sg.n + e] = vis ? 1 : 0;
}

// ------------------------------------------------------------ launchers
hipError_t launch_flux_items(const DeviceBatch& d, hipStream_t stream, const SideStream* ss) {
  if (d.n_items == 0) return hipSuccess;
  // Host order: items without beaming / lensing and without staged eclipse
  // intervals, then those with staged intervals, then the beamed / lensed
  // ones.
  const uint32_t nf = d.n_items_plain, ne = d.n_items_ecl, nx = d.n_items - nf - ne;
  // The three item kernels write disjoint partial rows, so the other two
  // may run beside the plain one.
  const bool fork = ss && ss->side && nf && (ne || nx);
  hipStream_t s2 = fork ? ss->side : stream;
  if (fork) {
    hipError_t e = hipEventRecord(ss->fork, stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(ss->side, ss->fork, 0);
    if (e != hipSuccess) return e;
  }
  if (nf)
    hipLaunchKernelGGL((flux_items_kernel<false, false>), dim3(nf), dim3(TILE), 0, stream, d.items, nf, d.segs,
                       d.sets, d.recs, d.moms, d.imoms, d.lane_phi, d.lane_e, d.lane_meta, d.lane_masks, d.partials);
  if (ne)
    hipLaunchKernelGGL((flux_items_kernel<false, true>), dim3(ne), dim3(TILE), 0, s2, d.items + nf, ne,
                       d.segs, d.sets, d.recs, d.moms, d.imoms, d.lane_phi, d.lane_e, d.lane_meta, d.lane_masks,
                       d.partials);
  if (nx)
    hipLaunchKernelGGL((flux_items_kernel<true, false>), dim3(nx), dim3(TILE), 0, s2, d.items + nf + ne, nx,
                       d.segs, d.sets, d.recs, d.moms, d.imoms, d.lane_phi, d.lane_e, d.lane_meta, d.lane_masks,
                       d.partials);
  if (fork) {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(ss->join, ss->side);
    if (e == hipSuccess) e = hipStreamWaitEvent(stream, ss->join, 0);
    if (e != hipSuccess) return e;
  }
  // After the items in stream order: limb chunks add to their rows.
  if (d.n_limb_chunks)
    hipLaunchKernelGGL(limb_kernel, dim3(d.n_limb_chunks), dim3(LIMB_WAVES * TILE), 0, stream, d);
  return hipGetLastError();
}

hipError_t launch_assemble(const DeviceBatch& d, int scale, hipStream_t stream) {
  if (d.n_sets == 0) return hipSuccess;
  const uint32_t gy = (d.max_set_lanes + ABLOCK - 1) / ABLOCK;
  if (gy > 65535u) return hipErrorInvalidValue;
  if (gy > 0) {
    hipLaunchKernelGGL(lane_kernel, dim3(d.n_sets, gy), dim3(ABLOCK), 0, stream, d);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(assemble_kernel, dim3(d.n_sets), dim3(ABLOCK), 0, stream, d, scale);
  return hipGetLastError();
}

hipError_t launch_visibility(const DeviceBatch& d, const SegDev& sg, const double* phases,
                             uint32_t nphase, uint8_t* mask, hipStream_t stream) {
  uint64_t total = static_cast<uint64_t>(sg.n) * nphase;
  if (total == 0) return hipSuccess;
  uint32_t blocks = static_cast<uint32_t>((total + 255) / 256);
  hipLaunchKernelGGL(visibility_kernel, dim3(blocks), dim3(256), 0, stream, d.recs, sg, phases,
                     nphase, mask);
  return hipGetLastError();
}

}  // namespace lch
