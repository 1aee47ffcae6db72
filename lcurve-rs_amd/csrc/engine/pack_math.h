// Arithmetic of the per-set pack shared by the host packer (batch_build.cpp)
// and the GPU packer (pack_kernels.hip): the spatial key that orders a
// segment's elements into patches, the bounding cone of a patch's normals
// and the per-(patch, tile) work-code decision. Only IEEE-exact operations
// (+ - * /, sqrt, comparisons) appear, in a fixed order, and the device
// translation unit compiles them without FMA contraction, so host and GPU
// packs of the same grid are identical bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "layout.h"

namespace lch {

// Position of a normal along a Hilbert curve over its octahedral map,
// 8 bits per axis: consecutive keys are close on the sphere, so runs of 64
// elements sorted by key have narrow normal cones. The FP32 normal the
// records hold is the input.
__host__ __device__ inline uint32_t direction_key16(float fx, float fy, float fz) {
  const double x0 = fx, y0 = fy, z0 = fz;
  const double ax = x0 < 0.0 ? -x0 : x0, ay = y0 < 0.0 ? -y0 : y0, az = z0 < 0.0 ? -z0 : z0;
  const double s = ax + ay + az;
  if (!(s > 0.0)) return 0u;
  double u = x0 / s, v = y0 / s;
  if (z0 < 0.0) {
    const double au = u < 0.0 ? -u : u, av = v < 0.0 ? -v : v;
    const double uu = (1.0 - av) * (u >= 0.0 ? 1.0 : -1.0);
    const double vv = (1.0 - au) * (v >= 0.0 ? 1.0 : -1.0);
    u = uu;
    v = vv;
  }
  constexpr int N = 1 << 8;
  auto q = [](double t) {
    double w = (t + 1.0) * 0.5 * static_cast<double>(N - 1);
    w = w < 0.0 ? 0.0 : (w > static_cast<double>(N - 1) ? static_cast<double>(N - 1) : w);
    return static_cast<int>(w);
  };
  int x = q(u), y = q(v);
  uint32_t key = 0;
  for (int sq = N / 2; sq > 0; sq /= 2) {
    const int rx = (x & sq) ? 1 : 0, ry = (y & sq) ? 1 : 0;
    key += static_cast<uint32_t>(sq) * static_cast<uint32_t>(sq) * static_cast<uint32_t>((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = sq - 1 - x;
        y = sq - 1 - y;
      }
      const int t = x;
      x = y;
      y = t;
    }
  }
  return key;
}

// Sums over the 64 slots of a patch are taken in one fixed order: the
// xor butterfly of a wavefront (offsets 32, 16, ..., 1, each lane adding
// its partner's value), after which every lane holds the same total. Lane
// 0's chain alone is the tree below: 63 additions, the lower half at each
// offset (a[i] + a[i + off] is lane i's addition; the upper lanes repeat
// the lower ones' values). `a` is overwritten.
__host__ __device__ inline double tree_sum64(double* a) {
  for (int off = 32; off > 0; off >>= 1)
    for (int i = 0; i < off; ++i) a[i] = a[i] + a[i + off];
  return a[0];
}

// Normal cone of a patch: unit axis (x, y, z) and cosw, a lower bound on
// the cosine of the angle between the axis and any normal of the patch
// (-1 or below: no bound). An empty patch or a zero mean has cosw < -1.
struct Cone {
  float x, y, z, cosw;
};

// A patch's cone from the sum of its FP32 normals (tree_sum64 order over
// the 64 slots, padding slots adding 0): set_sum(), axis(), then bound()
// for every element (a minimum, in any order), then finish().
struct ConeBuilder {
  double mx = 0.0, my = 0.0, mz = 0.0;
  double ux = 0.0, uy = 0.0, uz = 0.0;
  double mn = 1.0;
  bool any = false, ok = false;
  __host__ __device__ void set_sum(double sx, double sy, double sz, bool nonempty) {
    mx = sx;
    my = sy;
    mz = sz;
    any = nonempty;
  }
  __host__ __device__ void axis() {
    const double len = sqrt(mx * mx + my * my + mz * mz);
    ok = any && len > 0.0;
    if (ok) {
      const double inv = 1.0 / len;
      ux = inv * mx;
      uy = inv * my;
      uz = inv * mz;
    }
  }
  // Cosine between the axis and one normal (clamped to [-1, 1]), or 1
  // (no constraint) for a zero normal or a cone without an axis.
  __host__ __device__ double cosine(float x, float y, float z) const {
    if (!ok) return 1.0;
    const double dx = x, dy = y, dz = z;
    const double dl = sqrt(dx * dx + dy * dy + dz * dz);
    if (!(dl > 0.0)) return 1.0;
    const double ca = (ux * dx + uy * dy + uz * dz) / dl;
    return ca < -1.0 ? -1.0 : (ca > 1.0 ? 1.0 : ca);
  }
  __host__ __device__ void bound(float x, float y, float z) {
    const double ca = cosine(x, y, z);
    mn = ca < mn ? ca : mn;
  }
  __host__ __device__ Cone finish() const {
    Cone c{0.0f, 0.0f, 0.0f, -2.0f};
    if (!ok) return c;
    c.x = static_cast<float>(ux);
    c.y = static_cast<float>(uy);
    c.z = static_cast<float>(uz);
    // A small slack for the rounding of the bound itself, then rounded down
    // to FP32.
    double cp = mn - 2e-6;
    cp = cp < -1.0 ? -1.0 : cp;
    const float f = static_cast<float>(cp);
    c.cosw = static_cast<double>(f) > cp ? nextafterf(f, -4.0f) : f;
    return c;
  }
};

// Cosine between a tile's line of sight e (unit, rounded to FP32) and a
// patch axis p (unit, rounded to FP32) for the code decision: FP32
// products and sums in this order. Its rounding (a few 1e-7) is far inside
// the decision's 2e-3 rad margin.
__host__ __device__ inline float los_cosine(float ex, float ey, float ez, float px, float py, float pz) {
  return ex * px + ey * py + ez * pz;
}

// Unit axis of a cone (from its FP32 axis), with plen = 0 for a cone that
// bounds nothing.
__host__ __device__ inline void cone_axis(const Cone& c, double* px, double* py, double* pz, double* plen) {
  double x = c.x, y = c.y, z = c.z;
  const double len = sqrt(x * x + y * y + z * z);
  if (len > 0.0) {
    const double inv = 1.0 / len;
    x = inv * x;
    y = inv * y;
    z = inv * z;
  }
  *px = x;
  *py = y;
  *pz = z;
  *plen = len;
}

// Per-segment constants of the code decision. An element adds nothing
// unless mu > thr; it is seen whole when mu > thr_front = max(thr, 0) for
// the whole patch. With a = acos(thr) + margin and b = acos(thr_front) -
// margin (margin 2e-3 rad covers FP32 rounding of mu and of the cone):
//  * back-facing (skip) iff the patch half-angle w < pi - 1e-3 - a, i.e.
//    cosw > c_skip, and max over the tile of e.axis < cos(a + w)
//    = ca cosw - sa sinw;
//  * seen whole iff w < b - 1e-3, i.e. cosw > c_front, and min over the
//    tile of e.axis > cos(b - w) = cb cosw + sb sinw.
// c_skip / c_front = 2 disables the test. The host evaluates the
// transcendental constants once per segment.
struct CodeConsts {
  double ca, sa, c_skip;
  double cb, sb, c_front;
};

__host__ __device__ inline double cone_sin(float cosw) {
  const double c = cosw;
  const double s2 = 1.0 - c * c;
  return s2 > 0.0 ? sqrt(s2) : 0.0;
}

// cos(a + w) and cos(b - w) of CodeConsts for a cone of half-angle w.
__host__ __device__ inline double cos_skip(const CodeConsts& k, float cosw) {
  return k.ca * static_cast<double>(cosw) - k.sa * cone_sin(cosw);
}
__host__ __device__ inline double cos_front(const CodeConsts& k, float cosw) {
  return k.cb * static_cast<double>(cosw) + k.sb * cone_sin(cosw);
}

// The key that orders a segment's elements into patches: the top
// DIR_KEY_BITS bits of direction_key16, a Hilbert curve over 64 x 64 cells
// of the octahedral map. Ties keep the grid order. With the eclipse-pair
// count above it, a grid's sort key has 15 bits: two 8-bit radix passes.
constexpr uint32_t DIR_KEY_BITS = 12;
__host__ __device__ inline uint32_t direction_key(float fx, float fy, float fz) {
  return direction_key16(fx, fy, fz) >> (16u - DIR_KEY_BITS);
}

// One record's terms of a patch's direction moments (layout of MOM_STRIDE
// in layout.h) from its normal d and weight w: w d, the coefficients of
// w (e.d)^2 and of w (e.d)^3 in the monomials of e, and |w|. A record of
// zero weight (padding) has all terms +0. The moments are the tree_sum64
// sums of the terms over a patch's 64 slots.
__host__ __device__ inline void moment_terms(double* T, float fx, float fy, float fz, float fw) {
  const double x = fx, y = fy, z = fz, w = fw;
  if (w == 0.0) {
    for (uint32_t c = 0; c < MOM_STRIDE; ++c) T[c] = 0.0;
    return;
  }
  T[0] = w * x;
  T[1] = w * y;
  T[2] = w * z;
  T[3] = w * x * x;
  T[4] = w * y * y;
  T[5] = w * z * z;
  T[6] = 2.0 * w * x * y;
  T[7] = 2.0 * w * x * z;
  T[8] = 2.0 * w * y * z;
  T[9] = w * x * x * x;
  T[10] = w * y * y * y;
  T[11] = w * z * z * z;
  T[12] = 3.0 * w * x * x * y;
  T[13] = 3.0 * w * x * x * z;
  T[14] = 3.0 * w * y * y * x;
  T[15] = 3.0 * w * y * y * z;
  T[16] = 3.0 * w * z * z * x;
  T[17] = 3.0 * w * z * z * y;
  T[18] = 6.0 * w * x * y * z;
  T[19] = w < 0.0 ? -w : w;  // size bound for the limb lanes' fixed point
}

// Point::visible over a tile's phase range [lo, hi] (types.rs:50-58, the
// same FP64 comparisons) for one element's interval triple
// (ingress, egress, egress - 1): misses = no lane of the range eclipsed by
// it, holds = every lane eclipsed.
__host__ __device__ inline void interval_vs_range(const double* e3, double lo, double hi, bool* misses,
                                                  bool* holds) {
  *misses = (hi < e3[0] || lo > e3[1]) && lo > e3[2];
  *holds = (lo >= e3[0] && hi <= e3[1]) || hi <= e3[2];
}

}  // namespace lch
