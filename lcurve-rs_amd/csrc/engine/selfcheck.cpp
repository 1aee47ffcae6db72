// CPU self-check of the per-(patch, tile) work codes of a packed set.
//
// Every (element, lane) pair the flux kernel would see is evaluated twice in
// FP64 from the packed records: once as the codes direct (PT_SKIP adds
// nothing, PT_MOM uses the patch moments, PT_FREE drops the eclipse test,
// PT_FULL keeps it) and once by brute force with the eclipse test on every
// element. The two sums per (lane, segment) must agree; a patch skipped
// although some element adds to some lane, an eclipse test dropped although
// some element is eclipsed, or a wrong moment shows up as a difference.
// Beamed / lensed star segments use an indicator term (nonzero exactly when
// the reference adds the element) since only the skip decisions apply there.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>

#include "../setup/grid.h"
#include "../setup/roche.h"

#include "batch_build.h"
#include "selfcheck.h"

namespace lch {

namespace {

struct Lane {
  double phif;
  double ex, ey, ez;
};

bool visible(const uint8_t* e, uint32_t necl, double phif) {
  for (uint32_t k = 0; k < necl; ++k) {
    double e3[3];
    std::memcpy(e3, e + 24 * k, sizeof e3);
    if ((phif >= e3[0] && phif <= e3[1]) || phif <= e3[2]) return false;
  }
  return true;
}

// The reference's term for one element and lane (flux.rs:36-106), FP64 on
// the packed FP32 values.
double term(const SegDev& sg, const SetDev& d, const uint8_t* r, const Lane& L, bool eclipse_test) {
  float f[4];
  std::memcpy(f, r, sizeof f);
  const double mu = L.ex * f[0] + L.ey * f[1] + L.ez * f[2];
  const double w = f[3];
  const uint32_t eoff = (sg.flags & SEG_POSN) ? 32u : 16u;
  if (eclipse_test && !visible(r + eoff, sg.necl, L.phif)) return 0.0;
  if (sg.kind <= KIND_STAR2) {
    const bool s1 = sg.kind == KIND_STAR1;
    const double mucrit = s1 ? d.mucrit1 : d.mucrit2;
    if (!(mu > mucrit)) return 0.0;
    if (sg.flags & SEG_POSN) return w * (1.0 + std::fabs(mu));
    if (!(mu > 0.0)) return 0.0;
    const float* a = s1 ? d.ldc1 : d.ldc2;
    const double m = std::min(mu, 1.0);
    double im;
    if ((s1 ? d.ltype1 : d.ltype2) == 1u) {
      const double s = std::sqrt(m);
      im = 1.0 - (double(a[0]) + a[1] + a[2] + a[3]) + s * (a[0] + s * (a[1] + s * (a[2] + s * double(a[3]))));
    } else {
      const double o = 1.0 - m;
      im = 1.0 - o * (a[0] + o * (a[1] + o * (a[2] + o * double(a[3]))));
    }
    return mu * w * im;
  }
  if (!(mu > 0.0)) return 0.0;
  if (sg.kind == KIND_SPOT) return mu * w;
  const double o = 1.0 - mu;
  return mu * w * (1.0 - o * (double(d.lin_disc) + double(d.quad_disc) * o));
}

double moment_term(const SegDev& sg, const SetDev& d, const double* M, const Lane& L) {
  const bool star = sg.kind <= KIND_STAR2;
  const float* a = sg.kind == KIND_STAR1 ? d.ldc1 : d.ldc2;
  const double A = star ? a[0] : (sg.kind == KIND_SPOT ? 0.0 : d.lin_disc);
  const double B = star ? a[1] : (sg.kind == KIND_SPOT ? 0.0 : d.quad_disc);
  const double x = L.ex, y = L.ey, z = L.ez;
  const double s1 = M[0] * x + M[1] * y + M[2] * z;
  const double s2 = M[3] * x * x + M[4] * y * y + M[5] * z * z + M[6] * x * y + M[7] * x * z + M[8] * y * z;
  const double s3 = M[9] * x * x * x + M[10] * y * y * y + M[11] * z * z * z + M[12] * x * x * y +
                    M[13] * x * x * z + M[14] * y * y * x + M[15] * y * y * z + M[16] * z * z * x +
                    M[17] * z * z * y + M[18] * x * y * z;
  return (1.0 - A - B) * s1 + (A + 2.0 * B) * s2 - B * s3;
}

const SetPack::LaneSplit* find_split(const SetPack& S, uint32_t pt) {
  auto it = std::lower_bound(S.lsplit.begin(), S.lsplit.end(), pt,
                             [](const SetPack::LaneSplit& a, uint32_t v) { return a.pt < v; });
  return it != S.lsplit.end() && it->pt == pt ? &*it : nullptr;
}

}  // namespace

SelfCheck selfcheck_codes(const SetPack& S) {
  SelfCheck out;
  const SetDev& d = S.dev;
  const uint32_t ntile = S.n_tiles;
  for (const SegDev& sg : S.segs) {
    const uint32_t stride = rec_stride(sg.necl, (sg.flags & SEG_POSN) != 0);
    for (uint32_t t = 0; t < ntile; ++t) {
      for (uint32_t q = 0; q < sg.n / PATCH; ++q)
        ++out.counts[S.ptcode[static_cast<size_t>(sg.patch_off + q) * ntile + t]];
      for (uint32_t l = t * TILE; l < (t + 1) * TILE; ++l) {
        const uint32_t m = S.lane_meta[l];
        if (!(m & META_VALID) || !(sg.use_mask & cp_bit(meta_class(m), meta_ptype(m)))) continue;
        const double ph = S.lane_phi[l];
        Lane L;
        L.phif = ph - std::floor(ph);
        L.ex = S.lane_e[l].x;
        L.ey = S.lane_e[l].y;
        L.ez = S.lane_e[l].z;
        double coded = 0.0, brute = 0.0, scale = 0.0;
        for (uint32_t q = 0; q < sg.n / PATCH; ++q) {
          const uint8_t code = S.ptcode[static_cast<size_t>(sg.patch_off + q) * ntile + t];
          const uint8_t* pr = S.recs.data() + sg.rec_off + static_cast<uint64_t>(q) * PATCH * stride;
          double pb = 0.0;
          for (uint32_t i = 0; i < PATCH; ++i) {
            const double v = term(sg, d, pr + static_cast<size_t>(i) * stride, L, true);
            pb += v;
            scale += std::fabs(v);
          }
          brute += pb;
          const SetPack::LaneSplit* sp = code == PT_FREE ? find_split(S, (sg.patch_off + q) * ntile + t) : nullptr;
          if (sp) {
            const uint64_t bit = 1ull << (l - t * TILE);
            if (sp->all & bit) {
              ++out.split_lanes[0];
              coded += moment_term(sg, d, &S.moms[static_cast<size_t>(sg.patch_off + q) * MOM_STRIDE], L);
            } else if (sp->some & bit) {
              ++out.split_lanes[1];
              for (uint32_t i = 0; i < PATCH; ++i)
                coded += term(sg, d, pr + static_cast<size_t>(i) * stride, L, false);
            } else {
              ++out.split_lanes[2];
            }
          } else if (code == PT_MOM) {
            coded += moment_term(sg, d, &S.moms[static_cast<size_t>(sg.patch_off + q) * MOM_STRIDE], L);
          } else if (code != PT_SKIP) {
            for (uint32_t i = 0; i < PATCH; ++i)
              coded += term(sg, d, pr + static_cast<size_t>(i) * stride, L, code == PT_FULL);
          }
        }
        const double diff = std::fabs(coded - brute);
        const double rel = scale > 0.0 ? diff / scale : (diff > 0.0 ? 1.0 : 0.0);
        out.max_rel = std::max(out.max_rel, rel);
        ++out.lane_segments;
      }
    }
  }
  return out;
}

SelfCheck selfcheck_model(const lc::Model& model, const lc::Data& data) {
  auto P = lc::prepare_model(model, data);
  auto S = pack_set(*P, data, false);
  return selfcheck_codes(*S);
}

namespace {
bool same_bits(const void* a, const void* b, size_t n) { return std::memcmp(a, b, n) == 0; }

bool same_point(const lc::Point& a, const lc::Point& b) {
  if (!same_bits(&a.posn, &b.posn, sizeof a.posn) || !same_bits(&a.dirn, &b.dirn, sizeof a.dirn) ||
      !same_bits(&a.area, &b.area, sizeof a.area) ||
      !same_bits(&a.gravity, &b.gravity, sizeof a.gravity) || a.eclipse.size() != b.eclipse.size())
    return false;
  for (size_t k = 0; k < a.eclipse.size(); ++k)
    if (!same_bits(&a.eclipse[k], &b.eclipse[k], sizeof a.eclipse[k])) return false;
  return true;
}

struct PrescreenMode {
  bool old;
  explicit PrescreenMode(bool on) : old(lc::set_eclipse_prescreen(on)) {}
  ~PrescreenMode() { lc::set_eclipse_prescreen(old); }
};
}  // namespace

PrescreenCheck selfcheck_prescreen(const lc::Model& model) {
  using clock = std::chrono::steady_clock;
  auto ms = [](clock::time_point a, clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  PrescreenCheck out;
  for (lc::Star star : {lc::Star::Primary, lc::Star::Secondary}) {
    for (bool fine : {true, false}) {
      std::vector<lc::Point> plain, screened;
      long long s0, e0, s1, e1;
      const auto t0 = clock::now();
      {
        PrescreenMode mode(false);
        plain = lc::set_star_grid(model, star, fine);
      }
      const auto t1 = clock::now();
      lc::eclipse_prescreen_counts(&s0, &e0);
      {
        PrescreenMode mode(true);
        screened = lc::set_star_grid(model, star, fine);
      }
      const auto t2 = clock::now();
      lc::eclipse_prescreen_counts(&s1, &e1);
      out.ms_plain += ms(t0, t1);
      out.ms_screened += ms(t1, t2);
      out.skipped += s1 - s0;
      out.evaluated += e1 - e0;
      out.elements += static_cast<long long>(plain.size());
      if (plain.size() != screened.size()) {
        out.mismatches += static_cast<long long>(std::max(plain.size(), screened.size()));
        continue;
      }
      for (size_t i = 0; i < plain.size(); ++i)
        if (!same_point(plain[i], screened[i])) ++out.mismatches;
    }
  }
  return out;
}

}  // namespace lch
