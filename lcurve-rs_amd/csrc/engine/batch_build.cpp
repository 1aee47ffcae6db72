// Packs CPU-setup output into the device layout of layout.h.
//
// Per set:
//  * every grid is split into segments of equal eclipse-pair count (stable
//    order within a segment), so the kernel's interval loop has a fixed trip
//    count and segments with no eclipses skip the test;
//  * one lane per exposure sub-phase (exposure_weight, flux.rs:378-387),
//    then the calibration lanes at the Ginterp phases of
//    orchestration.rs:93-131 and the white-dwarf lane at phase 0.5
//    (orchestration.rs:256-259);
//  * per-time smearing counts, polynomial factors and observations.
#include "batch_build.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>

namespace lch {

namespace {

uint32_t mask_of(uint32_t cls, std::initializer_list<uint32_t> ptypes) {
  uint32_t m = 0;
  for (uint32_t p : ptypes) m |= cp_bit(cls, p);
  return m;
}

const std::initializer_list<uint32_t> kAll = {1u, 2u, 3u};


}  // namespace

// Per-segment constants of the code decision (CodeConsts), with the cone
// margin of 2e-3 rad.
CodeConsts code_consts(bool cull, double thr, bool front_ok, double thr_front) {
  constexpr double margin = 2e-3;
  CodeConsts k{1.0, 0.0, 2.0, 1.0, 0.0, 2.0};
  if (cull) {
    const double a = std::acos(thr) + margin;
    k.ca = std::cos(a);
    k.sa = std::sin(a);
    k.c_skip = lc::PI - 1e-3 - a > 0.0 ? std::cos(lc::PI - 1e-3 - a) : 2.0;
  }
  if (front_ok) {
    const double b = std::acos(thr_front) - margin;
    k.cb = std::cos(b);
    k.sb = std::sin(b);
    k.c_front = b - 1e-3 > 0.0 ? std::cos(b - 1e-3) : 2.0;
  }
  return k;
}

int moment_degree(const SegDev& sg, const SetDev& d) {
  if (sg.kind == KIND_SPOT) return 1;
  if (sg.kind == KIND_DISC || sg.kind == KIND_EDGE) return 3;
  if (sg.flags & SEG_POSN) return 0;  // beaming / lensing
  const bool s1 = sg.kind == KIND_STAR1;
  if ((s1 ? d.ltype1 : d.ltype2) != 0u) return 0;
  const float* a = s1 ? d.ldc1 : d.ldc2;
  if (a[2] != 0.0f || a[3] != 0.0f) return 0;
  return a[1] != 0.0f ? 3 : 2;
}

SegCodeRule seg_code_rule(const SegDev& sg, const SetDev& d, const lc::PreparedModel& P) {
  SegCodeRule r;
  r.mom_ok = moment_degree(sg, d) > 0;
  // An element adds nothing unless mu > thr: the star's mucrit (mu <= 0
  // also gives 0 through imu unless beaming evaluates imu at another mu),
  // or 0 for disc, edge and spot (flux.rs:36-106).
  r.thr = 0.0;
  if (sg.kind <= KIND_STAR2) {
    const bool s1 = sg.kind == KIND_STAR1;
    const double mc = s1 ? P.ldc1.mucrit : P.ldc2.mucrit;
    const bool beam = s1 ? P.beam_factor1 != 0.0 : P.beam_factor2 != 0.0;
    r.thr = beam ? mc : std::max(mc, 0.0);
  }
  r.cull = std::fabs(r.thr) < 0.99;
  // Seen whole: mu > max(thr, 0) for every element and lane, so that the
  // reference's term is the contribution polynomial itself.
  r.thr_front = std::max(r.thr, 0.0);
  r.front_ok = r.mom_ok && r.thr_front < 0.99;
  r.K = code_consts(r.cull, r.thr, r.front_ok, r.thr_front);
  return r;
}

bool linear_law_segment(const SegDev& sg, const SetDev& d) {
  if (sg.kind > KIND_STAR2 || (sg.flags & (SEG_POSN | SEG_LINW))) return false;
  const bool s1 = sg.kind == KIND_STAR1;
  if ((s1 ? d.ltype1 : d.ltype2) != 0u) return false;
  const float* a = s1 ? d.ldc1 : d.ldc2;
  return !(a[1] != 0.0f || a[2] != 0.0f || a[3] != 0.0f || (s1 ? d.mucrit1 : d.mucrit2) > 0.0f);
}

namespace {

// Direction moments of one patch (layout of MOM_STRIDE) from its records:
// each component the tree_sum64 sum of the records' terms.
void patch_moments(const uint8_t* pr, uint32_t stride, double* M) {
  double T[PATCH][MOM_STRIDE];
  for (uint32_t i = 0; i < PATCH; ++i) {
    float f[4];
    std::memcpy(f, pr + static_cast<size_t>(i) * stride, sizeof f);
    moment_terms(T[i], f[0], f[1], f[2], f[3]);
  }
  double a[PATCH];
  for (uint32_t c = 0; c < MOM_STRIDE; ++c) {
    for (uint32_t i = 0; i < PATCH; ++i) a[i] = T[i][c];
    M[c] = tree_sum64(a);
  }
}

// Cone of one patch of records, nr of them real (the rest padding).
Cone patch_cone(const uint8_t* pr, uint32_t stride, uint32_t nr) {
  double sx[PATCH], sy[PATCH], sz[PATCH];
  for (uint32_t i = 0; i < PATCH; ++i) {
    float f[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (i < nr) std::memcpy(f, pr + static_cast<size_t>(i) * stride, sizeof f);
    sx[i] = i < nr ? static_cast<double>(f[0]) : 0.0;
    sy[i] = i < nr ? static_cast<double>(f[1]) : 0.0;
    sz[i] = i < nr ? static_cast<double>(f[2]) : 0.0;
  }
  ConeBuilder cb;
  cb.set_sum(tree_sum64(sx), tree_sum64(sy), tree_sum64(sz), nr > 0);
  cb.axis();
  for (uint32_t i = 0; i < nr; ++i) {
    float f[4];
    std::memcpy(f, pr + static_cast<size_t>(i) * stride, sizeof f);
    cb.bound(f[0], f[1], f[2]);
  }
  return cb.finish();
}

void push_lane(SetPack& S, double phi, uint32_t cls, const lc::Ginterp& gi) {
  S.lane_phi.push_back(phi);
  S.lane_meta.push_back(cls | (static_cast<uint32_t>(gi.grid_type(phi)) << 2) | META_VALID);
}

}  // namespace

std::unique_ptr<SetPack> invalid_set(const std::string& err, const lc::Data& data) {
  auto sp = std::make_unique<SetPack>();
  SetPack& S = *sp;
  S.host.valid = false;
  S.host.error = err;
  S.host.n_times = static_cast<uint32_t>(data.size());
  S.dev.flags = 0;
  S.dev.n_times = static_cast<uint32_t>(data.size());
  S.dev.n_lanes = 0;
  S.t_lane0.assign(data.size(), 0u);
  S.t_ndiv.assign(data.size(), 0);
  S.t_slfac.assign(data.size(), 0.0);
  S.t_phase.assign(data.size(), 0.0);
  S.t_y.assign(data.size(), 0.0);
  S.t_ferr.assign(data.size(), 1.0);
  S.t_w.assign(data.size(), 0.0);
  return sp;
}

void set_scalars(SetPack& S, const lc::PreparedModel& P) {
  SetDev& d = S.dev;
  SetHost& H = S.host;
  H.valid = true;
  H.logg1 = P.logg1;
  H.logg2 = P.logg2;
  H.rv1 = P.rv1;
  H.rv2 = P.rv2;
  H.star1_alias = P.star1_alias;
  H.star2_alias = P.star2_alias;

  const bool beam1 = P.beam_factor1 != 0.0, beam2 = P.beam_factor2 != 0.0, lens = P.glens1;
  d.flags = SET_VALID | (P.need_cal1 ? SET_CAL1 : 0u) | (P.need_cal2 ? SET_CAL2 : 0u) |
            (beam1 ? SET_BEAM1 : 0u) | (beam2 ? SET_BEAM2 : 0u) | (lens ? SET_LENS : 0u) |
            (P.iscale ? SET_ISCALE : 0u) | (P.ncol == 5 ? SET_NCOL5 : 0u);
  d.ltype1 = static_cast<uint32_t>(P.ldc1.ltype);
  d.ltype2 = static_cast<uint32_t>(P.ldc2.ltype);
  const double l1[4] = {P.ldc1.ldc1, P.ldc1.ldc2, P.ldc1.ldc3, P.ldc1.ldc4};
  const double l2[4] = {P.ldc2.ldc1, P.ldc2.ldc2, P.ldc2.ldc3, P.ldc2.ldc4};
  for (int i = 0; i < 4; ++i) {
    d.ldc1[i] = static_cast<float>(l1[i]);
    d.ldc2[i] = static_cast<float>(l2[i]);
  }
  d.mucrit1 = static_cast<float>(P.ldc1.mucrit);
  d.mucrit2 = static_cast<float>(P.ldc2.mucrit);
  d.lin_disc = static_cast<float>(P.lin_limb_disc);
  d.quad_disc = static_cast<float>(P.quad_limb_disc);
  d.beam1 = static_cast<float>(P.beam_factor1);
  d.beam2 = static_cast<float>(P.beam_factor2);
  d.vfac = static_cast<float>(P.vfac);
  d.xcofm = static_cast<float>(P.xcofm);
  d.spin1 = static_cast<float>(P.spin1);
  d.spin2 = static_cast<float>(P.spin2);
  d.rlens1 = static_cast<float>(P.rlens1);
  d.cosi_lc = P.cosi_lc;
  d.sini_lc = P.sini_lc;
  d.cosi_rad = P.cosi_rad;
  d.sini_rad = P.sini_rad;
  d.phase1 = P.phase1;
  d.phase2 = P.phase2;
  d.third = P.third;
  for (double& g : d.gscl) g = 1.0;
}

std::vector<GridRef> grid_refs(const lc::PreparedModel& P) {
  const bool beam1 = P.beam_factor1 != 0.0, beam2 = P.beam_factor2 != 0.0, lens = P.glens1;
  // Which lanes add which grid (class x grid type; see LaneClass).
  const bool cc = P.cal_copies;
  std::vector<GridRef> grids;
  if (P.star1_alias) {
    grids.push_back({0, &P.star1f, KIND_STAR1, mask_of(CLS_NORMAL, kAll) | mask_of(CLS_WD, kAll), beam1});
  } else {
    uint32_t cal_f = (P.need_cal1 && !cc) ? mask_of(CLS_CAL1, {1u}) : 0u;
    uint32_t cal_c = (P.need_cal1 && !cc) ? mask_of(CLS_CAL1, {2u, 3u}) : 0u;
    grids.push_back({0, &P.star1f, KIND_STAR1, mask_of(CLS_NORMAL, {1u}) | mask_of(CLS_WD, {1u}) | cal_f, beam1});
    grids.push_back({1, &P.star1c, KIND_STAR1, mask_of(CLS_NORMAL, {2u, 3u}) | mask_of(CLS_WD, {2u, 3u}) | cal_c, beam1});
  }
  const bool posn2 = beam2 || lens;
  if (P.star2_alias) {
    grids.push_back({2, &P.star2f, KIND_STAR2, mask_of(CLS_NORMAL, kAll), posn2});
  } else {
    uint32_t cal_f = (P.need_cal2 && !cc) ? mask_of(CLS_CAL2, {3u}) : 0u;
    uint32_t cal_c = (P.need_cal2 && !cc) ? mask_of(CLS_CAL2, {1u, 2u}) : 0u;
    grids.push_back({2, &P.star2f, KIND_STAR2, mask_of(CLS_NORMAL, {3u}) | cal_f, posn2});
    grids.push_back({3, &P.star2c, KIND_STAR2, mask_of(CLS_NORMAL, {1u, 2u}) | cal_c, posn2});
  }
  grids.push_back({4, &P.disc, KIND_DISC, mask_of(CLS_NORMAL, kAll), false});
  grids.push_back({5, &P.edge, KIND_EDGE, mask_of(CLS_NORMAL, kAll), false});
  grids.push_back({6, &P.spot, KIND_SPOT, mask_of(CLS_NORMAL, kAll), false});
  if (cc) {
    // Calibration grids without the opaque-disc eclipses.
    if (P.need_cal1) {
      grids.push_back({7, &P.star1f_cal, KIND_STAR1, mask_of(CLS_CAL1, {1u}), beam1});
      grids.push_back({8, &P.star1c_cal, KIND_STAR1, mask_of(CLS_CAL1, {2u, 3u}), beam1});
    }
    if (P.need_cal2) {
      grids.push_back({9, &P.star2f_cal, KIND_STAR2, mask_of(CLS_CAL2, {3u}), posn2});
      grids.push_back({10, &P.star2c_cal, KIND_STAR2, mask_of(CLS_CAL2, {1u, 2u}), posn2});
    }
  }
  return grids;
}

uint64_t lay_out_segments(SetPack& S, const std::vector<GridRef>& grids, std::vector<uint32_t>* n_real,
                          std::vector<uint32_t>* seg_ref) {
  SetHost& H = S.host;
  S.segs.clear();
  S.seg_grid.clear();
  n_real->clear();
  seg_ref->clear();
  uint64_t off = 0;
  uint32_t patches = 0;
  for (size_t gi = 0; gi < grids.size(); ++gi) {
    const GridRef& g = grids[gi];
    const std::vector<lc::Point>& pts = *g.pts;
    if (g.id < 7) H.grids[g.id].n = static_cast<uint32_t>(pts.size());
    if (pts.empty()) continue;
    uint32_t hist[MAX_ECL + 1] = {};
    for (const lc::Point& pt : pts) {
      if (pt.eclipse.size() > static_cast<size_t>(MAX_ECL))
        throw lc::LcError("element with " + std::to_string(pt.eclipse.size()) +
                          " eclipse intervals exceeds the kernel limit of " + std::to_string(MAX_ECL));
      ++hist[pt.eclipse.size()];
    }
    for (uint32_t k = 0; k <= static_cast<uint32_t>(MAX_ECL); ++k) {
      if (!hist[k]) continue;
      const uint32_t n_pad = (hist[k] + PATCH - 1) / PATCH * PATCH;
      off = (off + 15) / 16 * 16;
      SegDev sg{};
      sg.rec_off = off;
      sg.n = n_pad;
      sg.use_mask = g.use;
      sg.kind = g.kind;
      sg.necl = k;
      sg.flags = g.posn ? SEG_POSN : 0u;
      sg.patch_off = patches;
      off += static_cast<uint64_t>(rec_stride(k, g.posn)) * n_pad;
      patches += n_pad / PATCH;
      S.segs.push_back(sg);
      S.seg_grid.push_back(g.id);
      n_real->push_back(hist[k]);
      seg_ref->push_back(static_cast<uint32_t>(gi));
    }
  }
  return off;
}

std::unique_ptr<SetPack> pack_records(const lc::PreparedModel& P, bool keep_maps) {
  auto sp = std::make_unique<SetPack>();
  SetPack& S = *sp;
  set_scalars(S, P);
  SetDev& d = S.dev;
  const std::vector<GridRef> grids = grid_refs(P);
  std::vector<uint32_t> n_real, seg_ref;
  S.recs.assign(lay_out_segments(S, grids, &n_real, &seg_ref), 0);
  // Records, one segment per (grid, eclipse count): the segment's elements
  // in spatial order (stable), so consecutive 64-element patches have
  // compact normal cones, which lets the kernel skip back-facing patches.
  std::vector<Cone>& cones = S.cones;  // normal cone per patch
  for (size_t si = 0; si < S.segs.size(); ++si) {
    const SegDev& sg = S.segs[si];
    const GridRef& g = grids[seg_ref[si]];
    const std::vector<lc::Point>& pts = *g.pts;
    const uint32_t k = sg.necl;
    std::vector<uint32_t> idx;
    for (uint32_t e = 0; e < pts.size(); ++e)
      if (pts[e].eclipse.size() == k) idx.push_back(e);
    std::vector<uint32_t> key(pts.size());
    for (uint32_t e : idx)
      key[e] = direction_key(static_cast<float>(pts[e].dirn.x), static_cast<float>(pts[e].dirn.y),
                             static_cast<float>(pts[e].dirn.z));
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
    const uint32_t nr = static_cast<uint32_t>(idx.size());
    if (nr != n_real[si] || cones.size() != sg.patch_off) throw lc::LcError("segment layout out of step");
    const uint32_t n_pad = sg.n;
    const uint32_t stride = rec_stride(k, g.posn);
    const uint32_t eoff = g.posn ? 32u : 16u;
    uint8_t* base = S.recs.data() + sg.rec_off;
    for (uint32_t i = 0; i < n_pad; ++i) {
      uint8_t* r = base + static_cast<size_t>(i) * stride;
      if (i >= nr) {
        // Padding: zero normal and flux contribute exactly 0; its
        // intervals can never contain a phase in [0, 1).
        std::memset(r, 0, eoff);
        for (uint32_t j = 0; j < k; ++j) {
          const double e3[3] = {2.0, -1.0, -2.0};
          std::memcpy(r + eoff + 24 * j, e3, 24);
        }
        continue;
      }
      const lc::Point& pt = pts[idx[i]];
      const float a4[4] = {static_cast<float>(pt.dirn.x), static_cast<float>(pt.dirn.y),
                           static_cast<float>(pt.dirn.z), pt.flux};
      std::memcpy(r, a4, 16);
      if (g.posn) {
        const float b4[4] = {static_cast<float>(pt.posn.x), static_cast<float>(pt.posn.y),
                             static_cast<float>(pt.posn.z), 0.0f};
        std::memcpy(r + 16, b4, 16);
      }
      for (uint32_t j = 0; j < k; ++j) {
        const double e3[3] = {pt.eclipse[j].first, pt.eclipse[j].second, pt.eclipse[j].second - 1.0};
        std::memcpy(r + eoff + 24 * j, e3, 24);
      }
    }
    for (uint32_t p0 = 0; p0 < n_pad; p0 += PATCH) {
      // The cone of the FP32 normals the records hold.
      const uint32_t pn = nr > p0 ? std::min(PATCH, nr - p0) : 0u;
      cones.push_back(patch_cone(base + static_cast<size_t>(p0) * stride, stride, pn));
    }
    if (keep_maps) {
      idx.resize(n_pad, UINT32_MAX);
      S.seg_orig.push_back(std::move(idx));
    }
  }

  // Direction moments of the patches of moment-capable segments.
  S.moms.assign(cones.size() * MOM_STRIDE, 0.0);
  for (const SegDev& sg : S.segs) {
    if (moment_degree(sg, d) <= 0) continue;
    const uint32_t stride = rec_stride(sg.necl, (sg.flags & SEG_POSN) != 0);
    for (uint32_t q = 0; q < sg.n / PATCH; ++q)
      patch_moments(S.recs.data() + sg.rec_off + static_cast<uint64_t>(q) * PATCH * stride, stride,
                    &S.moms[static_cast<size_t>(sg.patch_off + q) * MOM_STRIDE]);
  }
  return sp;
}

void pack_lane_tables(SetPack& S, const lc::PreparedModel& P, const lc::Data& data) {
  SetDev& d = S.dev;
  SetHost& H = S.host;
  S.lane_phi.clear();
  S.lane_meta.clear();
  // Lanes.
  lc::Ginterp gi;
  gi.phase1 = P.phase1;
  gi.phase2 = P.phase2;
  const size_t nt = P.phase.size();
  S.t_lane0.resize(nt);
  S.t_ndiv.resize(nt);
  for (size_t t = 0; t < nt; ++t) {
    S.t_lane0[t] = static_cast<uint32_t>(S.lane_phi.size());
    S.t_ndiv[t] = P.ndiv[t];
    for (int nd = 0; nd < P.ndiv[t]; ++nd) {
      double phi, wgt;
      lc::exposure_weight(P.phase[t], P.expose[t], nd, P.ndiv[t], &phi, &wgt);
      push_lane(S, phi, CLS_NORMAL, gi);
    }
  }
  // The calibration and white-dwarf lanes take a tile of their own, so that
  // the time lanes' tiles (and their work codes) do not depend on them.
  while (S.lane_phi.size() % TILE) {
    S.lane_phi.push_back(0.0);
    S.lane_meta.push_back(0u);
  }
  d.special0 = static_cast<uint32_t>(S.lane_phi.size());
  if (P.need_cal1) {
    push_lane(S, 0.9999999999 * P.phase1, CLS_CAL1, gi);
    push_lane(S, 1.0000000001 * P.phase1, CLS_CAL1, gi);
    push_lane(S, 1.0 - 0.9999999999 * P.phase1, CLS_CAL1, gi);
    push_lane(S, 1.0 - 1.0000000001 * P.phase1, CLS_CAL1, gi);
  }
  if (P.need_cal2) {
    push_lane(S, 1.0 - 1.0000000001 * P.phase2, CLS_CAL2, gi);
    push_lane(S, 1.0 - 0.9999999999 * P.phase2, CLS_CAL2, gi);
    push_lane(S, 1.0000000001 * P.phase2, CLS_CAL2, gi);
    push_lane(S, 0.9999999999 * P.phase2, CLS_CAL2, gi);
  }
  push_lane(S, 0.5, CLS_WD, gi);
  while (S.lane_phi.size() % TILE) {
    S.lane_phi.push_back(0.0);
    S.lane_meta.push_back(0u);
  }
  d.n_lanes = static_cast<uint32_t>(S.lane_phi.size());
  d.n_times = static_cast<uint32_t>(nt);

  // Earth vector per lane, FP32 as the kernel reads it: comp_light lanes
  // use the TWOPI*iangle/360 inclination, the single-star lanes
  // iangle.to_radians().
  S.lane_e.resize(S.lane_phi.size());
  for (size_t l = 0; l < S.lane_phi.size(); ++l) {
    const uint32_t m = S.lane_meta[l];
    // iscale time lanes follow the single-component functions, which use
    // iangle.to_radians() (flux.rs:121, 167, 211, 238, 264).
    const bool lcl = meta_class(m) == CLS_NORMAL && !P.iscale;
    const lc::Vec3 e = lc::set_earth(lcl ? P.cosi_lc : P.cosi_rad, lcl ? P.sini_lc : P.sini_rad, S.lane_phi[l]);
    S.lane_e[l] = make_float4(static_cast<float>(e.x), static_cast<float>(e.y), static_cast<float>(e.z), 0.0f);
  }

  // Element-visits in the reference's loop order (flux.rs:24-107 and the
  // single-star calls of orchestration.rs), used for throughput figures.
  const uint64_t n1f = P.star1f.size(), n1c = P.star1_alias ? n1f : P.star1c.size();
  const uint64_t n2f = P.star2f.size(), n2c = P.star2_alias ? n2f : P.star2c.size();
  const uint64_t nother = P.disc.size() + P.edge.size() + P.spot.size();
  uint64_t visits = 0;
  for (size_t l = 0; l < S.lane_meta.size(); ++l) {
    const uint32_t m = S.lane_meta[l];
    if (!(m & META_VALID)) continue;
    const uint32_t cls = meta_class(m), pt = meta_ptype(m);
    const uint64_t s1 = pt == 1 ? n1f : n1c, s2 = pt == 3 ? n2f : n2c;
    if (cls == CLS_NORMAL) visits += s1 + s2 + nother;
    else if (cls == CLS_CAL2) visits += s2;
    else visits += s1;
  }
  H.visits = visits;
  H.n_times = static_cast<uint32_t>(nt);

  // Per-time terms and observations.
  S.t_slfac = P.slfac;
  S.t_phase = P.phase;
  S.t_y.resize(nt);
  S.t_ferr.resize(nt);
  S.t_w.resize(nt);
  for (size_t t = 0; t < nt; ++t) {
    S.t_y[t] = data[t].flux;
    S.t_ferr[t] = data[t].ferr;
    S.t_w[t] = data[t].weight;
  }
  }

void pack_codes(SetPack& S, const SetPack& R, const lc::PreparedModel& P, bool elem_check) {
  SetDev& d = S.dev;
  const std::vector<Cone>& cones = R.cones;
  if (R.segs.size() != S.segs.size() || R.cones.size() * MOM_STRIDE != R.moms.size())
    throw lc::LcError("records and segments out of step");
  if (S.lane_e.size() != S.lane_phi.size()) throw lc::LcError("lane tables out of step");
  S.lsplit.clear();
  // Per (patch, tile) work codes. For each tile: its valid lanes' earth
  // vectors (the FP32 vectors the kernel reads: comp_light lanes use the
  // TWOPI*iangle/360 inclination, the single-star lanes iangle.to_radians()),
  // normalised, and the range of their visibility phases.
  const uint32_t ntile = d.n_lanes / TILE;
  struct TileInfo {
    bool any = false;
    double lo = 0.0, hi = 0.0;  // min / max of phi - floor(phi)
    std::vector<float> ex, ey, ez;  // unit lines of sight of the valid lanes, FP32
  };
  std::vector<TileInfo> tiles(ntile);
  for (uint32_t t = 0; t < ntile; ++t) {
    TileInfo& ti = tiles[t];
    for (uint32_t l = t * TILE; l < (t + 1) * TILE; ++l) {
      const uint32_t m = S.lane_meta[l];
      if (!(m & META_VALID)) continue;
      const float4 e = S.lane_e[l];
      lc::Vec3 v(e.x, e.y, e.z);
      const double len = v.length();
      if (len > 0.0) v = (1.0 / len) * v;
      ti.ex.push_back(static_cast<float>(v.x));
      ti.ey.push_back(static_cast<float>(v.y));
      ti.ez.push_back(static_cast<float>(v.z));
      const double phif = S.lane_phi[l] - std::floor(S.lane_phi[l]);
      ti.lo = ti.any ? std::min(ti.lo, phif) : phif;
      ti.hi = ti.any ? std::max(ti.hi, phif) : phif;
      ti.any = true;
    }
  }
  S.n_tiles = ntile;
  // Unit line of sight per lane rounded to FP32 (the vectors of the code
  // decision), for the per-lane split.
  std::vector<float> lex(S.lane_e.size()), ley(S.lane_e.size()), lez(S.lane_e.size());
  for (size_t l = 0; l < S.lane_e.size(); ++l) {
    lc::Vec3 v(S.lane_e[l].x, S.lane_e[l].y, S.lane_e[l].z);
    const double len = v.length();
    if (len > 0.0) v = (1.0 / len) * v;
    lex[l] = static_cast<float>(v.x);
    ley[l] = static_cast<float>(v.y);
    lez[l] = static_cast<float>(v.z);
  }
  S.ptcode.assign(cones.size() * ntile, PT_FULL);
  for (size_t g = 0; g < S.segs.size(); ++g) {
    const SegDev& sg = S.segs[g];
    const uint32_t stride = rec_stride(sg.necl, (sg.flags & SEG_POSN) != 0);
    const uint32_t eoff = (sg.flags & SEG_POSN) ? 32u : 16u;
    const SegCodeRule rule = seg_code_rule(sg, d, P);
    const double thr = rule.thr, thr_front = rule.thr_front;
    const bool cull = rule.cull, front_ok = rule.front_ok;
    const CodeConsts& K = rule.K;
    // Per-lane version of the skip / seen-whole tests for a PT_FREE pair:
    // every element of the patch is free of eclipses at every lane of the
    // tile, so a lane whose line of sight has cosine above cf with the
    // patch axis sees it whole and one below cg sees none of it (the FP32
    // cosine of the code decision; the cone constants' margin covers its
    // rounding).
    auto split_lanes = [&](uint32_t patch, uint32_t t, const lc::Vec3& pa, double plen, float cosw) {
      if (!front_ok || !(plen > 0.0)) return;
      const bool can_all = cosw > K.c_front, can_none = cull && cosw > K.c_skip;
      const double cf = can_all ? cos_front(K, cosw) : 2.0, cg = can_none ? cos_skip(K, cosw) : -2.0;
      uint64_t all = 0, some = 0;
      const float fx = static_cast<float>(pa.x), fy = static_cast<float>(pa.y), fz = static_cast<float>(pa.z);
      for (uint32_t j = 0; j < TILE; ++j) {
        const uint32_t l = t * TILE + j;
        const uint32_t m = S.lane_meta[l];
        if (!(m & META_VALID) || !(sg.use_mask & cp_bit(meta_class(m), meta_ptype(m)))) continue;
        const double c = los_cosine(lex[l], ley[l], lez[l], fx, fy, fz);
        if (c > cf) all |= 1ull << j;
        else if (!(c < cg)) some |= 1ull << j;
      }
      // The cone leaves a band of lanes undecided; with the element-level
      // check, decide them from the elements themselves: mu of every
      // element against the cut-offs with a 1e-6 margin, which covers the
      // FP32 rounding of mu here and in the kernel (a few 1e-7). Without it
      // they stay limb lanes (evaluated element by element).
      if (some && elem_check) {
        float lx[TILE], ly[TILE], lz[TILE], mn[TILE], mx[TILE];
        uint32_t js[TILE];
        uint32_t n = 0;
        for (uint64_t b = some; b; b &= b - 1ull) {
          const uint32_t j = static_cast<uint32_t>(__builtin_ctzll(b));
          const float4 e = S.lane_e[t * TILE + j];
          js[n] = j;
          lx[n] = e.x;
          ly[n] = e.y;
          lz[n] = e.z;
          mn[n] = 4.0f;
          mx[n] = -4.0f;
          ++n;
        }
        const uint8_t* pr = R.recs.data() + sg.rec_off + static_cast<uint64_t>(patch - sg.patch_off) * PATCH * stride;
        for (uint32_t i = 0; i < PATCH; ++i) {
          float f[4];
          std::memcpy(f, pr + static_cast<size_t>(i) * stride, sizeof f);
          if (f[3] == 0.0f) continue;  // zero flux adds 0 either way
          for (uint32_t k = 0; k < n; ++k) {
            const float mu = lx[k] * f[0] + ly[k] * f[1] + lz[k] * f[2];
            mn[k] = std::min(mn[k], mu);
            mx[k] = std::max(mx[k], mu);
          }
        }
        const float whole = static_cast<float>(thr_front + 1e-6), none = static_cast<float>(thr - 1e-6);
        for (uint32_t k = 0; k < n; ++k) {
          const uint64_t bit = 1ull << js[k];
          if (mn[k] > whole) {
            all |= bit;
            some &= ~bit;
          } else if (mx[k] < none) {
            some &= ~bit;
          }
        }
      }
      S.lsplit.push_back({patch * ntile + t, all, some});
    };
    for (uint32_t q = 0; q < sg.n / PATCH; ++q) {
      const Cone& pc = cones[sg.patch_off + q];
      lc::Vec3 pa;
      double plen = 0.0;
      cone_axis(pc, &pa.x, &pa.y, &pa.z, &plen);
      const uint8_t* pr = R.recs.data() + sg.rec_off + static_cast<uint64_t>(q) * PATCH * stride;
      for (uint32_t t = 0; t < ntile; ++t) {
        const TileInfo& ti = tiles[t];
        uint8_t& code = S.ptcode[static_cast<size_t>(sg.patch_off + q) * ntile + t];
        if (!ti.any) {
          code = PT_SKIP;
          continue;
        }
        // Cosines between the patch axis and the tile's lines of sight span
        // [emin, emax]; a normal of the patch has cosine >= cosw with the
        // axis.
        float emin = 1.0f, emax = -1.0f;
        const float fx = static_cast<float>(pa.x), fy = static_cast<float>(pa.y), fz = static_cast<float>(pa.z);
        for (size_t k = 0; k < ti.ex.size(); ++k) {
          const float c = los_cosine(ti.ex[k], ti.ey[k], ti.ez[k], fx, fy, fz);
          emin = c < emin ? c : emin;
          emax = c > emax ? c : emax;
        }
        // Back-facing: every normal is more than acos(thr) + margin from
        // every line of sight, so every FP32 mu <= thr. The margin (2e-3
        // rad) covers FP32 rounding of mu and of the patch cone.
        if (cull && plen > 0.0 && pc.cosw > K.c_skip && emax < cos_skip(K, pc.cosw)) {
          code = PT_SKIP;
          continue;
        }
        // Seen whole: every normal is less than acos(thr_front) - margin
        // from every line of sight.
        const bool front = front_ok && plen > 0.0 && pc.cosw > K.c_front && emin > cos_front(K, pc.cosw);
        if (sg.necl == 0) {
          code = front ? PT_MOM : PT_FREE;
          if (code == PT_FREE) split_lanes(sg.patch_off + q, t, pa, plen, pc.cosw);
          continue;
        }
        // Point::visible (types.rs:50-58) over the tile's phase range
        // [lo, hi], with the same FP64 comparisons: no lane eclipsed when
        // every interval misses the range; every lane eclipsed when some
        // interval holds it.
        bool all_vis = true, all_ecl = true;
        for (uint32_t i = 0; i < PATCH; ++i) {
          const uint8_t* r = pr + static_cast<size_t>(i) * stride;
          float w;
          std::memcpy(&w, r + 12, sizeof w);
          if (w == 0.0f) continue;  // zero flux (padding) adds 0 either way
          const uint8_t* e = r + eoff;
          bool vis = true, ecl = false;
          for (uint32_t k = 0; k < sg.necl; ++k) {
            double e3[3];
            std::memcpy(e3, e + 24 * k, sizeof e3);
            bool misses, holds;
            interval_vs_range(e3, ti.lo, ti.hi, &misses, &holds);
            vis = vis && misses;
            ecl = ecl || holds;
          }
          all_vis = all_vis && vis;
          all_ecl = all_ecl && ecl;
        }
        code = all_ecl ? PT_SKIP : (all_vis ? (front ? PT_MOM : PT_FREE) : PT_FULL);
        if (code == PT_FREE) split_lanes(sg.patch_off + q, t, pa, plen, pc.cosw);
      }
    }
  }

}

void pack_lanes(SetPack& S, const SetPack& R, const lc::PreparedModel& P, const lc::Data& data, bool elem_check) {
  pack_lane_tables(S, P, data);
  pack_codes(S, R, P, elem_check);
}

std::unique_ptr<SetPack> pack_set(const lc::PreparedModel& P, const lc::Data& data, bool keep_maps,
                                  bool elem_check) {
  auto sp = pack_records(P, keep_maps);
  pack_lanes(*sp, *sp, P, data, elem_check);
  return sp;
}

void scale_linear_records(SetPack& S) {
  const SetDev& d = S.dev;
  if (!(d.flags & SET_VALID)) return;
  for (SegDev& sg : S.segs) {
    if (!linear_law_segment(sg, d)) continue;
    const uint32_t stride = rec_stride(sg.necl, false);
    uint8_t* base = S.recs.data() + sg.rec_off;
    if (sg.rec_off + static_cast<uint64_t>(stride) * sg.n > S.recs.size()) throw lc::LcError("segment records out of range");
    float wmax = 0.0f;
    bool ok = true;
    for (uint32_t i = 0; i < sg.n && ok; ++i) {
      float f[4];
      std::memcpy(f, base + static_cast<size_t>(i) * stride, sizeof f);
      ok = f[3] >= 0.0f && std::isfinite(f[3]);
      if (ok) wmax = std::max(wmax, f[3]);
    }
    if (!ok || wmax == 0.0f) continue;
    int k = 0;
    std::frexp(std::sqrt(static_cast<double>(wmax)), &k);  // sqrt(wmax) < 2^k
    if (k < -60 || k > 60) continue;                       // keep 4^k a normal float
    const double inv = std::ldexp(1.0, -k);
    for (uint32_t i = 0; i < sg.n; ++i) {
      uint8_t* r = base + static_cast<size_t>(i) * stride;
      float f[4];
      std::memcpy(f, r, sizeof f);
      const double sc = std::sqrt(static_cast<double>(f[3])) * inv;  // in [0, 1)
      const float g[4] = {static_cast<float>(f[0] * sc), static_cast<float>(f[1] * sc),
                          static_cast<float>(f[2] * sc), static_cast<float>(sc)};
      std::memcpy(r, g, sizeof g);
    }
    sg.flags |= SEG_LINW | (static_cast<uint32_t>(k + 128) << 8);
  }
}

std::vector<std::unique_ptr<SetPack>> build_packs(const std::vector<lc::Model>& models,
                                                  const std::vector<lc::Data>& data, int nthreads,
                                                  bool keep_maps, bool elem_check) {
  const size_t ns = models.size();
  std::vector<std::unique_ptr<SetPack>> packs(ns);
  if (data.empty()) throw lc::LcError("no data");
  auto data_of = [&](size_t s) -> const lc::Data& { return data.size() == 1 ? data[0] : data[s]; };
  // Default: all cores, or LCURVE_HIP_THREADS when set (the CPU share a
  // process may use on a shared node).
  int nt = nthreads;
  if (nt <= 0) {
    const char* env = std::getenv("LCURVE_HIP_THREADS");
    nt = env ? std::atoi(env) : 0;
  }
  if (nt <= 0) nt = static_cast<int>(std::thread::hardware_concurrency());
  if (nt < 1) nt = 1;
  nt = static_cast<int>(std::min<size_t>(static_cast<size_t>(nt), std::max<size_t>(ns, 1)));
  std::atomic<size_t> next{0};
  std::exception_ptr fatal;
  std::mutex mu;
  auto worker = [&]() {
    for (;;) {
      size_t s = next.fetch_add(1);
      if (s >= ns) break;
      try {
        try {
          auto P = lc::prepare_model(models[s], data_of(s));
          packs[s] = pack_set(*P, data_of(s), keep_maps, elem_check);
          scale_linear_records(*packs[s]);
        } catch (const lc::LcError& e) {
          packs[s] = invalid_set(e.what(), data_of(s));
        }
      } catch (...) {
        std::lock_guard<std::mutex> lk(mu);
        if (!fatal) fatal = std::current_exception();
        next.store(ns);
      }
    }
  };
  if (nt == 1) {
    worker();
  } else {
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i) th.emplace_back(worker);
    for (auto& t : th) t.join();
  }
  if (fatal) std::rethrow_exception(fatal);
  return packs;
}

}  // namespace lch
