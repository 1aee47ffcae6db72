// Resident caller grids and their flux evaluation (grid_model.h).
#include "grid_model.h"

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace lch {

namespace {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

// One grid of Points from its arrays.
std::vector<lc::Point> points_of(const lch_grid_soa& g, int id) {
  const std::string name = "grid " + std::to_string(id);
  if (g.n < 0) throw std::invalid_argument(name + ": negative element count");
  std::vector<lc::Point> pts;
  if (g.n == 0) return pts;
  if (!g.posn || !g.dirn || !g.flux) throw std::invalid_argument(name + ": null posn, dirn or flux");
  if (g.ecl_off) {
    if (g.ecl_off[0] != 0) throw std::invalid_argument(name + ": ecl_off[0] must be 0");
    if (g.ecl_off[g.n] > 0 && (!g.ingress || !g.egress)) throw std::invalid_argument(name + ": null eclipse arrays");
  }
  pts.resize(static_cast<size_t>(g.n));
  for (long long i = 0; i < g.n; ++i) {
    lc::Point& p = pts[static_cast<size_t>(i)];
    p.posn = lc::Vec3(g.posn[3 * i], g.posn[3 * i + 1], g.posn[3 * i + 2]);
    p.dirn = lc::Vec3(g.dirn[3 * i], g.dirn[3 * i + 1], g.dirn[3 * i + 2]);
    p.area = g.area ? g.area[i] : 0.0f;
    p.gravity = g.gravity ? g.gravity[i] : 1.0f;
    p.flux = g.flux[i];
    if (g.ecl_off) {
      const long long a = g.ecl_off[i], b = g.ecl_off[i + 1];
      if (b < a) throw std::invalid_argument(name + ": ecl_off must not decrease");
      if (b - a > MAX_ECL)
        throw std::invalid_argument(name + ": element " + std::to_string(i) + " has " + std::to_string(b - a) +
                                    " eclipse intervals; at most " + std::to_string(MAX_ECL) + " are supported");
      for (long long k = a; k < b; ++k) p.eclipse.emplace_back(g.ingress[k], g.egress[k]);
    }
  }
  return pts;
}

}  // namespace

std::unique_ptr<lc::PreparedModel> prepared_from_grids(const lch_grid_soa* grids, const lch_flux_params& fp) {
  if (!grids) throw std::invalid_argument("null grids");
  if (fp.limb1 < 0 || fp.limb1 > 1 || fp.limb2 < 0 || fp.limb2 > 1)
    throw std::invalid_argument("limb law must be 0 (Poly) or 1 (Claret)");
  if (fp.gint_mode != 0 && fp.gint_mode != 1) throw std::invalid_argument("gint_mode must be 0 or 1");
  if (!(fp.q > 0.0)) throw std::invalid_argument("q must be > 0");
  auto pm = std::make_unique<lc::PreparedModel>();
  lc::PreparedModel& P = *pm;
  std::vector<lc::Point> g[LCH_NGRID];
  for (int i = 0; i < LCH_NGRID; ++i) g[i] = points_of(grids[i], i);
  P.star1f = std::move(g[LCH_GRID_STAR1F]);
  P.star2f = std::move(g[LCH_GRID_STAR2F]);
  // An empty coarse grid is the fine grid, as the reference clones it.
  P.star1_alias = g[LCH_GRID_STAR1C].empty();
  P.star2_alias = g[LCH_GRID_STAR2C].empty();
  if (!P.star1_alias) P.star1c = std::move(g[LCH_GRID_STAR1C]);
  if (!P.star2_alias) P.star2c = std::move(g[LCH_GRID_STAR2C]);
  P.disc = std::move(g[LCH_GRID_DISC]);
  P.edge = std::move(g[LCH_GRID_EDGE]);
  P.spot = std::move(g[LCH_GRID_SPOT]);
  // Ginterp: the caller's factors, or calibration lanes on the GPU
  // (orchestration.rs:85-131) for each star with a separate coarse grid.
  P.need_cal1 = fp.gint_mode == 1 && !P.star1_alias;
  P.need_cal2 = fp.gint_mode == 1 && !P.star2_alias;
  const bool any_cal = !g[7].empty() || !g[8].empty() || !g[9].empty() || !g[10].empty();
  if (any_cal) {
    if (fp.gint_mode != 1) throw std::invalid_argument("calibration grids are read only with gint_mode 1");
    if ((P.need_cal1 && (g[7].empty() || g[8].empty())) || (P.need_cal2 && (g[9].empty() || g[10].empty())))
      throw std::invalid_argument("calibration grids must come in fine / coarse pairs for each calibrated star");
    P.cal_copies = true;
    P.star1f_cal = std::move(g[LCH_GRID_STAR1F_CAL]);
    P.star1c_cal = std::move(g[LCH_GRID_STAR1C_CAL]);
    P.star2f_cal = std::move(g[LCH_GRID_STAR2F_CAL]);
    P.star2c_cal = std::move(g[LCH_GRID_STAR2C_CAL]);
  }
  auto ldc = [](int law, const double* a, double mucrit) {
    lc::LDC l;
    l.ltype = law == 1 ? lc::LdcType::Claret : lc::LdcType::Poly;
    l.ldc1 = a[0];
    l.ldc2 = a[1];
    l.ldc3 = a[2];
    l.ldc4 = a[3];
    l.mucrit = mucrit;
    return l;
  };
  P.ldc1 = ldc(fp.limb1, fp.ldc1, fp.mucrit1);
  P.ldc2 = ldc(fp.limb2, fp.ldc2, fp.mucrit2);
  P.lin_limb_disc = fp.lin_limb_disc;
  P.quad_limb_disc = fp.quad_limb_disc;
  P.beam_factor1 = fp.beam_factor1;
  P.beam_factor2 = fp.beam_factor2;
  P.spin1 = fp.spin1;
  P.spin2 = fp.spin2;
  // flux.rs:19-22: q / (1 + q); vscale is an f32 argument of comp_light.
  P.xcofm = fp.q / (1.0 + fp.q);
  P.vfac = static_cast<double>(static_cast<float>(fp.velocity_scale)) / (lc::C_LIGHT / 1e3);
  P.glens1 = fp.glens1 != 0;
  P.rlens1 = fp.rlens1;
  P.phase1 = fp.phase1;
  P.phase2 = fp.phase2;
  P.iangle = fp.iangle;
  P.cosi_lc = std::cos(lc::TWOPI * fp.iangle / 360.0);
  P.sini_lc = std::sin(lc::TWOPI * fp.iangle / 360.0);
  const double ri = lc::to_radians(fp.iangle);
  P.cosi_rad = std::cos(ri);
  P.sini_rad = std::sin(ri);
  P.iscale = false;
  P.ncol = 5;
  return pm;
}

GridModel::GridModel(Context& ctx, const lch_grid_soa* grids, const lch_flux_params& fp) : ctx_(ctx) {
  P_ = prepared_from_grids(grids, fp);
  rec_ = pack_records(*P_, false);
  if (fp.gint_mode == 0) {
    const double s[4] = {fp.scale11, fp.scale12, fp.scale21, fp.scale22};
    for (int i = 0; i < 4; ++i) rec_->dev.gscl[i] = s[i];
  }
  // Device copy of the records with the linear-law rescaling; the host
  // keeps the plain ones, from which every call's work codes are made.
  SetPack dev;
  dev.dev = rec_->dev;
  dev.segs = rec_->segs;
  dev.recs = rec_->recs;
  scale_linear_records(dev);
  segs_dev_ = dev.segs;
  rec_bytes_ = dev.recs.size();
  hip_check(hipSetDevice(ctx_.device()), "hipSetDevice");
  hip_check(hipMalloc(reinterpret_cast<void**>(&d_recs_), std::max<uint64_t>(rec_bytes_, 16)), "hipMalloc records");
  if (rec_bytes_) {
    hipError_t e = hipMemcpy(d_recs_, dev.recs.data(), rec_bytes_, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(d_recs_);
      d_recs_ = nullptr;
      hip_check(e, "hipMemcpy records");
    }
  }
}

GridModel::~GridModel() {
  cache_.reset();
  if (d_recs_) {
    (void)hipSetDevice(ctx_.device());
    (void)hipDeviceSynchronize();  // before the free, as in Batch::release_all
    (void)hipFree(d_recs_);
  }
}

std::unique_ptr<SetPack> GridModel::lanes_pack(const double* phase, const double* expose, const int* ndiv,
                                               const double* slfac, double third, int n) {
  lc::PreparedModel& P = *P_;
  P.phase.assign(phase, phase + n);
  P.expose.assign(expose, expose + n);
  P.ndiv.assign(ndiv, ndiv + n);
  P.slfac.assign(slfac, slfac + n);
  P.third = third;
  lc::Data data(static_cast<size_t>(n));  // no observations: weight 0
  for (int i = 0; i < n; ++i) {
    data[i].ferr = 1.0;
    data[i].ndiv = ndiv[i];
  }
  auto S = std::make_unique<SetPack>();
  S->dev = rec_->dev;
  S->dev.third = third;
  S->host = rec_->host;
  S->segs = rec_->segs;
  S->seg_grid = rec_->seg_grid;
  S->moms = rec_->moms;
  pack_lanes(*S, *rec_, P, data);
  S->segs = segs_dev_;
  return S;
}

void GridModel::eval(const double* phase, const double* expose, const int* ndiv, const double* slfac, double third,
                     double* out, int n, const BatchMemory* arena) {
  if (n <= 0) throw std::invalid_argument("n must be > 0");
  if (!phase || !expose || !ndiv || !slfac || !out) throw std::invalid_argument("null argument");
  for (int i = 0; i < n; ++i)
    if (ndiv[i] > 1 << 16) throw std::invalid_argument("ndiv too large");
  if (arena) {
    BatchMemory mem = *arena;
    mem.recs = d_recs_;
    mem.rec_bytes = rec_bytes_;
    std::vector<std::unique_ptr<SetPack>> packs;
    packs.push_back(lanes_pack(phase, expose, ndiv, slfac, third, n));
    Batch b(ctx_, std::move(packs), static_cast<size_t>(n), &mem);
    last_arena_ = b.arena_used();
    b.run(0, nullptr, nullptr, nullptr, nullptr, out);
    return;
  }
  const bool same = cache_ && key_phase_.size() == static_cast<size_t>(n) &&
                    std::memcmp(key_phase_.data(), phase, n * sizeof(double)) == 0 &&
                    std::memcmp(key_expose_.data(), expose, n * sizeof(double)) == 0 &&
                    std::memcmp(key_ndiv_.data(), ndiv, n * sizeof(int)) == 0;
  if (!same) {
    cache_.reset();
    std::vector<std::unique_ptr<SetPack>> packs;
    packs.push_back(lanes_pack(phase, expose, ndiv, slfac, third, n));
    BatchMemory mem;
    mem.recs = d_recs_;
    mem.rec_bytes = rec_bytes_;
    cache_ = std::make_unique<Batch>(ctx_, std::move(packs), static_cast<size_t>(n), &mem);
    key_phase_.assign(phase, phase + n);
    key_expose_.assign(expose, expose + n);
    key_ndiv_.assign(ndiv, ndiv + n);
  } else {
    cache_->set_curve_terms(slfac, third);
  }
  cache_->run(0, nullptr, nullptr, nullptr, nullptr, out);
}

}  // namespace lch
