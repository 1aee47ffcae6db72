// CPU setup for one model: crates/lcurve/src/orchestration.rs:36-179 and the
// per-time phase / polynomial-factor terms of orchestration.rs:172-241.
#include "prepare.h"

namespace lc {

size_t PreparedModel::n_elements() const {
  return star1f.size() + star1c.size() + star2f.size() + star2c.size() + disc.size() +
         edge.size() + spot.size();
}

// crates/lcurve/src/orchestration.rs:231-237
double time_to_phase(const Model& mdl, double time) {
  double phase = (time - mdl.t0.value) / mdl.period.value;
  for (int i = 0; i < 4; ++i) {
    phase -= (mdl.t0.value + phase * (mdl.period.value + mdl.pdot.value * phase) - time) /
             (mdl.period.value + 2.0 * mdl.pdot.value * phase);
  }
  phase += mdl.deltat.value / mdl.period.value / 2.0 * (std::cos(2.0 * PI * phase) - 1.0);
  return phase;
}

std::unique_ptr<PreparedModel> prepare_model(const Model& mdl, const Data& data) {
  if (data.empty()) throw LcError("no data points");
  auto pm = std::make_unique<PreparedModel>();
  PreparedModel& P = *pm;

  // orchestration.rs:42-46
  double r1, r2;
  mdl.get_r1r2(&r1, &r2);
  double rl2 = 1.0 - xl12(mdl.q.value, mdl.spin2.value);
  if (r2 < 0.0) r2 = rl2;
  else if (r2 > rl2) throw LcError("light_curve_comp: secondary larger than Roche lobe");

  P.ldc1 = mdl.get_ldc1();
  P.ldc2 = mdl.get_ldc2();

  // orchestration.rs:52-58
  P.glens1 = mdl.glens1;
  if (mdl.glens1) {
    double gm = powi3(1000.0 * mdl.velocity_scale.value) * mdl.tperiod * DAY / TWOPI;
    double a = std::pow(gm / sqr(TWOPI / DAY / mdl.tperiod), 1.0 / 3.0);
    P.rlens1 = 4.0 * gm / (1.0 + mdl.q.value) / a / sqr(C_LIGHT);
  }

  // orchestration.rs:61-83
  P.star1f = set_star_grid(mdl, Star::Primary, true);
  P.star2f = set_star_grid(mdl, Star::Secondary, true);
  set_star_continuum(mdl, P.star1f, P.star2f);

  P.star1_alias = mdl.nlat1f == mdl.nlat1c;
  if (!P.star1_alias) P.star1c = set_star_grid(mdl, Star::Primary, false);

  bool copy2 = mdl.nlat2f == mdl.nlat2c &&
               (!mdl.npole || r1 >= r2 || (mdl.nlatfill == 0 && mdl.nlngfill == 0));
  P.star2_alias = copy2;
  if (!copy2) P.star2c = set_star_grid(mdl, Star::Secondary, false);

  if (mdl.nlat1c != mdl.nlat1f || !copy2) {
    // The reference recomputes brightness on the coarse pair, one of which
    // may be a clone of its fine grid; a clone would receive the values it
    // already has, and the two stars are processed independently, so only
    // the separate coarse grids are passed.
    std::vector<Point> none1, none2;
    set_star_continuum(mdl, P.star1_alias ? none1 : P.star1c, P.star2_alias ? none2 : P.star2c);
  }

  // Ginterp phases; scale factors are found on the GPU (orchestration.rs:86-131).
  P.phase1 = mdl.phase1;
  P.phase2 = mdl.phase2;
  P.need_cal1 = mdl.nlat1c != mdl.nlat1f;
  P.need_cal2 = !copy2;

  // orchestration.rs:134-170
  if (mdl.add_disc) {
    P.disc = set_disc_grid(mdl);
    P.edge = set_disc_edge(mdl, true);
    double rdisc1 = mdl.rdisc1.value > 0.0 ? mdl.rdisc1.value : r1;
    double rdisc2 = mdl.rdisc2.value > 0.0 ? mdl.rdisc2.value : mdl.radius_spot.value;
    if (mdl.opaque) {
      if (P.need_cal1 || P.need_cal2) {
        P.cal_copies = true;
        P.star1f_cal = P.star1f;
        P.star1c_cal = P.star1c;
        P.star2f_cal = P.star2f;
        P.star2c_cal = P.star2c;
      }
      for (std::vector<Point>* g : {&P.star1f, &P.star1c, &P.star2f, &P.star2c}) {
        for (Point& pt : *g) {
          EclipseVec e = disc_eclipse(mdl.iangle.value, rdisc1, rdisc2, mdl.beta_disc.value,
                                      mdl.height_disc.value, pt.posn);
          for (const auto& pr : e) pt.eclipse.push_back(pr);
        }
      }
    }
    set_disc_continuum(rdisc2, mdl.temp_disc.value, mdl.texp_disc.value, mdl.wavelength, P.disc);
    set_edge_continuum(mdl.temp_edge.value, r2, std::fabs(mdl.t2.value), mdl.absorb_edge.value,
                       mdl.wavelength, P.edge);
  }
  if (mdl.add_spot) P.spot = set_bright_spot_grid(mdl);

  // Flux-path scalars (flux.rs:7-22, orchestration.rs:243-251).
  P.lin_limb_disc = mdl.lin_limb_disc.value;
  P.quad_limb_disc = mdl.quad_limb_disc.value;
  P.beam_factor1 = mdl.beam_factor1.value;
  P.beam_factor2 = mdl.beam_factor2.value;
  P.spin1 = mdl.spin1.value;
  P.spin2 = mdl.spin2.value;
  P.xcofm = mdl.q.value / (1.0 + mdl.q.value);
  P.vfac = static_cast<double>(static_cast<float>(mdl.velocity_scale.value)) / (C_LIGHT / 1e3);
  P.iangle = mdl.iangle.value;
  P.cosi_lc = std::cos(TWOPI * mdl.iangle.value / 360.0);
  P.sini_lc = std::sin(TWOPI * mdl.iangle.value / 360.0);
  double ri = to_radians(mdl.iangle.value);
  P.cosi_rad = std::cos(ri);
  P.sini_rad = std::sin(ri);
  P.third = mdl.third.value;
  P.iscale = mdl.iscale;
  P.ncol = mdl.t2.value > 0.0 ? 5 : 4;

  // Per-time terms (orchestration.rs:172-241).
  double xmin = data[0].time, xmax = data[0].time;
  for (size_t i = 1; i < data.size(); ++i) {
    if (data[i].time < xmin) xmin = data[i].time;
    if (data[i].time > xmax) xmax = data[i].time;
  }
  double middle = (xmin + xmax) / 2.0;
  double range = (xmax - xmin) / 2.0;
  size_t n = data.size();
  P.phase.resize(n);
  P.expose.resize(n);
  P.ndiv.resize(n);
  P.slfac.resize(n);
  for (size_t i = 0; i < n; ++i) {
    P.phase[i] = time_to_phase(mdl, data[i].time);
    P.expose[i] = data[i].expose / mdl.period.value;
    P.ndiv[i] = data[i].ndiv;
    double frac = (data[i].time - middle) / range;
    P.slfac[i] = 1.0 + frac * (mdl.slope.value + frac * (mdl.quad.value + frac * mdl.cube.value));
  }

  // Per-curve scalars (orchestration.rs:312-315).
  P.logg1 = comp_gravity1(mdl, P.star1f);
  P.logg2 = comp_gravity2(mdl, P.star2f);
  P.rv1 = comp_radius1(P.star1f);
  P.rv2 = comp_radius2(P.star2f);
  return pm;
}

}  // namespace lc
