// Grid generation and brightness, restated from crates/lcurve/src/grid.rs,
// crates/lcurve/src/brightness.rs and the per-curve scalar sums of
// crates/lcurve/src/flux.rs:282-375.
#include "grid.h"

#include <algorithm>

namespace lc {

namespace {

// crates/lcurve/src/grid.rs:49-73
void star_eclipse(double q, double r, double spin, double ffac, double iangle, const Vec3& posn,
                  double delta, bool roche, Star star, EclipseVec& eclipses) {
  double ri = to_radians(iangle);
  double cosi = std::cos(ri);
  double sini = std::sin(ri);
  Vec3 cofm = star == Star::Primary ? Vec3(0, 0, 0) : Vec3(1.0, 0.0, 0.0);
  if (roche) {
    double ing, eg;
    if (ingress_egress(q, star, spin, ffac, iangle, delta, posn, &ing, &eg))
      eclipses.push_back({ing, eg});
  } else {
    double phi1, phi2, lam1, lam2;
    if (sphere_eclipse_phase(cosi, sini, posn, cofm, r, &phi1, &phi2, &lam1, &lam2))
      eclipses.push_back({phi1, phi2});
  }
}

// crates/lcurve/src/grid.rs:324-333
void envelope(double rangle, double lambda, double r1, double* x, double* y) {
  double sini = std::sin(rangle);
  double cosi = std::cos(rangle);
  double sinl = std::sin(lambda);
  double cosl = std::cos(lambda);
  double norm = std::sqrt(cosi * cosi + sini * sini * cosl * cosl);
  *x = sinl + r1 * cosi * sinl / norm;
  *y = -cosi * cosl - r1 * cosl / norm;
}

struct StarCtx {
  const Model* mdl;
  Star which;
  double r1, r2, rref1, rref2, pref1, pref2, ffac1, ffac2;
  bool eclipse;
  double gref, acc;
};

// crates/lcurve/src/grid.rs:201-322
void add_faces(std::vector<Point>& star, double tlo, double thi, double dtheta, int nlatfill,
               int nlngfill, const StarCtx& c) {
  const Model& mdl = *c.mdl;
  const Vec3 cofm1(0, 0, 0), cofm2(1.0, 0.0, 0.0);
  bool infill = nlatfill > 0 || nlngfill > 0;
  int nlat, nlat1, nlat2;
  if (infill) {
    nlat1 = as_i32(std::ceil(static_cast<double>(1 + nlatfill) * (thi - tlo) / dtheta));
    nlat2 = as_i32(std::ceil((thi - tlo) / dtheta));
    nlat = nlat1 + nlat2;
  } else {
    nlat1 = 0;
    nlat2 = 0;
    nlat = as_i32(std::ceil((thi - tlo) / dtheta));
  }

  for (int nt = 0; nt < nlat; ++nt) {
    double theta, phi1, phi2;
    int nphi, nl;
    if (infill) {
      if (nt < nlat1) {
        theta = tlo + (thi - tlo) * (static_cast<double>(nt) + 0.5) / static_cast<double>(nlat1);
        double sint = std::sin(theta);
        nphi = std::max(8, as_i32(PI * sint * static_cast<double>(1 + nlngfill) / dtheta));
        nl = nlat1;
        if (c.which == Star::Primary) { phi1 = -PI / 2.0; phi2 = PI / 2.0; }
        else { phi1 = PI / 2.0; phi2 = 3.0 * PI / 2.0; }
      } else {
        theta = tlo + (thi - tlo) * (static_cast<double>(nt - nlat1) + 0.5) /
                          static_cast<double>(nlat2);
        double sint = std::sin(theta);
        nphi = std::max(8, as_i32(PI * sint / dtheta));
        nl = nlat2;
        if (c.which == Star::Primary) { phi1 = PI / 2.0; phi2 = 3.0 * PI / 2.0; }
        else { phi1 = -PI / 2.0; phi2 = PI / 2.0; }
      }
    } else {
      theta = tlo + (thi - tlo) * (static_cast<double>(nt) + 0.5) / static_cast<double>(nlat);
      double sint = std::sin(theta);
      nphi = std::max(16, as_i32(TWOPI * sint / dtheta));
      nl = nlat;
      phi1 = 0.0;
      phi2 = TWOPI;
    }

    double sint = std::sin(theta);
    double cost = std::cos(theta);
    for (int np = 0; np < nphi; ++np) {
      double phi = phi1 + (phi2 - phi1) * (static_cast<double>(np) + 0.5) / static_cast<double>(nphi);
      double sinp = std::sin(phi);
      double cosp = std::cos(phi);
      Vec3 dirn = mdl.npole ? Vec3(sint * cosp, sint * sinp, cost) : Vec3(cost, sint * cosp, sint * sinp);

      Vec3 posn, dvec;
      double rad, gravity;
      if (c.which == Star::Primary && mdl.roche1) {
        face(mdl.q.value, Star::Primary, mdl.spin1.value, dirn, c.rref1, c.pref1, c.acc, &posn, &dvec,
             &rad, &gravity);
      } else if (c.which == Star::Secondary && mdl.roche2) {
        face(mdl.q.value, Star::Secondary, mdl.spin2.value, dirn, c.rref2, c.pref2, c.acc, &posn,
             &dvec, &rad, &gravity);
      } else if (c.which == Star::Primary) {
        rad = c.r1;
        posn = cofm1 + rad * dirn;
        dvec = dirn;
        gravity = 1.0;
      } else {
        rad = c.r2;
        posn = cofm2 + rad * dirn;
        dvec = dirn;
        gravity = 1.0;
      }

      double area = ((phi2 - phi1) / static_cast<double>(nphi) * rad * sint) *
                    ((thi - tlo) / static_cast<double>(nl) * rad) / dot(dirn, dvec);

      EclipseVec eclipses;
      if (c.eclipse) {
        if (c.which == Star::Primary)
          star_eclipse(mdl.q.value, c.r2, mdl.spin2.value, c.ffac2, mdl.iangle.value, posn,
                       mdl.delta_phase, mdl.roche2, Star::Secondary, eclipses);
        else
          star_eclipse(mdl.q.value, c.r1, mdl.spin1.value, c.ffac1, mdl.iangle.value, posn,
                       mdl.delta_phase, mdl.roche1, Star::Primary, eclipses);
      }
      star.emplace_back(posn, dvec, area, gravity / c.gref, std::move(eclipses));
    }
  }
}

// crates/lcurve/src/grid.rs:157-198
std::vector<Point> set_star_grid_inner(const StarCtx& c0, int nlat, bool infill, double thelo,
                                       double thehi, int nlatfill, int nlngfill) {
  StarCtx c = c0;
  const Model& mdl = *c.mdl;
  c.acc = mdl.delta_phase / 10.0;
  double dtheta = PI / static_cast<double>(nlat);
  if (c.which == Star::Primary && mdl.roche1) {
    Vec3 p, d; double r, g;
    face(mdl.q.value, Star::Primary, mdl.spin1.value, Vec3(-1.0, 0.0, 0.0), c.rref1, c.pref1, c.acc,
         &p, &d, &r, &g);
    c.gref = g;
  } else if (c.which == Star::Secondary && mdl.roche2) {
    Vec3 p, d; double r, g;
    face(mdl.q.value, Star::Secondary, mdl.spin2.value, Vec3(1.0, 0.0, 0.0), c.rref2, c.pref2,
         c.acc, &p, &d, &r, &g);
    c.gref = g;
  } else {
    c.gref = 1.0;
  }
  std::vector<Point> star;
  if (infill) {
    add_faces(star, 0.0, thelo, dtheta, 0, 0, c);
    add_faces(star, thelo, thehi, dtheta, nlatfill, nlngfill, c);
    add_faces(star, thehi, PI, dtheta, 0, 0, c);
  } else {
    add_faces(star, 0.0, PI, dtheta, 0, 0, c);
  }
  return star;
}

}  // namespace

// crates/lcurve/src/grid.rs:76-155
std::vector<Point> set_star_grid(const Model& mdl, Star which, bool fine) {
  double r1, r2;
  mdl.get_r1r2(&r1, &r2);
  bool eclipse = which == Star::Primary ? mdl.eclipse1 : mdl.eclipse2;
  int nlat = which == Star::Primary ? (fine ? mdl.nlat1f : mdl.nlat1c)
                                    : (fine ? mdl.nlat2f : mdl.nlat2c);
  int nlatfill = which == Star::Primary ? 0 : (fine ? mdl.nlatfill : 0);
  int nlngfill = which == Star::Primary ? 0 : (fine ? mdl.nlngfill : 0);

  double rl1 = xl11(mdl.q.value, mdl.spin1.value);
  if (r1 < 0.0) r1 = rl1;
  else if (r1 > rl1) throw LcError("set_star_grid: primary star larger than Roche lobe");
  double rl2 = 1.0 - xl12(mdl.q.value, mdl.spin2.value);
  if (r2 < 0.0) r2 = rl2;
  else if (r2 > rl2) throw LcError("set_star_grid: secondary star larger than Roche lobe");

  StarCtx c{};
  c.mdl = &mdl;
  c.which = which;
  c.r1 = r1;
  c.r2 = r2;
  c.ffac1 = r1 / rl1;
  ref_sphere(mdl.q.value, Star::Primary, mdl.spin1.value, c.ffac1, &c.rref1, &c.pref1);
  c.ffac2 = r2 / rl2;
  ref_sphere(mdl.q.value, Star::Secondary, mdl.spin2.value, c.ffac2, &c.rref2, &c.pref2);
  c.eclipse = eclipse;

  bool infill = mdl.npole && which == Star::Secondary && (nlatfill > 0 || nlngfill > 0) && r2 > r1;
  double thelo = 0.0, thehi = 0.0;
  if (infill) {
    double rangle = to_radians(mdl.iangle.value);
    double cosi = std::cos(rangle);
    double ratio = (cosi + r1) / r2;
    if (ratio >= 1.0) {
      thehi = PI / 2.0 + rangle;
    } else {
      double llo_val = 0.0;
      double lhi_val = PI / 2.0;
      while (lhi_val > llo_val + 1e-7) {
        double lmid = (llo_val + lhi_val) / 2.0;
        double x, y;
        envelope(rangle, lmid, r1, &x, &y);
        if (x * x + y * y < r2 * r2) llo_val = lmid;
        else lhi_val = lmid;
      }
      double sini = std::sin(rangle);
      double x, y;
      envelope(rangle, (llo_val + lhi_val) / 2.0, r1, &x, &y);
      thehi = fmaxr(PI / 2.0 - to_radians(mdl.llo),
                    fminr(PI / 2.0 + rangle, std::acos(y * sini / r2) + to_radians(mdl.lfudge)));
    }
    double ratio2 = (cosi - r1) / r2;
    if (ratio2 >= 1.0) {
      return set_star_grid_inner(c, nlat, false, 0.0, 0.0, 0, 0);
    } else if (ratio2 <= -1.0) {
      thelo = 0.0;
    } else {
      thelo = fminr(PI / 2.0 - to_radians(mdl.lhi),
                    fmaxr(PI / 2.0 - std::acos(ratio2) + rangle - to_radians(mdl.lfudge), 0.0));
    }
  }
  return set_star_grid_inner(c, nlat, infill, thelo, thehi, nlatfill, nlngfill);
}

// crates/lcurve/src/grid.rs:336-392
std::vector<Point> set_disc_grid(const Model& mdl) {
  const double EFAC = 1.0000001;
  double r1, r2;
  mdl.get_r1r2(&r1, &r2);
  double rl1 = xl11(mdl.q.value, mdl.spin1.value);
  if (r1 < 0.0) r1 = rl1;
  double rl2 = 1.0 - xl12(mdl.q.value, mdl.spin2.value);
  if (r2 < 0.0) r2 = rl2;
  double rdisc1 = mdl.rdisc1.value > 0.0 ? mdl.rdisc1.value : r1;
  double rdisc2 = mdl.rdisc2.value > 0.0 ? mdl.rdisc2.value : mdl.radius_spot.value;
  double ffac1 = r1 / rl1;
  double ffac2 = r2 / rl2;
  double drad = (rdisc2 - rdisc1) / static_cast<double>(mdl.nrad);
  double drrad = rdisc2 / static_cast<double>(mdl.nrad);

  std::vector<Point> disc;
  for (int i = 0; i < mdl.nrad; ++i) {
    double rad = rdisc1 + (rdisc2 - rdisc1) * (static_cast<double>(i) + 0.5) /
                              static_cast<double>(mdl.nrad);
    int ntheta = std::max(8, as_i32(std::ceil(TWOPI * rad / drrad)));
    double h = EFAC * mdl.height_disc.value * std::pow(rad, mdl.beta_disc.value);
    double tanp = mdl.beta_disc.value * h / rad;
    double cosp = 1.0 / std::sqrt(1.0 + sqr(tanp));
    double sinp = tanp * cosp;
    double area = TWOPI / static_cast<double>(ntheta) * rad * drad;
    for (int j = 0; j < ntheta; ++j) {
      double theta = TWOPI * static_cast<double>(j) / static_cast<double>(ntheta);
      double sint = std::sin(theta);
      double cost = std::cos(theta);
      Vec3 posn(rad * cost, rad * sint, h);
      Vec3 dirn(-cost * sinp, -sint * sinp, cosp);
      EclipseVec ecl;
      if (mdl.opaque)
        ecl = disc_eclipse(mdl.iangle.value, rdisc1, rdisc2, mdl.beta_disc.value,
                           mdl.height_disc.value, posn);
      if (mdl.eclipse1)
        star_eclipse(mdl.q.value, r1, mdl.spin1.value, ffac1, mdl.iangle.value, posn,
                     mdl.delta_phase, mdl.roche1, Star::Primary, ecl);
      if (mdl.eclipse2)
        star_eclipse(mdl.q.value, r2, mdl.spin2.value, ffac2, mdl.iangle.value, posn,
                     mdl.delta_phase, mdl.roche2, Star::Secondary, ecl);
      disc.emplace_back(posn, dirn, area, 1.0, std::move(ecl));
    }
  }
  return disc;
}

// crates/lcurve/src/grid.rs:395-481
std::vector<Point> set_disc_edge(const Model& mdl, bool outer) {
  const double EFAC = 1.0000001;
  double r1, r2;
  mdl.get_r1r2(&r1, &r2);
  double rl1 = xl11(mdl.q.value, mdl.spin1.value);
  if (r1 < 0.0) r1 = rl1;
  double rl2 = 1.0 - xl12(mdl.q.value, mdl.spin2.value);
  if (r2 < 0.0) r2 = rl2;
  double rdisc1 = mdl.rdisc1.value > 0.0 ? mdl.rdisc1.value : r1;
  double rdisc2 = mdl.rdisc2.value > 0.0 ? mdl.rdisc2.value : mdl.radius_spot.value;
  double size = rdisc2 / static_cast<double>(mdl.nrad);
  double ffac1 = r1 / rl1;
  double ffac2 = r2 / rl2;
  double rad_val = outer ? rdisc2 : rdisc1;
  double h = mdl.height_disc.value * std::pow(rad_val, mdl.beta_disc.value);
  int nout = as_i32(2.0 * h / size) + 2;
  int ntheta = std::max(8, as_i32(std::ceil(TWOPI * rad_val / size)));
  double area = TWOPI / static_cast<double>(ntheta) * rad_val * 2.0 * h / static_cast<double>(nout - 1);

  auto eclipses_for = [&](const Vec3& posn) {
    EclipseVec ecl;
    if (mdl.opaque)
      ecl = disc_eclipse(mdl.iangle.value, rdisc1, rdisc2, mdl.beta_disc.value,
                         mdl.height_disc.value, posn);
    if (mdl.eclipse1) {
      // The reference ignores errors from this call (grid.rs:439-441).
      try {
        star_eclipse(mdl.q.value, r1, mdl.spin1.value, ffac1, mdl.iangle.value, posn,
                     mdl.delta_phase, mdl.roche1, Star::Primary, ecl);
      } catch (const LcError&) {
      }
    }
    if (mdl.eclipse2)
      star_eclipse(mdl.q.value, r2, mdl.spin2.value, ffac2, mdl.iangle.value, posn,
                   mdl.delta_phase, mdl.roche2, Star::Secondary, ecl);
    return ecl;
  };

  std::vector<Point> edge;
  for (int i = 0; i < ntheta; ++i) {
    double theta = TWOPI * static_cast<double>(i) / static_cast<double>(ntheta);
    double sint = std::sin(theta);
    double cost = std::cos(theta);
    Vec3 posn = outer ? Vec3(EFAC * rad_val * cost, EFAC * rad_val * sint, EFAC * h)
                      : Vec3(rad_val * cost / EFAC, rad_val * sint / EFAC, EFAC * h);
    Vec3 dirn(cost, sint, 0.0);
    edge.emplace_back(posn, dirn, area / 2.0, 1.0, eclipses_for(posn));
    for (int j = 0; j < nout - 1; ++j) {
      double z = -h + 2.0 * h * static_cast<double>(j) / static_cast<double>(nout - 1);
      Vec3 p2, d2;
      if (outer) {
        p2 = Vec3(EFAC * rad_val * cost, EFAC * rad_val * sint, z);
        d2 = Vec3(cost, sint, 0.0);
      } else {
        p2 = Vec3(rad_val * cost / EFAC, rad_val * sint / EFAC, z);
        d2 = Vec3(-cost, -sint, 0.0);
      }
      edge.emplace_back(p2, d2, area, 1.0, eclipses_for(p2));
    }
  }
  return edge;
}

// crates/lcurve/src/grid.rs:484-553
std::vector<Point> set_bright_spot_grid(const Model& mdl) {
  double r1, r2;
  mdl.get_r1r2(&r1, &r2);
  double rl1 = xl11(mdl.q.value, mdl.spin1.value);
  if (r1 < 0.0) r1 = rl1;
  else if (r1 > rl1) throw LcError("set_bright_spot_grid: primary larger than Roche lobe");
  double rl2 = 1.0 - xl12(mdl.q.value, mdl.spin2.value);
  if (r2 < 0.0) r2 = rl2;
  else if (r2 > rl2) throw LcError("set_bright_spot_grid: secondary larger than Roche lobe");
  double ffac1 = r1 / rl1;
  double ffac2 = r2 / rl2;

  Vec3 bspot, v;
  strinit(mdl.q.value, &bspot, &v);
  stradv(mdl.q.value, &bspot, &v, mdl.radius_spot.value, 1e-10, 1e-3);

  double theta = std::atan2(bspot.y, bspot.x) + TWOPI / 4.0 + to_radians(mdl.angle_spot.value);
  double alpha = to_radians(mdl.yaw_spot.value);
  double tilt = to_radians(mdl.tilt_spot.value);
  Vec3 bvec(std::cos(theta), std::sin(theta), 0.0);
  Vec3 pvec(0.0, 0.0, 1.0);
  Vec3 tvec(std::sin(tilt) * std::sin(theta + alpha), -std::sin(tilt) * std::cos(theta + alpha),
            std::cos(tilt));

  double bmax = std::pow(mdl.expon_spot.value / mdl.epow_spot.value, 1.0 / mdl.epow_spot.value);
  double sfac = 20.0 + bmax;
  double area = sfac * mdl.length_spot.value * mdl.height_spot.value /
                static_cast<double>(mdl.nspot - 1);
  double bright_ref = planck(mdl.wavelength, mdl.temp_spot.value);

  std::vector<Point> spot;
  spot.reserve(mdl.nspot > 0 ? 2 * static_cast<size_t>(mdl.nspot) : 0);
  for (int i = 0; i < mdl.nspot; ++i) {
    double dist = sfac * static_cast<double>(i) / static_cast<double>(mdl.nspot - 1);
    Vec3 posn = bspot + mdl.length_spot.value * (dist - bmax) * bvec;
    EclipseVec ecl;
    if (mdl.eclipse1)
      star_eclipse(mdl.q.value, r1, mdl.spin1.value, ffac1, mdl.iangle.value, posn,
                   mdl.delta_phase, mdl.roche1, Star::Primary, ecl);
    if (mdl.eclipse2)
      star_eclipse(mdl.q.value, r2, mdl.spin2.value, ffac2, mdl.iangle.value, posn,
                   mdl.delta_phase, mdl.roche2, Star::Secondary, ecl);
    double bright = bright_ref * std::pow(dist / bmax, mdl.expon_spot.value) *
                    std::exp(mdl.expon_spot.value / mdl.epow_spot.value -
                             std::pow(dist, mdl.epow_spot.value));
    Point tilt_pt(posn, tvec, area, 1.0, ecl);
    tilt_pt.flux = static_cast<float>(bright * (1.0 - mdl.cfrac_spot.value) * area);
    spot.push_back(std::move(tilt_pt));
    Point par_pt(posn, pvec, area, 1.0, std::move(ecl));
    par_pt.flux = static_cast<float>(bright * mdl.cfrac_spot.value * area);
    spot.push_back(std::move(par_pt));
  }
  return spot;
}

// ---------------------------------------------------------------- brightness
namespace {
const double GAUSS_EFAC = 2.3548200450309493;  // sqrt(8 ln 2), brightness.rs:8

// crates/lcurve/src/brightness.rs:150-174
void irradiation(double t_base, double gravity, double gdcbol, double absorb, double t_irr,
                 double mu, double r, double r_irr, double* geom_out, double* temp_out) {
  double t_grav = t_base * std::pow(gravity, gdcbol);
  if (mu >= r_irr) {
    double geom = sqr(r_irr / r) * mu;
    *geom_out = geom;
    *temp_out = std::pow(powi4(t_grav) + absorb * powi4(t_irr) * geom, 0.25);
  } else if (mu > -r_irr) {
    double x0 = -mu / r_irr;
    double geom = sqr(r_irr / r) * r_irr *
                  (std::sqrt(1.0 - x0 * x0) * (2.0 + x0 * x0) / 3.0 - x0 * (PI / 2.0 - std::asin(x0))) /
                  PI;
    *geom_out = geom;
    *temp_out = std::pow(powi4(t_grav) + absorb * powi4(t_irr) * geom, 0.25);
  } else {
    *geom_out = 0.0;
    *temp_out = t_grav;
  }
}

// crates/lcurve/src/brightness.rs:177-183
double angular_dist(const Vec3& posn, const Vec3& dir, double posn_len) {
  if (posn_len <= 0.0) return 180.0;
  double cos_ang = clampd(dot(posn, dir) / posn_len, -1.0, 1.0);
  return to_degrees(std::acos(cos_ang));
}

struct SpotInfo {
  bool on = false;
  Vec3 dir;
  double fwhm = 0, tcen = 0;
};

// crates/lcurve/src/brightness.rs:187-201
SpotInfo spot_info(const Pparam& lo, const Pparam& la, const Pparam& fw, const Pparam& tc) {
  SpotInfo s;
  if (!lo.defined || !la.defined || !fw.defined || !tc.defined) return s;
  double clong = std::cos(to_radians(lo.value));
  double slong = std::sin(to_radians(lo.value));
  double clat = std::cos(to_radians(la.value));
  double slat = std::sin(to_radians(la.value));
  s.on = true;
  s.dir = Vec3(clat * clong, clat * slong, slat);
  s.fwhm = fw.value;
  s.tcen = tc.value;
  return s;
}
}  // namespace

// crates/lcurve/src/brightness.rs:13-147
void set_star_continuum(const Model& mdl, std::vector<Point>& star1, std::vector<Point>& star2) {
  double r1, r2;
  mdl.get_r1r2(&r1, &r2);
  double rl1 = xl1(mdl.q.value);
  if (r1 < 0.0) r1 = rl1;
  double rl2 = 1.0 - rl1;
  if (r2 < 0.0) r2 = rl2;
  const Vec3 cofm2(1.0, 0.0, 0.0);

  double gdcbol1 = mdl.gdark_bolom1 ? mdl.gravity_dark1.value
                                    : mdl.gravity_dark1.value / dlpdlt(mdl.wavelength, mdl.t1.value);
  SpotInfo s11 = spot_info(mdl.stsp11_long, mdl.stsp11_lat, mdl.stsp11_fwhm, mdl.stsp11_tcen);
  SpotInfo s12 = spot_info(mdl.stsp12_long, mdl.stsp12_lat, mdl.stsp12_fwhm, mdl.stsp12_tcen);
  SpotInfo s13 = spot_info(mdl.stsp13_long, mdl.stsp13_lat, mdl.stsp13_fwhm, mdl.stsp13_tcen);

  bool is_uespot = mdl.uesp_long1.defined && mdl.uesp_long2.defined && mdl.uesp_lathw.defined &&
                   mdl.uesp_taper.defined && mdl.uesp_temp.defined;
  Vec3 uespot_dir(0, 0, 0);
  double longhw = 0.0;
  if (is_uespot) {
    longhw = (mdl.uesp_long2.value - mdl.uesp_long1.value) / 2.0;
    double lcen = to_radians((mdl.uesp_long1.value + mdl.uesp_long2.value) / 2.0);
    uespot_dir = Vec3(std::cos(lcen), std::sin(lcen), 0.0);
  }

  for (Point& pt : star1) {
    Vec3 vec = cofm2 - pt.posn;
    double r = vec.length();
    double mu = dot(pt.dirn, vec) / r;
    double t1 = mdl.t1.value;
    for (const SpotInfo* s : {&s11, &s12, &s13}) {
      if (!s->on) continue;
      double dist = angular_dist(pt.posn, s->dir, pt.posn.length());
      t1 += (s->tcen - mdl.t1.value) * std::exp(-sqr(dist / (s->fwhm / GAUSS_EFAC)) / 2.0);
    }
    if (is_uespot) {
      Vec3 evec(pt.posn.x, pt.posn.y, 0.0);
      double evecl = evec.length();
      if (evecl > 0.0) {
        double rd = pt.posn.length();
        double clat = clampd(dot(pt.posn, evec) / evecl / rd, -1.0, 1.0);
        double dlat = to_degrees(std::acos(clat));
        double clong = clampd(dot(uespot_dir, evec) / evecl, -1.0, 1.0);
        double dlong = to_degrees(std::acos(clong));
        if (dlat <= mdl.uesp_lathw.value && dlong <= longhw) {
          t1 = mdl.uesp_temp.value;
        } else if (dlat <= mdl.uesp_lathw.value) {
          t1 += (mdl.uesp_temp.value - mdl.t1.value) * std::exp(-(dlong - longhw) / mdl.uesp_taper.value);
        } else if (dlong <= longhw) {
          t1 += (mdl.uesp_temp.value - mdl.t1.value) *
                std::exp(-(dlat - mdl.uesp_lathw.value) / mdl.uesp_taper.value);
        } else {
          t1 += (mdl.uesp_temp.value - mdl.t1.value) *
                std::exp(-(dlat - mdl.uesp_lathw.value) / mdl.uesp_taper.value) *
                std::exp(-(dlong - longhw) / mdl.uesp_taper.value);
        }
      }
    }
    double geom, temp;
    irradiation(t1, static_cast<double>(pt.gravity), gdcbol1, mdl.absorb.value, mdl.t2.value, mu, r,
                r2, &geom, &temp);
    pt.flux = static_cast<float>(static_cast<double>(pt.area) * planck(mdl.wavelength, temp));
    if (mdl.mirror)
      pt.flux += static_cast<float>(static_cast<double>(pt.area) * geom *
                                    planck(mdl.wavelength, std::fabs(mdl.t2.value)));
  }

  const Vec3 cofm1(0, 0, 0);
  double t2_abs = std::fabs(mdl.t2.value);
  double gdcbol2 = mdl.gdark_bolom2 ? mdl.gravity_dark2.value
                                    : mdl.gravity_dark2.value / dlpdlt(mdl.wavelength, t2_abs);
  SpotInfo s21 = spot_info(mdl.stsp21_long, mdl.stsp21_lat, mdl.stsp21_fwhm, mdl.stsp21_tcen);
  SpotInfo s22 = spot_info(mdl.stsp22_long, mdl.stsp22_lat, mdl.stsp22_fwhm, mdl.stsp22_tcen);

  for (Point& pt : star2) {
    Vec3 vec = cofm1 - pt.posn;
    double r = vec.length();
    double mu = dot(pt.dirn, vec) / r;
    double t2 = t2_abs;
    for (const SpotInfo* s : {&s21, &s22}) {
      if (!s->on) continue;
      Vec3 off = pt.posn - cofm2;
      double dist = angular_dist(off, s->dir, off.length());
      t2 += (s->tcen - t2) * std::exp(-sqr(dist / (s->fwhm / GAUSS_EFAC)) / 2.0);
    }
    double geom, temp;
    irradiation(t2, static_cast<double>(pt.gravity), gdcbol2, mdl.absorb.value, mdl.t1.value, mu, r,
                r1, &geom, &temp);
    pt.flux = static_cast<float>(static_cast<double>(pt.area) * planck(mdl.wavelength, temp));
    if (mdl.mirror)
      pt.flux += static_cast<float>(static_cast<double>(pt.area) * geom *
                                    planck(mdl.wavelength, mdl.t1.value));
  }
}

// crates/lcurve/src/brightness.rs:204-212
void set_disc_continuum(double rdisc, double tdisc, double texp, double wave,
                        std::vector<Point>& disc) {
  double bright = planck(wave, tdisc);
  for (Point& pt : disc) {
    double r = pt.posn.length();
    pt.flux = static_cast<float>(bright * std::pow(r / rdisc, texp) * static_cast<double>(pt.area));
  }
}

// crates/lcurve/src/brightness.rs:215-239
void set_edge_continuum(double tedge, double r2, double t2, double absorb, double wave,
                        std::vector<Point>& edge) {
  const Vec3 cofm2(1.0, 0.0, 0.0);
  for (Point& pt : edge) {
    Vec3 vec = cofm2 - pt.posn;
    double r = vec.length();
    double mu = dot(pt.dirn, vec) / r;
    double temp;
    if (mu >= r2) {
      double geom = sqr(r2 / r) * mu;
      temp = std::pow(powi4(tedge) + absorb * powi4(t2) * geom, 0.25);
    } else if (mu > -r2) {
      double x0 = -mu / r2;
      double geom = sqr(r2 / r) * r2 *
                    (std::sqrt(1.0 - x0 * x0) * (2.0 + x0 * x0) / 3.0 - x0 * (PI / 2.0 - std::asin(x0))) /
                    PI;
      temp = std::pow(powi4(tedge) + absorb * powi4(t2) * geom, 0.25);
    } else {
      temp = tedge;
    }
    pt.flux = static_cast<float>(static_cast<double>(pt.area) * planck(wave, temp));
  }
}

// --------------------------------------------- per-curve scalar summaries
// crates/lcurve/src/flux.rs:282-315
double comp_gravity1(const Model& mdl, const std::vector<Point>& star1) {
  double gm = powi3(1000.0 * mdl.velocity_scale.value) * mdl.tperiod * DAY / TWOPI;
  double a = std::pow(gm / sqr(TWOPI / DAY / mdl.tperiod), 1.0 / 3.0);
  double gscale = 100.0 * gm / sqr(a);
  double r1, r2;
  mdl.get_r1r2(&r1, &r2);
  double rl1;
  try {
    rl1 = xl12(mdl.q.value, mdl.spin1.value);
  } catch (const LcError&) {
    rl1 = 0.5;
  }
  double gref;
  if (mdl.roche1) {
    double acc = mdl.delta_phase / 10.0;
    double ffac1 = r1 / rl1;
    double rref1, pref1;
    ref_sphere(mdl.q.value, Star::Primary, mdl.spin1.value, ffac1, &rref1, &pref1);
    Vec3 p, d; double rad, g;
    face(mdl.q.value, Star::Primary, mdl.spin1.value, Vec3(-1.0, 0.0, 0.0), rref1, pref1, acc, &p, &d,
         &rad, &g);
    gref = g * gscale;
  } else {
    gref = gscale / (1.0 + mdl.q.value) / sqr(r1);
  }
  double sumfg = 0.0, sumf = 0.0;
  for (const Point& pt : star1) {
    sumfg += static_cast<double>(pt.flux) * static_cast<double>(pt.gravity);
    sumf += static_cast<double>(pt.flux);
  }
  if (gref > 0.0 && sumfg > 0.0 && sumf > 0.0) return std::log10(gref * sumfg / sumf);
  return 0.0;
}

// crates/lcurve/src/flux.rs:318-349
double comp_gravity2(const Model& mdl, const std::vector<Point>& star2) {
  double gm = powi3(1000.0 * mdl.velocity_scale.value) * mdl.tperiod * DAY / TWOPI;
  double a = std::pow(gm / sqr(TWOPI / DAY / mdl.tperiod), 1.0 / 3.0);
  double gscale = 100.0 * gm / sqr(a);
  double r1, r2;
  mdl.get_r1r2(&r1, &r2);
  double rl2;
  try {
    rl2 = 1.0 - xl12(mdl.q.value, mdl.spin2.value);
  } catch (const LcError&) {
    rl2 = 1.0 - 0.5;
  }
  double gref;
  if (mdl.roche2) {
    double acc = mdl.delta_phase / 10.0;
    double ffac2 = r2 / rl2;
    double rref2, pref2;
    ref_sphere(mdl.q.value, Star::Secondary, mdl.spin2.value, ffac2, &rref2, &pref2);
    Vec3 p, d; double rad, g;
    face(mdl.q.value, Star::Secondary, mdl.spin2.value, Vec3(1.0, 0.0, 0.0), rref2, pref2, acc, &p,
         &d, &rad, &g);
    gref = g * gscale;
  } else {
    gref = gscale * mdl.q.value / (1.0 + mdl.q.value) / sqr(r2);
  }
  double sumfg = 0.0, sumf = 0.0;
  for (const Point& pt : star2) {
    sumfg += static_cast<double>(pt.flux) * static_cast<double>(pt.gravity);
    sumf += static_cast<double>(pt.flux);
  }
  if (gref > 0.0 && sumfg > 0.0 && sumf > 0.0) return std::log10(gref * sumfg / sumf);
  return 0.0;
}

// crates/lcurve/src/flux.rs:352-375
static double comp_radius(const std::vector<Point>& star, const Vec3& cofm) {
  double sumsa = 0.0, sumvol = 0.0;
  for (const Point& pt : star) {
    Vec3 vec = pt.posn - cofm;
    double r = vec.length();
    double rcosa = dot(pt.dirn, vec);
    sumsa += static_cast<double>(pt.area) * rcosa / powi3(r);
    sumvol += static_cast<double>(pt.area) * rcosa;
  }
  return std::pow(sumvol / sumsa, 1.0 / 3.0);
}
double comp_radius1(const std::vector<Point>& star1) { return comp_radius(star1, Vec3(0, 0, 0)); }
double comp_radius2(const std::vector<Point>& star2) { return comp_radius(star2, Vec3(1.0, 0, 0)); }

}  // namespace lc
