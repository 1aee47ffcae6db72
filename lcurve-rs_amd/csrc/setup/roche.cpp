// Roche geometry restated from crates/roche/src and crates/subs/src/numerical.
// Operation order follows the reference line by line (see the file:line
// notes) so that grids and eclipse phases come out identical.
#include "roche.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>

namespace lc {

namespace {
constexpr int NMAX = 1000;
constexpr double LEPS = 1.0e-12;

// Certified short cuts in face() and in ingress_egress's phase scan: on by
// default, switchable per thread for self-checks; counters of the scan.
thread_local bool tl_prescreen = true;
thread_local long long tl_skipped = 0, tl_evaluated = 0;

// Newton iteration on the quintic for an L1-like point.
// crates/roche/src/lagrange.rs:8-31 (same loop shape in xl11 / xl12)
double quintic_root(double x, double a1, double a2, double a3, double a4, double a5, double a6,
                    const char* who) {
  double d1 = a2, d2 = 2.0 * a3, d3 = 3.0 * a4, d4 = 4.0 * a5, d5 = 5.0 * a6;
  for (int i = 0; i < NMAX; ++i) {
    double xold = x;
    double f = x * (x * (x * (x * (x * a6 + a5) + a4) + a3) + a2) + a1;
    double df = x * (x * (x * (x * d5 + d4) + d3) + d2) + d1;
    x -= f / df;
    if (std::fabs(x - xold) <= LEPS * std::fabs(x)) return x;
  }
  throw LcError(std::string(who) + ": exceeded maximum iterations");
}

void check_q(double q, const char* who) {
  if (q <= 0.0) throw LcError(std::string(who) + ": q <= 0");
}
}  // namespace

// crates/roche/src/lagrange.rs:8-31
double xl1(double q) {
  check_q(q, "xl1");
  double mu = q / (1.0 + q);
  return quintic_root(1.0 / (1.0 + q), -1.0 + mu, 2.0 - 2.0 * mu, -1.0 + mu, 1.0 + 2.0 * mu,
                      -2.0 - mu, 1.0, "xl1");
}

// crates/roche/src/lagrange.rs:89-114
double xl11(double q, double spin) {
  check_q(q, "xl11");
  double ssq = spin * spin;
  double mu = q / (1.0 + q);
  return quintic_root(1.0 / (1.0 + q), -1.0 + mu, 2.0 - 2.0 * mu, -1.0 + mu, ssq + 2.0 * mu,
                      -2.0 * ssq - mu, ssq, "xl11");
}

// crates/roche/src/lagrange.rs:117-142
double xl12(double q, double spin) {
  check_q(q, "xl12");
  double ssq = spin * spin;
  double mu = q / (1.0 + q);
  return quintic_root(1.0 / (1.0 + q), -1.0 + mu, 2.0 - 2.0 * mu, -ssq + mu,
                      3.0 * ssq + 2.0 * mu - 2.0, 1.0 - mu - 3.0 * ssq, ssq, "xl12");
}

// crates/roche/src/potential.rs:8-19
double rpot(double q, const Vec3& p) {
  check_q(q, "rpot");
  double mu = q / (1.0 + q);
  double comp = 1.0 - mu;
  double x2y2 = sqr(p.x) + sqr(p.y);
  double r1sq = x2y2 + sqr(p.z);
  double r1 = std::sqrt(r1sq);
  double r2 = std::sqrt(r1sq + 1.0 - 2.0 * p.x);
  return -comp / r1 - mu / r2 - (x2y2 + mu * (mu - 2.0 * p.x)) / 2.0;
}

namespace {
// Potentials without the q > 0 check: callers validate q once (every
// setup path calls xl11 / xl12 first, which reject q <= 0).
// crates/roche/src/potential.rs:22-33
inline double rpot1_u(double q, double spin, const Vec3& p) {
  double mu = q / (1.0 + q);
  double comp = 1.0 - mu;
  double x2y2 = sqr(p.x) + sqr(p.y);
  double r1sq = x2y2 + sqr(p.z);
  double r1 = std::sqrt(r1sq);
  double r2 = std::sqrt(r1sq + 1.0 - 2.0 * p.x);
  return -comp / r1 - mu / r2 - sqr(spin) * x2y2 / 2.0 + mu * p.x;
}

// crates/roche/src/potential.rs:36-47
inline double rpot2_u(double q, double spin, const Vec3& p) {
  double mu = q / (1.0 + q);
  double comp = 1.0 - mu;
  double x2y2 = sqr(p.x) + sqr(p.y);
  double r1sq = x2y2 + sqr(p.z);
  double r1 = std::sqrt(r1sq);
  double r2 = std::sqrt(r1sq + 1.0 - 2.0 * p.x);
  return -comp / r1 - mu / r2 - sqr(spin) * (0.5 + 0.5 * x2y2 - p.x) - comp * p.x;
}

// crates/roche/src/potential.rs:68-86
inline Vec3 drpot1_u(double q, double spin, const Vec3& p) {
  double r1sq = p.sqr();
  double r2sq = r1sq + 1.0 - 2.0 * p.x;
  double r1 = std::sqrt(r1sq);
  double r2 = std::sqrt(r2sq);
  double mu = q / (1.0 + q);
  double mu1 = mu / r2 / r2sq;
  double comp = (1.0 - mu) / r1 / r1sq;
  double ssq = sqr(spin);
  return {comp * p.x + mu1 * (p.x - 1.0) - ssq * p.x + mu, comp * p.y + mu1 * p.y - ssq * p.y,
          comp * p.z + mu1 * p.z};
}

// crates/roche/src/potential.rs:88-106
inline Vec3 drpot2_u(double q, double spin, const Vec3& p) {
  double r1sq = p.sqr();
  double r2sq = r1sq + 1.0 - 2.0 * p.x;
  double r1 = std::sqrt(r1sq);
  double r2 = std::sqrt(r2sq);
  double mu = q / (1.0 + q);
  double mu1 = mu / r2 / r2sq;
  double comp = (1.0 - mu) / r1 / r1sq;
  double ssq = sqr(spin);
  return {comp * p.x + mu1 * (p.x - 1.0) - ssq * (p.x - 1.0) + mu - 1.0,
          comp * p.y + mu1 * p.y - ssq * p.y, comp * p.z + mu1 * p.z};
}
}  // namespace

double rpot1(double q, double spin, const Vec3& p) {
  check_q(q, "rpot1");
  return rpot1_u(q, spin, p);
}
double rpot2(double q, double spin, const Vec3& p) {
  check_q(q, "rpot2");
  return rpot2_u(q, spin, p);
}
Vec3 drpot1(double q, double spin, const Vec3& p) {
  check_q(q, "drpot1");
  return drpot1_u(q, spin, p);
}
Vec3 drpot2(double q, double spin, const Vec3& p) {
  check_q(q, "drpot2");
  return drpot2_u(q, spin, p);
}

namespace {
inline Vec3 star_cofm(Star s) { return s == Star::Primary ? Vec3(0, 0, 0) : Vec3(1.0, 0, 0); }

inline double pot_star(double q, Star s, double spin, const Vec3& p) {
  return s == Star::Primary ? rpot1_u(q, spin, p) : rpot2_u(q, spin, p);
}

// Potential and its line-of-sight derivative at p + lam * earth
// (crates/roche/src/potential.rs:109-126), specialised on the star.
template <Star S>
inline double rpot_along_los(double q, double spin, const Vec3& earth, const Vec3& p, double lam) {
  Vec3 point = p + lam * earth;
  return S == Star::Primary ? rpot1_u(q, spin, point) : rpot2_u(q, spin, point);
}
template <Star S>
inline double drpot_along_los(double q, double spin, const Vec3& earth, const Vec3& p, double lam) {
  Vec3 point = p + lam * earth;
  Vec3 g = S == Star::Primary ? drpot1_u(q, spin, point) : drpot2_u(q, spin, point);
  return dot(g, earth);
}

inline double sign_of(double a, double b) { return b >= 0.0 ? std::fabs(a) : -std::fabs(a); }

// Brent minimisation with derivatives, early exit below fref.
// crates/subs/src/numerical/dbrent.rs:20-165. Returns false on too many
// iterations (the reference's Err).
template <class F, class D>
bool dbrent(double ax, double bx, double cx, F&& func, D&& dfunc, double acc, bool stopfast,
            double fref, double* fmin_out, double* xmin_out) {
  const int ITMAX = 100;
  double a = fminr(ax, cx);
  double b = fmaxr(ax, cx);
  double x = bx, w = bx, v = bx;
  double fx = func(x);
  double fw = fx, fv = fx;
  if (stopfast && fx < fref) { *fmin_out = fx; *xmin_out = x; return true; }
  double dx = dfunc(x);
  double dw = dx, dv = dx;
  double e = 0.0, d = 0.0;
  for (int it = 0; it < ITMAX; ++it) {
    double xm = 0.5 * (a + b);
    double tol1 = acc;
    double tol2 = 2.0 * tol1;
    if (std::fabs(x - xm) <= tol2 - 0.5 * (b - a)) { *fmin_out = fx; *xmin_out = x; return true; }
    if (std::fabs(e) > tol1) {
      double d1 = 2.0 * (b - a);
      double d2 = d1;
      if (dw != dx) d1 = (w - x) * dx / (dx - dw);
      if (dv != dx) d2 = (v - x) * dx / (dx - dv);
      double u1 = x + d1;
      double u2 = x + d2;
      bool ok1 = (a - u1) * (u1 - b) > 0.0 && dx * d1 <= 0.0;
      bool ok2 = (a - u2) * (u2 - b) > 0.0 && dx * d2 <= 0.0;
      double olde = e;
      e = d;
      if (ok1 || ok2) {
        if (ok1 && ok2) d = std::fabs(d1) < std::fabs(d2) ? d1 : d2;
        else if (ok1) d = d1;
        else d = d2;
        if (std::fabs(d) <= std::fabs(0.5 * olde)) {
          double u = x + d;
          if (u - a < tol2 || b - u < tol2) d = sign_of(tol1, xm - x);
        } else {
          e = dx >= 0.0 ? a - x : b - x;
          d = 0.5 * e;
        }
      } else {
        e = dx >= 0.0 ? a - x : b - x;
        d = 0.5 * e;
      }
    } else {
      e = dx >= 0.0 ? a - x : b - x;
      d = 0.5 * e;
    }
    double u, fu;
    if (std::fabs(d) >= tol1) {
      u = x + d;
      fu = func(u);
      if (stopfast && fu < fref) { *fmin_out = fu; *xmin_out = u; return true; }
    } else {
      u = x + sign_of(tol1, d);
      fu = func(u);
      if (stopfast && fu < fref) { *fmin_out = fu; *xmin_out = u; return true; }
      if (fu > fx) { *fmin_out = fx; *xmin_out = x; return true; }
    }
    double du = dfunc(u);
    if (fu <= fx) {
      if (u >= x) a = x; else b = x;
      v = w; fv = fw; dv = dw;
      w = x; fw = fx; dw = dx;
      x = u; fx = fu; dx = du;
    } else {
      if (u < x) a = u; else b = u;
      if (fu <= fw || w == x) {
        v = w; fv = fw; dv = dw;
        w = u; fw = fu; dw = du;
      } else if (fu < fv || v == x || v == w) {
        v = u; fv = fu; dv = du;
      }
    }
  }
  (void)fv;
  return false;
}

// crates/roche/src/eclipse.rs:202-224
bool sphere_eclipse_los(const Vec3& earth, const Vec3& r, const Vec3& c, double rsphere,
                        double* lam1, double* lam2) {
  Vec3 d = r - c;
  double bquad = dot(earth, d);
  if (bquad >= 0.0) return false;
  double cquad = d.sqr() - sqr(rsphere);
  double fac = sqr(bquad) - cquad;
  if (fac <= 0.0) return false;
  fac = std::sqrt(fac);
  *lam2 = -bquad + fac;
  *lam1 = fmaxr(cquad / *lam2, 0.0);
  return true;
}
}  // namespace

// crates/roche/src/surface.rs:14-39
void ref_sphere(double q, Star star, double spin, double ffac, double* rref, double* pref) {
  double rl1 = star == Star::Primary ? xl11(q, spin) : 1.0 - xl12(q, spin);
  double tref = rl1;
  double surf_r = ffac * tref;
  Vec3 p = star == Star::Primary ? Vec3(surf_r, 0.0, 0.0) : Vec3(1.0 - surf_r, 0.0, 0.0);
  *pref = star == Star::Primary ? rpot1(q, spin, p) : rpot2(q, spin, p);
  *rref = fminr(1.001 * ffac, 1.0) * tref;
}

// crates/roche/src/surface.rs:54-121
//
// The binary chop for the surface radius compares the potential at each
// midpoint with pref. Once the bracket is narrower than kFaceTaylor, the
// chop takes a Newton step to a centre c near the root and evaluates the
// potential F and its radial derivative D there; after that a midpoint's
// comparison is decided without evaluating the potential whenever
//   F + D h + K h^2 / 2 (h = mid - c, K the bound of the radial second
// derivative in the direction that matters) clears pref by far more than
// the rounding of one potential evaluation. Along the ray from the star's
// own centre the second derivative is exactly -2 m_own / rho^3 from its own
// mass, -ssq (dx^2 + dy^2) from the rotation and within
// [-2 m_oth, m_oth] / r_oth^3 from the other star (r_oth >= 1 - hi). A
// decided comparison is the one the evaluation would give, so the chop
// visits the same midpoints and returns the same radius to the bit;
// undecided midpoints (within ~1e-11 of the root) are evaluated.
namespace {
constexpr double kFaceTaylor = 1e-2;
}

void face(double q, Star star, double spin, const Vec3& dirn, double rref, double pref, double acc,
          Vec3* pvec_out, Vec3* dvec_out, double* radius_out, double* gravity_out) {
  Vec3 cofm = star_cofm(star);
  auto pot_fn = [&](double radius) {
    Vec3 p = cofm + radius * dirn;
    return pot_star(q, star, spin, p);
  };
  auto dpot_fn = [&](double radius) {
    Vec3 p = cofm + radius * dirn;
    Vec3 g = star == Star::Primary ? drpot1_u(q, spin, p) : drpot2_u(q, spin, p);
    return dot(g, dirn);
  };
  double lo = rref / 2.0;
  double hi = rref;
  double tref = pref + 1.0;
  for (int i = 0; i < 30; ++i) {
    if (tref <= pref) break;
    lo = hi / 2.0;
    tref = pot_fn(lo);
    if (tref > pref) hi = lo;
  }
  // Taylor centre state: 0 not set, 1 set, -1 unusable.
  int taylor = tl_prescreen ? 0 : -1;
  double c0 = 0.0, F0 = 0.0, D0 = 0.0, klo = 0.0, khi = 0.0, tol = 0.0;
  for (int i = 0; i < 100; ++i) {
    if (hi - lo <= acc) break;
    double mid = (lo + hi) / 2.0;
    if (taylor == 0 && hi - lo <= kFaceTaylor && lo > 0.0 && hi < 1.0) {
      const double Fm = pot_fn(mid), Dm = dpot_fn(mid);
      double c = Dm > 0.0 ? mid - (Fm - pref) / Dm : mid;
      c = c < lo ? lo : (c > hi ? hi : c);
      F0 = pot_fn(c);
      D0 = dpot_fn(c);
      c0 = c;
      const double mu = q / (1.0 + q);
      const double mown = star == Star::Primary ? 1.0 - mu : mu;
      const double moth = 1.0 - mown;
      const double roth = 1.0 - hi;
      const double cen = sqr(spin) * (sqr(dirn.x) + sqr(dirn.y));
      klo = -2.0 * mown / (lo * lo * lo) - 2.0 * moth / (roth * roth * roth) - cen;
      khi = -2.0 * mown / (hi * hi * hi) + moth / (roth * roth * roth) - cen;
      tol = 1e-12 * (1.0 + mown / lo + moth / roth + sqr(spin) + std::fabs(pref));
      taylor = D0 > 0.0 && std::isfinite(F0) && std::isfinite(klo) ? 1 : -1;
    }
    bool below;
    int known = -1;
    if (taylor == 1) {
      const double h = mid - c0;
      const double lin = F0 + D0 * h;
      if (lin + 0.5 * khi * h * h + tol < pref) known = 1;
      else if (lin + 0.5 * klo * h * h - tol > pref) known = 0;
    }
    if (known >= 0) {
      below = known == 1;
    } else {
      double pot_mid = pot_fn(mid);
      below = pot_mid < pref;
    }
    if (below) lo = mid;
    else hi = mid;
  }
  double radius = (lo + hi) / 2.0;
  Vec3 pvec = cofm + radius * dirn;
  Vec3 grad = star == Star::Primary ? drpot1(q, spin, pvec) : drpot2(q, spin, pvec);
  double gravity = grad.length();
  Vec3 dvec = grad;
  dvec.unit();
  *pvec_out = pvec;
  *dvec_out = dvec;
  *radius_out = radius;
  *gravity_out = gravity;
}

// crates/roche/src/eclipse.rs:229-265
bool sphere_eclipse_phase(double cosi, double sini, const Vec3& r, const Vec3& c, double rsphere,
                          double* phi1, double* phi2, double* lam1, double* lam2) {
  Vec3 d = r - c;
  double pdist = std::sqrt(sqr(d.x) + sqr(d.y));
  double bquad = d.z * cosi - pdist * sini;
  if (bquad >= 0.0) return false;
  double cquad = d.sqr() - sqr(rsphere);
  double fac = sqr(bquad) - cquad;
  if (fac <= 0.0) return false;
  fac = std::sqrt(fac);
  *lam2 = -bquad + fac;
  *lam1 = fmaxr(cquad / *lam2, 0.0);
  if (cquad < 0.0) {
    *phi1 = 0.0;
    *phi2 = 1.0;
  } else {
    double delta = std::acos((cosi * d.z + std::sqrt(cquad)) / (sini * pdist));
    double phi = std::atan2(d.y, -d.x);
    double p1 = (phi - delta) / TWOPI;
    p1 -= std::floor(p1);
    *phi1 = p1;
    *phi2 = p1 + 2.0 * delta / TWOPI;
  }
  return true;
}

namespace {
// fblink's body up to its Brent search (crates/roche/src/eclipse.rs:123-181),
// with the reference sphere passed in: ref_sphere is a pure function of
// (q, star, spin, ffac), so ingress_egress computes it once for its many
// fblink calls. Returns 0 or 1 when this part decides; 2 when a minimum of
// the potential along the line of sight is bracketed at *lam (potential
// *flam) inside the chord [*lam1, *lam2] through the reference sphere.
template <Star S>
int fblink_bracket(double q, double spin, double acc, const Vec3& earth, const Vec3& p,
                   double rref, double pref, double* lam1_out, double* lam2_out, double* lam_out,
                   double* flam_out) {
  const Vec3 cofm = star_cofm(S);
  double lam1, lam2;
  if (!sphere_eclipse_los(earth, p, cofm, rref, &lam1, &lam2)) return 0;
  if (lam1 == 0.0) return 1;

  auto func = [&](double lam) { return rpot_along_los<S>(q, spin, earth, p, lam); };
  size_t nstep = 1;
  double step_size = lam2 - lam1;
  double f1 = 0.0, f2 = 0.0, flam = 1.0;
  double lam = lam1;
  while (step_size > acc) {
    lam = lam1 + step_size / 2.0;
    bool bracketted = false;
    for (size_t n = 0; n < nstep; ++n) {
      flam = func(lam + static_cast<double>(n) * step_size);
      if (flam <= pref) return 1;
      if (nstep == 1) {
        f1 = func(lam1);
        f2 = func(lam2);
      }
      if (flam < f1 && flam < f2) {
        lam = lam + static_cast<double>(n) * step_size;
        bracketted = true;
        break;
      }
    }
    if (bracketted) break;
    step_size /= 2.0;
    nstep *= 2;
  }
  if (flam < f1 && flam < f2) {
    *lam1_out = lam1;
    *lam2_out = lam2;
    *lam_out = lam;
    *flam_out = flam;
    return 2;
  }
  return 0;
}

// fblink's Brent search from the bracket (eclipse.rs:182-196).
template <Star S>
bool fblink_min(double q, double spin, double acc, const Vec3& earth, const Vec3& p, double pref,
                double lam1, double lam, double lam2) {
  auto func = [&](double x) { return rpot_along_los<S>(q, spin, earth, p, x); };
  auto dfunc = [&](double x) { return drpot_along_los<S>(q, spin, earth, p, x); };
  double fmin, xmin;
  if (dbrent(lam1, lam, lam2, func, dfunc, acc, true, pref, &fmin, &xmin)) return fmin < pref;
  return false;
}

// crates/roche/src/eclipse.rs:123-197
template <Star S>
bool fblink_ref(double q, double spin, double acc, const Vec3& earth, const Vec3& p, double rref,
                double pref) {
  double lam1, lam2, lam, flam;
  const int k = fblink_bracket<S>(q, spin, acc, earth, p, rref, pref, &lam1, &lam2, &lam, &flam);
  if (k != 2) return k == 1;
  return fblink_min<S>(q, spin, acc, earth, p, pref, lam1, lam, lam2);
}

template <Star S>
bool fblink_t(double q, double spin, double ffac, double acc, const Vec3& earth, const Vec3& p) {
  double rref, pref;
  ref_sphere(q, S, spin, ffac, &rref, &pref);
  return fblink_ref<S>(q, spin, acc, earth, p, rref, pref);
}

// Certified short cut for the phase scan of ingress_egress.
//
// Most scan steps' lines of sight pass beside the eclipsing lobe, and for
// each fblink runs its grid search, brackets the potential minimum along the
// chord [lam1, lam2] through the reference sphere and then runs dbrent, the
// costly part, only to find the minimum above pref. The scan runs
// fblink_bracket exactly as fblink does and skips dbrent when its result is
// certain to be false:
//
//  (A) a rigorous lower bound of the potential over the part of the line of
//      sight inside the ball B(c, rx), rx = rref + 1e-6, exceeds pref by far
//      more than the rounding of one potential evaluation, so no point of
//      the chord gives a potential <= pref; and
//  (B) the bound over [lam1, lam1 + acc] and over [lam2 - acc, lam2] exceeds
//      flam, the value at the bracketing point, likewise.
//
// Why that suffices: dbrent's best point x starts at the bracketing point
// and only moves to trial points with a potential <= its own, so by (B) it
// never lies within acc of a chord end. Its bracket starts as [lam1, lam2]
// and only takes values of x and of trial points. Its trial points are
// parabolic steps strictly inside the bracket, midpoints between x and a
// bracket end, or steps of acc from x; with x at least acc inside the
// chord, all of them lie in the chord. So dbrent evaluates the potential
// only on the chord, where by (A) every value is above pref: it returns a
// minimum above pref (or fails), and fblink is false either way. Steps not
// certified run dbrent as before, so the scan finds the same first step and
// ingress and egress are unchanged to the bit.
//
// (A) covers a block of consecutive phases at once: every line of sight of
// the block lies within |lam| * de of the line for the block's centre phase
// (de bounds |earth(phi) - earth(phi_c)| <= 2 pi sin(i) |phi - phi_c|).
// Along a piece [la, lb] of that central line the terms are bounded
// separately: -comp/r1 and -mu/r2 by the closest approach of the widened
// piece to each star, and the polynomial part -ssq/2 rho^2 + linear, which
// is concave along a line, by its smaller end value less a Lipschitz
// widening term. Pieces that do not clear pref are halved, up to a budget.
template <Star S>
struct ScanBound {
  double mu, comp, ssq, pref, acc, rx, lam_lo, lam_hi, cosi, sini;
  Vec3 r, c;
  bool usable;

  ScanBound(double q, double spin, double cosi_, double sini_, const Vec3& r_, double acc_,
            double rref, double pref_)
      : cosi(cosi_), sini(sini_), r(r_), c(star_cofm(S)) {
    mu = q / (1.0 + q);
    comp = 1.0 - mu;
    ssq = spin * spin;
    pref = pref_;
    acc = acc_;
    rx = rref + 2.0 * acc + 1e-6;
    const double dist = (r - c).length();
    lam_lo = (dist - rx) * (1.0 - 1e-12) - 1e-12;
    lam_hi = (dist + rx) * (1.0 + 1e-12) + 1e-12;
    // An element inside (or touching) the ball is left to the plain scan.
    usable = lam_lo > 1e-9 && std::isfinite(pref) && std::isfinite(lam_hi);
  }

  double seg_dist(const Vec3& o, const Vec3& e, double la, double lb) const {
    Vec3 d = r - o;
    double t = -dot(d, e);
    t = t < la ? la : (t > lb ? lb : t);
    return (d + t * e).length();
  }

  // Lower bound of the potential over the points within `fat` of
  // r + lam * e, lam in [la, lb]; *mag bounds the size of its terms.
  double piece_bound(const Vec3& e, double fat, double la, double lb, double* mag) const {
    const double d1 = seg_dist(Vec3(0.0, 0.0, 0.0), e, la, lb) - fat - 1e-12;
    const double d2 = seg_dist(Vec3(1.0, 0.0, 0.0), e, la, lb) - fat - 1e-12;
    if (!(d1 > 0.0 && d2 > 0.0)) {
      *mag = 1.0;
      return -HUGE_VAL;
    }
    const Vec3 A = r + la * e, B = r + lb * e;
    const double ax = S == Star::Primary ? 0.0 : 1.0;  // spin axis x
    const double rhoA = std::sqrt(sqr(A.x - ax) + sqr(A.y));
    const double rhoB = std::sqrt(sqr(B.x - ax) + sqr(B.y));
    // Linear term: +mu x about star 1, -comp x about star 2.
    const double lin = S == Star::Primary ? mu : -comp;
    const double PA = -0.5 * ssq * sqr(rhoA) + lin * A.x;
    const double PB = -0.5 * ssq * sqr(rhoB) + lin * B.x;
    const double rho = std::fmax(rhoA, rhoB);
    const double widen = ssq * (rho * fat + 0.5 * fat * fat) + std::fabs(lin) * fat;
    const double t1 = comp / d1, t2 = mu / d2;
    *mag = 1.0 + t1 + t2 + 0.5 * ssq * sqr(rho + fat) +
           std::fabs(lin) * (std::fmax(std::fabs(A.x), std::fabs(B.x)) + fat);
    return std::fmin(PA, PB) - widen - t1 - t2;
  }

  // (A) for one piece: no point within `fat` of it inside B(c, rx) has a
  // potential <= pref.
  bool piece_clear(const Vec3& e, double de, double la, double lb) const {
    const double fat = std::fmax(std::fabs(la), std::fabs(lb)) * de + 1e-12;
    if (seg_dist(c, e, la, lb) - fat > rx) return true;
    double mag;
    const double bound = piece_bound(e, fat, la, lb, &mag);
    return bound - 1e-9 * mag > pref;
  }

  // (A) for the phases within hw of phi_c.
  bool phases_clear(double phi_c, double hw, int budget) const {
    const Vec3 e = set_earth(cosi, sini, phi_c);
    const double de = TWOPI * std::fabs(sini) * (hw + 1e-14) + 1e-14;
    double stk[48][2];
    int n = 0;
    // Split first at the closest approach to the eclipsing star.
    const double tc = dot(c - r, e);
    if (tc > lam_lo && tc < lam_hi) {
      stk[n][0] = tc; stk[n][1] = lam_hi; ++n;
      stk[n][0] = lam_lo; stk[n][1] = tc; ++n;
    } else {
      stk[n][0] = lam_lo; stk[n][1] = lam_hi; ++n;
    }
    while (n > 0) {
      --n;
      const double a = stk[n][0], b = stk[n][1];
      if (piece_clear(e, de, a, b)) continue;
      if (--budget <= 0 || n + 2 > 48) return false;
      const double m = 0.5 * (a + b);
      stk[n][0] = m; stk[n][1] = b; ++n;
      stk[n][0] = a; stk[n][1] = m; ++n;
    }
    return true;
  }

  // (B) at one chord end `end` (inward direction in = +1 at lam1, -1 at
  // lam2; fend the potential fblink computes there; fm an upper bound of
  // dbrent's best value, fm <= fend). Along the line,
  // Phi(end + in t) >= fend + in D t - K t^2 / 2, with D the derivative at
  // the end and -K a lower bound of the second derivative on the zone (for
  // a point mass m, e.H.e = m (1 - 3 cos^2) / r^3 >= -2 m / r^3; the
  // centrifugal part contributes >= -ssq). Required: that bound exceeds fm
  // for t in [dl, acc + dl], where dl = 0 if fend > fm and dl = acc / 1000
  // otherwise. No dbrent trial point lands within dl of the end while its
  // best point stays out of the zone (trial points sit 2 acc inside the
  // bracket, halfway to an end, or acc from the best point), so the best
  // point, whose value is below fm, never enters the zone.
  bool end_clear(const Vec3& e, double q, double spin, double end, double in, double fend,
                 double fm) const {
    const double dl = fend > fm ? 0.0 : 1e-3 * acc;
    const double T = acc + dl;
    const double la = in > 0.0 ? end : end - T, lb = la + T;
    const double r1 = seg_dist(Vec3(0.0, 0.0, 0.0), e, la, lb) * (1.0 - 1e-12) - 1e-12;
    const double r2 = seg_dist(Vec3(1.0, 0.0, 0.0), e, la, lb) * (1.0 - 1e-12) - 1e-12;
    if (!(r1 > 0.0 && r2 > 0.0)) return false;
    const double K = 2.0 * comp / (r1 * r1 * r1) + 2.0 * mu / (r2 * r2 * r2) + ssq;
    const double D = in * drpot_along_los<S>(q, spin, e, r, end);
    // Rounding of potentials and of D, with a wide margin.
    const double tol = 1e-10 * (1.0 + comp / r1 + mu / r2 + ssq + std::fabs(fend));
    auto g = [&](double t) { return (fend - fm) + D * t - 0.5 * K * t * t; };
    if (g(dl) > tol && g(T) > tol) return true;
    // Or the potential rises outward over [end - acc, end + acc] (|Phi''| <=
    // K): a step of acc from a best point within acc of the end to beyond
    // it is then worse by more than rounding, and dbrent returns at once.
    return -(D + K * acc) * acc > tol;
  }

  // (B) for one step that has (A): true when fblink is certain to be false.
  bool step_clear(double q, double spin, const Vec3& e, double rref) const {
    double lam1, lam2;
    if (!sphere_eclipse_los(e, r, c, rref, &lam1, &lam2)) return true;
    if (lam1 == 0.0) return false;
    if (lam2 - lam1 <= acc) return true;
    // fblink's first three potentials. A minimum bracketed by the chord's
    // midpoint keeps dbrent's best point at or below fmid; one bracketed
    // deeper in the grid search, below min(f1, f2).
    const double f1 = rpot_along_los<S>(q, spin, e, r, lam1);
    const double f2 = rpot_along_los<S>(q, spin, e, r, lam2);
    const double fmid = rpot_along_los<S>(q, spin, e, r, lam1 + (lam2 - lam1) / 2.0);
    const double v = fmid < f1 && fmid < f2 ? fmid : std::fmin(f1, f2);
    return end_clear(e, q, spin, lam1, 1.0, f1, v) && end_clear(e, q, spin, lam2, -1.0, f2, v);
  }
};

// crates/roche/src/eclipse.rs:347-373. The steps are visited in the
// reference's order and the first one whose fblink is true is returned;
// for steps that ScanBound certifies, dbrent's false is taken without
// running it.
template <Star S>
bool pot_min_search(double q, double cosi, double sini, double spin, const Vec3& r, double phi1,
                    double phi2, double acc, double rref, double pref, const ScanBound<S>& sb,
                    bool screen, double* phi_out) {
  size_t nsteps = as_usize(std::ceil((phi2 - phi1) / (acc / TWOPI)));
  nsteps = std::min<size_t>(std::max<size_t>(nsteps, 10), 1000);
  double dphi = (phi2 - phi1) / static_cast<double>(nsteps);
  // Steps < clear_end have (A). Blocks double while they clear, up to 16
  // steps (wider ones rarely clear: the fan of lines of sight gets too
  // wide); a failed block halves the size and caps it for the rest of the
  // scan, which runs toward the lobe.
  // After a single step fails, the next (A) attempts wait 1, 2, 4, ...
  // steps: near the lobe they keep failing.
  size_t blk = 1, cap = 16, clear_end = 0, retry_at = 0, wait = 1;
  for (size_t i = 0; i <= nsteps; ++i) {
    if (screen && i >= clear_end && i >= retry_at) {
      for (;;) {
        const size_t j = std::min(i + blk - 1, nsteps);
        // Steps i..j: phases phi1 + k dphi, within hw of the centre.
        const double ic = 0.5 * (static_cast<double>(i) + static_cast<double>(j));
        const double hw = 0.5 * static_cast<double>(j - i) * dphi;
        if (sb.phases_clear(phi1 + ic * dphi, hw, blk == 1 ? 4 : 12)) {
          clear_end = j + 1;
          blk = std::min<size_t>(blk * 2, cap);
          wait = 1;
          break;
        }
        if (blk == 1) {
          retry_at = i + wait;
          wait = std::min<size_t>(wait * 2, 16);
          break;
        }
        blk /= 2;
        cap = blk;
      }
    }
    double phi = phi1 + static_cast<double>(i) * dphi;
    Vec3 earth = set_earth(cosi, sini, phi);
    if (screen && i < clear_end && sb.step_clear(q, spin, earth, rref)) {
      ++tl_skipped;
      continue;
    }
    ++tl_evaluated;
    const bool hit = fblink_ref<S>(q, spin, acc, earth, r, rref, pref);
    if (hit) {
      *phi_out = phi;
      return true;
    }
  }
  return false;
}

// crates/roche/src/eclipse.rs:270-338
template <Star S>
bool ingress_egress_t(double q, double spin, double ffac, double iangle, double delta,
                      const Vec3& r, double* ingress_out, double* egress_out) {
  double rref, pref;
  ref_sphere(q, S, spin, ffac, &rref, &pref);
  double irad = to_radians(iangle);
  double cosi = std::cos(irad);
  double sini = std::sin(irad);
  Vec3 cofm = star_cofm(S);
  double phi1, phi2, lam1, lam2;
  if (!sphere_eclipse_phase(cosi, sini, r, cofm, rref, &phi1, &phi2, &lam1, &lam2)) return false;
  double acc = 2.0 * std::sqrt(2.0 * TWOPI * (lam2 - lam1) * delta);
  const ScanBound<S> sb(q, spin, cosi, sini, r, acc, rref, pref);
  const bool screen = tl_prescreen && sb.usable;
  double phi;
  if (!pot_min_search<S>(q, cosi, sini, spin, r, phi1, phi2, acc, rref, pref, sb, screen, &phi))
    return false;

  // The binary chops (eclipse.rs:302-332): their midpoints close in on the
  // lobe's limb, where nothing can be certified, so fblink runs on each.
  double pin = phi, pout = phi1;
  while (std::fabs(pin - pout) > delta) {
    double pmid = (pin + pout) / 2.0;
    Vec3 earth = set_earth(cosi, sini, pmid);
    if (fblink_ref<S>(q, spin, acc, earth, r, rref, pref)) pin = pmid;
    else pout = pmid;
  }
  double ingress = (pin + pout) / 2.0;
  ingress -= std::floor(ingress);

  pin = phi;
  pout = phi2;
  while (std::fabs(pin - pout) > delta) {
    double pmid = (pin + pout) / 2.0;
    Vec3 earth = set_earth(cosi, sini, pmid);
    if (fblink_ref<S>(q, spin, acc, earth, r, rref, pref)) pin = pmid;
    else pout = pmid;
  }
  double egress = (pin + pout) / 2.0;
  egress -= std::floor(egress);
  if (egress < ingress) egress += 1.0;
  *ingress_out = ingress;
  *egress_out = egress;
  return true;
}
}  // namespace

// crates/roche/src/eclipse.rs:123-197
bool fblink(double q, Star star, double spin, double ffac, double acc, const Vec3& earth,
            const Vec3& p) {
  check_q(q, "fblink");
  return star == Star::Primary ? fblink_t<Star::Primary>(q, spin, ffac, acc, earth, p)
                               : fblink_t<Star::Secondary>(q, spin, ffac, acc, earth, p);
}

bool set_eclipse_prescreen(bool on) {
  bool old = tl_prescreen;
  tl_prescreen = on;
  return old;
}

void eclipse_prescreen_counts(long long* skipped, long long* evaluated) {
  *skipped = tl_skipped;
  *evaluated = tl_evaluated;
}

bool ingress_egress(double q, Star star, double spin, double ffac, double iangle, double delta,
                    const Vec3& r, double* ingress_out, double* egress_out) {
  check_q(q, "ingress_egress");
  return star == Star::Primary
             ? ingress_egress_t<Star::Primary>(q, spin, ffac, iangle, delta, r, ingress_out, egress_out)
             : ingress_egress_t<Star::Secondary>(q, spin, ffac, iangle, delta, r, ingress_out,
                                                 egress_out);
}

// ------------------------------------------------------------ disc eclipse
namespace {
enum class Circle { Above, Inside, Outside, Separate, Crossing };

// crates/roche/src/disc_eclipse.rs:29-32
double cut_phase(double rxy, double rcone, double radius) {
  double cos_val = (sqr(rxy) + sqr(rcone) - sqr(radius)) / (2.0 * rcone * rxy);
  return std::acos(clampd(cos_val, -1.0, 1.0)) / TWOPI;
}

// crates/roche/src/disc_eclipse.rs:44-70
Circle circle_eclipse(double rxy, double z, double zcirc, double radius, double tani, double* phase) {
  *phase = 0.0;
  if (z >= zcirc) return Circle::Above;
  double rcone = tani * (zcirc - z);
  if (rcone >= rxy + radius) return Circle::Outside;
  if (rxy >= rcone + radius) return Circle::Separate;
  if (rxy + rcone <= radius) return Circle::Inside;
  *phase = cut_phase(rxy, rcone, radius);
  return Circle::Crossing;
}

// Applies one more rim circle to the running appear phase.
// Shape shared by disc_eclipse.rs:226-253 / 266-305.
void apply_rim(double rxy, double z, double zc, double rad, double tani, double* appear) {
  if (*appear > 0.0) {
    double ph;
    Circle res = circle_eclipse(rxy, z, zc, rad, tani, &ph);
    if (res == Circle::Crossing) *appear = fminr(*appear, ph);
    else if (res != Circle::Inside) *appear = -1.0;
  }
}
}  // namespace

// crates/roche/src/disc_eclipse.rs:84-330
EclipseVec disc_eclipse(double iangle, double rdisc1, double rdisc2, double beta, double height,
                        const Vec3& r) {
  EclipseVec temp;
  double hout = height * std::pow(rdisc2, beta);
  if (r.z >= hout) return temp;

  double irad = iangle * PI / 180.0;
  double sini = std::sin(irad);
  double cosi = std::cos(irad);

  if (cosi == 0.0) {
    if (std::fabs(r.z) < hout) {
      double rxy = std::sqrt(sqr(r.x) + sqr(r.y));
      if (rxy <= rdisc2) {
        temp.push_back({0.0, 1.0});
      } else {
        double subtend = std::asin(rdisc2 / rxy) / TWOPI;
        double pcen = std::atan2(r.y, -r.x) / TWOPI;
        double ingress = pcen - subtend;
        ingress = ingress - std::floor(ingress);
        double egress = ingress + 2.0 * subtend;
        temp.push_back({ingress, egress});
      }
    }
    return temp;
  }

  double rxy = std::sqrt(sqr(r.x) + sqr(r.y));
  if (rdisc1 < rxy && rxy < rdisc2 && std::fabs(r.z) < height * std::pow(rxy, beta)) {
    temp.push_back({0.0, 1.1});
    return temp;
  }

  double tani = sini / cosi;

  if (rxy < rdisc2 && r.z >= height * std::pow(fmaxr(rdisc1, rxy), beta)) {
    double phase;
    Circle result = circle_eclipse(rxy, r.z, hout, rdisc2, tani, &phase);
    if (result == Circle::Outside) {
      temp.push_back({0.0, 1.1});
    } else if (result == Circle::Crossing) {
      double phi0 = std::atan2(r.y, -r.x) / TWOPI;
      double ingress = phi0 + phase;
      ingress = ingress - std::floor(ingress);
      double egress = ingress + 1.0 - 2.0 * phase;
      temp.push_back({ingress, egress});
    }
    return temp;
  }

  double rcone_lo = fmaxr(tani * (-hout - r.z), 0.0);
  if (rcone_lo >= rxy + rdisc2) return temp;
  double rcone_hi = tani * (hout - r.z);
  if (rxy >= rcone_hi + rdisc2) return temp;

  double eclipse_phase;
  if (rxy + rcone_lo <= rdisc2) {
    eclipse_phase = 0.5;
  } else if (rxy <= rdisc2) {
    eclipse_phase = cut_phase(rxy, rcone_lo, rdisc2);
  } else {
    if (sqr(rcone_hi) + sqr(rdisc2) >= sqr(rxy) && sqr(rcone_lo) + sqr(rdisc2) <= sqr(rxy)) {
      eclipse_phase = std::asin(rdisc2 / rxy) / TWOPI;
    } else if (sqr(rcone_hi) + sqr(rdisc2) < sqr(rxy)) {
      eclipse_phase = cut_phase(rxy, rcone_hi, rdisc2);
    } else {
      eclipse_phase = cut_phase(rxy, rcone_lo, rdisc2);
    }
  }

  double hin = height * std::pow(rdisc1, beta);
  double appear_phase = -1.0;

  if (r.z < -hout) {
    double ph;
    Circle res = circle_eclipse(rxy, r.z, -hout, rdisc2, tani, &ph);
    if (res == Circle::Inside) appear_phase = 0.5;
    else if (res == Circle::Crossing) appear_phase = ph;
    apply_rim(rxy, r.z, -hin, rdisc1, tani, &appear_phase);
    apply_rim(rxy, r.z, hin, rdisc1, tani, &appear_phase);
    apply_rim(rxy, r.z, hout, rdisc2, tani, &appear_phase);
  } else if (rxy < rdisc1) {
    if (r.z < -hin) {
      double ph;
      Circle res = circle_eclipse(rxy, r.z, -hin, rdisc1, tani, &ph);
      if (res == Circle::Inside) appear_phase = 0.5;
      else if (res == Circle::Crossing) appear_phase = ph;
      apply_rim(rxy, r.z, hin, rdisc1, tani, &appear_phase);
      apply_rim(rxy, r.z, hout, rdisc2, tani, &appear_phase);
    } else if (r.z < hin) {
      double ph;
      Circle res = circle_eclipse(rxy, r.z, hin, rdisc1, tani, &ph);
      if (res == Circle::Inside) appear_phase = 0.5;
      else if (res == Circle::Crossing) appear_phase = ph;
      apply_rim(rxy, r.z, hout, rdisc2, tani, &appear_phase);
    }
  }

  double phi0 = std::atan2(r.y, -r.x) / TWOPI;
  if (appear_phase <= 0.0) {
    double ingress = phi0 - eclipse_phase;
    ingress = ingress - std::floor(ingress);
    double egress = ingress + 2.0 * eclipse_phase;
    temp.push_back({ingress, egress});
  } else if (appear_phase < eclipse_phase) {
    double ingress = phi0 - eclipse_phase;
    ingress = ingress - std::floor(ingress);
    double egress = ingress + (eclipse_phase - appear_phase);
    temp.push_back({ingress, egress});
    ingress = phi0 + appear_phase;
    ingress = ingress - std::floor(ingress);
    egress = ingress + (eclipse_phase - appear_phase);
    temp.push_back({ingress, egress});
  }
  return temp;
}

// ---------------------------------------------------------------- misc
// crates/roche/src/misc.rs:4-19
Vec3 rocacc(double q, const Vec3& r, const Vec3& v) {
  double f1 = 1.0 / (1.0 + q);
  double f2 = f1 * q;
  double yzsq = sqr(r.y) + sqr(r.z);
  double r1sq = sqr(r.x) + yzsq;
  double r2sq = sqr(r.x - 1.0) + yzsq;
  double fm1 = f1 / (r1sq * std::sqrt(r1sq));
  double fm2 = f2 / (r2sq * std::sqrt(r2sq));
  double fm3 = fm1 + fm2;
  return {-fm3 * r.x + fm2 + 2.0 * v.y + r.x - f2, -fm3 * r.y - 2.0 * v.x + r.y, -fm3 * r.z};
}

// crates/roche/src/misc.rs:22-29
double jacobi(double q, const Vec3& r, const Vec3& v) {
  double f1 = 1.0 / (1.0 + q);
  double f2 = f1 * q;
  double yzsq = sqr(r.y) + sqr(r.z);
  return (sqr(v.x) + sqr(v.y) + sqr(v.z) - sqr(r.y) - sqr(r.x - f2)) / 2.0 -
         f1 / std::sqrt(sqr(r.x) + yzsq) - f2 / std::sqrt(sqr(r.x - 1.0) + yzsq);
}

// ------------------------------------------------- Bulirsch-Stoer stepper
namespace {
constexpr int KMAXX = 8;
constexpr int IMAXX = KMAXX + 1;
constexpr int NSEQ[IMAXX] = {2, 4, 6, 8, 10, 12, 14, 16, 18};
constexpr double SAFE1 = 0.25, SAFE2 = 0.7, REDMAX = 1.0e-5, REDMIN = 0.7, TINY = 1.0e-30,
                 SCALMX = 0.1;

// crates/subs/src/numerical/bsstep.rs:15-41
struct BsState {
  bool first = true;
  int kmax = 0;
  int kopt = 0;
  double epsold = -1.0;
  double xnew = -1.0e29;
  double a[IMAXX] = {0};
  double alf[KMAXX][KMAXX] = {{0}};
};

constexpr int NV = 6;
using Deriv = void (*)(double q, const double* y, double* dydt);

// crates/roche/src/stream.rs:65-77
void rocder(double q, const double* y, double* dydt) {
  Vec3 r(y[0], y[1], y[2]);
  Vec3 v(y[3], y[4], y[5]);
  dydt[0] = v.x; dydt[1] = v.y; dydt[2] = v.z;
  Vec3 a = rocacc(q, r, v);
  dydt[3] = a.x; dydt[4] = a.y; dydt[5] = a.z;
}

// crates/subs/src/numerical/bsstep.rs:54-75
void mmid(const double* y, const double* dydx, double xs, double htot, int nstep, double* yout,
          double q) {
  double h = htot / static_cast<double>(nstep);
  double ym[NV], yn[NV];
  for (int i = 0; i < NV; ++i) {
    ym[i] = y[i];
    yn[i] = y[i] + h * dydx[i];
  }
  double x = xs + h;
  (void)x;
  rocder(q, yn, yout);
  double h2 = 2.0 * h;
  for (int n = 1; n < nstep; ++n) {
    for (int i = 0; i < NV; ++i) {
      double swap = ym[i] + h2 * yout[i];
      ym[i] = yn[i];
      yn[i] = swap;
    }
    x += h;
    rocder(q, yn, yout);
  }
  for (int i = 0; i < NV; ++i) yout[i] = 0.5 * (ym[i] + yn[i] + h * yout[i]);
}

// crates/subs/src/numerical/bsstep.rs:78-131
void pzextr(int iest, double xest, const double* yest, double* yz, double* dy,
            double d_tab[NV][KMAXX], double* x_tab) {
  x_tab[iest] = xest;
  for (int j = 0; j < NV; ++j) {
    dy[j] = yest[j];
    yz[j] = yest[j];
  }
  if (iest == 0) {
    for (int j = 0; j < NV; ++j) d_tab[j][0] = yest[j];
  } else {
    double c[NV];
    for (int j = 0; j < NV; ++j) c[j] = yest[j];
    for (int k1 = 0; k1 < iest; ++k1) {
      double delta = 1.0 / (x_tab[iest - k1 - 1] - xest);
      double f1 = xest * delta;
      double f2 = x_tab[iest - k1 - 1] * delta;
      for (int j = 0; j < NV; ++j) {
        double qv = d_tab[j][k1];
        d_tab[j][k1] = dy[j];
        double delta_val = c[j] - qv;
        dy[j] = f1 * delta_val;
        c[j] = f2 * delta_val;
        yz[j] += dy[j];
      }
    }
    for (int j = 0; j < NV; ++j) d_tab[j][iest] = dy[j];
  }
}

inline void chk(bool ok) {
  if (!ok) throw LcError("bsstep: extrapolation table index out of range");
}

// crates/subs/src/numerical/bsstep.rs:141-306 (including its quirks).
void bsstep(double* y, const double* dydx, double* xx, double htry, double eps,
            const double* yscal, BsState& st, double q, double* hdid_out, double* hnext_out,
            bool* underflow) {
  double yerr[NV], ysav[NV], yseq[NV];
  double err[KMAXX] = {0};
  double d_tab[NV][KMAXX] = {{0}};
  double x_tab[KMAXX] = {0};
  *underflow = false;

  if (eps != st.epsold) {
    st.xnew = -1.0e29;
    double eps1 = SAFE1 * eps;
    st.a[0] = static_cast<double>(NSEQ[0]) + 1.0;
    for (int k = 0; k < KMAXX; ++k) st.a[k + 1] = st.a[k] + static_cast<double>(NSEQ[k + 1]);
    for (int iq = 1; iq < KMAXX; ++iq) {
      for (int k = 0; k < iq; ++k) {
        st.alf[k][iq] = std::pow(eps1, (st.a[k + 1] - st.a[iq + 1]) /
                                            ((st.a[iq + 1] - st.a[0] + 1.0) *
                                             static_cast<double>(2 * k + 3)));
      }
    }
    st.epsold = eps;
    st.kopt = 1;
    for (int kopt = 1; kopt < KMAXX - 1; ++kopt) {
      if (st.a[kopt + 1] > st.a[kopt] * st.alf[kopt - 1][kopt]) {
        st.kopt = kopt;
        break;
      }
      st.kopt = kopt;
    }
    st.kmax = st.kopt;
  }

  double h = htry;
  for (int i = 0; i < NV; ++i) ysav[i] = y[i];
  if (*xx != st.xnew || h != st.xnew) {
    st.first = true;
    st.kopt = st.kmax;
  }

  bool reduct = false;
  int km = 0;
  for (;;) {
    bool exitflag = false;
    for (int k = 0; k <= st.kmax; ++k) {
      st.xnew = *xx + h;
      if (st.xnew == *xx) {
        *hdid_out = 0.0;
        *hnext_out = 0.0;
        *underflow = true;
        return;
      }
      mmid(ysav, dydx, *xx, h, NSEQ[k], yseq, q);
      double xest = sqr(h / static_cast<double>(NSEQ[k]));
      pzextr(k, xest, yseq, y, yerr, d_tab, x_tab);
      if (k != 0) {
        double errmax = TINY;
        for (int i = 0; i < NV; ++i) errmax = fmaxr(errmax, std::fabs(yerr[i] / yscal[i]));
        errmax /= eps;
        km = k - 1;
        err[km] = std::pow(errmax / SAFE1, 1.0 / static_cast<double>(2 * km + 3));
      }
      int koptm1 = st.kopt > 0 ? st.kopt - 1 : 0;
      if (k != 0 && (k >= koptm1 || st.first)) {
        if (err[km] != 0.0) {
          double errmax = TINY;
          for (int i = 0; i < NV; ++i) errmax = fmaxr(errmax, std::fabs(yerr[i] / yscal[i]));
          errmax /= eps;
          if (errmax < 1.0) {
            exitflag = true;
            break;
          }
        }
        if (k == st.kmax || k == st.kopt + 1) break;
        // Same short-circuit order as the reference; indices are checked
        // only where the reference would index.
        if (k == st.kopt) {
          chk(st.kopt >= 1 && st.kopt < KMAXX);
          if (st.alf[st.kopt - 1][st.kopt] < err[km]) break;
        }
        if (st.kopt == st.kmax) {
          chk(st.kmax >= 1 && st.kmax - 1 < KMAXX);
          if (st.alf[km][st.kmax - 1] < err[km]) break;
        }
        chk(st.kopt < KMAXX);
        if (st.alf[km][st.kopt] < err[km]) break;
      }
    }
    if (exitflag) break;
    double red = err[km] > 0.0 ? fminr(fmaxr(SAFE2 / err[km], REDMAX), REDMIN) : REDMIN;
    h *= red;
    reduct = true;
  }

  *xx = st.xnew;
  double hdid = h;
  st.first = false;

  double wrkmin = 1.0e35;
  double scale = 1.0;
  for (int kk = 0; kk <= km; ++kk) {
    double fact = fmaxr(err[kk], SCALMX);
    double work = fact * st.a[kk + 1];
    if (work < wrkmin) {
      scale = fact;
      wrkmin = work;
      st.kopt = kk + 1;
    }
  }
  double hnext = h / scale;
  int k_final = km + 1;
  if (st.kopt >= k_final && st.kopt != st.kmax && !reduct) {
    chk(st.kopt >= 1 && st.kopt + 1 < IMAXX);
    double fact = fmaxr(scale / st.alf[st.kopt - 1][st.kopt], SCALMX);
    if (st.a[st.kopt + 1] * fact <= wrkmin) {
      hnext = h / fact;
      st.kopt += 1;
    }
  }
  *hdid_out = hdid;
  *hnext_out = hnext;
}

// crates/roche/src/stream.rs:26-62
void gsint(double q, Vec3* r, Vec3* v, double ttry, double* time, double eps, BsState& st,
           double* tdid, double* tnext) {
  double y[NV] = {r->x, r->y, r->z, v->x, v->y, v->z};
  double dydt[NV];
  rocder(q, y, dydt);
  const double yscal[NV] = {1, 1, 1, 1, 1, 1};
  bool underflow;
  bsstep(y, dydt, time, ttry, eps, yscal, st, q, tdid, tnext, &underflow);
  *r = Vec3(y[0], y[1], y[2]);
  *v = Vec3(y[3], y[4], y[5]);
}
}  // namespace

// crates/roche/src/stream.rs:8-20
void strinit(double q, Vec3* r, Vec3* v) {
  const double SMALL = 1.0e-5;
  double rl1 = xl1(q);
  double mu = q / (1.0 + q);
  double a = (1.0 - mu) / powi3(rl1) + mu / powi3(1.0 - rl1);
  double lambda1 = std::sqrt(((a - 2.0) + std::sqrt(a * (9.0 * a - 8.0))) / 2.0);
  double m1 = (lambda1 * lambda1 - 2.0 * a - 1.0) / (2.0 * lambda1);
  *r = Vec3(rl1 - SMALL, -m1 * SMALL, 0.0);
  *v = Vec3(-lambda1 * SMALL, -lambda1 * m1 * SMALL, 0.0);
}

// crates/roche/src/stream.rs:82-145
double stradv(double q, Vec3* r, Vec3* v, double rad, double acc, double smax) {
  const double EPS = 1.0e-8;
  const double TMAX = 10.0;
  BsState st;
  double rinit = r->length();
  double rnow = rinit;
  double time = 0.0;
  double tnext = 1.0e-2;
  // Guard against a stalled integrator (the reference loops until TMAX).
  for (long iter = 0; iter < 100000000L; ++iter) {
    Vec3 ro = *r, vo = *v;
    double ttry = fminr(tnext, smax);
    double tdid, tn;
    gsint(q, r, v, ttry, &time, EPS, st, &tdid, &tn);
    tnext = tn;
    rnow = r->length();
    if ((rinit > rad && rnow <= rad) || (rinit < rad && rnow >= rad)) {
      double lo = 0.0, hi = tdid;
      double rlo = ro.length(), rhi = rnow;
      double to = time - tdid;
      while (std::fabs(rhi - rlo) > acc) {
        double tmid = (lo + hi) / 2.0;
        *r = ro;
        *v = vo;
        double t = to;
        BsState st2;
        double d1, d2;
        gsint(q, r, v, tmid, &t, EPS, st2, &d1, &d2);
        rnow = r->length();
        if ((rhi > rad && rnow > rad) || (rhi < rad && rnow < rad)) {
          rhi = rnow;
          hi = tmid;
        } else {
          rlo = rnow;
          lo = tmid;
        }
      }
      return time;
    }
    if (time > TMAX) throw LcError("stradv: took too long without crossing radius");
  }
  throw LcError("stradv: integrator stalled");
}

}  // namespace lc
