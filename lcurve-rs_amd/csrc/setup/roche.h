// Roche geometry used by grid setup: Lagrange points, potentials and their
// gradients, equipotential surface search, eclipse phases by a Roche lobe,
// a sphere or a flared disc, and the gas-stream integration that places the
// bright spot. Restated from crates/roche/src and crates/subs/src/numerical
// with the reference's floating-point operation order.
#pragma once

#include "lc_core.h"

namespace lc {

enum class Star { Primary, Secondary };

// crates/roche/src/lagrange.rs
double xl1(double q);
double xl11(double q, double spin);
double xl12(double q, double spin);

// crates/roche/src/potential.rs
double rpot(double q, const Vec3& p);
double rpot1(double q, double spin, const Vec3& p);
double rpot2(double q, double spin, const Vec3& p);
Vec3 drpot1(double q, double spin, const Vec3& p);
Vec3 drpot2(double q, double spin, const Vec3& p);

// crates/roche/src/surface.rs
void ref_sphere(double q, Star star, double spin, double ffac, double* rref, double* pref);
// Returns surface position, outward normal, radius and gravity.
void face(double q, Star star, double spin, const Vec3& dirn, double rref, double pref, double acc,
          Vec3* pvec, Vec3* dvec, double* radius, double* gravity);

// crates/roche/src/eclipse.rs
bool sphere_eclipse_phase(double cosi, double sini, const Vec3& r, const Vec3& c, double rsphere,
                          double* phi1, double* phi2, double* lam1, double* lam2);
bool fblink(double q, Star star, double spin, double ffac, double acc, const Vec3& earth,
            const Vec3& p);
bool ingress_egress(double q, Star star, double spin, double ffac, double iangle, double delta,
                    const Vec3& r, double* ingress, double* egress);

// Certified short cuts: the phase scan of ingress_egress skips steps whose
// line of sight provably misses the eclipsing lobe (a certified lower bound
// of the potential lies above the lobe's potential), and face()'s chop
// decides midpoints from a certified Taylor bound. Results are identical
// with and without them; the switch is per thread, for self-checks.
// Returns the previous setting.
bool set_eclipse_prescreen(bool on);
// Scan steps skipped / scan steps run through fblink on this thread.
void eclipse_prescreen_counts(long long* skipped, long long* evaluated);

// crates/roche/src/disc_eclipse.rs
EclipseVec disc_eclipse(double iangle, double rdisc1, double rdisc2, double beta, double height,
                        const Vec3& r);

// crates/roche/src/stream.rs
void strinit(double q, Vec3* r, Vec3* v);
double stradv(double q, Vec3* r, Vec3* v, double rad, double acc, double smax);

// crates/roche/src/misc.rs
Vec3 rocacc(double q, const Vec3& r, const Vec3& v);
double jacobi(double q, const Vec3& r, const Vec3& v);

}  // namespace lc
