// Core numeric types of the light-curve setup: 3-vectors, physical
// constants, the Planck function, surface elements and the small helper
// types (limb darkening, grid interpolation, data points).
//
// Every routine here is a restatement of the reference's Rust code and keeps
// its floating-point operation order, because the setup output (element
// directions, fluxes and eclipse phases) feeds the bit-exact FP64 visibility
// test on the GPU. Build with -ffp-contract=off so no FMA is fused in.
//
// Reference: crates/subs/src/{vec3,constants,planck}.rs,
//            crates/lcurve/src/types.rs
#pragma once

#include <cmath>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace lc {

// Error raised by setup code; the C ABI turns it into an error code.
struct LcError : public std::runtime_error {
  explicit LcError(const std::string& m) : std::runtime_error(m) {}
};

// ---------------------------------------------------------------- constants
// crates/subs/src/constants.rs:5-95
constexpr double C_LIGHT = 2.99792458e8;   // m/s
constexpr double H_PLANCK = 6.6262e-34;    // MKS
constexpr double K_BOLTZ = 1.3806e-23;     // MKS
constexpr double DAY = 86400.0;
constexpr double PI = 3.14159265358979323846;
constexpr double TWOPI = 2.0 * PI;

inline double sqr(double x) { return x * x; }
// Rust f64::powi with a constant exponent expands to repeated squaring.
inline double powi3(double x) { return x * (x * x); }
inline double powi4(double x) { double x2 = x * x; return x2 * x2; }
// Rust f64::to_radians / to_degrees.
inline double to_radians(double d) { return d * (PI / 180.0); }
inline double to_degrees(double r) { return r * (180.0 / PI); }
// Rust f64::clamp (NaN passes through).
inline double clampd(double x, double lo, double hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
// Rust f64::max / f64::min (IEEE maxNum / minNum).
inline double fmaxr(double a, double b) { return std::fmax(a, b); }
inline double fminr(double a, double b) { return std::fmin(a, b); }
// Rust `x as i32`: saturating, NaN -> 0.
inline int as_i32(double x) {
  if (std::isnan(x)) return 0;
  if (x >= 2147483647.0) return 2147483647;
  if (x <= -2147483648.0) return -2147483647 - 1;
  return static_cast<int>(x);
}
// Rust `x as usize`: saturating at 0, NaN -> 0.
inline size_t as_usize(double x) {
  if (std::isnan(x) || x <= 0.0) return 0;
  if (x >= 1.8446744073709552e19) return SIZE_MAX;
  return static_cast<size_t>(x);
}

// -------------------------------------------------------------------- Vec3
// crates/subs/src/vec3.rs:7-188
struct Vec3 {
  double x = 0.0, y = 0.0, z = 0.0;
  Vec3() = default;
  Vec3(double x_, double y_, double z_) : x(x_), y(y_), z(z_) {}
  double sqr() const { return x * x + y * y + z * z; }
  double length() const { return std::sqrt(sqr()); }
  // Normalise in place; returns false for the null vector (left unchanged).
  bool unit() {
    double norm = sqr();
    if (norm > 0.0) {
      norm = std::sqrt(norm);
      x /= norm; y /= norm; z /= norm;
      return true;
    }
    return false;
  }
};
inline Vec3 operator+(const Vec3& a, const Vec3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator-(const Vec3& a, const Vec3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator*(double s, const Vec3& v) { return {s * v.x, s * v.y, s * v.z}; }
inline Vec3 operator*(const Vec3& v, double s) { return {v.x * s, v.y * s, v.z * s}; }
inline double dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// ----------------------------------------------------------------- Planck
// crates/subs/src/planck.rs:5-43
constexpr double PLANCK_FAC1 = 2.0e27 * H_PLANCK * C_LIGHT;
constexpr double PLANCK_FAC2 = 1.0e9 * H_PLANCK * C_LIGHT / K_BOLTZ;

inline double planck(double wave, double temp) {
  double efac = PLANCK_FAC2 / (wave * temp);
  if (efac > 40.0) return PLANCK_FAC1 * std::exp(-efac) / (wave * sqr(wave));
  return PLANCK_FAC1 / (std::exp(efac) - 1.0) / (wave * sqr(wave));
}
// d ln B_nu / d ln T.
inline double dlpdlt(double wave, double temp) {
  double efac = PLANCK_FAC2 / (wave * temp);
  return efac / (1.0 - std::exp(-efac));
}

// ------------------------------------------------------------------ Point
// crates/lcurve/src/types.rs:4-59
using EclipseVec = std::vector<std::pair<double, double>>;

struct Point {
  Vec3 posn;
  Vec3 dirn;
  float area = 0.0f;
  float gravity = 1.0f;
  EclipseVec eclipse;
  float flux = 0.0f;

  Point() = default;
  Point(const Vec3& p, const Vec3& d, double a, double g, EclipseVec e)
      : posn(p), dirn(d), area(static_cast<float>(a)),
        gravity(static_cast<float>(g)), eclipse(std::move(e)), flux(0.0f) {}

  // types.rs:50-58 — FP64 eclipse-interval test.
  bool visible(double phase) const {
    double phi = phase - std::floor(phase);
    for (const auto& pr : eclipse) {
      if ((phi >= pr.first && phi <= pr.second) || phi <= pr.second - 1.0) return false;
    }
    return true;
  }
};

// ----------------------------------------------------------- limb darkening
// crates/lcurve/src/types.rs:61-126
enum class LdcType { Poly = 0, Claret = 1 };

struct LDC {
  double ldc1 = 0, ldc2 = 0, ldc3 = 0, ldc4 = 0, mucrit = 0;
  LdcType ltype = LdcType::Poly;

  double imu(double mu) const {
    if (mu <= 0.0) return 0.0;
    mu = fminr(mu, 1.0);
    double ommu = 1.0 - mu;
    double im = 1.0;
    if (ltype == LdcType::Poly) {
      im -= ommu * (ldc1 + ommu * (ldc2 + ommu * (ldc3 + ommu * ldc4)));
    } else {
      im -= ldc1 + ldc2 + ldc3 + ldc4;
      double msq = std::sqrt(mu);
      im += msq * (ldc1 + msq * (ldc2 + msq * (ldc3 + msq * ldc4)));
    }
    return im;
  }
  bool see(double mu) const { return mu > mucrit; }
};

// ------------------------------------------------- fine/coarse interpolation
// crates/lcurve/src/types.rs:128-203
struct Ginterp {
  double phase1 = 0.05, phase2 = 0.45;
  double scale11 = 1.0, scale12 = 1.0, scale21 = 1.0, scale22 = 1.0;

  double scale1(double phase) const {
    double pnorm = phase - std::floor(phase);
    if (pnorm <= phase1 || pnorm >= 1.0 - phase1) return 1.0;
    return (scale11 * (1.0 - phase1 - pnorm) + scale12 * (pnorm - phase1)) / (1.0 - 2.0 * phase1);
  }
  double scale2(double phase) const {
    double pnorm = phase - std::floor(phase);
    if (pnorm >= phase2 && pnorm <= 1.0 - phase2) return 1.0;
    if (pnorm < 0.5)
      return (scale22 * (phase2 - pnorm) + scale21 * (pnorm + phase2)) / (2.0 * phase2);
    return (scale21 * (1.0 + phase2 - pnorm) + scale22 * (pnorm - 1.0 + phase2)) / (2.0 * phase2);
  }
  // 1 = fine star1 / coarse star2, 2 = coarse both, 3 = coarse star1 / fine star2
  int grid_type(double phase) const {
    double pnorm = phase - std::floor(phase);
    if (pnorm <= phase1 || pnorm >= 1.0 - phase1) return 1;
    if ((pnorm > phase1 && pnorm < phase2) || (pnorm > 1.0 - phase2 && pnorm < 1.0 - phase1)) return 2;
    return 3;
  }
};

// ------------------------------------------------------------------- Datum
// crates/lcurve/src/types.rs:205-265
struct Datum {
  double time = 0, expose = 0, flux = 0, ferr = 0, weight = 0;
  int ndiv = 1;
};
using Data = std::vector<Datum>;

Data read_data(const std::string& path);

// Sub-phase and trapezoid weight for exposure smearing.
// crates/lcurve/src/flux.rs:378-387
inline void exposure_weight(double phase, double expose, int nd, int ndiv, double* phi, double* wgt) {
  if (ndiv == 1) {
    *phi = phase;
    *wgt = 1.0;
  } else {
    *phi = phase + expose * (static_cast<double>(nd) - static_cast<double>(ndiv - 1) / 2.0) /
                       static_cast<double>(ndiv - 1);
    *wgt = (nd == 0 || nd == ndiv - 1) ? 0.5 : 1.0;
  }
}

// Earth-pointing unit vector. crates/roche/src/misc.rs:42-45
inline Vec3 set_earth(double cosi, double sini, double phase) {
  double phi = TWOPI * phase;
  return {sini * std::cos(phi), -sini * std::sin(phi), cosi};
}

}  // namespace lc
