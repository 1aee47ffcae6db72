// Model parameter file: 77 physical parameters with fitting metadata and
// 34 computational parameters. Restated from crates/lcurve/src/model.rs.
#pragma once

#include <string>
#include <vector>

#include "lc_core.h"

namespace lc {

// crates/lcurve/src/model.rs:6-39
struct Pparam {
  double value = 0.0, range = 0.0, dstep = 0.0;
  bool vary = false, defined = false;
};

// The physical parameter names, in the reference's declaration order
// (crates/lcurve/src/model.rs:194-214).
#define LCH_PPARAM_LIST(X)                                                                          \
  X(q) X(iangle) X(r1) X(r2) X(cphi3) X(cphi4) X(spin1) X(spin2) X(t1) X(t2) X(ldc1_1) X(ldc1_2) \
  X(ldc1_3) X(ldc1_4) X(ldc2_1) X(ldc2_2) X(ldc2_3) X(ldc2_4) X(velocity_scale) X(beam_factor1)   \
  X(beam_factor2) X(t0) X(period) X(pdot) X(deltat) X(gravity_dark1) X(gravity_dark2) X(absorb)    \
  X(slope) X(quad) X(cube) X(third) X(rdisc1) X(rdisc2) X(height_disc) X(beta_disc) X(temp_disc)   \
  X(texp_disc) X(lin_limb_disc) X(quad_limb_disc) X(temp_edge) X(absorb_edge) X(radius_spot)       \
  X(length_spot) X(height_spot) X(expon_spot) X(epow_spot) X(angle_spot) X(yaw_spot) X(temp_spot)  \
  X(tilt_spot) X(cfrac_spot) X(stsp11_long) X(stsp11_lat) X(stsp11_fwhm) X(stsp11_tcen)           \
  X(stsp12_long) X(stsp12_lat) X(stsp12_fwhm) X(stsp12_tcen) X(stsp13_long) X(stsp13_lat)          \
  X(stsp13_fwhm) X(stsp13_tcen) X(stsp21_long) X(stsp21_lat) X(stsp21_fwhm) X(stsp21_tcen)         \
  X(stsp22_long) X(stsp22_lat) X(stsp22_fwhm) X(stsp22_tcen) X(uesp_long1) X(uesp_long2)           \
  X(uesp_lathw) X(uesp_taper) X(uesp_temp)

// crates/lcurve/src/model.rs:65-182
struct Model {
#define LCH_DECL(n) Pparam n;
  LCH_PPARAM_LIST(LCH_DECL)
#undef LCH_DECL

  double delta_phase = 0;
  int nlat1f = 0, nlat2f = 0, nlat1c = 0, nlat2c = 0;
  bool npole = false;
  int nlatfill = 0, nlngfill = 0;
  double lfudge = 0, llo = 0, lhi = 0, phase1 = 0, phase2 = 0, wavelength = 0;
  bool roche1 = false, roche2 = false, eclipse1 = false, eclipse2 = false, glens1 = false;
  bool use_radii = false;
  double tperiod = 0;
  bool gdark_bolom1 = false, gdark_bolom2 = false;
  double mucrit1 = 0, mucrit2 = 0;
  LdcType limb1 = LdcType::Poly, limb2 = LdcType::Poly;
  bool mirror = false, add_disc = false;
  int nrad = 0;
  bool opaque = false, add_spot = false;
  int nspot = 0;
  bool iscale = false;

  // Parse a parameter file (model.rs:186-442) or the same text from memory.
  static Model from_file(const std::string& path);
  static Model from_text(const std::string& text);

  // model.rs:445-454
  void get_r1r2(double* r1, double* r2) const;
  // model.rs:459-493: returns false for an unknown name.
  bool set_param_value(const std::string& name, double value);
  // model.rs:498-532: returns nullptr for an unknown name.
  const Pparam* pparam(const std::string& name) const;
  Pparam* pparam(const std::string& name);
  LDC get_ldc1() const;
  LDC get_ldc2() const;
};

// Names of the physical parameters in declaration order.
const std::vector<std::string>& pparam_names();

}  // namespace lc
