// Per-model CPU setup: everything light_curve_comp does before its per-time
// flux loop (crates/lcurve/src/orchestration.rs:36-179), plus the per-time
// phases and polynomial factors (orchestration.rs:172-241). The Ginterp
// scale factors (orchestration.rs:85-131) are not computed here: they are
// flux sums, evaluated on the GPU from the calibration grids kept below.
#pragma once

#include <memory>

#include "grid.h"

namespace lc {

struct PreparedModel {
  // Star grids. When a coarse grid equals its fine grid (the reference
  // clones it) the *_alias flag is set and only the fine grid is kept.
  std::vector<Point> star1f, star1c, star2f, star2c, disc, edge, spot;
  bool star1_alias = true;  // nlat1f == nlat1c
  bool star2_alias = true;  // `copy2` in orchestration.rs:72-73
  bool need_cal1 = false;   // scale11/scale12 must be computed
  bool need_cal2 = false;   // scale21/scale22 must be computed
  // The reference computes the Ginterp scale factors before it appends the
  // opaque-disc eclipses to the star grids (orchestration.rs:93-157). When
  // both happen, the calibration sums use these copies taken before the
  // disc eclipses were added.
  bool cal_copies = false;
  std::vector<Point> star1f_cal, star1c_cal, star2f_cal, star2c_cal;

  // Flux-path scalars.
  LDC ldc1, ldc2;
  double lin_limb_disc = 0, quad_limb_disc = 0;
  double beam_factor1 = 0, beam_factor2 = 0, spin1 = 1, spin2 = 1;
  double vfac = 0;   // vscale (rounded to f32) / (c / 1e3), flux.rs:22
  double xcofm = 0;  // q / (1 + q)
  bool glens1 = false;
  double rlens1 = 0;
  double phase1 = 0.05, phase2 = 0.45;  // Ginterp switch phases
  double iangle = 0;
  double cosi_lc = 0, sini_lc = 0;    // cos/sin(TWOPI*iangle/360), flux.rs:20-21
  double cosi_rad = 0, sini_rad = 0;  // iangle.to_radians(), flux.rs:121-123
  double third = 0;
  // Per-component scale factors (orchestration.rs:181-227): the curve is
  // sum_j sfac[j] * column j over star 1, disc, edge, spot and, when
  // t2 > 0 (ncol = 5), star 2.
  bool iscale = false;
  int ncol = 5;

  // Per-curve scalars computed on the CPU (flux.rs:282-375).
  double logg1 = 0, logg2 = 0, rv1 = 0, rv2 = 0;

  // Per-time inputs of the flux kernel.
  std::vector<double> phase;      // orbital phase incl. pdot and Roemer terms
  std::vector<double> expose;     // exposure in phase units
  std::vector<int> ndiv;
  std::vector<double> slfac;      // 1 + frac*(slope + frac*(quad + frac*cube))

  size_t n_elements() const;
};

// Runs the CPU setup for one model against the data times. Throws LcError
// where the reference returns Err (e.g. a star larger than its Roche lobe).
std::unique_ptr<PreparedModel> prepare_model(const Model& mdl, const Data& data);

// Orbital phase for one time (orchestration.rs:231-237).
double time_to_phase(const Model& mdl, double time);

}  // namespace lc
