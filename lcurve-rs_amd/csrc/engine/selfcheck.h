// CPU self-check of the work codes that let the flux kernel skip patches,
// drop eclipse tests and sum whole patches from their moments.
#pragma once

#include "batch_build.h"

namespace lch {

struct SelfCheck {
  double max_rel = 0.0;          // max |coded - brute| / sum |terms| per (lane, segment)
  long long counts[4] = {0, 0, 0, 0};  // (patch, tile) pairs per code PT_FULL..PT_MOM
  long long lane_segments = 0;
  // Lanes of split PT_FREE pairs: seen whole (moments), in part (element
  // by element), not at all.
  long long split_lanes[3] = {0, 0, 0};
};

SelfCheck selfcheck_codes(const SetPack& S);
SelfCheck selfcheck_model(const lc::Model& model, const lc::Data& data);

// Star grids (fine and coarse, both stars) built with and without the
// certified pre-screen of the eclipse phase scan, compared bit for bit.
struct PrescreenCheck {
  long long elements = 0;    // grid elements compared
  long long mismatches = 0;  // elements whose position, normal, area, gravity or eclipses differ
  long long skipped = 0;     // scan steps the pre-screen certified
  long long evaluated = 0;   // scan steps run through fblink
  double ms_plain = 0.0, ms_screened = 0.0;
};
PrescreenCheck selfcheck_prescreen(const lc::Model& model);

}  // namespace lch
