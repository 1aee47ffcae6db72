// CPU self-check of the work codes that let the flux kernel skip patches,
// drop eclipse tests and sum whole patches from their moments.
#pragma once

#include "batch_build.h"

namespace lch {

struct SelfCheck {
  double max_rel = 0.0;          // max |coded - brute| / sum |terms| per (lane, segment)
  long long counts[4] = {0, 0, 0, 0};  // (patch, tile) pairs per code PT_FULL..PT_MOM
  long long lane_segments = 0;
};

SelfCheck selfcheck_codes(const SetPack& S);
SelfCheck selfcheck_model(const lc::Model& model, const lc::Data& data);

}  // namespace lch
