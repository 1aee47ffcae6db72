// Static device scratch of the entry points that allocate nothing.
#pragma once

#include <cstddef>
#include <mutex>

#include "../../../include/lcurve_hip.h"

namespace lch {

constexpr size_t kScratchBytes = static_cast<size_t>(LCH_STATIC_SCRATCH_BYTES);

// Device address of the scratch array on `device`, found once per device
// with that device current (hipGetSymbolAddress). Returns nullptr and sets
// *err when the runtime cannot give it.
void* scratch_device_ptr(int device, const char** err);
// Serialises the calls that use `device`'s scratch within this process.
std::mutex& scratch_lock(int device);

}  // namespace lch
