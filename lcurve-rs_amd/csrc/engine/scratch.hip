// Static device scratch of the entry points that allocate nothing
// (lch_flux_eval_static). Each GPU the library runs on holds its own copy of
// the array until the process ends; host code reaches it only through the
// per-device address hipGetSymbolAddress returns (the array's name is a
// host-side stand-in, not a device address).
#include <hip/hip_runtime.h>

#include "scratch.h"

namespace lch {

static __device__ __align__(256) char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
std::mutex g_call_mu[kMaxDevices];
}  // namespace

void* scratch_device_ptr(int device, const char** err) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (device < 0 || device >= kMaxDevices) {
    *err = "device id out of range for the static scratch";
    return nullptr;
  }
  if (!g_ptr[device]) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != device) {
      *err = "the device of the static scratch is not current";
      return nullptr;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *err = hipGetErrorString(e);
      return nullptr;
    }
    g_ptr[device] = p;
  }
  return g_ptr[device];
}

std::mutex& scratch_lock(int device) { return g_call_mu[device < 0 || device >= kMaxDevices ? 0 : device]; }

}  // namespace lch
