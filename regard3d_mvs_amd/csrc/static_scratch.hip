// Fixed device scratch for the entry points that allocate nothing.
//
// `scratch` is a file-scope device array: every GPU the library runs on holds
// its own copy, at its own address, until the process ends.  Host code never
// passes the array's name to a kernel or a copy (on the host it names a
// stand-in, not device memory): the device address is looked up once per
// device with hipGetSymbolAddress, with that device current, and kept.
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <mutex>

#include "kernels.hpp"

namespace r3dg {
namespace {

// Union-find parents and component sizes of the speckle filter (int32 each).
static __device__ __align__(256) std::int32_t scratch[2 * kStaticSpecklePixels];

constexpr int kMaxDevices = 64;
std::array<std::int32_t*, kMaxDevices> g_addr{};
std::array<std::once_flag, kMaxDevices> g_once;
std::array<hipError_t, kMaxDevices> g_status{};
std::array<std::mutex, kMaxDevices> g_use;

}  // namespace

hipError_t static_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::call_once(g_once[std::size_t(device)], [device] {
    void* p = nullptr;
    // The caller made `device` current.
    g_status[std::size_t(device)] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (g_status[std::size_t(device)] == hipSuccess) g_addr[std::size_t(device)] = static_cast<std::int32_t*>(p);
  });
  return g_status[std::size_t(device)];
}

std::int32_t* static_scratch(int device) {
  return device >= 0 && device < kMaxDevices ? g_addr[std::size_t(device)] : nullptr;
}

std::mutex& static_scratch_mutex(int device) { return g_use[std::size_t(device) % kMaxDevices]; }

}  // namespace r3dg
