// Target selection of the r3d API (reference: src/r3d/core/simd.hpp:8-31).
// The reference dispatches each kernel over the CPU SIMD targets compiled in;
// this build has one target, the gfx950 GPU, so the list holds that one and
// ScopedSimdTarget only accepts it.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace r3d {

[[nodiscard]] std::string simd_active_target_name();
[[nodiscard]] std::vector<std::int64_t> simd_available_targets();
[[nodiscard]] std::string simd_target_name(std::int64_t target);

class ScopedSimdTarget {
 public:
  explicit ScopedSimdTarget(std::int64_t target);
  ScopedSimdTarget(const ScopedSimdTarget&) = delete;
  ScopedSimdTarget& operator=(const ScopedSimdTarget&) = delete;
  ~ScopedSimdTarget();
};

}  // namespace r3d
