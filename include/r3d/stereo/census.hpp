// Census transform of the r3d API (reference: src/r3d/stereo/census.hpp:13-20),
// computed on the GPU.
#pragma once

#include <cstdint>

#include "r3d/core/image.hpp"

namespace r3d {

inline constexpr int kCensusWidth = 9;
inline constexpr int kCensusHeight = 7;
inline constexpr int kCensusBits = kCensusWidth * kCensusHeight - 1;

// 9x7 Census, 62 bits, first raster neighbour in bit 61, replicated borders.
// `scalar_reference` selects a CPU kernel in the reference; the GPU result is
// the same either way.
[[nodiscard]] Image<std::uint64_t> census_transform(ImageView<const std::uint16_t> img,
                                                    bool scalar_reference = false);

}  // namespace r3d
