// Census transform of the r3d API (reference: src/r3d/stereo/census.hpp:13-29),
// computed on the GPU.
#pragma once

#include <cstddef>
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

// Row kernel: rows[j] points to row y - 3 + j of an image padded by 4
// replicated pixels on both sides, at the padded start (pixel x of window row
// j is rows[j][x + 4]); writes out[0 .. n).  scalar::census_row is the same
// GPU kernel (the reference's scalar twin gives the same bits).
void census_row(const std::uint16_t* const rows[kCensusHeight], std::uint64_t* out, std::size_t n);

namespace scalar {
void census_row(const std::uint16_t* const rows[kCensusHeight], std::uint64_t* out, std::size_t n);
}  // namespace scalar

}  // namespace r3d
