// Gaussian pyramid reduction of the r3d API for 8- and 16-bit images
// (reference: src/r3d/core/pyramid.hpp:20), computed on the GPU:
// [1 4 6 4 1] / 16 separable filter, replicated borders, 2x decimation,
// exact integer rounding (sum + 128) >> 8; ceil(w / 2) x ceil(h / 2).
#pragma once

#include <cstdint>

#include "r3d/core/image.hpp"

namespace r3d {

[[nodiscard]] Image<std::uint8_t> pyramid_down(const ImageView<const std::uint8_t>& src);
[[nodiscard]] Image<std::uint16_t> pyramid_down(const ImageView<const std::uint16_t>& src);

template <typename T>
[[nodiscard]] Image<T> pyramid_down(const Image<T>& src) {
  return pyramid_down(src.view());
}

}  // namespace r3d
