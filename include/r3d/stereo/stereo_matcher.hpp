// r3d stereo matcher API, served by the MI355X (gfx950) implementation.
//
// Same declarations as the reference's src/r3d/stereo/stereo_matcher.hpp:23-130
// (types, defaults and functions), so the dense pipeline, bench_stereo and the
// stereo tests compile unchanged against this header and link libr3d_stereo.so.
// Every function runs on the process-wide GPU context (device 0, or the one
// set with set_stereo_device) through the C ABI of r3dg.h; errors surface as
// r3d::Error, cancellation as r3d::Cancelled.
#pragma once

#include <cstddef>
#include <cstdint>
#include <limits>
#include <vector>

#include "r3d/core/image.hpp"

namespace r3d {

enum class SubpixelMethod {
  none,
  parabola,     // fit through the aggregated costs at d - 1, d, d + 1
  equiangular,  // symmetric V fit (Shimizu & Okutomi 2001)
};

// stereo_matcher.hpp:30-57 (same defaults).
struct StereoOptions {
  int p1 = 28;
  int p2 = 200;
  int p2_adapt = 4;
  int invalid_cost = 31;
  SubpixelMethod subpixel = SubpixelMethod::equiangular;
  float uniqueness = 0.0F;
  int lr_max_diff = 1;
  bool lr_full_match = true;
  int speckle_size = 100;
  float speckle_diff = 1.0F;
  bool median = true;
  int levels = -1;
  int range_radius = 3;
  int range_margin = 4;
  int range_fill = 64;
  int tile_size = 512;
  int tile_margin = 32;
  int tile_blend = 32;
  std::size_t memory_budget = std::size_t{2} << 30;
  bool scalar_kernels = false;  // CPU kernel selection in the reference; no effect here
};

inline constexpr float kInvalidDisparity = std::numeric_limits<float>::infinity();

struct DisparityMap {
  ImageF32 disparity;
  ImageF32 confidence;
  int disparity_min = 0;
  int num_disparities = 0;
  int levels = 0;
  std::uint64_t cells = 0;
};

struct LabelRanges {
  Image<std::uint16_t> lo;
  Image<std::uint16_t> num;
};

[[nodiscard]] DisparityMap match_stereo(ImageView<const std::uint16_t> left, ImageView<const std::uint16_t> right,
                                        ImageView<const std::uint8_t> left_valid,
                                        ImageView<const std::uint8_t> right_valid, int disparity_min,
                                        int num_disparities, const StereoOptions& options = {});

void remove_speckles(ImageF32& disparity, int min_size, float max_diff);
[[nodiscard]] ImageF32 median_filter_3x3(const ImageF32& disparity);

// Selects the GPU of the process-wide context (before its first use).
void set_stereo_device(int device);

namespace detail {

[[nodiscard]] std::vector<std::uint16_t> sgm_cost_volume(ImageView<const std::uint16_t> left,
                                                         ImageView<const std::uint16_t> right, int disparity_min,
                                                         int num_disparities, const StereoOptions& options,
                                                         const LabelRanges* ranges = nullptr);

// The textbook volume of the reference (stereo_matcher.hpp:105-109) is CPU test
// code; tests get it from the oracle's shim (tests/cpp/reference_shim.cpp).
[[nodiscard]] std::vector<std::uint16_t> sgm_cost_volume_reference(ImageView<const std::uint16_t> left,
                                                                   ImageView<const std::uint16_t> right,
                                                                   int disparity_min, int num_disparities,
                                                                   const StereoOptions& options,
                                                                   const LabelRanges* ranges = nullptr);

[[nodiscard]] LabelRanges full_ranges(int width, int height, int num_disparities);

[[nodiscard]] LabelRanges ranges_from_coarse(const ImageF32& coarse, int width, int height,
                                             ImageView<const std::uint8_t> fine_valid,
                                             ImageView<const std::uint8_t> partner_valid, int partner_width,
                                             int partner_sign, int disparity_min, int num_disparities,
                                             const StereoOptions& options, bool mirrored = false);

[[nodiscard]] int automatic_levels(int width, int height, int num_disparities);

}  // namespace detail

}  // namespace r3d
