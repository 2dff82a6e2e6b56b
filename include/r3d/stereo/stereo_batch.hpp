// Several stereo pairs in one call, served by the MI355X implementation.
//
// No counterpart in the reference: its dense pipeline calls match_stereo once
// per partner (mvs/multi_stereo.cpp:264-293).  match_stereo_batch gives each
// pair exactly the DisparityMap match_stereo returns for it, and runs the
// pairs' pyramid levels together on the GPU (r3dg_match_batch).
#pragma once

#include <cstdint>
#include <vector>

#include "r3d/core/image.hpp"
#include "r3d/stereo/stereo_matcher.hpp"

namespace r3d {

struct StereoPair {
  ImageView<const std::uint16_t> left;
  ImageView<const std::uint16_t> right;
  ImageView<const std::uint8_t> left_valid;   // empty: all valid
  ImageView<const std::uint8_t> right_valid;  // empty: all valid
  int disparity_min = 0;
  int num_disparities = 0;
};

// One DisparityMap per pair, in order.  Throws r3d::Error for the first
// invalid pair (after the valid ones have been matched).
std::vector<DisparityMap> match_stereo_batch(const std::vector<StereoPair>& pairs, const StereoOptions& options = {});

}  // namespace r3d
