// CPU oracle of the Census-SGM matcher -- TEST INFRASTRUCTURE ONLY.
//
// A restatement, in plain C++, of the reference matcher
// (rhiestan/Regard3D_MVS, src/r3d/stereo/{census,sgm_kernels,stereo_matcher}.cpp
// and src/r3d/core/pyramid.cpp).  It is the checker of the GPU library and the
// CPU baseline of bench.py; only tests/, __graft_entry__.smoke() and bench.py's
// CPU-baseline leg may load it.  The product never links it.
//
// Pinning (how each part is checked against something other than itself):
//   census            KATs of tests/test_census.cpp:26-65, brute force in numpy
//   pyramid_down      closed form (u8 pyramid == rounded float pyramid,
//                     test_pyramid.cpp:143-176), flat images stay flat
//   aggregate_tile    == textbook_volume (an independent 8-sweep restatement of
//                     stereo_matcher.cpp:1280-1362), as test_stereo.cpp:87-151
//   wta_row           brute force (first argmin, second outside best +- 1)
//   ranges_from_coarse KATs of test_stereo.cpp:343-400
//   remove_speckles   KATs of test_stereo.cpp:451-469 + connected components
//   median_filter_3x3 KATs of test_stereo.cpp:451-469 + sorted-window median
//   match_stereo      behaviour thresholds of test_stereo.cpp:191-449
//   rectify_image     see rectify.cpp
//   combine_pairs     see combine.cpp
// Floating-point steps (sub-pixel fit, blending, LR rounding) are restated
// operation by operation and compiled with -ffp-contract=off.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "r3dg.h"

namespace oracle {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

template <typename T>
struct Img {
  int w = 0;
  int h = 0;
  std::vector<T> px;
  Img() = default;
  Img(int w_, int h_, T v = T{}) : w(w_), h(h_), px(static_cast<std::size_t>(w_) * static_cast<std::size_t>(h_), v) {}
  T& operator()(int x, int y) { return px[static_cast<std::size_t>(y) * w + x]; }
  const T& operator()(int x, int y) const { return px[static_cast<std::size_t>(y) * w + x]; }
  T* row(int y) { return px.data() + static_cast<std::size_t>(y) * w; }
  const T* row(int y) const { return px.data() + static_cast<std::size_t>(y) * w; }
  [[nodiscard]] bool empty() const { return w == 0 || h == 0; }
};
using ImgU8 = Img<std::uint8_t>;
using ImgU16 = Img<std::uint16_t>;
using ImgF = Img<float>;
using ImgU64 = Img<std::uint64_t>;

// Per-pixel label ranges (LabelRanges, stereo_matcher.hpp:73-76).
struct Ranges {
  ImgU16 lo;
  ImgU16 num;
};

constexpr float kInvalid = __builtin_huge_valf();

int align16(int n);
int floor_shift(int v, int k);
int ceil_shift(int v, int k);

// Census (census.cpp:20-35, 104-113, 121-164).
ImgU64 census(const ImgU16& img);

// pyramid_down (pyramid.cpp:34-85) and reduce_mask (stereo_matcher.cpp:892-900).
ImgU16 pyramid_down(const ImgU16& img);
ImgU8 pyramid_down(const ImgU8& img);
ImgU8 reduce_mask(const ImgU8& mask);

// A camera of rectify_image: CameraIntrinsics + Pose (core/camera_model.hpp:
// 32-41, core/camera.hpp:9-12); model 0..4 as r3dg_camera_model.
struct Camera {
  int model = 0;
  int width = 0;
  int height = 0;
  double fx = 0.0, fy = 0.0, cx = 0.0, cy = 0.0;
  double params[5] = {};
  double rotation[9] = {};
  double center[3] = {};
};

// combine_pairs (mvs/multi_stereo.cpp:44-216), per pixel (combine.cpp).
struct CombinePairIn {
  double grid_to_rect[9] = {};
  double fb = 0.0, offset = 0.0;
  double partner_center[3] = {};
  ImgF disparity;
  ImgF confidence;  // may be empty
};
struct CombineGridIn {
  int width = 0, height = 0;
  double inv_k[9] = {}, rot_t[9] = {}, center[3] = {};
};
struct CombineOptions {
  double inlier_px = 2.0, max_reprojection_px = 1.0, cos_min_angle = 0.0;
  int min_pairs = 1;
};
struct Combined {
  ImgF depth, confidence;
  Img<std::uint16_t> pair_mask;
};
float sample_disparity(const ImgF& d, double u, double v);
Combined combine_pairs(const CombineGridIn& g, const std::vector<CombinePairIn>& pairs, const CombineOptions& o);

// rectify_image (rectification.cpp:288-350), one channel (rectify.cpp).
template <typename T>
Img<T> rectify_image(const Img<T>& src, const Camera& cam, const Camera& target, ImgU8* valid);

int automatic_levels(int width, int height, int num_disparities);

// Aggregated volume of the whole image as one tile (sgm_cost_volume,
// stereo_matcher.cpp:1269-1278) through the row-pass dataflow.
std::vector<std::uint16_t> cost_volume(const ImgU16& left, const ImgU16& right, int dmin, int nd,
                                       const r3dg_options& o, const Ranges* ranges);
// Textbook volume (stereo_matcher.cpp:1280-1362): 8 separate path sweeps in int.
std::vector<std::uint16_t> textbook_volume(const ImgU16& left, const ImgU16& right, int dmin, int nd,
                                           const r3dg_options& o, const Ranges* ranges);

// WTA of one row (sgm_kernels.cpp:405-438).
void wta_row(const std::uint16_t* sum, int width, const std::int32_t* lo, const std::uint16_t* num,
             std::uint16_t* best, std::uint16_t* best_cost, std::uint16_t* second, std::uint16_t* right_cost,
             std::uint16_t* right_best, std::ptrdiff_t right_shift);

Ranges ranges_from_coarse(const ImgF& coarse, int width, int height, const ImgU8& fine_valid,
                          const ImgU8& partner_valid, int partner_width, int partner_sign, int dmin, int nd,
                          const r3dg_options& o, bool mirrored);

// A tile of blend_tiles: extent, core [c0, c1) x [r0, r1), and its maps
// (w x h, row-major; c and lab may be null).
struct BlendTile {
  int x0, y0, w, h;
  int core[4];
  const float* d;
  const float* c;
  const std::int16_t* lab;
};
void blend_tiles(int width, int height, int blend, const std::vector<BlendTile>& tiles, ImgF& disparity,
                 ImgF* confidence, Img<std::int16_t>* labels);
ImgU8 left_right_mismatch(const ImgF& map, const ImgF& other, int sign, int max_diff);
void remove_speckles(ImgF& d, int min_size, float max_diff);
ImgF median3x3(const ImgF& d);

struct MatchOutput {
  ImgF disparity;
  ImgF confidence;
  Img<std::int16_t> labels;  // level 0: dominant tile's integer disparity (INT16_MIN: none)
  ImgU8 lr_mask;             // level 0: 1 where the LR check removed a valid pixel
  int disparity_min = 0;
  int num_disparities = 0;
  int levels = 0;
  int tiles_level0 = 0;
  std::uint64_t cells = 0;
  std::uint64_t tile_pixels = 0;
};

MatchOutput match_stereo(const ImgU16& left, const ImgU16& right, const ImgU8& left_valid, const ImgU8& right_valid,
                         int dmin, int nd, const r3dg_options& o, int threads);

void check_options(const r3dg_options& o);

}  // namespace oracle
