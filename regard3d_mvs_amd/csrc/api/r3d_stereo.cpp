// The r3d C++ matcher API (include/r3d/...) over the C ABI of r3dg.h.
//
// Host-side only: converts ImageView / Image to r3dg_image descriptors, maps
// StereoOptions to r3dg_options and r3dg status codes to r3d::Error /
// r3d::Cancelled.  One process-wide GPU context serves every caller, as the
// reference's process-wide thread pool does (core/thread_pool.hpp:65-77);
// the C ABI serialises calls on it.
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "r3d/core/pyramid.hpp"
#include "r3d/stereo/census.hpp"
#include "r3d/stereo/stereo_batch.hpp"
#include "r3d/stereo/stereo_matcher.hpp"
#include "r3dg.h"

namespace r3d {
namespace {

std::mutex g_mutex;
r3dg_ctx* g_ctx = nullptr;
int g_device = 0;

[[noreturn]] void raise(int32_t status, r3dg_ctx* ctx) {
  char buf[1024];
  if (ctx != nullptr) {
    r3dg_last_error(ctx, buf, sizeof(buf));
  } else {
    r3dg_global_error(buf, sizeof(buf));
  }
  if (status == R3DG_ERROR_CANCELLED) throw Cancelled();
  throw Error(std::string(buf));
}

r3dg_ctx* context() {
  const std::lock_guard<std::mutex> lock(g_mutex);
  if (g_ctx == nullptr) {
    const int32_t rc = r3dg_ctx_create(g_device, 0, &g_ctx);
    if (rc != R3DG_OK) raise(rc, nullptr);
  }
  return g_ctx;
}

void check(int32_t rc) {
  if (rc != R3DG_OK) raise(rc, g_ctx);
}

template <typename T>
r3dg_image desc(ImageView<T> v) {
  r3dg_image d{};
  if (v.empty()) return d;
  if (v.channels() != 1) throw Error("stereo: expected single-channel images");
  d.data = const_cast<void*>(static_cast<const void*>(v.data()));
  d.width = v.width();
  d.height = v.height();
  d.stride = v.stride();
  d.memory = R3DG_HOST;
  return d;
}

template <typename T>
r3dg_image desc(Image<T>& img) {
  return desc(img.view());
}

r3dg_options options_of(const StereoOptions& o) {
  r3dg_options c{};
  c.p1 = o.p1;
  c.p2 = o.p2;
  c.p2_adapt = o.p2_adapt;
  c.invalid_cost = o.invalid_cost;
  c.subpixel = o.subpixel == SubpixelMethod::none       ? R3DG_SUBPIXEL_NONE
               : o.subpixel == SubpixelMethod::parabola ? R3DG_SUBPIXEL_PARABOLA
                                                        : R3DG_SUBPIXEL_EQUIANGULAR;
  c.uniqueness = o.uniqueness;
  c.lr_max_diff = o.lr_max_diff;
  c.lr_full_match = o.lr_full_match ? 1 : 0;
  c.speckle_size = o.speckle_size;
  c.speckle_diff = o.speckle_diff;
  c.median = o.median ? 1 : 0;
  c.levels = o.levels;
  c.range_radius = o.range_radius;
  c.range_margin = o.range_margin;
  c.range_fill = o.range_fill;
  c.tile_size = o.tile_size;
  c.tile_margin = o.tile_margin;
  c.tile_blend = o.tile_blend;
  c.memory_budget = o.memory_budget;
  return c;
}

int align16(int n) { return (n + 15) / 16 * 16; }

}  // namespace

void set_stereo_device(int device) {
  const std::lock_guard<std::mutex> lock(g_mutex);
  if (g_ctx != nullptr) throw Error("set_stereo_device: the GPU context is already in use");
  g_device = device;
}

DisparityMap match_stereo(ImageView<const std::uint16_t> left, ImageView<const std::uint16_t> right,
                          ImageView<const std::uint8_t> left_valid, ImageView<const std::uint8_t> right_valid,
                          int disparity_min, int num_disparities, const StereoOptions& options) {
  r3dg_ctx* ctx = context();
  DisparityMap out;
  out.disparity = ImageF32::uninitialized(left.width(), left.height(), 1);
  out.confidence = ImageF32::uninitialized(left.width(), left.height(), 1);
  const r3dg_options o = options_of(options);
  r3dg_match_info info{};
  check(r3dg_match(ctx, desc(left), desc(right), desc(left_valid), desc(right_valid), disparity_min, num_disparities,
                   &o, desc(out.disparity), desc(out.confidence), nullptr, &info));
  out.disparity_min = info.disparity_min;
  out.num_disparities = info.num_disparities;
  out.levels = info.levels;
  out.cells = info.cells;
  return out;
}

std::vector<DisparityMap> match_stereo_batch(const std::vector<StereoPair>& pairs, const StereoOptions& options) {
  std::vector<DisparityMap> out(pairs.size());
  std::vector<r3dg_pair> desc_pairs(pairs.size());
  for (std::size_t i = 0; i < pairs.size(); ++i) {
    const StereoPair& p = pairs[i];
    out[i].disparity = ImageF32::uninitialized(p.left.width(), p.left.height(), 1);
    out[i].confidence = ImageF32::uninitialized(p.left.width(), p.left.height(), 1);
    r3dg_pair& q = desc_pairs[i];
    q = r3dg_pair{};
    q.left = desc(p.left);
    q.right = desc(p.right);
    q.left_valid = desc(p.left_valid);
    q.right_valid = desc(p.right_valid);
    q.disparity_min = p.disparity_min;
    q.num_disparities = p.num_disparities;
    q.disparity = desc(out[i].disparity);
    q.confidence = desc(out[i].confidence);
  }
  const r3dg_options o = options_of(options);
  check(r3dg_match_batch(context(), desc_pairs.data(), static_cast<int32_t>(desc_pairs.size()), &o));
  for (std::size_t i = 0; i < pairs.size(); ++i) {
    out[i].disparity_min = desc_pairs[i].info.disparity_min;
    out[i].num_disparities = desc_pairs[i].info.num_disparities;
    out[i].levels = desc_pairs[i].info.levels;
    out[i].cells = desc_pairs[i].info.cells;
  }
  return out;
}

void remove_speckles(ImageF32& disparity, int min_size, float max_diff) {
  if (min_size <= 1 || disparity.empty()) return;
  check(r3dg_remove_speckles(context(), desc(disparity), min_size, max_diff));
}

ImageF32 median_filter_3x3(const ImageF32& disparity) {
  ImageF32 out = ImageF32::uninitialized(disparity.width(), disparity.height(), 1);
  if (disparity.empty()) return out;
  check(r3dg_median_filter_3x3(context(), desc(disparity.view()), desc(out)));
  return out;
}

Image<std::uint64_t> census_transform(ImageView<const std::uint16_t> img, bool /*scalar_reference*/) {
  if (img.channels() != 1) throw Error("census_transform: expected a single-channel image");
  Image<std::uint64_t> out = Image<std::uint64_t>::uninitialized(img.width(), img.height(), 1);
  if (img.empty()) return out;
  check(r3dg_census(context(), desc(img), desc(out)));
  return out;
}

Image<std::uint8_t> pyramid_down(const ImageView<const std::uint8_t>& src) {
  if (src.empty()) throw Error("pyramid_down: empty image");
  Image<std::uint8_t> out = Image<std::uint8_t>::uninitialized((src.width() + 1) / 2, (src.height() + 1) / 2, 1);
  check(r3dg_pyramid_down_u8(context(), desc(src), desc(out)));
  return out;
}

Image<std::uint16_t> pyramid_down(const ImageView<const std::uint16_t>& src) {
  if (src.empty()) throw Error("pyramid_down: empty image");
  Image<std::uint16_t> out = Image<std::uint16_t>::uninitialized((src.width() + 1) / 2, (src.height() + 1) / 2, 1);
  check(r3dg_pyramid_down_u16(context(), desc(src), desc(out)));
  return out;
}

namespace detail {

std::vector<std::uint16_t> sgm_cost_volume(ImageView<const std::uint16_t> left, ImageView<const std::uint16_t> right,
                                           int disparity_min, int num_disparities, const StereoOptions& options,
                                           const LabelRanges* ranges) {
  std::uint64_t cells = 0;
  r3dg_image lo{}, num{};
  if (ranges != nullptr) {
    if (ranges->num.width() != left.width() || ranges->num.height() != left.height())
      throw Error("stereo: label ranges of the wrong size");
    for (int y = 0; y < ranges->num.height(); ++y)
      for (int x = 0; x < ranges->num.width(); ++x) cells += ranges->num(x, y);
    lo = desc(ranges->lo.view());
    num = desc(ranges->num.view());
  } else {
    cells = std::uint64_t(left.width()) * std::uint64_t(left.height()) *
            std::uint64_t(align16(num_disparities > 0 ? num_disparities : 0));
  }
  std::vector<std::uint16_t> out(cells);
  const r3dg_options o = options_of(options);
  check(r3dg_cost_volume(context(), desc(left), desc(right), disparity_min, num_disparities, &o, lo, num, out.data(),
                         cells, R3DG_HOST));
  return out;
}

LabelRanges full_ranges(int width, int height, int num_disparities) {
  LabelRanges r{Image<std::uint16_t>(width, height, 1), Image<std::uint16_t>(width, height, 1)};
  r.num.fill(static_cast<std::uint16_t>(align16(num_disparities)));
  return r;
}

LabelRanges ranges_from_coarse(const ImageF32& coarse, int width, int height, ImageView<const std::uint8_t> fine_valid,
                               ImageView<const std::uint8_t> partner_valid, int partner_width, int partner_sign,
                               int disparity_min, int num_disparities, const StereoOptions& options, bool mirrored) {
  LabelRanges r{Image<std::uint16_t>::uninitialized(width, height, 1),
                Image<std::uint16_t>::uninitialized(width, height, 1)};
  const r3dg_options o = options_of(options);
  check(r3dg_ranges_from_coarse(context(), desc(coarse.view()), width, height, desc(fine_valid), desc(partner_valid),
                                partner_width, partner_sign, disparity_min, num_disparities, &o, mirrored ? 1 : 0,
                                desc(r.lo), desc(r.num)));
  return r;
}

int automatic_levels(int width, int height, int num_disparities) {
  return r3dg_automatic_levels(width, height, num_disparities);
}

}  // namespace detail
}  // namespace r3d
