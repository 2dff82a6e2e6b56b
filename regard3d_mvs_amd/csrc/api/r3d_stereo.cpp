// The r3d C++ matcher API (include/r3d/...) over the C ABI of r3dg.h.
//
// Host-side only: converts ImageView / Image to r3dg_image descriptors, maps
// StereoOptions to r3dg_options and r3dg status codes to r3d::Error /
// r3d::Cancelled.  One process-wide GPU context serves every caller, as the
// reference's process-wide thread pool does (core/thread_pool.hpp:65-77);
// the C ABI serialises calls on it.
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "r3d/core/cancel.hpp"
#include "r3d/core/pyramid.hpp"
#include "r3d/core/simd.hpp"
#include "r3d/core/thread_pool.hpp"
#include "r3d/stereo/census.hpp"
#include "r3d/stereo/sgm_kernels.hpp"
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

thread_local const CancellationToken* t_token = nullptr;
std::mutex g_call_mutex;

int32_t poll_token(void* user) { return static_cast<const CancellationToken*>(user)->requested() ? 1 : 0; }

// One call on the process-wide context with the calling thread's
// cancellation token attached (polled by the context between levels and
// chunks); calls of several threads run one after the other.
template <typename F>
void call(F&& f) {
  r3dg_ctx* ctx = context();
  const CancellationToken* token = t_token;
  if (token != nullptr) token->throw_if_requested();
  int32_t rc = R3DG_OK;
  {
    const std::lock_guard<std::mutex> lock(g_call_mutex);
    r3dg_set_cancel_poll(ctx, token != nullptr ? poll_token : nullptr, const_cast<CancellationToken*>(token));
    rc = f(ctx);
    r3dg_set_cancel_poll(ctx, nullptr, nullptr);
  }
  check(rc);
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
  DisparityMap out;
  out.disparity = ImageF32::uninitialized(left.width(), left.height(), 1);
  out.confidence = ImageF32::uninitialized(left.width(), left.height(), 1);
  const r3dg_options o = options_of(options);
  r3dg_match_info info{};
  call([&](r3dg_ctx* ctx) {
    return r3dg_match(ctx, desc(left), desc(right), desc(left_valid), desc(right_valid), disparity_min,
                      num_disparities, &o, desc(out.disparity), desc(out.confidence), nullptr, &info);
  });
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
  call([&](r3dg_ctx* ctx) { return r3dg_match_batch(ctx, desc_pairs.data(), static_cast<int32_t>(desc_pairs.size()), &o); });
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
  call([&](r3dg_ctx* ctx) { return r3dg_remove_speckles(ctx, desc(disparity), min_size, max_diff); });
}

ImageF32 median_filter_3x3(const ImageF32& disparity) {
  ImageF32 out = ImageF32::uninitialized(disparity.width(), disparity.height(), 1);
  if (disparity.empty()) return out;
  call([&](r3dg_ctx* ctx) { return r3dg_median_filter_3x3(ctx, desc(disparity.view()), desc(out)); });
  return out;
}

Image<std::uint64_t> census_transform(ImageView<const std::uint16_t> img, bool /*scalar_reference*/) {
  if (img.channels() != 1) throw Error("census_transform: expected a single-channel image");
  Image<std::uint64_t> out = Image<std::uint64_t>::uninitialized(img.width(), img.height(), 1);
  if (img.empty()) return out;
  call([&](r3dg_ctx* ctx) { return r3dg_census(ctx, desc(img), desc(out)); });
  return out;
}

Image<std::uint8_t> pyramid_down(const ImageView<const std::uint8_t>& src) {
  if (src.empty()) throw Error("pyramid_down: empty image");
  Image<std::uint8_t> out = Image<std::uint8_t>::uninitialized((src.width() + 1) / 2, (src.height() + 1) / 2, 1);
  call([&](r3dg_ctx* ctx) { return r3dg_pyramid_down_u8(ctx, desc(src), desc(out)); });
  return out;
}

Image<std::uint16_t> pyramid_down(const ImageView<const std::uint16_t>& src) {
  if (src.empty()) throw Error("pyramid_down: empty image");
  Image<std::uint16_t> out = Image<std::uint16_t>::uninitialized((src.width() + 1) / 2, (src.height() + 1) / 2, 1);
  call([&](r3dg_ctx* ctx) { return r3dg_pyramid_down_u16(ctx, desc(src), desc(out)); });
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
  call([&](r3dg_ctx* ctx) { return r3dg_cost_volume(ctx, desc(left), desc(right), disparity_min, num_disparities, &o, lo, num, out.data(),
                         cells, R3DG_HOST); });
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
  call([&](r3dg_ctx* ctx) { return r3dg_ranges_from_coarse(ctx, desc(coarse.view()), width, height, desc(fine_valid), desc(partner_valid),
                                partner_width, partner_sign, disparity_min, num_disparities, &o, mirrored ? 1 : 0,
                                desc(r.lo), desc(r.num)); });
  return r;
}

int automatic_levels(int width, int height, int num_disparities) {
  return r3dg_automatic_levels(width, height, num_disparities);
}

}  // namespace detail

// ---- row kernels (census.hpp:25-29, sgm_kernels.hpp:41-97)
void census_row(const std::uint16_t* const rows[kCensusHeight], std::uint64_t* out, std::size_t n) {
  call([&](r3dg_ctx* ctx) { return r3dg_census_row(ctx, rows, out, n); });
}

namespace scalar {
void census_row(const std::uint16_t* const rows[kCensusHeight], std::uint64_t* out, std::size_t n) {
  r3d::census_row(rows, out, n);
}
}  // namespace scalar

namespace sgm {
namespace {
r3dg_row_labels row_labels(const RowLabels& l) { return r3dg_row_labels{l.lo, l.num, l.offset}; }
}  // namespace

void census_costs(const std::uint64_t* left, const std::uint64_t* rev, std::size_t width, const RowLabels& labels,
                  std::int32_t rev_origin, std::uint16_t* cost) {
  const r3dg_row_labels l = row_labels(labels);
  call([&](r3dg_ctx* ctx) { return r3dg_census_costs(ctx, left, rev, width, &l, rev_origin, cost); });
}

void sgm_row(const SgmRowArgs& a) {
  r3dg_sgm_row_args c{};
  c.cost = a.cost;
  c.p2 = a.p2;
  c.prev = a.prev;
  c.prev_min = a.prev_min;
  c.cur = a.cur;
  c.cur_min = a.cur_min;
  c.scratch = a.scratch;
  c.sum = a.sum;
  c.labels = row_labels(a.labels);
  c.prev_labels = row_labels(a.prev_labels);
  c.width = a.width;
  c.max_num = a.max_num;
  c.p1 = a.p1;
  c.direction = a.direction;
  c.has_prev = a.has_prev ? 1 : 0;
  c.accumulate = a.accumulate ? 1 : 0;
  call([&](r3dg_ctx* ctx) { return r3dg_sgm_row(ctx, &c); });
}

void wta_row(const std::uint16_t* sum, std::size_t width, const RowLabels& labels, std::uint16_t* best,
             std::uint16_t* best_cost, std::uint16_t* second, std::uint16_t* right_cost, std::uint16_t* right_best,
             std::ptrdiff_t right_shift) {
  const r3dg_row_labels l = row_labels(labels);
  call([&](r3dg_ctx* ctx) {
    return r3dg_wta_row(ctx, sum, width, &l, best, best_cost, second, right_cost, right_best, right_shift);
  });
}

namespace scalar {
void census_costs(const std::uint64_t* left, const std::uint64_t* rev, std::size_t width, const RowLabels& labels,
                  std::int32_t rev_origin, std::uint16_t* cost) {
  sgm::census_costs(left, rev, width, labels, rev_origin, cost);
}
void sgm_row(const SgmRowArgs& args) { sgm::sgm_row(args); }
void wta_row(const std::uint16_t* sum, std::size_t width, const RowLabels& labels, std::uint16_t* best,
             std::uint16_t* best_cost, std::uint16_t* second, std::uint16_t* right_cost, std::uint16_t* right_best,
             std::ptrdiff_t right_shift) {
  sgm::wta_row(sum, width, labels, best, best_cost, second, right_cost, right_best, right_shift);
}
}  // namespace scalar
}  // namespace sgm

// ---- targets (core/simd.hpp): one, the GPU
namespace {
constexpr std::int64_t kGpuTarget = std::int64_t{1} << 40;
}  // namespace

std::string simd_active_target_name() { return "gfx950"; }
std::vector<std::int64_t> simd_available_targets() { return {kGpuTarget}; }
std::string simd_target_name(std::int64_t target) { return target == kGpuTarget ? "gfx950" : "unknown"; }
ScopedSimdTarget::ScopedSimdTarget(std::int64_t target) {
  if (target != kGpuTarget) throw Error("ScopedSimdTarget: target " + std::to_string(target) + " is not available");
}
ScopedSimdTarget::~ScopedSimdTarget() = default;

// ---- cancellation (core/cancel.hpp)
const CancellationToken* current_cancellation() noexcept { return t_token; }
void throw_if_cancelled() {
  if (t_token != nullptr) t_token->throw_if_requested();
}
CancellationScope::CancellationScope(const CancellationToken* token) noexcept : previous_(t_token) { t_token = token; }
CancellationScope::~CancellationScope() { t_token = previous_; }

// ---- thread counts (core/thread_pool.hpp)
namespace {
unsigned hardware_threads() { return std::max(1U, std::thread::hardware_concurrency()); }
std::mutex g_pool_mutex;
std::unique_ptr<ThreadPool> g_pool;
}  // namespace

ThreadPool::ThreadPool(unsigned num_threads) : num_threads_(num_threads == 0 ? hardware_threads() : num_threads) {}

ThreadPool& default_thread_pool() {
  const std::lock_guard<std::mutex> lock(g_pool_mutex);
  if (!g_pool) g_pool = std::make_unique<ThreadPool>(0);
  return *g_pool;
}

void set_default_thread_count(unsigned num_threads) {
  const std::lock_guard<std::mutex> lock(g_pool_mutex);
  g_pool = std::make_unique<ThreadPool>(num_threads);
}

}  // namespace r3d
