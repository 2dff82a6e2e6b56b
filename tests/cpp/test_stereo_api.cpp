// C++ tests of the r3d drop-in API (libr3d_stereo.so over the GPU C ABI),
// written the way the reference's tests/test_stereo.cpp and test_census.cpp
// call it.  Procedural scenes restate tests/test_util.hpp:22-39 and
// tests/test_stereo.cpp:19-73; the GPU results are also compared bit for bit
// with the CPU oracle.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <chrono>
#include <climits>
#include <cstring>
#include <functional>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../oracle/oracle.hpp"
#include "r3d/core/cancel.hpp"
#include "r3d/core/pyramid.hpp"
#include "r3d/core/simd.hpp"
#include "r3d/core/thread_pool.hpp"
#include "r3d/stereo/census.hpp"
#include "r3d/stereo/sgm_kernels.hpp"
#include "r3d/stereo/stereo_batch.hpp"
#include "r3d/stereo/stereo_matcher.hpp"

namespace {

int g_failed = 0, g_checks = 0;
#define CHECK(c)                                                            \
  do {                                                                      \
    ++g_checks;                                                             \
    if (!(c)) {                                                             \
      ++g_failed;                                                           \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);            \
    }                                                                       \
  } while (0)

struct Case {
  const char* name;
  std::function<void()> fn;
};
std::vector<Case>& cases() {
  static std::vector<Case> c;
  return c;
}
struct Reg {
  Reg(const char* n, std::function<void()> f) { cases().push_back({n, std::move(f)}); }
};
#define TEST(id, name) \
  static void id();    \
  static Reg reg_##id(name, id); \
  static void id()

// Portable random numbers (std::mt19937 + own distributions).
class Rng {
 public:
  explicit Rng(unsigned seed) : gen_(seed) {}
  double uniform(double lo, double hi) { return lo + (hi - lo) * ((gen_() + 0.5) / 4294967296.0); }
  int uniform_int(int lo, int hi) { return lo + static_cast<int>(gen_() % static_cast<std::uint32_t>(hi - lo + 1)); }

 private:
  std::mt19937 gen_;
};

struct Texture {
  std::vector<double> fx, fy, ph, amp;
  explicit Texture(unsigned seed) {
    Rng rng(seed);
    for (int i = 0; i < 24; ++i) {
      const double a = rng.uniform(0.05, 0.9);
      fx.push_back(a * rng.uniform(-1.0, 1.0));
      const double b = rng.uniform(0.05, 0.9);
      fy.push_back(b * rng.uniform(-1.0, 1.0));
      ph.push_back(rng.uniform(0.0, 6.283));
      amp.push_back(1.0 / (1.0 + i * 0.2));
    }
  }
  double operator()(double x, double y) const {
    double v = 0.0, n = 0.0;
    for (std::size_t i = 0; i < fx.size(); ++i) {
      v += amp[i] * std::sin(fx[i] * x + fy[i] * y + ph[i]);
      n += amp[i];
    }
    return 32768.0 + 30000.0 * v / n * 2.0;
  }
};

std::uint16_t to_u16(double v) { return static_cast<std::uint16_t>(std::lround(std::clamp(v, 0.0, 65535.0))); }

template <typename D>
void render(const Texture& t, int w, int h, D&& d, r3d::ImageU16& l, r3d::ImageU16& r, r3d::ImageF32& gt) {
  l = r3d::ImageU16(w, h, 1);
  r = r3d::ImageU16(w, h, 1);
  gt = r3d::ImageF32(w, h, 1);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      l(x, y) = to_u16(t(x, y));
      gt(x, y) = static_cast<float>(d(x, y));
      double xl = x;
      for (int it = 0; it < 20; ++it) xl = x + d(xl, y);
      r(x, y) = to_u16(t(xl, y));
    }
}

template <typename T>
oracle::Img<T> to_oracle(const r3d::Image<T>& img) {
  oracle::Img<T> o(img.width(), img.height());
  for (int y = 0; y < img.height(); ++y) std::memcpy(o.row(y), img.row(y), sizeof(T) * std::size_t(img.width()));
  return o;
}

bool same_bits(const r3d::ImageF32& a, const oracle::ImgF& b) {
  if (a.width() != b.w || a.height() != b.h) return false;
  for (int y = 0; y < a.height(); ++y)
    if (std::memcmp(a.row(y), b.row(y), sizeof(float) * std::size_t(a.width())) != 0) return false;
  return true;
}

bool same_bits(const r3d::ImageF32& a, const r3d::ImageF32& b) {
  if (a.width() != b.width() || a.height() != b.height()) return false;
  for (int y = 0; y < a.height(); ++y)
    if (std::memcmp(a.row(y), b.row(y), sizeof(float) * std::size_t(a.width())) != 0) return false;
  return true;
}

r3dg_options oracle_options(const r3d::StereoOptions& s) {
  r3dg_options o{};
  r3dg_default_options(&o);
  o.p1 = s.p1;
  o.p2 = s.p2;
  o.p2_adapt = s.p2_adapt;
  o.invalid_cost = s.invalid_cost;
  o.subpixel = s.subpixel == r3d::SubpixelMethod::none ? 0 : (s.subpixel == r3d::SubpixelMethod::parabola ? 1 : 2);
  o.uniqueness = s.uniqueness;
  o.lr_max_diff = s.lr_max_diff;
  o.lr_full_match = s.lr_full_match ? 1 : 0;
  o.speckle_size = s.speckle_size;
  o.speckle_diff = s.speckle_diff;
  o.median = s.median ? 1 : 0;
  o.levels = s.levels;
  o.tile_size = s.tile_size;
  o.tile_margin = s.tile_margin;
  o.tile_blend = s.tile_blend;
  return o;
}

TEST(census_order, "Census bits follow the documented order") {
  r3d::ImageU16 img(9, 7, 1);
  img.fill(100);
  CHECK(r3d::census_transform(img.view())(4, 3) == 0);
  img(4, 3) = 200;
  CHECK(r3d::census_transform(img.view())(4, 3) == (std::uint64_t{1} << 62) - 1);
  img.fill(100);
  img(0, 0) = 50;
  CHECK(r3d::census_transform(img.view())(4, 3) == std::uint64_t{1} << 61);
  img(0, 0) = 100;
  img(8, 6) = 50;
  CHECK(r3d::census_transform(img.view())(4, 3) == 1);
  img(8, 6) = 100;
  img(5, 3) = 50;
  CHECK(r3d::census_transform(img.view())(4, 3) == std::uint64_t{1} << 30);
}

TEST(volume_textbook, "SGM cost volume equals the textbook implementation") {
  const Texture tex(3);
  r3d::ImageU16 left, right;
  r3d::ImageF32 gt;
  render(tex, 41, 23, [](double x, double) { return 5.0 + 0.1 * x; }, left, right, gt);
  Rng rng(9);
  for (int y = 0; y < 23; ++y)
    for (int x = 0; x < 41; ++x) right(x, y) = static_cast<std::uint16_t>(right(x, y) ^ (rng.uniform_int(0, 65535) & 0x0F00));
  for (const int adapt : {0, 12})
    for (const auto& [dmin, nd] : {std::pair{-3, 16}, std::pair{2, 48}}) {
      r3d::StereoOptions o;
      o.p2_adapt = adapt;
      o.p2 = 120;
      CHECK(r3d::detail::sgm_cost_volume(left.view(), right.view(), dmin, nd, o) ==
            r3d::detail::sgm_cost_volume_reference(left.view(), right.view(), dmin, nd, o));
    }
}

TEST(volume_ranges, "SGM cost volume with per-pixel label ranges equals the textbook implementation") {
  const Texture tex(13);
  r3d::ImageU16 left, right;
  r3d::ImageF32 gt;
  const int w = 37, h = 19, nd = 96;
  render(tex, w, h, [](double x, double y) { return 20.0 + 0.5 * x - 0.3 * y; }, left, right, gt);
  Rng rng(31);
  r3d::LabelRanges ranges{r3d::Image<std::uint16_t>(w, h, 1), r3d::Image<std::uint16_t>(w, h, 1)};
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const int n = 16 * rng.uniform_int(1, 3);
      ranges.num(x, y) = static_cast<std::uint16_t>(n);
      ranges.lo(x, y) = static_cast<std::uint16_t>(rng.uniform_int(0, nd - n));
    }
  r3d::StereoOptions o;
  o.p2 = 150;
  o.p2_adapt = 8;
  CHECK(r3d::detail::sgm_cost_volume(left.view(), right.view(), -2, nd, o, &ranges) ==
        r3d::detail::sgm_cost_volume_reference(left.view(), right.view(), -2, nd, o, &ranges));
  const r3d::LabelRanges full = r3d::detail::full_ranges(w, h, nd);
  CHECK(r3d::detail::sgm_cost_volume(left.view(), right.view(), -2, nd, o, &full) ==
        r3d::detail::sgm_cost_volume(left.view(), right.view(), -2, nd, o));
}

TEST(slanted, "Stereo matching recovers a slanted plane and equals the oracle") {
  const Texture tex(11);
  r3d::ImageU16 left, right;
  r3d::ImageF32 gt;
  render(tex, 200, 120, [](double x, double y) { return 12.3 + 0.04 * x + 0.02 * y; }, left, right, gt);
  for (const auto m : {r3d::SubpixelMethod::parabola, r3d::SubpixelMethod::equiangular}) {
    r3d::StereoOptions o;
    o.subpixel = m;
    const r3d::DisparityMap r = r3d::match_stereo(left.view(), right.view(), {}, {}, 0, 32, o);
    CHECK(r.num_disparities == 32);
    int n = 0, est = 0, bad = 0;
    double err = 0.0;
    for (int y = 0; y < 120; ++y)
      for (int x = 30; x < 200; ++x) {
        ++n;
        const float d = r.disparity(x, y);
        if (!std::isfinite(d)) continue;
        ++est;
        const double e = std::abs(double(d) - gt(x, y));
        err += e;
        bad += e > 1.0 ? 1 : 0;
      }
    CHECK(est > 0.95 * n);
    CHECK(bad < 0.01 * est);
    CHECK(err / est < 0.2);
    const auto c = oracle::match_stereo(to_oracle(left), to_oracle(right), {}, {}, 0, 32, oracle_options(o), 8);
    CHECK(same_bits(r.disparity, c.disparity));
    CHECK(same_bits(r.confidence, c.confidence));
    CHECK(r.cells == c.cells);
  }
}

TEST(batch, "match_stereo_batch gives every pair the map of match_stereo") {
  const Texture tex(23);
  r3d::ImageU16 l1, r1, l2, r2;
  r3d::ImageF32 gt;
  render(tex, 200, 120, [](double x, double y) { return 12.3 + 0.04 * x + 0.02 * y; }, l1, r1, gt);
  render(tex, 150, 90, [](double x, double) { return 5.0 + 0.03 * x; }, l2, r2, gt);
  std::vector<r3d::StereoPair> pairs(3);
  pairs[0].left = l1.view();
  pairs[0].right = r1.view();
  pairs[0].num_disparities = 32;
  pairs[1].left = l2.view();
  pairs[1].right = r2.view();
  pairs[1].disparity_min = -4;
  pairs[1].num_disparities = 24;
  pairs[2] = pairs[0];
  pairs[2].num_disparities = 48;
  const auto got = r3d::match_stereo_batch(pairs);
  CHECK(got.size() == 3);
  for (std::size_t i = 0; i < pairs.size(); ++i) {
    const auto& p = pairs[i];
    const r3d::DisparityMap one =
        r3d::match_stereo(p.left, p.right, p.left_valid, p.right_valid, p.disparity_min, p.num_disparities);
    CHECK(same_bits(got[i].disparity, one.disparity));
    CHECK(same_bits(got[i].confidence, one.confidence));
    CHECK(got[i].cells == one.cells && got[i].levels == one.levels);
  }
}

TEST(masks, "Masks: invalid left pixels stay invalid, matches into invalid right pixels are rejected") {
  const Texture tex(8);
  r3d::ImageU16 left, right;
  r3d::ImageF32 gt;
  render(tex, 120, 60, [](double, double) { return 10.0; }, left, right, gt);
  r3d::ImageU8 lm(120, 60, 1), rm(120, 60, 1);
  lm.fill(255);
  rm.fill(255);
  for (int y = 0; y < 60; ++y) {
    for (int x = 100; x < 120; ++x) lm(x, y) = 0;
    for (int x = 0; x < 40; ++x) rm(x, y) = 0;
  }
  const r3d::DisparityMap r = r3d::match_stereo(left.view(), right.view(), lm.view(), rm.view(), 0, 32);
  bool ok = true;
  int good = 0;
  for (int y = 0; y < 60; ++y) {
    for (int x = 100; x < 120; ++x) ok = ok && !std::isfinite(r.disparity(x, y));
    for (int x = 0; x < 40; ++x) ok = ok && !std::isfinite(r.disparity(x, y));
    for (int x = 55; x < 95; ++x) good += std::abs(r.disparity(x, y) - 10.0F) < 0.25F ? 1 : 0;
  }
  CHECK(ok);
  CHECK(good > 0.95 * 60 * 40);
  const auto c = oracle::match_stereo(to_oracle(left), to_oracle(right), to_oracle(lm), to_oracle(rm), 0, 32,
                                      oracle_options({}), 8);
  CHECK(same_bits(r.disparity, c.disparity));
}

TEST(filters, "Speckle and median filters") {
  r3d::ImageF32 d(20, 10, 1);
  d.fill(5.0F);
  d(3, 3) = 20.0F;
  d(10, 5) = r3d::kInvalidDisparity;
  for (int x = 15; x < 20; ++x) d(x, 0) = 30.0F;
  r3d::ImageF32 s = d;
  r3d::remove_speckles(s, 6, 1.0F);
  CHECK(!std::isfinite(s(3, 3)));
  CHECK(!std::isfinite(s(17, 0)));
  CHECK(s(0, 0) == 5.0F);
  s = d;
  r3d::remove_speckles(s, 5, 1.0F);
  CHECK(s(17, 0) == 30.0F);
  const r3d::ImageF32 m = r3d::median_filter_3x3(d);
  CHECK(m(3, 3) == 5.0F);
  CHECK(!std::isfinite(m(10, 5)));
  CHECK(m(11, 5) == 5.0F);
}

TEST(ranges, "Per-pixel label ranges from a coarser level") {
  const int wc = 40, hc = 20;
  r3d::ImageF32 coarse(wc, hc, 1);
  for (int y = 0; y < hc; ++y)
    for (int x = 0; x < wc; ++x) coarse(x, y) = x < wc / 2 ? 10.0F : 30.0F;
  coarse(5, 5) = r3d::kInvalidDisparity;
  for (int y = 0; y < hc; ++y)
    for (int x = 30; x < wc; ++x) coarse(x, y) = r3d::kInvalidDisparity;
  const int w = 2 * wc, h = 2 * hc, dmin = 4, nd = 96;
  r3d::ImageU8 valid(w, h, 1);
  valid.fill(255);
  valid(1, 1) = 0;
  r3d::StereoOptions o;
  o.range_radius = 1;
  o.range_margin = 3;
  o.range_fill = 4;
  const r3d::LabelRanges r = r3d::detail::ranges_from_coarse(coarse, w, h, valid.view(), {}, 1000, 1, dmin, nd, o);
  auto covers = [&](int x, int y, int lo, int hi) { return r.lo(x, y) <= lo && r.lo(x, y) + r.num(x, y) - 1 >= hi; };
  CHECK(covers(10, 30, 20 - dmin - 3, 20 - dmin + 3));
  CHECK(r.num(10, 30) == 16);
  CHECK(covers(10, 10, 20 - dmin - 3, 20 - dmin + 3));
  CHECK(covers(39, 20, 20 - dmin - 3, 60 - dmin + 3));
  CHECK(r.num(1, 1) == 16);
  CHECK(covers(62, 20, 60 - dmin - 3, 60 - dmin + 3));
  CHECK(r.lo(78, 20) == 0);
  CHECK(r.num(78, 20) == nd);
  CHECK(r3d::detail::automatic_levels(6048, 4032, 512) == 4);
}

TEST(pyramid, "Pyramid of a flat image stays flat") {
  r3d::ImageU16 flat(9, 9, 1);
  flat.fill(54321);
  const r3d::ImageU16 down = r3d::pyramid_down(flat);
  CHECK(down.width() == 5 && down.height() == 5);
  CHECK(down(0, 0) == 54321 && down(4, 4) == 54321);
}

TEST(errors, "Invalid options raise r3d::Error") {
  r3d::ImageU16 img(16, 16, 1);
  r3d::StereoOptions o;
  o.p2 = 9000;
  bool thrown = false;
  try {
    (void)r3d::match_stereo(img.view(), img.view(), {}, {}, 0, 16, o);
  } catch (const r3d::Error&) {
    thrown = true;
  }
  CHECK(thrown);
  thrown = false;
  try {
    (void)r3d::match_stereo(img.view(), img.view(), {}, {}, 0, 0);
  } catch (const r3d::Error&) {
    thrown = true;
  }
  CHECK(thrown);
}


// ---- row kernels, driven the way the reference's tile aggregation drives them
// (stereo_matcher.cpp:253-424): Census rows, cost rows through census_costs
// with the right Census reversed, P2 per path from the intensity LUT, two
// passes of sgm_row with prev / cur swapped per row.  The volume must equal
// the textbook volume.
std::uint16_t p2_of(const r3d::StereoOptions& o, int di) {
  if (o.p2_adapt <= 0) return static_cast<std::uint16_t>(o.p2);
  return static_cast<std::uint16_t>(std::max(o.p1 + 1, o.p2 * o.p2_adapt / (o.p2_adapt + di)));
}

std::vector<std::uint16_t> row_kernel_volume(const r3d::ImageU16& left, const r3d::ImageU16& right, int dmin, int nd,
                                             const r3d::StereoOptions& o, const r3d::LabelRanges* ranges) {
  namespace sgm = r3d::sgm;
  const int w = left.width(), h = left.height(), wr = right.width();
  const int D = (nd + 15) / 16 * 16;
  // Census of both images, the left one row by row through census_row.
  r3d::Image<std::uint64_t> cl(w, h, 1);
  auto padded = std::vector<std::vector<std::uint16_t>>(std::size_t(h), std::vector<std::uint16_t>(std::size_t(w) + 8));
  for (int y = 0; y < h; ++y)
    for (int x = -4; x < w + 4; ++x) padded[std::size_t(y)][std::size_t(x + 4)] = left(std::clamp(x, 0, w - 1), y);
  for (int y = 0; y < h; ++y) {
    const std::uint16_t* rows[r3d::kCensusHeight];
    for (int j = 0; j < r3d::kCensusHeight; ++j) rows[j] = padded[std::size_t(std::clamp(y - 3 + j, 0, h - 1))].data();
    r3d::census_row(rows, cl.row(y), std::size_t(w));
  }
  const r3d::Image<std::uint64_t> cr = r3d::census_transform(right.view());
  // Packed labels per row.
  auto lo = std::vector<std::vector<std::int32_t>>(std::size_t(h), std::vector<std::int32_t>(std::size_t(w)));
  auto num = std::vector<std::vector<std::uint16_t>>(std::size_t(h), std::vector<std::uint16_t>(std::size_t(w)));
  auto off = std::vector<std::vector<std::size_t>>(std::size_t(h), std::vector<std::size_t>(std::size_t(w) + 1, 0));
  std::vector<std::size_t> row_start(std::size_t(h) + 1, 0);
  std::size_t max_num = 16, max_row = 0;
  for (int y = 0; y < h; ++y) {
    for (int x = 0; x < w; ++x) {
      lo[y][x] = ranges != nullptr ? ranges->lo(x, y) : 0;
      num[y][x] = ranges != nullptr ? ranges->num(x, y) : static_cast<std::uint16_t>(D);
      off[y][x + 1] = off[y][x] + num[y][x];
      max_num = std::max<std::size_t>(max_num, num[y][x]);
    }
    row_start[y + 1] = row_start[y] + off[y][w];
    max_row = std::max(max_row, off[y][w]);
  }
  auto labels = [&](int y) { return sgm::RowLabels{lo[y].data(), num[y].data(), off[y].data()}; };
  std::vector<std::uint16_t> volume(row_start[std::size_t(h)]), cost(max_row), p2(4 * std::size_t(w));
  const std::size_t path = max_row + std::size_t(w) * 2 * sgm::kLabelOffset;
  std::vector<std::uint16_t> prev(3 * path), cur(3 * path), prev_min(3 * std::size_t(w)), cur_min(3 * std::size_t(w));
  std::vector<std::uint16_t> scratch(sgm::kScratchBuffers * sgm::path_stride(max_num), 0);
  auto cost_row = [&](int y) {
    const sgm::RowLabels rl = labels(y);
    std::int32_t origin = INT32_MAX, end = INT32_MIN;
    for (int x = 0; x < w; ++x) {
      origin = std::min(origin, w - 1 - x + rl.lo[x]);
      end = std::max(end, w - 1 - x + rl.lo[x] + std::int32_t(rl.num[x]));
    }
    std::vector<std::uint64_t> rev(std::size_t(end - origin));
    const int base = w - 1 - dmin - origin;
    for (int k = 0; k < end - origin; ++k) {
      const int xr = base - k;
      rev[std::size_t(k)] = xr >= 0 && xr < wr ? cr(xr, y) : 0;
    }
    sgm::census_costs(cl.row(y), rev.data(), std::size_t(w), rl, origin, cost.data());
    for (int x = 0; x < w; ++x) {
      const int n = rl.num[x];
      std::uint16_t* c = cost.data() + rl.offset[x];
      const int a = std::clamp(x - dmin - (wr - 1) - rl.lo[x], 0, n);
      const int b = std::clamp(x - dmin + 1 - rl.lo[x], a, n);
      std::fill(c, c + a, static_cast<std::uint16_t>(o.invalid_cost));
      std::fill(c + b, c + n, static_cast<std::uint16_t>(o.invalid_cost));
    }
  };
  auto penalty_row = [&](int y, int dir) {
    const int yp = y - dir;
    const bool has = yp >= 0 && yp < h;
    auto pen = [&](std::uint16_t a, std::uint16_t b) { return p2_of(o, std::abs(int(a) - int(b)) >> 8); };
    for (int x = 0; x < w; ++x) {
      const int xa = x - dir, xb = x + dir;
      const std::uint16_t v = left(x, y);
      p2[x] = xa >= 0 && xa < w ? pen(v, left(xa, y)) : static_cast<std::uint16_t>(o.p2);
      p2[w + x] = has ? pen(v, left(x, yp)) : static_cast<std::uint16_t>(o.p2);
      p2[2 * w + x] = has && xa >= 0 && xa < w ? pen(v, left(xa, yp)) : static_cast<std::uint16_t>(o.p2);
      p2[3 * w + x] = has && xb >= 0 && xb < w ? pen(v, left(xb, yp)) : static_cast<std::uint16_t>(o.p2);
    }
  };
  for (const int dir : {1, -1}) {
    for (int i = 0; i < h; ++i) {
      const int y = dir > 0 ? i : h - 1 - i;
      cost_row(y);
      penalty_row(y, dir);
      sgm::SgmRowArgs a;
      a.cost = cost.data();
      a.p2 = p2.data();
      a.prev = prev.data();
      a.prev_min = prev_min.data();
      a.cur = cur.data();
      a.cur_min = cur_min.data();
      a.scratch = scratch.data();
      a.sum = volume.data() + row_start[std::size_t(y)];
      a.labels = labels(y);
      a.has_prev = i > 0;
      if (a.has_prev) a.prev_labels = labels(y - dir);
      a.width = std::size_t(w);
      a.max_num = max_num;
      a.p1 = static_cast<std::uint16_t>(o.p1);
      a.direction = dir;
      a.accumulate = dir < 0;
      (i % 2 == 0 ? sgm::sgm_row : sgm::scalar::sgm_row)(a);
      std::swap(prev, cur);
      std::swap(prev_min, cur_min);
    }
  }
  return volume;
}

TEST(row_kernels_textbook, "Row kernels driven like the tile aggregation equal the textbook volume") {
  const Texture tex(3);
  r3d::ImageU16 left, right;
  r3d::ImageF32 gt;
  render(tex, 41, 23, [](double x, double) { return 5.0 + 0.1 * x; }, left, right, gt);
  Rng rng(9);
  for (int y = 0; y < 23; ++y)
    for (int x = 0; x < 41; ++x) right(x, y) = static_cast<std::uint16_t>(right(x, y) ^ (rng.uniform_int(0, 65535) & 0x0F00));
  for (const int adapt : {0, 12})
    for (const auto& [dmin, nd] : {std::pair{-3, 16}, std::pair{2, 48}}) {
      r3d::StereoOptions o;
      o.p2_adapt = adapt;
      o.p2 = 120;
      CHECK(row_kernel_volume(left, right, dmin, nd, o, nullptr) ==
            r3d::detail::sgm_cost_volume_reference(left.view(), right.view(), dmin, nd, o));
    }
  // Per-pixel ranges (test_stereo.cpp:119-151): neighbours overlap, nest or are disjoint.
  const Texture tex2(13);
  const int w = 37, h = 19, nd = 96;
  render(tex2, w, h, [](double x, double y) { return 20.0 + 0.5 * x - 0.3 * y; }, left, right, gt);
  Rng r2(31);
  r3d::LabelRanges ranges{r3d::Image<std::uint16_t>(w, h, 1), r3d::Image<std::uint16_t>(w, h, 1)};
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const int n = 16 * r2.uniform_int(1, 3);
      ranges.num(x, y) = static_cast<std::uint16_t>(n);
      ranges.lo(x, y) = static_cast<std::uint16_t>(r2.uniform_int(0, nd - n));
    }
  r3d::StereoOptions o;
  o.p2 = 150;
  o.p2_adapt = 8;
  CHECK(row_kernel_volume(left, right, -2, nd, o, &ranges) ==
        r3d::detail::sgm_cost_volume_reference(left.view(), right.view(), -2, nd, o, &ranges));
}

TEST(wta_targets, "WTA row kernel: every target is bit-exact with scalar and the oracle") {
  // test_stereo.cpp:153-189 with the same construction.
  Rng rng(17);
  for (const int max_n : {1, 2, 3, 5}) {
    const std::size_t w = 37;
    std::vector<std::int32_t> lo(w);
    std::vector<std::uint16_t> num(w);
    std::vector<std::size_t> offset(w + 1, 0);
    std::ptrdiff_t shift = 0, top = 0;
    for (std::size_t x = 0; x < w; ++x) {
      num[x] = static_cast<std::uint16_t>(16 * rng.uniform_int(1, max_n));
      lo[x] = rng.uniform_int(0, 100);
      offset[x + 1] = offset[x] + num[x];
      shift = std::max<std::ptrdiff_t>(shift, lo[x] + num[x] - 1 - static_cast<std::ptrdiff_t>(x));
      top = std::max<std::ptrdiff_t>(top, static_cast<std::ptrdiff_t>(x) - lo[x]);
    }
    const r3d::sgm::RowLabels labels{lo.data(), num.data(), offset.data()};
    const auto nr = static_cast<std::size_t>(top + shift + 1);
    std::vector<std::uint16_t> sum(offset[w]);
    for (auto& v : sum) v = static_cast<std::uint16_t>(rng.uniform_int(0, 40));
    std::vector<std::uint16_t> best(w), cost(w), second(w), rc(nr, 0xFFFF), rb(nr, 0xFFFF);
    r3d::sgm::scalar::wta_row(sum.data(), w, labels, best.data(), cost.data(), second.data(), rc.data(), rb.data(),
                              shift);
    std::vector<std::uint16_t> ob(w), oc(w), os(w), orc(nr, 0xFFFF), orb(nr, 0xFFFF);
    oracle::wta_row(sum.data(), static_cast<int>(w), lo.data(), num.data(), ob.data(), oc.data(), os.data(),
                    orc.data(), orb.data(), shift);
    CHECK(best == ob && cost == oc && second == os && rc == orc && rb == orb);
    for (const std::int64_t target : r3d::simd_available_targets()) {
      const r3d::ScopedSimdTarget scoped(target);
      CHECK(r3d::simd_target_name(target) == "gfx950");
      std::vector<std::uint16_t> best2(w), cost2(w), second2(w), rc2(nr, 0xFFFF), rb2(nr, 0xFFFF);
      r3d::sgm::wta_row(sum.data(), w, labels, best2.data(), cost2.data(), second2.data(), rc2.data(), rb2.data(),
                        shift);
      CHECK(best2 == best);
      CHECK(cost2 == cost);
      CHECK(second2 == second);
      CHECK(rc2 == rc);
      CHECK(rb2 == rb);
    }
  }
}

TEST(census_targets, "Census: every target is bit-exact with scalar, and census_row with the image transform") {
  // test_census.cpp:67-85, with portable random images.
  for (const int w : {1, 5, 16, 31, 64, 77}) {
    Rng rng(static_cast<unsigned>(w));
    r3d::ImageU16 img(w, 13, 1);
    for (int y = 0; y < 13; ++y)
      for (int x = 0; x < w; ++x) img(x, y) = static_cast<std::uint16_t>(rng.uniform_int(0, w % 2 == 0 ? 7 : 65535));
    const r3d::Image<std::uint64_t> expected = r3d::census_transform(img.view(), true);
    const oracle::ImgU64 o = oracle::census(to_oracle(img));
    for (const std::int64_t target : r3d::simd_available_targets()) {
      const r3d::ScopedSimdTarget scoped(target);
      const r3d::Image<std::uint64_t> actual = r3d::census_transform(img.view());
      bool equal = true;
      for (int y = 0; y < 13; ++y)
        for (int x = 0; x < w; ++x) equal = equal && actual(x, y) == expected(x, y) && actual(x, y) == o(x, y);
      CHECK(equal);
    }
    // Row kernel on the padded rows (census.cpp:121-164 builds them the same way).
    auto pad = std::vector<std::vector<std::uint16_t>>(13, std::vector<std::uint16_t>(std::size_t(w) + 8));
    for (int y = 0; y < 13; ++y)
      for (int x = -4; x < w + 4; ++x) pad[y][std::size_t(x + 4)] = img(std::clamp(x, 0, w - 1), y);
    bool equal = true;
    for (int y = 0; y < 13; ++y) {
      const std::uint16_t* rows[r3d::kCensusHeight];
      for (int j = 0; j < r3d::kCensusHeight; ++j) rows[j] = pad[std::size_t(std::clamp(y - 3 + j, 0, 12))].data();
      auto a = std::vector<std::uint64_t>(std::size_t(w)), b = std::vector<std::uint64_t>(std::size_t(w));
      r3d::census_row(rows, a.data(), std::size_t(w));
      r3d::scalar::census_row(rows, b.data(), std::size_t(w));
      for (int x = 0; x < w; ++x) equal = equal && a[x] == expected(x, y) && b[x] == expected(x, y);
    }
    CHECK(equal);
  }
}

TEST(threads, "Tiles, thread counts and targets give the same map") {
  // test_stereo.cpp:191-234 (two planes with a depth step across tile borders).
  const Texture tex(5);
  r3d::ImageU16 left, right;
  r3d::ImageF32 gt;
  render(tex, 150, 90, [](double x, double) { return x < 70 ? 6.0 : 14.5; }, left, right, gt);
  r3d::StereoOptions options;
  options.tile_size = 48;
  options.tile_margin = 8;
  options.tile_blend = 8;
  options.scalar_kernels = true;
  const r3d::DisparityMap reference = r3d::match_stereo(left.view(), right.view(), {}, {}, 0, 32, options);
  options.scalar_kernels = false;
  for (const std::int64_t target : r3d::simd_available_targets()) {
    const r3d::ScopedSimdTarget scoped(target);
    CHECK(same_bits(r3d::match_stereo(left.view(), right.view(), {}, {}, 0, 32, options).disparity,
                    reference.disparity));
  }
  r3d::set_default_thread_count(1);
  CHECK(r3d::default_thread_pool().num_threads() == 1);
  const r3d::DisparityMap single = r3d::match_stereo(left.view(), right.view(), {}, {}, 0, 32, options);
  r3d::set_default_thread_count(0);
  CHECK(same_bits(single.disparity, reference.disparity));
}

TEST(cancel, "A requested token stops the matcher with r3d::Cancelled; a reset one lets it run") {
  const Texture tex(8);
  r3d::ImageU16 left, right;
  r3d::ImageF32 gt;
  render(tex, 160, 96, [](double x, double) { return 10.0 + 0.02 * x; }, left, right, gt);
  r3d::CancellationToken token;
  const r3d::CancellationScope scope(&token);
  CHECK(r3d::current_cancellation() == &token);
  token.request();
  bool cancelled = false;
  try {
    (void)r3d::match_stereo(left.view(), right.view(), {}, {}, 0, 32);
  } catch (const r3d::Cancelled&) {
    cancelled = true;
  }
  CHECK(cancelled);
  cancelled = false;
  try {
    r3d::throw_if_cancelled();
  } catch (const r3d::Cancelled&) {
    cancelled = true;
  }
  CHECK(cancelled);
  token.reset();
  const r3d::DisparityMap m = r3d::match_stereo(left.view(), right.view(), {}, {}, 0, 32);
  CHECK(m.levels >= 0 && m.cells > 0);
  // Requested from another thread while a large match runs: the call stops
  // at its next poll point (between levels and tile chunks).
  r3d::ImageU16 bl, br;
  render(tex, 1600, 1200, [](double x, double y) { return 40.0 + 0.02 * x + 0.01 * y; }, bl, br, gt);
  std::thread t([&] {
    std::this_thread::sleep_for(std::chrono::milliseconds(2));
    token.request();
  });
  cancelled = false;
  try {
    r3d::StereoOptions o;
    o.levels = 0;
    o.tile_size = 128;
    o.memory_budget = std::size_t{64} << 20;
    for (int i = 0; i < 50 && !cancelled; ++i) (void)r3d::match_stereo(bl.view(), br.view(), {}, {}, 0, 256, o);
  } catch (const r3d::Cancelled&) {
    cancelled = true;
  }
  t.join();
  CHECK(cancelled);
  token.reset();
  {
    const r3d::CancellationScope none(nullptr);
    CHECK(r3d::current_cancellation() == nullptr);
  }
  CHECK(r3d::current_cancellation() == &token);
}

}  // namespace

int main() {
  for (const Case& c : cases()) {
    const int before = g_failed;
    try {
      c.fn();
    } catch (const std::exception& e) {
      ++g_failed;
      std::printf("FAILED %s: exception %s\n", c.name, e.what());
    }
    std::printf("%s: %s\n", g_failed == before ? "ok" : "FAIL", c.name);
  }
  std::printf("%d checks, %d failed\n", g_checks, g_failed);
  return g_failed == 0 ? 0 : 1;
}
