// C++ tests of the r3d drop-in API (libr3d_stereo.so over the GPU C ABI),
// written the way the reference's tests/test_stereo.cpp and test_census.cpp
// call it.  Procedural scenes restate tests/test_util.hpp:22-39 and
// tests/test_stereo.cpp:19-73; the GPU results are also compared bit for bit
// with the CPU oracle.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "../../oracle/oracle.hpp"
#include "r3d/core/pyramid.hpp"
#include "r3d/stereo/census.hpp"
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
