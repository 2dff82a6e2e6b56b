// Host AddressSanitizer + UndefinedBehaviorSanitizer run of the oracle
// (make -C oracle sanitize): every oracle entry point on small seeded inputs,
// including one-pixel and odd widths, per-pixel ranges, masks, unequal
// widths, both left-right checks, several levels, tile blending, the
// filters, rectification and combine_pairs.  Also checks that the row-pass
// volume equals the textbook volume on each case, so a run that sanitizes
// clean is also a consistent one.  Test infrastructure only.
#include <cmath>
#include <cstdio>
#include <random>

#include "oracle.hpp"

namespace {

int g_failed = 0;
#define EXPECT(c)                                                 \
  do {                                                            \
    if (!(c)) {                                                   \
      ++g_failed;                                                 \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);  \
    }                                                             \
  } while (0)

r3dg_options defaults() {
  r3dg_options o{};
  o.p1 = 28;
  o.p2 = 200;
  o.p2_adapt = 4;
  o.invalid_cost = 31;
  o.subpixel = R3DG_SUBPIXEL_EQUIANGULAR;
  o.uniqueness = 0.0F;
  o.lr_max_diff = 1;
  o.lr_full_match = 1;
  o.speckle_size = 100;
  o.speckle_diff = 1.0F;
  o.median = 1;
  o.levels = -1;
  o.range_radius = 3;
  o.range_margin = 4;
  o.range_fill = 64;
  o.tile_size = 512;
  o.tile_margin = 32;
  o.tile_blend = 32;
  o.memory_budget = std::uint64_t{2} << 30;
  return o;
}

oracle::ImgU16 noise(int w, int h, unsigned seed) {
  std::mt19937 g(seed);
  oracle::ImgU16 img(w, h);
  for (auto& v : img.px) v = static_cast<std::uint16_t>((g() & 255U) * 257U);
  return img;
}

// Right image: the left shifted by `d` with fresh noise at the border.
oracle::ImgU16 shifted(const oracle::ImgU16& l, int wr, int d, unsigned seed) {
  std::mt19937 g(seed);
  oracle::ImgU16 r(wr, l.h);
  for (int y = 0; y < l.h; ++y)
    for (int x = 0; x < wr; ++x) {
      const int xl = x + d;
      r(x, y) = xl < l.w ? l(xl, y) : static_cast<std::uint16_t>((g() & 255U) * 257U);
    }
  return r;
}

void volumes(int w, int h, int wr, int dmin, int nd, bool with_ranges, unsigned seed) {
  const oracle::ImgU16 l = noise(w, h, seed), r = shifted(l, wr, 3, seed + 1);
  r3dg_options o = defaults();
  oracle::Ranges rg;
  const int D = oracle::align16(nd);
  if (with_ranges) {
    std::mt19937 g(seed + 2);
    rg.lo = oracle::ImgU16(w, h);
    rg.num = oracle::ImgU16(w, h);
    for (std::size_t i = 0; i < rg.lo.px.size(); ++i) {
      const int n = 16 * static_cast<int>(1 + g() % static_cast<unsigned>(D / 16));
      rg.num.px[i] = static_cast<std::uint16_t>(n);
      rg.lo.px[i] = static_cast<std::uint16_t>(g() % static_cast<unsigned>(D - n + 1));
    }
  }
  const auto a = oracle::cost_volume(l, r, dmin, nd, o, with_ranges ? &rg : nullptr);
  const auto b = oracle::textbook_volume(l, r, dmin, nd, o, with_ranges ? &rg : nullptr);
  EXPECT(a == b);
}

void match(int w, int h, int wr, int dmin, int nd, int levels, bool masks, bool full_lr, int subpixel, unsigned seed) {
  const oracle::ImgU16 l = noise(w, h, seed), r = shifted(l, wr, 5, seed + 1);
  r3dg_options o = defaults();
  o.levels = levels;
  o.lr_full_match = full_lr ? 1 : 0;
  o.subpixel = subpixel;
  o.tile_size = 64;  // several tiles, so the blend runs
  oracle::ImgU8 lv, rv;
  if (masks) {
    lv = oracle::ImgU8(w, h, 255);
    rv = oracle::ImgU8(wr, h, 255);
    for (int y = 0; y < h; ++y) {
      lv(0, y) = 0;
      rv(wr - 1, y) = 0;
    }
  }
  const auto m = oracle::match_stereo(l, r, lv, rv, dmin, nd, o, 2);
  EXPECT(m.disparity.w == w && m.disparity.h == h);
  int finite = 0;
  for (float v : m.disparity.px) finite += std::isfinite(v) ? 1 : 0;
  EXPECT(finite > 0 || w * h < 64);
}

}  // namespace

int main() {
  for (int w = 1; w <= 9; ++w) (void)oracle::census(noise(w, 3, 10U + static_cast<unsigned>(w)));
  (void)oracle::pyramid_down(noise(13, 7, 1));
  (void)oracle::reduce_mask(oracle::ImgU8(13, 7, 200));
  volumes(41, 23, 41, 0, 32, false, 20);
  volumes(37, 19, 30, -3, 48, true, 21);
  volumes(1, 5, 1, 0, 16, false, 22);
  match(120, 70, 110, 0, 32, 0, false, true, R3DG_SUBPIXEL_EQUIANGULAR, 30);
  match(150, 90, 150, 0, 48, 1, true, true, R3DG_SUBPIXEL_PARABOLA, 31);
  match(150, 90, 140, -2, 48, 0, true, false, R3DG_SUBPIXEL_NONE, 32);
  match(300, 200, 300, 0, 96, -1, false, true, R3DG_SUBPIXEL_EQUIANGULAR, 33);
  {
    oracle::ImgF d(60, 40, 1.0F);
    for (int y = 10; y < 14; ++y)
      for (int x = 10; x < 14; ++x) d(x, y) = 9.0F;
    d(30, 30) = oracle::kInvalid;
    oracle::remove_speckles(d, 20, 1.0F);
    EXPECT(!std::isfinite(d(11, 11)) && std::isfinite(d(0, 0)));
    (void)oracle::median3x3(d);
    (void)oracle::left_right_mismatch(d, d, 1, 1);
  }
  {
    oracle::Camera cam;
    cam.width = 64;
    cam.height = 48;
    cam.fx = cam.fy = 60.0;
    cam.cx = 31.5;
    cam.cy = 23.5;
    cam.rotation[0] = cam.rotation[4] = cam.rotation[8] = 1.0;
    oracle::Camera tgt = cam;
    tgt.width = 70;
    tgt.height = 50;
    oracle::ImgU8 valid;
    const auto out = oracle::rectify_image(noise(64, 48, 40), cam, tgt, &valid);
    EXPECT(out.w == 70 && valid.h == 50);
  }
  {
    oracle::CombineGridIn g;
    g.width = 32;
    g.height = 24;
    g.inv_k[0] = g.inv_k[4] = 1.0 / 30.0;
    g.inv_k[2] = -0.5;
    g.inv_k[5] = -0.4;
    g.inv_k[8] = 1.0;
    g.rot_t[0] = g.rot_t[4] = g.rot_t[8] = 1.0;
    oracle::CombinePairIn p;
    p.grid_to_rect[0] = p.grid_to_rect[4] = p.grid_to_rect[8] = 1.0;
    p.fb = 30.0;
    p.partner_center[0] = 1.0;
    p.disparity = oracle::ImgF(32, 24, 6.0F);
    oracle::CombineOptions o;
    const auto c = oracle::combine_pairs(g, {p}, o);
    EXPECT(c.depth.w == 32);
  }
  std::printf("sanitize_check: %s\n", g_failed == 0 ? "ok" : "FAILED");
  return g_failed == 0 ? 0 : 1;
}
