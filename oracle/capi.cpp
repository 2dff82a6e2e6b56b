// extern "C" entry points of the CPU oracle -- TEST INFRASTRUCTURE ONLY.
// Contiguous arrays (stride == width); the Python side is oracle/oracle.py.
#include <cstring>
#include <string>

#include "oracle.hpp"

namespace {
thread_local std::string g_error;

template <typename T>
oracle::Img<T> copy_in(const T* p, int w, int h) {
  oracle::Img<T> img;
  if (p == nullptr || w <= 0 || h <= 0) return img;
  img = oracle::Img<T>(w, h);
  std::memcpy(img.px.data(), p, sizeof(T) * img.px.size());
  return img;
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

oracle::Ranges ranges_in(const std::uint16_t* lo, const std::uint16_t* num, int w, int h) {
  return oracle::Ranges{copy_in(lo, w, h), copy_in(num, w, h)};
}
}  // namespace

template <typename T>
static oracle::Img<T> copy_strided(const T* p, int w, int h, std::int64_t stride) {
  oracle::Img<T> m(w, h);
  for (int y = 0; y < h; ++y) std::memcpy(m.row(y), p + std::int64_t(y) * stride, sizeof(T) * std::size_t(w));
  return m;
}

static oracle::Camera camera_of(const r3dg_camera* c) {
  oracle::Camera o;
  o.model = c->model;
  o.width = c->width;
  o.height = c->height;
  o.fx = c->fx;
  o.fy = c->fy;
  o.cx = c->cx;
  o.cy = c->cy;
  std::memcpy(o.params, c->params, sizeof(o.params));
  std::memcpy(o.rotation, c->rotation, sizeof(o.rotation));
  std::memcpy(o.center, c->center, sizeof(o.center));
  return o;
}

template <typename T>
static int rectify_impl(const T* src, int w, int h, const r3dg_camera* cam, const r3dg_camera* target, T* out,
                        std::uint8_t* valid) {
  return guarded([&] {
    const oracle::Camera c = camera_of(cam), t = camera_of(target);
    const auto s = copy_in(src, w, h);
    oracle::ImgU8 v;
    const auto r = oracle::rectify_image(s, c, t, &v);
    std::memcpy(out, r.px.data(), sizeof(T) * r.px.size());
    if (valid != nullptr) std::memcpy(valid, v.px.data(), v.px.size());
  });
}

extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }

int oracle_census(const std::uint16_t* img, int w, int h, std::uint64_t* out) {
  return guarded([&] {
    const auto c = oracle::census(copy_in(img, w, h));
    std::memcpy(out, c.px.data(), sizeof(std::uint64_t) * c.px.size());
  });
}

int oracle_pyramid_down_u16(const std::uint16_t* img, int w, int h, std::uint16_t* out) {
  return guarded([&] {
    const auto p = oracle::pyramid_down(copy_in(img, w, h));
    std::memcpy(out, p.px.data(), sizeof(std::uint16_t) * p.px.size());
  });
}

int oracle_pyramid_down_u8(const std::uint8_t* img, int w, int h, std::uint8_t* out) {
  return guarded([&] {
    const auto p = oracle::pyramid_down(copy_in(img, w, h));
    std::memcpy(out, p.px.data(), p.px.size());
  });
}

int oracle_reduce_mask(const std::uint8_t* img, int w, int h, std::uint8_t* out) {
  return guarded([&] {
    const auto p = oracle::reduce_mask(copy_in(img, w, h));
    std::memcpy(out, p.px.data(), p.px.size());
  });
}

int oracle_automatic_levels(int w, int h, int nd) { return oracle::automatic_levels(w, h, nd); }

static int volume(bool textbook, const std::uint16_t* left, int wl, int h, const std::uint16_t* right, int wr,
                  int dmin, int nd, const r3dg_options* o, const std::uint16_t* lo, const std::uint16_t* num,
                  std::uint16_t* out, std::uint64_t cells) {
  return guarded([&] {
    const auto L = copy_in(left, wl, h);
    const auto R = copy_in(right, wr, h);
    const bool has = lo != nullptr && num != nullptr;
    const oracle::Ranges r = has ? ranges_in(lo, num, wl, h) : oracle::Ranges{};
    const auto S = textbook ? oracle::textbook_volume(L, R, dmin, nd, *o, has ? &r : nullptr)
                            : oracle::cost_volume(L, R, dmin, nd, *o, has ? &r : nullptr);
    if (S.size() != cells) throw oracle::Error("oracle: output size does not match the cell count");
    std::memcpy(out, S.data(), sizeof(std::uint16_t) * S.size());
  });
}

int oracle_cost_volume(const std::uint16_t* left, int wl, int h, const std::uint16_t* right, int wr, int dmin,
                       int nd, const r3dg_options* o, const std::uint16_t* lo, const std::uint16_t* num,
                       std::uint16_t* out, std::uint64_t cells) {
  return volume(false, left, wl, h, right, wr, dmin, nd, o, lo, num, out, cells);
}

int oracle_textbook_volume(const std::uint16_t* left, int wl, int h, const std::uint16_t* right, int wr, int dmin,
                           int nd, const r3dg_options* o, const std::uint16_t* lo, const std::uint16_t* num,
                           std::uint16_t* out, std::uint64_t cells) {
  return volume(true, left, wl, h, right, wr, dmin, nd, o, lo, num, out, cells);
}

int oracle_wta_row(const std::uint16_t* sum, int w, const std::int32_t* lo, const std::uint16_t* num,
                   std::uint16_t* best, std::uint16_t* best_cost, std::uint16_t* second, std::uint16_t* right_cost,
                   std::uint16_t* right_best, std::int64_t right_shift) {
  return guarded([&] { oracle::wta_row(sum, w, lo, num, best, best_cost, second, right_cost, right_best, right_shift); });
}

int oracle_ranges_from_coarse(const float* coarse, int wc, int hc, int width, int height, const std::uint8_t* fine_valid,
                              const std::uint8_t* partner_valid, int partner_width, int partner_sign, int dmin,
                              int nd, const r3dg_options* o, int mirrored, std::uint16_t* lo, std::uint16_t* num) {
  return guarded([&] {
    const auto r = oracle::ranges_from_coarse(copy_in(coarse, wc, hc), width, height, copy_in(fine_valid, width, height),
                                              copy_in(partner_valid, partner_width, height), partner_width,
                                              partner_sign, dmin, nd, *o, mirrored != 0);
    std::memcpy(lo, r.lo.px.data(), sizeof(std::uint16_t) * r.lo.px.size());
    std::memcpy(num, r.num.px.data(), sizeof(std::uint16_t) * r.num.px.size());
  });
}

int oracle_remove_speckles(float* d, int w, int h, int min_size, float max_diff) {
  return guarded([&] {
    auto img = copy_in(d, w, h);
    oracle::remove_speckles(img, min_size, max_diff);
    if (!img.px.empty()) std::memcpy(d, img.px.data(), sizeof(float) * img.px.size());
  });
}

int oracle_median3x3(const float* d, int w, int h, float* out) {
  return guarded([&] {
    const auto m = oracle::median3x3(copy_in(d, w, h));
    if (!m.px.empty()) std::memcpy(out, m.px.data(), sizeof(float) * m.px.size());
  });
}

// ext: per tile x0, y0, w, h, core c0, c1, r0, r1; d[i] / c[i]: tile i's maps
// (c may be null: no confidence).
int oracle_blend(int width, int height, int blend, int ntiles, const int* ext, const float* const* d,
                 const float* const* c, float* out_d, float* out_c) {
  return guarded([&] {
    std::vector<oracle::BlendTile> tiles(static_cast<std::size_t>(ntiles));
    for (int i = 0; i < ntiles; ++i) {
      const int* e = ext + 8 * i;
      tiles[std::size_t(i)] = oracle::BlendTile{e[0], e[1], e[2], e[3], {e[4], e[5], e[6], e[7]}, d[i],
                                                c != nullptr ? c[i] : nullptr, nullptr};
    }
    oracle::ImgF disp, conf;
    oracle::blend_tiles(width, height, blend, tiles, disp, c != nullptr ? &conf : nullptr, nullptr);
    std::memcpy(out_d, disp.px.data(), sizeof(float) * disp.px.size());
    if (c != nullptr) std::memcpy(out_c, conf.px.data(), sizeof(float) * conf.px.size());
  });
}

int oracle_lr_mismatch(const float* map, int w, int h, const float* other, int wo, int sign, int max_diff,
                       std::uint8_t* out) {
  return guarded([&] {
    const auto m = oracle::left_right_mismatch(copy_in(map, w, h), copy_in(other, wo, h), sign, max_diff);
    std::memcpy(out, m.px.data(), m.px.size());
  });
}

int oracle_match(const std::uint16_t* left, int wl, int h, const std::uint16_t* right, int wr,
                 const std::uint8_t* left_valid, const std::uint8_t* right_valid, int dmin, int nd,
                 const r3dg_options* o, int threads, float* disparity, float* confidence, std::int16_t* labels,
                 std::uint8_t* lr_mask, r3dg_match_info* info) {
  return guarded([&] {
    const auto r = oracle::match_stereo(copy_in(left, wl, h), copy_in(right, wr, h), copy_in(left_valid, wl, h),
                                        copy_in(right_valid, wr, h), dmin, nd, *o, threads);
    const std::size_t n = std::size_t(wl) * std::size_t(h);
    std::memcpy(disparity, r.disparity.px.data(), sizeof(float) * n);
    if (confidence != nullptr) std::memcpy(confidence, r.confidence.px.data(), sizeof(float) * n);
    if (labels != nullptr) std::memcpy(labels, r.labels.px.data(), sizeof(std::int16_t) * n);
    if (lr_mask != nullptr) std::memcpy(lr_mask, r.lr_mask.px.data(), n);
    if (info != nullptr) {
      info->disparity_min = r.disparity_min;
      info->num_disparities = r.num_disparities;
      info->levels = r.levels;
      info->tiles_level0 = r.tiles_level0;
      info->cells = r.cells;
      info->tile_pixels = r.tile_pixels;
      info->seconds = 0.0;
    }
  });
}

int oracle_rectify_u8(const std::uint8_t* src, int w, int h, const r3dg_camera* cam, const r3dg_camera* target, std::uint8_t* out,
                      std::uint8_t* valid) {
  return rectify_impl(src, w, h, cam, target, out, valid);
}
int oracle_rectify_u16(const std::uint16_t* src, int w, int h, const r3dg_camera* cam, const r3dg_camera* target,
                       std::uint16_t* out, std::uint8_t* valid) {
  return rectify_impl(src, w, h, cam, target, out, valid);
}
int oracle_rectify_f32(const float* src, int w, int h, const r3dg_camera* cam, const r3dg_camera* target, float* out,
                       std::uint8_t* valid) {
  return rectify_impl(src, w, h, cam, target, out, valid);
}

float oracle_sample_disparity(const float* d, int w, int h, double u, double v) {
  float r = 0.0F;
  guarded([&] { r = oracle::sample_disparity(copy_in(d, w, h), u, v); });
  return r;
}

int oracle_combine_pairs(const r3dg_grid* grid, const r3dg_pair_geometry* pairs, int count,
                         const r3dg_multi_options* options, float* depth, float* confidence, std::uint16_t* mask) {
  return guarded([&] {
    oracle::CombineGridIn g;
    g.width = grid->width;
    g.height = grid->height;
    std::memcpy(g.inv_k, grid->inv_k, sizeof(g.inv_k));
    std::memcpy(g.rot_t, grid->rot_t, sizeof(g.rot_t));
    std::memcpy(g.center, grid->center, sizeof(g.center));
    std::vector<oracle::CombinePairIn> ps(static_cast<std::size_t>(count));
    for (int i = 0; i < count; ++i) {
      const r3dg_pair_geometry& q = pairs[i];
      oracle::CombinePairIn& p = ps[static_cast<std::size_t>(i)];
      std::memcpy(p.grid_to_rect, q.grid_to_rect, sizeof(p.grid_to_rect));
      p.fb = q.fb;
      p.offset = q.offset;
      std::memcpy(p.partner_center, q.partner_center, sizeof(p.partner_center));
      p.disparity = copy_strided(static_cast<const float*>(q.disparity.data), q.disparity.width, q.disparity.height,
                                 q.disparity.stride);
      if (q.confidence.data != nullptr && q.confidence.width > 0 && q.confidence.height > 0)
        p.confidence = copy_strided(static_cast<const float*>(q.confidence.data), q.confidence.width,
                                    q.confidence.height, q.confidence.stride);
    }
    oracle::CombineOptions o;
    o.inlier_px = options->inlier_px;
    o.max_reprojection_px = options->max_reprojection_px;
    o.cos_min_angle = options->cos_min_angle;
    o.min_pairs = options->min_pairs;
    const auto r = oracle::combine_pairs(g, ps, o);
    std::memcpy(depth, r.depth.px.data(), sizeof(float) * r.depth.px.size());
    std::memcpy(confidence, r.confidence.px.data(), sizeof(float) * r.confidence.px.size());
    std::memcpy(mask, r.pair_mask.px.data(), sizeof(std::uint16_t) * r.pair_mask.px.size());
  });
}

}  // extern "C"
