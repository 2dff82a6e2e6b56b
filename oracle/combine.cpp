// CPU oracle of combine_pairs -- TEST INFRASTRUCTURE ONLY.
//
// Restates rhiestan/Regard3D_MVS src/r3d/mvs/multi_stereo.cpp:44-72
// (sample_disparity) and :103-213 (the per-pixel part of combine_pairs) for
// the per-pair geometry the caller computes (:84-96).  Double precision in
// the reference's order of operations; compiled with -ffp-contract=off.
//
// Pinning (tests/test_oracle_combine.py): the sample_disparity KATs of
// tests/test_mvs.cpp:217-227; a slanted plane seen by three pairs, one of them
// 6 pixels off, triangulates to the plane within 1e-3 with the wrong pair
// outvoted (test_mvs.cpp:130-172); ties, min_pairs and narrow triangulation
// angles are rejected (test_mvs.cpp:174-215); the geometry is built in numpy
// from cameras, independently of this file.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>

#include "oracle.hpp"

namespace oracle {

float sample_disparity(const ImgF& d, double u, double v) {
  const int w = d.w, h = d.h;
  if (!(u > -0.5 && v > -0.5 && u < w - 0.5 && v < h - 0.5)) return std::numeric_limits<float>::infinity();
  const int x0 = std::clamp(static_cast<int>(std::floor(u)), 0, w - 1);
  const int y0 = std::clamp(static_cast<int>(std::floor(v)), 0, h - 1);
  const int x1 = std::min(x0 + 1, w - 1);
  const int y1 = std::min(y0 + 1, h - 1);
  const float d00 = d(x0, y0), d10 = d(x1, y0), d01 = d(x0, y1), d11 = d(x1, y1);
  if (std::isfinite(d00) && std::isfinite(d10) && std::isfinite(d01) && std::isfinite(d11)) {
    const float lo = std::min({d00, d10, d01, d11});
    const float hi = std::max({d00, d10, d01, d11});
    if (hi - lo <= 1.0F) {
      const auto fx = static_cast<float>(std::clamp(u - x0, 0.0, 1.0));
      const auto fy = static_cast<float>(std::clamp(v - y0, 0.0, 1.0));
      const float top = d00 + fx * (d10 - d00);
      const float bottom = d01 + fx * (d11 - d01);
      return top + fy * (bottom - top);
    }
  }
  const int xn = std::clamp(static_cast<int>(std::lround(u)), 0, w - 1);
  const int yn = std::clamp(static_cast<int>(std::lround(v)), 0, h - 1);
  return d(xn, yn);
}

namespace {

struct Cand {
  double w = 0.0, k = 0.0;
  float confidence = 0.0F;
  int pair = 0;
};

void normalize(double* a) {
  const double n2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
  if (n2 > 0.0) {
    const double n = std::sqrt(n2);
    for (int i = 0; i < 3; ++i) a[i] = a[i] / n;
  }
}

}  // namespace

Combined combine_pairs(const CombineGridIn& g, const std::vector<CombinePairIn>& pairs, const CombineOptions& o) {
  if (pairs.size() > 16) throw Error("combine_pairs: at most 16 pairs per view");
  Combined out;
  out.depth = ImgF(g.width, g.height);
  out.confidence = ImgF(g.width, g.height);
  out.pair_mask = Img<std::uint16_t>(g.width, g.height);
  const int min_pairs = std::max(1, o.min_pairs);
  Cand cand[16];
  int used[16];
  for (int yy = 0; yy < g.height; ++yy) {
    for (int x = 0; x < g.width; ++x) {
      int n = 0;
      for (std::size_t i = 0; i < pairs.size(); ++i) {
        const CombinePairIn& p = pairs[i];
        const double* m = p.grid_to_rect;
        const double rx = m[0] * x + m[1] * yy + m[2] * 1.0;
        const double ry = m[3] * x + m[4] * yy + m[5] * 1.0;
        const double c = m[6] * x + m[7] * yy + m[8] * 1.0;
        if (!(c > 1e-9)) continue;
        const double u = rx / c, v = ry / c;
        const float d = sample_disparity(p.disparity, u, v);
        if (!std::isfinite(d) || d <= p.offset) continue;
        Cand& cd = cand[n++];
        cd.k = p.fb / c;
        cd.w = (d - p.offset) / cd.k;
        cd.pair = static_cast<int>(i);
        cd.confidence = 0.0F;
        if (!p.confidence.empty()) {
          const int xn = std::clamp(static_cast<int>(std::lround(u)), 0, p.confidence.w - 1);
          const int yn = std::clamp(static_cast<int>(std::lround(v)), 0, p.confidence.h - 1);
          cd.confidence = p.confidence(xn, yn);
        }
      }
      if (n < min_pairs) continue;
      int best = -1, best_support = 0;
      float best_conf = -1.0F;
      for (int i = 0; i < n; ++i) {
        int support = 0;
        float conf = 0.0F;
        for (int j = 0; j < n; ++j) {
          if (std::abs(cand[j].k * (cand[j].w - cand[i].w)) <= o.inlier_px) {
            ++support;
            conf += cand[j].confidence;
          }
        }
        if (support > best_support || (support == best_support && conf > best_conf)) {
          best = i;
          best_support = support;
          best_conf = conf;
        }
      }
      int m = 0;
      for (int j = 0; j < n; ++j)
        if (std::abs(cand[j].k * (cand[j].w - cand[best].w)) <= o.inlier_px) used[m++] = j;
      double w_opt = 0.0;
      while (m > 0) {
        double sw = 0.0, swx = 0.0;
        for (int t = 0; t < m; ++t) {
          const Cand& cd = cand[used[t]];
          sw += cd.k * cd.k;
          swx += cd.k * cd.k * cd.w;
        }
        w_opt = swx / sw;
        int worst = -1;
        double worst_res = o.max_reprojection_px;
        for (int t = 0; t < m; ++t) {
          const Cand& cd = cand[used[t]];
          const double res = std::abs(cd.k * (cd.w - w_opt));
          if (res > worst_res) {
            worst_res = res;
            worst = t;
          }
        }
        if (worst < 0) break;
        used[worst] = used[--m];
      }
      const bool majority = m > n - m || (m >= 2 && m >= n - m);
      if (m < min_pairs || !majority || !(w_opt > 0.0)) continue;
      const double z = 1.0 / w_opt;
      const double pix[3] = {static_cast<double>(x), static_cast<double>(yy), 1.0};
      double t[3], xw[3], tb[3];
      for (int i = 0; i < 3; ++i) t[i] = g.inv_k[3 * i] * pix[0] + g.inv_k[3 * i + 1] * pix[1] + g.inv_k[3 * i + 2] * pix[2];
      for (int i = 0; i < 3; ++i) {
        const double ray = g.rot_t[3 * i] * t[0] + g.rot_t[3 * i + 1] * t[1] + g.rot_t[3 * i + 2] * t[2];
        xw[i] = g.center[i] + z * ray;
        tb[i] = g.center[i] - xw[i];
      }
      normalize(tb);
      bool wide = false;
      float conf = 0.0F;
      std::uint16_t mask = 0;
      for (int k = 0; k < m; ++k) {
        const Cand& cd = cand[used[k]];
        mask = static_cast<std::uint16_t>(mask | (1U << cd.pair));
        const double* pc = pairs[static_cast<std::size_t>(cd.pair)].partner_center;
        double tp[3] = {pc[0] - xw[0], pc[1] - xw[1], pc[2] - xw[2]};
        normalize(tp);
        wide = wide || tb[0] * tp[0] + tb[1] * tp[1] + tb[2] * tp[2] <= o.cos_min_angle;
        conf += cd.confidence;
      }
      if (!wide) continue;
      out.depth(x, yy) = static_cast<float>(z);
      out.confidence(x, yy) = conf / static_cast<float>(m);
      out.pair_mask(x, yy) = mask;
    }
  }
  return out;
}

}  // namespace oracle
