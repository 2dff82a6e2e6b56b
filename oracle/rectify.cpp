// CPU oracle of rectify_image -- TEST INFRASTRUCTURE ONLY.
//
// Restates rhiestan/Regard3D_MVS src/r3d/rectify/rectification.cpp:288-350
// (rectify_image), :152-164 (pyramid_camera), src/r3d/core/camera_model.cpp:
// 68-103 (distort), src/r3d/core/sampling.hpp:16-55 (inside_image,
// cubic_weights, sample_bicubic) and src/r3d/core/pyramid.cpp:34-106
// (pyramid_down, float accumulators for float images) for one channel.
// Compiled with -ffp-contract=off: every double and float operation is a
// separate IEEE operation, in the reference's order.
//
// Pinning (tests/test_oracle_rectify.py): identity and integer-shift cameras
// reproduce the source exactly; Catmull-Rom reproduces linear ramps; the
// distortion models agree with a Newton inverse restated in numpy; the
// reference's own behaviour test (a Gaussian blob lands on its rectified
// centre to 0.05 px, tests/test_rectification.cpp:128-166) is restated; the
// anti-aliasing level matches a float pyramid computed in numpy.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <type_traits>

#include "oracle.hpp"

namespace oracle {
namespace {

void fail(const char* m) { throw Error(m); }

// pyramid.cpp:34-85 with float sums for float images.
ImgF pyramid_float(const ImgF& s) {
  const int w = s.w, h = s.h, dw = (w + 1) / 2, dh = (h + 1) / 2;
  ImgF tmp(dw, h);
  for (int y = 0; y < h; ++y) {
    const float* in = s.row(y);
    for (int x = 0; x < dw; ++x) {
      const int sx = 2 * x;
      tmp(x, y) = in[std::max(sx - 2, 0)] + in[std::min(sx + 2, w - 1)] +
                  4.0F * (in[std::max(sx - 1, 0)] + in[std::min(sx + 1, w - 1)]) + 6.0F * in[sx];
    }
  }
  ImgF out(dw, dh);
  for (int y = 0; y < dh; ++y) {
    const int sy = 2 * y;
    for (int x = 0; x < dw; ++x) {
      const float sum = tmp(x, std::max(sy - 2, 0)) + tmp(x, std::min(sy + 2, h - 1)) +
                        4.0F * (tmp(x, std::max(sy - 1, 0)) + tmp(x, std::min(sy + 1, h - 1))) + 6.0F * tmp(x, sy);
      out(x, y) = sum * (1.0F / 256.0F);
    }
  }
  return out;
}

template <typename T>
Img<T> down(const Img<T>& s) {
  if constexpr (std::is_same_v<T, float>) {
    return pyramid_float(s);
  } else {
    return pyramid_down(s);
  }
}

// camera_model.cpp:68-103 (OpenMVG formulas).
void distort(const Camera& c, double ux, double uy, double& dx, double& dy) {
  const double r2 = ux * ux + uy * uy;
  const double* p = c.params;
  if (c.model == 1) {
    const double f = 1.0 + p[0] * r2;
    dx = ux * f;
    dy = uy * f;
  } else if (c.model == 2) {
    const double f = 1.0 + r2 * (p[0] + r2 * (p[1] + r2 * p[2]));
    dx = ux * f;
    dy = uy * f;
  } else if (c.model == 3) {
    const double k1 = p[0], k2 = p[1], k3 = p[2], t1 = p[3], t2 = p[4];
    const double radial = r2 * (k1 + r2 * (k2 + r2 * k3));
    const double tx = t2 * (r2 + 2.0 * ux * ux) + 2.0 * t1 * ux * uy;
    const double ty = t1 * (r2 + 2.0 * uy * uy) + 2.0 * t2 * ux * uy;
    dx = ux + ux * radial + tx;
    dy = uy + uy * radial + ty;
  } else if (c.model == 4 && std::sqrt(r2) > 1e-8) {
    const double r = std::sqrt(r2);
    const double theta = std::atan(r);
    const double t2 = theta * theta;
    const double theta_d = theta * (1.0 + t2 * (p[0] + t2 * (p[1] + t2 * (p[2] + t2 * p[3]))));
    dx = ux * (theta_d / r);
    dy = uy * (theta_d / r);
  } else {
    dx = ux;
    dy = uy;
  }
}

// sampling.hpp:21-28
void cubic_weights(float t, float w[4]) {
  const float t2 = t * t;
  const float t3 = t2 * t;
  w[0] = -0.5F * t3 + t2 - 0.5F * t;
  w[1] = 1.5F * t3 - 2.5F * t2 + 1.0F;
  w[2] = -1.5F * t3 + 2.0F * t2 + 0.5F * t;
  w[3] = 0.5F * t3 - 0.5F * t2;
}

// rectification.cpp:140-148
template <typename T>
T to_pixel(float v) {
  if constexpr (std::is_floating_point_v<T>) {
    return static_cast<T>(v);
  } else {
    return static_cast<T>(std::clamp(std::round(v), 0.0F, static_cast<float>(std::numeric_limits<T>::max())));
  }
}

}  // namespace

template <typename T>
Img<T> rectify_image(const Img<T>& src, const Camera& cam, const Camera& target, ImgU8* valid) {
  if (src.w != cam.width || src.h != cam.height) fail("rectify_image: image size differs from the camera's");
  const double* a = cam.center;
  const double* b = target.center;
  const double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
  if (std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) > 1e-9 * (1.0 + std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])))
    fail("rectify_image: the target camera must have the same centre as the source");
  int level = 0;
  while (cam.fx / target.fx > 1.5 * std::ldexp(1.0, level) && (cam.width >> (level + 1)) >= 16 &&
         (cam.height >> (level + 1)) >= 16)
    ++level;
  Img<T> source = src;
  for (int l = 0; l < level && !source.empty(); ++l) source = down(source);
  Camera sc = cam;
  const double s = std::ldexp(1.0, -level);
  sc.fx *= s;
  sc.fy *= s;
  sc.cx *= s;
  sc.cy *= s;
  // m = R_source * R_target^T (row i of R_source against row j of R_target).
  double m[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      m[3 * i + j] = cam.rotation[3 * i] * target.rotation[3 * j] + cam.rotation[3 * i + 1] * target.rotation[3 * j + 1] +
                     cam.rotation[3 * i + 2] * target.rotation[3 * j + 2];
  Img<T> out(target.width, target.height);
  if (valid != nullptr) *valid = ImgU8(target.width, target.height);
  for (int y = 0; y < target.height; ++y) {
    const double ry = (y - target.cy) / target.fy;
    for (int x = 0; x < target.width; ++x) {
      const double vx = (x - target.cx) / target.fx;
      const double r0 = m[0] * vx + m[1] * ry + m[2] * 1.0;
      const double r1 = m[3] * vx + m[4] * ry + m[5] * 1.0;
      const double r2 = m[6] * vx + m[7] * ry + m[8] * 1.0;
      bool inside = r2 > 1e-9 && !source.empty();  // an empty source has no pixel inside
      double px = 0.0, py = 0.0;
      if (inside) {
        double dx, dy;
        distort(sc, r0 / r2, r1 / r2, dx, dy);
        px = sc.fx * dx + sc.cx;
        py = sc.fy * dy + sc.cy;
        inside = px >= -0.5 && py >= -0.5 && px <= source.w - 0.5 && py <= source.h - 0.5;
      }
      T v{};
      if (inside) {
        const double fx = std::floor(px), fy = std::floor(py);
        float wx[4], wy[4];
        cubic_weights(static_cast<float>(px - fx), wx);
        cubic_weights(static_cast<float>(py - fy), wy);
        const int x0 = static_cast<int>(fx) - 1, y0 = static_cast<int>(fy) - 1;
        float value = 0.0F;
        for (int j = 0; j < 4; ++j) {
          const T* row = source.row(std::clamp(y0 + j, 0, source.h - 1));
          float acc = 0.0F;
          for (int i = 0; i < 4; ++i) acc += wx[i] * static_cast<float>(row[std::clamp(x0 + i, 0, source.w - 1)]);
          value += wy[j] * acc;
        }
        v = to_pixel<T>(value);
      }
      out(x, y) = v;
      if (valid != nullptr) (*valid)(x, y) = inside ? 255 : 0;
    }
  }
  return out;
}

template ImgU8 rectify_image(const ImgU8&, const Camera&, const Camera&, ImgU8*);
template ImgU16 rectify_image(const ImgU16&, const Camera&, const Camera&, ImgU8*);
template ImgF rectify_image(const ImgF&, const Camera&, const Camera&, ImgU8*);

}  // namespace oracle
