// CPU oracle of the Census-SGM matcher -- TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// Restates, step by step, the reference's scalar semantics.  Citations are
// to rhiestan/Regard3D_MVS files; "matcher" = src/r3d/stereo/stereo_matcher.cpp,
// "kernels" = src/r3d/stereo/sgm_kernels.cpp.
#include "oracle.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <bit>
#include <climits>
#include <cmath>
#include <cstring>
#include <thread>

namespace oracle {

int align16(int n) { return (n + 15) / 16 * 16; }
// matcher:889-890
int floor_shift(int v, int k) { return v >= 0 ? v >> k : -((-v + (1 << k) - 1) >> k); }
int ceil_shift(int v, int k) { return -floor_shift(-v, k); }

static void fail(const std::string& m) { throw Error(m); }

// matcher:81-95
void check_options(const r3dg_options& o) {
  if (o.p1 < 0 || o.p2 < o.p1 || o.p2 > 8000) fail("stereo: invalid penalties (0 <= P1 <= P2 <= 8000)");
  if (o.invalid_cost < 0 || o.invalid_cost > 255) fail("stereo: invalid_cost outside [0, 255]");
  if (o.tile_size < 16 || o.tile_margin < 0 || o.tile_blend < 0) fail("stereo: invalid tiling");
  if (o.levels < -1 || o.levels > 12 || o.range_radius < 0 || o.range_margin < 0 || o.range_fill < 0)
    fail("stereo: invalid hierarchy options");
}

// ---------------------------------------------------------------- census
// census.cpp:20-35 (raster-order neighbours, centre skipped), :104-113
// (bit = neighbour < centre, first neighbour most significant), :137-155
// (columns replicated, rows clamped).
ImgU64 census(const ImgU16& img) {
  ImgU64 out(img.w, img.h);
  for (int y = 0; y < img.h; ++y) {
    for (int x = 0; x < img.w; ++x) {
      const std::uint16_t c = img(x, y);
      std::uint64_t bits = 0;
      for (int j = 0; j < 7; ++j) {
        const std::uint16_t* r = img.row(std::clamp(y - 3 + j, 0, img.h - 1));
        for (int i = 0; i < 9; ++i) {
          if (j == 3 && i == 4) continue;
          bits = (bits << 1) | (r[std::clamp(x - 4 + i, 0, img.w - 1)] < c ? 1U : 0U);
        }
      }
      out(x, y) = bits;
    }
  }
  return out;
}

// ---------------------------------------------------------------- pyramid
// pyramid.cpp:34-85: horizontal [1 4 6 4 1] of every row into uint32, then
// vertical, (sum + 128) >> 8; replicated borders; ceil(w/2) x ceil(h/2).
template <typename T>
static Img<T> pyramid_impl(const Img<T>& s) {
  if (s.empty()) fail("pyramid_down: empty image");
  const int w = s.w, h = s.h, dw = (w + 1) / 2, dh = (h + 1) / 2;
  Img<std::uint32_t> tmp(dw, h);
  for (int y = 0; y < h; ++y) {
    const T* in = s.row(y);
    for (int x = 0; x < dw; ++x) {
      const int sx = 2 * x;
      const std::uint32_t a = in[std::max(sx - 2, 0)], b = in[std::max(sx - 1, 0)], c = in[sx],
                          d = in[std::min(sx + 1, w - 1)], e = in[std::min(sx + 2, w - 1)];
      tmp(x, y) = a + e + 4U * (b + d) + 6U * c;
    }
  }
  Img<T> out(dw, dh);
  for (int y = 0; y < dh; ++y) {
    const int sy = 2 * y;
    const std::uint32_t* a = tmp.row(std::max(sy - 2, 0));
    const std::uint32_t* b = tmp.row(std::max(sy - 1, 0));
    const std::uint32_t* c = tmp.row(sy);
    const std::uint32_t* d = tmp.row(std::min(sy + 1, h - 1));
    const std::uint32_t* e = tmp.row(std::min(sy + 2, h - 1));
    for (int x = 0; x < dw; ++x) {
      out(x, y) = static_cast<T>((a[x] + e[x] + 4U * (b[x] + d[x]) + 6U * c[x] + 128U) >> 8U);
    }
  }
  return out;
}
ImgU16 pyramid_down(const ImgU16& img) { return pyramid_impl(img); }
ImgU8 pyramid_down(const ImgU8& img) { return pyramid_impl(img); }

// matcher:892-900
ImgU8 reduce_mask(const ImgU8& mask) {
  ImgU8 m = pyramid_impl(mask);
  for (auto& v : m.px) v = v >= 128 ? 255 : 0;
  return m;
}

// matcher:1105-1122
int automatic_levels(int width, int height, int num_disparities) {
  int levels = 0, w = width, h = height, n = num_disparities;
  while (levels < 8) {
    const int nw = (w + 1) / 2, nh = (h + 1) / 2;
    if (std::min(nw, nh) < 96 || n <= 48) break;
    w = nw;
    h = nh;
    n = (n + 1) / 2;
    ++levels;
  }
  return levels;
}

// ---------------------------------------------------------------- context
struct Interval {
  int first = 0;
  int last = -1;
};

// matcher:43-67: first / last column with mask >= 128 (interior holes ignored).
static std::vector<Interval> valid_intervals(const ImgU8& mask, int width, int height) {
  std::vector<Interval> rows(static_cast<std::size_t>(height), Interval{0, width - 1});
  if (mask.empty()) return rows;
  for (int y = 0; y < height; ++y) {
    const std::uint8_t* m = mask.row(y);
    Interval& r = rows[static_cast<std::size_t>(y)];
    r = Interval{};
    int x = 0;
    while (x < width && m[x] < 128) ++x;
    if (x == width) continue;
    r.first = x;
    x = width - 1;
    while (m[x] < 128) --x;
    r.last = x;
  }
  return rows;
}

// matcher:69-79
static std::array<std::uint16_t, 256> p2_table(const r3dg_options& o) {
  std::array<std::uint16_t, 256> lut{};
  for (int di = 0; di < 256; ++di) {
    int p2 = o.p2;
    if (o.p2_adapt > 0) p2 = std::max(o.p1 + 1, o.p2 * o.p2_adapt / (o.p2_adapt + di));
    lut[static_cast<std::size_t>(di)] = static_cast<std::uint16_t>(p2);
  }
  return lut;
}

// matcher:97-111
static void check_ranges(const Ranges& r, int width, int height, int D) {
  if (r.lo.w != width || r.lo.h != height || r.num.w != width || r.num.h != height)
    fail("stereo: label ranges of the wrong size");
  for (std::size_t i = 0; i < r.lo.px.size(); ++i) {
    const int lo = r.lo.px[i], n = r.num.px[i];
    if (n <= 0 || n % 16 != 0 || lo + n > D) fail("stereo: invalid label range");
  }
}

struct Ctx {
  const ImgU16* left = nullptr;
  const ImgU16* right = nullptr;
  ImgU64 cl, cr;  // whole-image Census (equal to the per-tile regions, matcher:226-250)
  std::vector<Interval> lrows, rrows;
  int dmin = 0;
  int D = 0;
  const Ranges* ranges = nullptr;
  std::array<std::uint16_t, 256> lut{};
  const r3dg_options* o = nullptr;
  bool diagonal_lr = false;
  bool confidence = true;
  bool want_labels = false;
};

// matcher:128-163
static Ctx make_context(const ImgU16& left, const ImgU16& right, const ImgU8& lv, const ImgU8& rv, int dmin, int nd,
                        const Ranges* ranges, const r3dg_options& o) {
  check_options(o);
  if (left.h != right.h || left.empty() || right.empty())
    fail("stereo: rectified images must be non-empty with equal height");
  if (!lv.empty() && (lv.w != left.w || lv.h != left.h)) fail("stereo: left mask size differs from the image");
  if (!rv.empty() && (rv.w != right.w || rv.h != right.h)) fail("stereo: right mask size differs from the image");
  if (nd <= 0) fail("stereo: no disparities to search");
  Ctx c;
  c.left = &left;
  c.right = &right;
  c.cl = census(left);
  c.cr = census(right);
  c.lrows = valid_intervals(lv, left.w, left.h);
  c.rrows = valid_intervals(rv, right.w, right.h);
  c.dmin = dmin;
  c.D = align16(nd);
  if (ranges != nullptr) check_ranges(*ranges, left.w, left.h, c.D);
  c.ranges = ranges;
  c.lut = p2_table(o);
  c.o = &o;
  return c;
}

struct Tile {
  int x0 = 0, y0 = 0, w = 0, h = 0;
};

// Packed label layout of a tile (matcher:172-224): pixel (x, y) owns the cells
// [off(x, y), off(x, y) + num) of the tile volume, rows one after the other.
struct TileLabels {
  int w = 0, h = 0;
  std::vector<std::int32_t> lo;
  std::vector<std::uint16_t> num;
  std::vector<std::uint64_t> off;  // [h][w + 1]
  int max_num = 0;
  std::size_t max_row_cells = 0;

  TileLabels(const Ctx& c, const Tile& t) : w(t.w), h(t.h), lo(std::size_t(w) * h), num(std::size_t(w) * h),
                                            off(std::size_t(w + 1) * h) {
    std::uint64_t acc = 0;
    for (int y = 0; y < h; ++y) {
      const std::uint64_t row0 = acc;
      for (int x = 0; x < w; ++x) {
        const std::size_t i = std::size_t(y) * w + x;
        if (c.ranges != nullptr) {
          lo[i] = c.ranges->lo(t.x0 + x, t.y0 + y);
          num[i] = c.ranges->num(t.x0 + x, t.y0 + y);
        } else {
          lo[i] = 0;
          num[i] = static_cast<std::uint16_t>(c.D);
        }
        off[std::size_t(y) * (w + 1) + x] = acc;
        acc += num[i];
        max_num = std::max<int>(max_num, num[i]);
      }
      off[std::size_t(y) * (w + 1) + w] = acc;
      max_row_cells = std::max<std::size_t>(max_row_cells, acc - row0);
    }
  }
  [[nodiscard]] std::uint64_t at(int x, int y) const { return off[std::size_t(y) * (w + 1) + x]; }
  [[nodiscard]] std::uint64_t cells() const { return h > 0 ? off[std::size_t(h - 1) * (w + 1) + w] : 0; }
};

// Matching costs of tile row y (matcher:305-353): popcount(CL ^ CR) for right
// columns inside the right row's valid interval, invalid_cost outside it, and
// 0 for every label of a left pixel outside the left row's valid interval.
static void cost_row(const Ctx& c, const Tile& t, const TileLabels& L, int y, std::uint16_t* cost) {
  const int Y = t.y0 + y;
  const Interval lv = c.lrows[std::size_t(Y)], rv = c.rrows[std::size_t(Y)];
  const std::uint64_t* cl = c.cl.row(Y);
  const std::uint64_t* cr = c.cr.row(Y);
  const auto invalid = static_cast<std::uint16_t>(c.o->invalid_cost);
  const std::uint64_t row0 = L.at(0, y);
  for (int x = 0; x < t.w; ++x) {
    const int X = t.x0 + x;
    const std::size_t i = std::size_t(y) * t.w + x;
    const int n = L.num[i];
    std::uint16_t* out = cost + (L.at(x, y) - row0);
    if (X < lv.first || X > lv.last) {
      std::fill(out, out + n, std::uint16_t{0});
      continue;
    }
    const std::uint64_t a = cl[X];
    const int xr0 = X - c.dmin - L.lo[i];
    for (int j = 0; j < n; ++j) {
      const int xr = xr0 - j;
      out[j] = xr >= rv.first && xr <= rv.last ? static_cast<std::uint16_t>(std::popcount(a ^ cr[xr])) : invalid;
    }
  }
}

// Predecessor path costs of one path aligned to the labels lo - 1 .. lo + n of
// the current pixel (kernels:37-53): labels outside q's range read 0xFFFF.
// `q` points at q's leading sentinel (label q_lo - 1).
static const std::uint16_t* align_pred(const std::uint16_t* q, int q_lo, int q_num, int lo, int n,
                                       std::uint16_t* buf) {
  if (lo >= q_lo && lo + n <= q_lo + q_num) return q + (lo - q_lo);
  for (int i = 0; i < n + 2; ++i) {
    const int k = lo - 1 + i - q_lo;
    buf[i] = k >= 0 && k < q_num ? q[1 + k] : std::uint16_t{0xFFFF};
  }
  return buf;
}

// SGM aggregation of one tile into its packed volume S (the two 4-path passes
// of matcher:286-289, 377-408 with the per-label recursion of kernels:387-403):
//   L(p, d) = C(p, d) + min(L(q, d), sat(min(L(q, d-1), L(q, d+1)) + P1),
//                          sat(min_q + P2)) - min_q        (mod 2^16)
// with L = C at the first pixel of a path; pass +1 stores S, pass -1 adds.
static void aggregate_tile(const Ctx& c, const Tile& t, const TileLabels& L, std::uint16_t* S) {
  const int w = t.w, h = t.h;
  const std::uint32_t P1 = static_cast<std::uint32_t>(c.o->p1);
  // Paths 1..3 of a row: pixel x stored as [sentinel, labels..., sentinel] at
  // (cell offset within the row) + 2x.
  std::size_t row_path = 0;
  for (int y = 0; y < h; ++y) row_path = std::max<std::size_t>(row_path, (L.at(w, y) - L.at(0, y)) + 2 * std::size_t(w));
  std::vector<std::uint16_t> prev(3 * row_path), cur(3 * row_path);
  std::vector<std::uint16_t> prev_min(3 * std::size_t(w)), cur_min(3 * std::size_t(w));
  std::vector<std::uint16_t> cost(L.max_row_cells);
  const std::size_t buf_n = std::size_t(L.max_num) + 2;
  std::vector<std::uint16_t> hbuf(2 * buf_n), abuf(4 * buf_n);
  const ImgU16& img = *c.left;

  for (const int dir : {1, -1}) {
    for (int i = 0; i < h; ++i) {
      const int y = dir > 0 ? i : h - 1 - i;
      cost_row(c, t, L, y, cost.data());
      const bool has_prev = i > 0;
      const int yq = y - dir;
      const std::uint64_t row0 = L.at(0, y);
      const std::uint64_t qrow0 = has_prev ? L.at(0, yq) : 0;
      const std::uint16_t* I = img.row(t.y0 + y) + t.x0;
      const std::uint16_t* Iq = has_prev ? img.row(t.y0 + yq) + t.x0 : nullptr;
      std::uint16_t hmin = 0;
      for (int ix = 0; ix < w; ++ix) {
        const int x = dir > 0 ? ix : w - 1 - ix;
        const std::size_t pi = std::size_t(y) * w + x;
        const int lo = L.lo[pi], n = L.num[pi];
        const std::uint16_t* cst = cost.data() + (L.at(x, y) - row0);
        std::uint16_t* s = S + L.at(x, y);

        const std::uint16_t* g[4] = {nullptr, nullptr, nullptr, nullptr};  // null: first pixel of the path
        std::uint16_t minq[4] = {0, 0, 0, 0}, pen[4] = {0, 0, 0, 0};
        auto penalty = [&](std::uint16_t mq, std::uint16_t a, std::uint16_t b) {
          const std::uint32_t p = std::uint32_t(mq) + c.lut[std::size_t(std::abs(int(a) - int(b)) >> 8)];
          return static_cast<std::uint16_t>(std::min<std::uint32_t>(p, 0xFFFFU));
        };
        std::uint16_t* hcur = hbuf.data() + std::size_t(ix & 1) * buf_n;
        if (ix > 0) {
          const int xq = x - dir;
          const std::size_t qi = std::size_t(y) * w + xq;
          const std::uint16_t* hq = hbuf.data() + std::size_t((ix + 1) & 1) * buf_n;
          g[0] = align_pred(hq, L.lo[qi], L.num[qi], lo, n, abuf.data());
          minq[0] = hmin;
          pen[0] = penalty(hmin, I[x], I[xq]);
        }
        if (has_prev) {
          const int xqs[3] = {x, x - dir, x + dir};
          for (int k = 0; k < 3; ++k) {
            const int xq = xqs[k];
            if (xq < 0 || xq >= w) continue;
            const std::size_t qi = std::size_t(yq) * w + xq;
            const std::uint16_t* q = prev.data() + std::size_t(k) * row_path + (L.at(xq, yq) - qrow0) + 2 * std::size_t(xq);
            g[k + 1] = align_pred(q, L.lo[qi], L.num[qi], lo, n, abuf.data() + std::size_t(k + 1) * buf_n);
            minq[k + 1] = prev_min[std::size_t(k) * w + xq];
            pen[k + 1] = penalty(minq[k + 1], I[x], Iq[xq]);
          }
        }
        std::uint16_t* out[4];
        out[0] = hcur + 1;
        const std::size_t pos = (L.at(x, y) - row0) + 2 * std::size_t(x);
        for (int k = 1; k < 4; ++k) {
          std::uint16_t* base = cur.data() + std::size_t(k - 1) * row_path + pos;
          base[0] = 0xFFFF;
          base[n + 1] = 0xFFFF;
          out[k] = base + 1;
        }
        hcur[0] = 0xFFFF;
        hcur[n + 1] = 0xFFFF;
        std::uint16_t mins[4] = {0xFFFF, 0xFFFF, 0xFFFF, 0xFFFF};
        for (int k = 0; k < 4; ++k) {
          std::uint16_t* o = out[k];
          std::uint16_t m = 0xFFFF;
          if (g[k] == nullptr) {
            for (int j = 0; j < n; ++j) {
              o[j] = cst[j];
              m = std::min(m, cst[j]);
            }
          } else {
            const std::uint16_t* gq = g[k];
            const std::uint16_t mq = minq[k], pk = pen[k];
            for (int j = 0; j < n; ++j) {
              const std::uint32_t nb = std::min<std::uint32_t>(std::min(gq[j], gq[j + 2]) + P1, 0xFFFFU);
              const std::uint32_t best = std::min<std::uint32_t>(std::min<std::uint32_t>(gq[j + 1], nb), pk);
              const auto l = static_cast<std::uint16_t>(best - mq + cst[j]);
              o[j] = l;
              m = std::min(m, l);
            }
          }
          mins[k] = m;
        }
        if (dir > 0) {
          for (int j = 0; j < n; ++j)
            s[j] = static_cast<std::uint16_t>(out[0][j] + out[1][j] + out[2][j] + out[3][j]);
        } else {
          for (int j = 0; j < n; ++j)
            s[j] = static_cast<std::uint16_t>(s[j] + out[0][j] + out[1][j] + out[2][j] + out[3][j]);
        }
        hmin = mins[0];
        for (int k = 1; k < 4; ++k) cur_min[std::size_t(k - 1) * w + x] = mins[k];
      }
      prev.swap(cur);
      prev_min.swap(cur_min);
    }
  }
}

std::vector<std::uint16_t> cost_volume(const ImgU16& left, const ImgU16& right, int dmin, int nd,
                                       const r3dg_options& o, const Ranges* ranges) {
  const ImgU8 none;
  const Ctx c = make_context(left, right, none, none, dmin, nd, ranges, o);
  const Tile t{0, 0, left.w, left.h};
  const TileLabels L(c, t);
  std::vector<std::uint16_t> S(L.cells());
  aggregate_tile(c, t, L, S.data());
  return S;
}

// matcher:1280-1362: the textbook volume, restated with its own data layout:
// integer costs and path costs, kInf outside ranges, 8 independent sweeps.
std::vector<std::uint16_t> textbook_volume(const ImgU16& left, const ImgU16& right, int dmin, int nd,
                                           const r3dg_options& o, const Ranges* ranges) {
  check_options(o);
  const int w = left.w, h = left.h, wr = right.w, D = align16(nd);
  if (ranges != nullptr) check_ranges(*ranges, w, h, D);
  auto first = [&](int x, int y) { return ranges != nullptr ? int(ranges->lo(x, y)) : 0; };
  auto count = [&](int x, int y) { return ranges != nullptr ? int(ranges->num(x, y)) : D; };
  const ImgU64 cl = census(left), cr = census(right);
  const auto lut = p2_table(o);
  auto at = [&](int x, int y, int d) { return (std::size_t(y) * w + x) * std::size_t(D) + std::size_t(d); };
  std::vector<int> cost(std::size_t(w) * h * D);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x)
      for (int d = 0; d < D; ++d) {
        const int xr = x - dmin - d;
        cost[at(x, y, d)] = xr >= 0 && xr < wr ? std::popcount(cl(x, y) ^ cr(xr, y)) : o.invalid_cost;
      }
  constexpr int kInf = 1 << 28;
  std::vector<int> total(cost.size(), 0), path(cost.size(), kInf);
  const int dirs[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, -1}, {-1, 1}, {1, -1}};
  for (const auto& r : dirs) {
    const int dx = r[0], dy = r[1];
    for (int iy = 0; iy < h; ++iy) {
      const int y = dy >= 0 ? iy : h - 1 - iy;
      for (int ix = 0; ix < w; ++ix) {
        const int x = dx >= 0 ? ix : w - 1 - ix;
        const int qx = x - dx, qy = y - dy;
        const int d0 = first(x, y), d1 = d0 + count(x, y);
        if (qx < 0 || qy < 0 || qx >= w || qy >= h) {
          for (int d = d0; d < d1; ++d) path[at(x, y, d)] = cost[at(x, y, d)];
          continue;
        }
        const int q0 = first(qx, qy), q1 = q0 + count(qx, qy);
        int min_q = kInf;
        for (int d = q0; d < q1; ++d) min_q = std::min(min_q, path[at(qx, qy, d)]);
        auto lq = [&](int d) { return d >= q0 && d < q1 ? path[at(qx, qy, d)] : kInf; };
        const int p2 = lut[std::size_t(std::abs(int(left(x, y)) - int(left(qx, qy))) >> 8)];
        for (int d = d0; d < d1; ++d) {
          const int best = std::min({lq(d), lq(d - 1) + o.p1, lq(d + 1) + o.p1, min_q + p2});
          path[at(x, y, d)] = cost[at(x, y, d)] + best - min_q;
        }
      }
    }
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x)
        for (int d = first(x, y); d < first(x, y) + count(x, y); ++d) total[at(x, y, d)] += path[at(x, y, d)];
  }
  std::vector<std::uint16_t> out;
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x)
      for (int d = first(x, y); d < first(x, y) + count(x, y); ++d)
        out.push_back(static_cast<std::uint16_t>(total[at(x, y, d)]));
  return out;
}

// ---------------------------------------------------------------- selection
// kernels:405-438 (scalar WTA): first argmin, second best outside best +- 1,
// optional right-image WTA along x - label.
void wta_row(const std::uint16_t* sum, int width, const std::int32_t* lo, const std::uint16_t* num,
             std::uint16_t* best, std::uint16_t* best_cost, std::uint16_t* second, std::uint16_t* right_cost,
             std::uint16_t* right_best, std::ptrdiff_t right_shift) {
  std::size_t off = 0;
  for (int x = 0; x < width; ++x) {
    const std::uint16_t* s = sum + off;
    const int n = num[x];
    int kb = 0;
    for (int k = 1; k < n; ++k)
      if (s[k] < s[kb]) kb = k;
    std::uint16_t sec = 0xFFFF;
    for (int k = 0; k < n; ++k)
      if (k + 1 < kb || k > kb + 1) sec = std::min(sec, s[k]);
    best[x] = static_cast<std::uint16_t>(lo[x] + kb);
    best_cost[x] = s[kb];
    second[x] = sec;
    if (right_cost != nullptr) {
      for (int k = 0; k < n; ++k) {
        const std::ptrdiff_t i = std::ptrdiff_t(x) - lo[x] - k + right_shift;
        if (s[k] < right_cost[i]) {
          right_cost[i] = s[k];
          right_best[i] = static_cast<std::uint16_t>(lo[x] + k);
        }
      }
    }
    off += std::size_t(n);
  }
}

// matcher:426-436
static float subpixel_offset(int method, float sm, float s0, float sp) {
  float delta = 0.0F;
  if (method == R3DG_SUBPIXEL_PARABOLA) {
    const float denom = sm - 2.0F * s0 + sp;
    delta = denom > 0.0F ? (sm - sp) / (2.0F * denom) : 0.0F;
  } else if (method == R3DG_SUBPIXEL_EQUIANGULAR) {
    const float slope = std::max(sm - s0, sp - s0);
    delta = slope > 0.0F ? (sm - sp) / (2.0F * slope) : 0.0F;
  }
  return std::clamp(delta, -0.5F, 0.5F);
}

struct TileResult {
  Tile t;
  std::vector<float> d, c;
  std::vector<std::int16_t> lab;
  std::uint64_t cells = 0;
};

// matcher:446-525
static void select_disparities(const Ctx& ctx, const TileLabels& L, const std::uint16_t* S, TileResult& out) {
  const Tile& t = out.t;
  const r3dg_options& o = *ctx.o;
  const int w = t.w;
  out.d.assign(std::size_t(w) * t.h, kInvalid);
  if (ctx.confidence) out.c.assign(std::size_t(w) * t.h, 0.0F);
  if (ctx.want_labels) out.lab.assign(std::size_t(w) * t.h, INT16_MIN);
  const auto W = static_cast<std::size_t>(w);
  std::vector<std::uint16_t> rcost, rbest, bl(W), bv(W), sv(W);
  for (int y = 0; y < t.h; ++y) {
    const std::int32_t* lo = L.lo.data() + std::size_t(y) * w;
    const std::uint16_t* num = L.num.data() + std::size_t(y) * w;
    const int Y = t.y0 + y;
    std::ptrdiff_t shift = 0;
    if (ctx.diagonal_lr) {
      std::ptrdiff_t top = 0;
      shift = PTRDIFF_MIN;
      for (int x = 0; x < w; ++x) {
        shift = std::max<std::ptrdiff_t>(shift, lo[x] + num[x] - 1 - x);
        top = std::max<std::ptrdiff_t>(top, x - lo[x]);
      }
      rcost.assign(std::size_t(top + shift + 1), 0xFFFF);
      rbest.assign(rcost.size(), 0xFFFF);
    }
    const Interval lv = ctx.lrows[std::size_t(Y)], rv = ctx.rrows[std::size_t(Y)];
    const int xb = std::clamp(lv.first - t.x0, 0, w);
    const int xe = std::clamp(lv.last + 1 - t.x0, xb, w);
    if (xb < xe) {
      wta_row(S + L.at(xb, y), xe - xb, lo + xb, num + xb, bl.data() + xb, bv.data() + xb, sv.data() + xb,
              ctx.diagonal_lr ? rcost.data() + xb : nullptr, ctx.diagonal_lr ? rbest.data() + xb : nullptr, shift);
    }
    for (int x = xb; x < xe; ++x) {
      const int X = t.x0 + x;
      const std::uint16_t* s = S + L.at(x, y);
      const int best = bl[std::size_t(x)];
      const int j = best - lo[x];
      const int n = num[x];
      const unsigned second = sv[std::size_t(x)];
      const int xr = X - ctx.dmin - best;
      if (xr < rv.first || xr > rv.last) continue;
      if ((j == 0 && lo[x] > 0) || (j == n - 1 && lo[x] + n < ctx.D)) continue;
      if (ctx.diagonal_lr && std::abs(int(rbest[std::size_t(std::ptrdiff_t(x) - best + shift)]) - best) > o.lr_max_diff)
        continue;
      const float best_cost = bv[std::size_t(x)];
      if (o.uniqueness > 0.0F && second != 0xFFFFU && best_cost > (1.0F - o.uniqueness) * static_cast<float>(second))
        continue;
      float delta = 0.0F;
      if (j > 0 && j < n - 1) delta = subpixel_offset(o.subpixel, s[j - 1], best_cost, s[j + 1]);
      const std::size_t i = std::size_t(y) * w + x;
      out.d[i] = static_cast<float>(ctx.dmin + best) + delta;
      if (ctx.confidence)
        out.c[i] = second != 0xFFFFU && second > 0 ? 1.0F - best_cost / static_cast<float>(second) : 0.0F;
      if (ctx.want_labels) out.lab[i] = static_cast<std::int16_t>(ctx.dmin + best);
    }
  }
}

// matcher:530-543
static float axis_weight(int v, int c0, int c1, int size, int blend_px) {
  const auto blend = static_cast<float>(blend_px);
  float w = 1.0F;
  const float pos = static_cast<float>(v) + 0.5F;
  if (c0 > 0)
    w = std::min(w, blend_px > 0 ? std::clamp((pos - static_cast<float>(c0) + 0.5F * blend) / blend, 0.0F, 1.0F)
                                 : (v >= c0 ? 1.0F : 0.0F));
  if (c1 < size)
    w = std::min(w, blend_px > 0 ? std::clamp((static_cast<float>(c1) + 0.5F * blend - pos) / blend, 0.0F, 1.0F)
                                 : (v < c1 ? 1.0F : 0.0F));
  return w;
}

// matcher:545-552
static std::vector<int> split(int size, int tile) {
  const int n = std::max(1, (size + tile - 1) / tile);
  std::vector<int> b(std::size_t(n) + 1);
  for (int i = 0; i <= n; ++i) b[std::size_t(i)] = static_cast<int>(static_cast<long long>(size) * i / n);
  return b;
}

// matcher:555-577
static std::uint64_t count_cells(const Ctx& c, int x0, int y0, int x1, int y1) {
  if (c.ranges == nullptr) return std::uint64_t(x1 - x0) * std::uint64_t(y1 - y0) * std::uint64_t(c.D);
  std::uint64_t cells = 0;
  for (int y = y0; y < y1; ++y)
    for (int x = x0; x < x1; ++x) cells += c.ranges->num(x, y);
  return cells;
}
static std::size_t estimated_tile_bytes(const Ctx& c, int x0, int y0, int x1, int y1, std::size_t right_width) {
  const auto cells = static_cast<std::size_t>(count_cells(c, x0, y0, x1, y1));
  const auto w = static_cast<std::size_t>(x1 - x0), h = static_cast<std::size_t>(y1 - y0);
  const std::size_t row_cells = cells / std::max<std::size_t>(h, 1);
  return cells * 2 + row_cells * 2 + 6 * (row_cells + w * 2 * 8) * 2 + w * h * 32 + (w + right_width + 16) * (h + 6) * 8;
}

struct TilesOut {
  ImgF disparity, confidence;
  Img<std::int16_t> labels;
  std::uint64_t cells = 0, tile_pixels = 0;
  int tiles = 0;
};

template <typename F>
static void parallel_tiles(std::size_t n, int threads, F&& f) {
  const int nt = std::max(1, std::min<int>(threads, static_cast<int>(n)));
  if (nt <= 1) {
    for (std::size_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<std::size_t> next{0};
  std::exception_ptr err;
  std::atomic<bool> failed{false};
  std::vector<std::thread> pool;
  for (int k = 0; k < nt; ++k)
    pool.emplace_back([&] {
      for (;;) {
        const std::size_t i = next.fetch_add(1);
        if (i >= n || failed.load()) return;
        try {
          f(i);
        } catch (...) {
          if (!failed.exchange(true)) err = std::current_exception();
        }
      }
    });
  for (auto& t : pool) t.join();
  if (err) std::rethrow_exception(err);
}

// matcher:613-771: tile layout, per-tile aggregation + selection, blending.
static TilesOut match_tiles(const Ctx& ctx, int threads) {
  const r3dg_options& o = *ctx.o;
  const int width = ctx.left->w, height = ctx.left->h;
  const int ext = o.tile_margin + (o.tile_blend + 1) / 2;
  struct Laid {
    Tile t;
    int core[4];
  };
  auto layout = [&](int ts) {
    std::vector<Laid> tiles;
    const auto bx = split(width, ts), by = split(height, ts);
    for (std::size_t j = 0; j + 1 < by.size(); ++j)
      for (std::size_t i = 0; i + 1 < bx.size(); ++i) {
        const int x0 = std::max(0, bx[i] - ext), x1 = std::min(width, bx[i + 1] + ext);
        const int y0 = std::max(0, by[j] - ext), y1 = std::min(height, by[j + 1] + ext);
        tiles.push_back({Tile{x0, y0, x1 - x0, y1 - y0}, {bx[i], bx[i + 1], by[j], by[j + 1]}});
      }
    return tiles;
  };
  auto max_bytes = [&](const std::vector<Laid>& tiles) {
    std::size_t m = 0;
    for (const auto& l : tiles)
      m = std::max(m, estimated_tile_bytes(ctx, l.t.x0, l.t.y0, l.t.x0 + l.t.w, l.t.y0 + l.t.h,
                                           static_cast<std::size_t>(ctx.right->w)));
    return m;
  };
  int ts = o.tile_size;
  auto laid = layout(ts);
  while (ts > 64 && max_bytes(laid) > o.memory_budget) {
    ts /= 2;
    laid = layout(ts);
  }
  std::vector<TileResult> res(laid.size());
  parallel_tiles(laid.size(), threads, [&](std::size_t i) {
    res[i].t = laid[i].t;
    const TileLabels L(ctx, laid[i].t);
    std::vector<std::uint16_t> S(L.cells());
    aggregate_tile(ctx, laid[i].t, L, S.data());
    select_disparities(ctx, L, S.data(), res[i]);
    res[i].cells = L.cells();
  });

  TilesOut out;
  out.tiles = static_cast<int>(laid.size());
  std::vector<BlendTile> bt(laid.size());
  for (std::size_t i = 0; i < laid.size(); ++i) {
    out.cells += res[i].cells;
    out.tile_pixels += std::uint64_t(laid[i].t.w) * std::uint64_t(laid[i].t.h);
    bt[i] = BlendTile{laid[i].t.x0, laid[i].t.y0, laid[i].t.w, laid[i].t.h,
                      {laid[i].core[0], laid[i].core[1], laid[i].core[2], laid[i].core[3]}, res[i].d.data(),
                      ctx.confidence ? res[i].c.data() : nullptr, ctx.want_labels ? res[i].lab.data() : nullptr};
  }
  blend_tiles(width, height, o.tile_blend, bt, out.disparity, ctx.confidence ? &out.confidence : nullptr,
              ctx.want_labels ? &out.labels : nullptr);
  return out;
}

// matcher:688-766: per row, the tiles in tile order; weighted mean where the
// valid weight is at least half of the total, the dominant tile's value where
// the mean is more than 1 away from it.
void blend_tiles(int width, int height, int blend, const std::vector<BlendTile>& tiles, ImgF& disparity,
                 ImgF* confidence, Img<std::int16_t>* labels) {
  std::vector<std::vector<float>> colw(tiles.size());
  for (std::size_t i = 0; i < tiles.size(); ++i) {
    colw[i].resize(std::size_t(tiles[i].w));
    for (int x = 0; x < tiles[i].w; ++x)
      colw[i][std::size_t(x)] = axis_weight(tiles[i].x0 + x, tiles[i].core[0], tiles[i].core[1], width, blend);
  }
  disparity = ImgF(width, height, kInvalid);
  if (confidence != nullptr) *confidence = ImgF(width, height, 0.0F);
  if (labels != nullptr) *labels = Img<std::int16_t>(width, height, INT16_MIN);
  const auto W = std::size_t(width);
  std::vector<float> wsum(W), wall(W), dsum(W), csum(W), wbest(W), dbest(W);
  std::vector<std::int16_t> lbest(W);
  for (int Y = 0; Y < height; ++Y) {
    std::fill(wsum.begin(), wsum.end(), 0.0F);
    std::fill(wall.begin(), wall.end(), 0.0F);
    std::fill(dsum.begin(), dsum.end(), 0.0F);
    std::fill(csum.begin(), csum.end(), 0.0F);
    std::fill(wbest.begin(), wbest.end(), 0.0F);
    std::fill(dbest.begin(), dbest.end(), kInvalid);
    std::fill(lbest.begin(), lbest.end(), std::int16_t{INT16_MIN});
    for (std::size_t i = 0; i < tiles.size(); ++i) {
      const BlendTile& t = tiles[i];
      if (Y < t.y0 || Y >= t.y0 + t.h) continue;
      const float wy = axis_weight(Y, t.core[2], t.core[3], height, blend);
      if (wy <= 0.0F) continue;
      const std::size_t row = std::size_t(Y - t.y0) * std::size_t(t.w);
      for (int x = 0; x < t.w; ++x) {
        const float wgt = wy * colw[i][std::size_t(x)];
        if (wgt <= 0.0F) continue;
        const std::size_t ti = row + std::size_t(x);
        const auto pi = std::size_t(t.x0 + x);
        wall[pi] += wgt;
        const float dv = t.d[ti];
        if (std::isfinite(dv)) {
          wsum[pi] += wgt;
          dsum[pi] += wgt * dv;
          if (confidence != nullptr) csum[pi] += wgt * t.c[ti];
          if (wgt > wbest[pi]) {
            wbest[pi] = wgt;
            dbest[pi] = dv;
            if (labels != nullptr) lbest[pi] = t.lab[ti];
          }
        }
      }
    }
    for (std::size_t x = 0; x < W; ++x) {
      float dv = kInvalid;
      if (wsum[x] > 0.0F && wsum[x] >= 0.5F * wall[x]) {
        dv = dsum[x] / wsum[x];
        if (std::abs(dv - dbest[x]) > 1.0F) dv = dbest[x];
        if (confidence != nullptr) (*confidence)(int(x), Y) = csum[x] / wsum[x];
        if (labels != nullptr) (*labels)(int(x), Y) = lbest[x];
      }
      disparity(int(x), Y) = dv;
    }
  }
}

// matcher:777-805
static bool left_right_differ(const ImgF& map, const ImgF& other, int x, int y, int sign, int max_diff) {
  const float d = map(x, y);
  const auto xo = static_cast<int>(std::lround(static_cast<float>(x) + static_cast<float>(sign) * d));
  const float dother = xo >= 0 && xo < other.w ? other(xo, y) : kInvalid;
  return !std::isfinite(dother) || std::abs(d - dother) > static_cast<float>(max_diff);
}

// matcher:797-805 over a whole map: 1 where a finite value fails the check.
ImgU8 left_right_mismatch(const ImgF& map, const ImgF& other, int sign, int max_diff) {
  ImgU8 out(map.w, map.h, 0);
  for (int y = 0; y < map.h; ++y)
    for (int x = 0; x < map.w; ++x)
      if (std::isfinite(map(x, y)) && left_right_differ(map, other, x, y, sign, max_diff)) out(x, y) = 1;
  return out;
}

// matcher:919-970: BFS over 4-neighbours with |d_j - d_i| <= max_diff.
void remove_speckles(ImgF& disparity, int min_size, float max_diff) {
  if (min_size <= 1 || disparity.empty()) return;
  const int w = disparity.w;
  const std::size_t n = disparity.px.size();
  const std::vector<float> d = disparity.px;
  std::vector<std::uint8_t> visited(n, 0);
  std::vector<std::uint32_t> seg;
  for (std::size_t start = 0; start < n; ++start) {
    if (visited[start] != 0 || !std::isfinite(d[start])) continue;
    seg.clear();
    seg.push_back(static_cast<std::uint32_t>(start));
    visited[start] = 1;
    for (std::size_t head = 0; head < seg.size(); ++head) {
      const std::size_t i = seg[head];
      const float v = d[i];
      const auto x = static_cast<int>(i % std::size_t(w));
      auto visit = [&](std::size_t j) {
        if (visited[j] == 0 && std::abs(d[j] - v) <= max_diff) {
          visited[j] = 1;
          seg.push_back(static_cast<std::uint32_t>(j));
        }
      };
      if (x > 0) visit(i - 1);
      if (x + 1 < w) visit(i + 1);
      if (i >= std::size_t(w)) visit(i - std::size_t(w));
      if (i + std::size_t(w) < n) visit(i + std::size_t(w));
    }
    if (static_cast<int>(seg.size()) < min_size)
      for (const std::uint32_t i : seg) disparity.px[i] = kInvalid;
  }
}

// matcher:903-915 (Paeth's median-of-9 network) and :972-1021.
static float median9(std::array<float, 9>& p) {
  auto s2 = [&](int i, int j) {
    const float a = p[std::size_t(i)], b = p[std::size_t(j)];
    p[std::size_t(i)] = std::min(a, b);
    p[std::size_t(j)] = std::max(a, b);
  };
  s2(1, 2); s2(4, 5); s2(7, 8); s2(0, 1); s2(3, 4); s2(6, 7);
  s2(1, 2); s2(4, 5); s2(7, 8); s2(0, 3); s2(5, 8); s2(4, 7);
  s2(3, 6); s2(1, 4); s2(2, 5); s2(4, 7); s2(4, 2); s2(6, 4);
  s2(4, 2);
  return p[4];
}

ImgF median3x3(const ImgF& d) {
  const int w = d.w, h = d.h;
  ImgF out(w, h);
  for (int y = 0; y < h; ++y) {
    for (int x = 0; x < w; ++x) {
      const float c = d(x, y);
      if (!std::isfinite(c)) {
        out(x, y) = c;
        continue;
      }
      if (x > 0 && y > 0 && x + 1 < w && y + 1 < h) {
        std::array<float, 9> a{};
        bool all = true;
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            const float v = d(x + dx, y + dy);
            a[std::size_t((dy + 1) * 3 + dx + 1)] = v;
            all = all && std::isfinite(v);
          }
        if (all) {
          out(x, y) = median9(a);
          continue;
        }
      }
      std::array<float, 9> v{};
      std::size_t n = 0;
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int qx = x + dx, qy = y + dy;
          if (qx >= 0 && qy >= 0 && qx < w && qy < h && std::isfinite(d(qx, qy))) v[n++] = d(qx, qy);
        }
      const auto mid = v.begin() + std::ptrdiff_t(n / 2);
      std::nth_element(v.begin(), mid, v.begin() + std::ptrdiff_t(n));
      float m = *mid;
      if (n % 2 == 0) m = 0.5F * (m + *std::max_element(v.begin(), mid));
      out(x, y) = m;
    }
  }
  return out;
}

// matcher:1124-1267
Ranges ranges_from_coarse(const ImgF& coarse, int width, int height, const ImgU8& fine_valid,
                          const ImgU8& partner_valid, int partner_width, int partner_sign, int dmin, int nd,
                          const r3dg_options& o, bool mirrored) {
  const int wc = coarse.w, hc = coarse.h;
  if (wc != (width + 1) / 2 || hc != (height + 1) / 2) fail("stereo: coarse map size does not match the level");
  constexpr float inf = kInvalid;
  const int r = o.range_radius;
  ImgF hmin(wc, hc), hmax(wc, hc);
  for (int y = 0; y < hc; ++y)
    for (int x = 0; x < wc; ++x) {
      float mn = inf, mx = -inf;
      for (int k = std::max(0, x - r); k <= std::min(wc - 1, x + r); ++k) {
        const float v = coarse(k, y);
        if (std::isfinite(v)) {
          mn = std::min(mn, v);
          mx = std::max(mx, v);
        }
      }
      hmin(x, y) = mn;
      hmax(x, y) = mx;
    }
  ImgF cmin(wc, hc), cmax(wc, hc);
  for (int y = 0; y < hc; ++y) {
    const int k0 = std::max(0, y - r), k1 = std::min(hc - 1, y + r);
    for (int x = 0; x < wc; ++x) {
      float mn = inf, mx = -inf;
      for (int k = k0; k <= k1; ++k) {
        mn = std::min(mn, hmin(x, k));
        mx = std::max(mx, hmax(x, k));
      }
      cmin(x, y) = mn;
      cmax(x, y) = mx;
    }
  }
  std::vector<ImgF> pmin{cmin}, pmax{cmax};
  while ((2 << pmin.size()) <= 2 * o.range_fill && pmin.back().w > 1) {
    const ImgF& a = pmin.back();
    const ImgF& b = pmax.back();
    const int nw = (a.w + 1) / 2, nh = (a.h + 1) / 2;
    ImgF mn(nw, nh), mx(nw, nh);
    for (int y = 0; y < nh; ++y)
      for (int x = 0; x < nw; ++x) {
        const int x1 = std::min(2 * x + 1, a.w - 1), y1 = std::min(2 * y + 1, a.h - 1);
        mn(x, y) = std::min(std::min(a(2 * x, 2 * y), a(x1, 2 * y)), std::min(a(2 * x, y1), a(x1, y1)));
        mx(x, y) = std::max(std::max(b(2 * x, 2 * y), b(x1, 2 * y)), std::max(b(2 * x, y1), b(x1, y1)));
      }
    pmin.push_back(std::move(mn));
    pmax.push_back(std::move(mx));
  }
  for (int y = 0; y < hc; ++y)
    for (int x = 0; x < wc; ++x) {
      if (pmin[0](x, y) != inf) continue;
      for (std::size_t k = 1; k < pmin.size(); ++k) {
        const ImgF& a = pmin[k];
        const ImgF& b = pmax[k];
        const int bx = x >> k, by = y >> k;
        float mn = inf, mx = -inf;
        for (int qy = std::max(0, by - 1); qy <= std::min(a.h - 1, by + 1); ++qy)
          for (int qx = std::max(0, bx - 1); qx <= std::min(a.w - 1, bx + 1); ++qx) {
            mn = std::min(mn, a(qx, qy));
            mx = std::max(mx, b(qx, qy));
          }
        if (mn != inf) {
          cmin(x, y) = mn;
          cmax(x, y) = mx;
          break;
        }
      }
    }
  const int n_all = align16(nd);
  Ranges out{ImgU16(width, height), ImgU16(width, height)};
  const auto prow = valid_intervals(partner_valid, partner_width, height);
  for (int y = 0; y < height; ++y) {
    const Interval pr = prow[std::size_t(y)];
    const int cy0 = y / 2, cy1 = std::min(cy0 + 1, hc - 1);
    for (int x = 0; x < width; ++x) {
      const int cx0 = x / 2, cx1 = std::min(cx0 + 1, wc - 1);
      const float mn = std::min(std::min(cmin(cx0, cy0), cmin(cx1, cy0)), std::min(cmin(cx0, cy1), cmin(cx1, cy1)));
      const float mx = std::max(std::max(cmax(cx0, cy0), cmax(cx1, cy0)), std::max(cmax(cx0, cy1), cmax(cx1, cy1)));
      int lo = 0, hi = n_all - 1;
      if (mn != inf) {
        lo = static_cast<int>(std::floor(2.0F * mn)) - dmin - o.range_margin;
        hi = static_cast<int>(std::ceil(2.0F * mx)) - dmin + o.range_margin;
      }
      const int match_lo = (partner_sign < 0 ? x - pr.last : pr.first - x) - dmin;
      const int match_hi = (partner_sign < 0 ? x - pr.first : pr.last - x) - dmin;
      lo = std::max(lo, match_lo);
      hi = std::min(hi, match_hi);
      if (lo > hi || pr.last < pr.first || (!fine_valid.empty() && fine_valid(x, y) < 128)) hi = lo;
      lo = std::clamp(lo, 0, n_all - 1);
      hi = std::clamp(hi, lo, n_all - 1);
      const int n = std::min(n_all, (hi - lo + 16) / 16 * 16);
      lo = std::clamp(lo - (n - (hi - lo + 1)) / 2, 0, n_all - n);
      const int xo = mirrored ? width - 1 - x : x;
      out.lo(xo, y) = static_cast<std::uint16_t>(lo);
      out.num(xo, y) = static_cast<std::uint16_t>(n);
    }
  }
  return out;
}

template <typename T>
static Img<T> mirror(const Img<T>& img) {
  Img<T> out(img.w, img.h);
  for (int y = 0; y < img.h; ++y) std::reverse_copy(img.row(y), img.row(y) + img.w, out.row(y));
  return out;
}

static void post_filter(ImgF& d, const r3dg_options& o, int speckle) {
  remove_speckles(d, speckle, o.speckle_diff);
  if (o.median != 0) d = median3x3(d);
}

struct LevelOut {
  TilesOut left;
  ImgF right;
  ImgU8 lr_mask;
};

// matcher:831-886
static LevelOut match_level(const ImgU16& L, const ImgU16& R, const ImgU8& mL, const ImgU8& mR, int dmin, int nd,
                            const Ranges* lranges, const Ranges* rranges, bool need_right, int speckle,
                            const r3dg_options& o, bool want_debug, int threads) {
  Ctx ctx = make_context(L, R, mL, mR, dmin, nd, lranges, o);
  const bool check_lr = o.lr_max_diff >= 0;
  ctx.diagonal_lr = check_lr && o.lr_full_match == 0;
  ctx.want_labels = want_debug;
  LevelOut out;
  out.left = match_tiles(ctx, threads);
  if (want_debug) out.lr_mask = ImgU8(L.w, L.h, 0);
  if (check_lr && o.lr_full_match != 0) {
    const int wl = L.w, wr = R.w;
    const ImgU16 mLi = mirror(L), mRi = mirror(R);
    const ImgU8 mLv = mL.empty() ? ImgU8() : mirror(mL);
    const ImgU8 mRv = mR.empty() ? ImgU8() : mirror(mR);
    Ctx rctx = make_context(mRi, mLi, mRv, mLv, dmin + wr - wl, nd, rranges, o);
    rctx.diagonal_lr = false;
    rctx.confidence = false;
    const TilesOut rmap = match_tiles(rctx, threads);
    out.left.cells += rmap.cells;
    out.left.tile_pixels += rmap.tile_pixels;
    out.right = ImgF(wr, R.h);
    for (int y = 0; y < R.h; ++y)
      for (int x = 0; x < wr; ++x) out.right(x, y) = rmap.disparity(wr - 1 - x, y) - static_cast<float>(wr - wl);
    ImgF& lm = out.left.disparity;
    std::vector<std::uint8_t> rbad;
    if (need_right) {
      rbad.assign(out.right.px.size(), 0);
      for (int y = 0; y < R.h; ++y)
        for (int x = 0; x < wr; ++x)
          if (std::isfinite(out.right(x, y)) && left_right_differ(out.right, lm, x, y, 1, o.lr_max_diff))
            rbad[std::size_t(y) * wr + x] = 1;
    }
    for (int y = 0; y < L.h; ++y)
      for (int x = 0; x < wl; ++x)
        if (std::isfinite(lm(x, y)) && left_right_differ(lm, out.right, x, y, -1, o.lr_max_diff)) {
          lm(x, y) = kInvalid;
          if (want_debug) out.lr_mask(x, y) = 1;
        }
    if (need_right) {
      for (std::size_t i = 0; i < rbad.size(); ++i)
        if (rbad[i] != 0) out.right.px[i] = kInvalid;
      post_filter(out.right, o, speckle);
    } else {
      out.right = ImgF();
    }
  }
  post_filter(out.left.disparity, o, speckle);
  return out;
}

// matcher:1023-1095
MatchOutput match_stereo(const ImgU16& left, const ImgU16& right, const ImgU8& left_valid, const ImgU8& right_valid,
                         int dmin, int nd, const r3dg_options& o, int threads) {
  check_options(o);
  if (nd <= 0) fail("stereo: no disparities to search");
  const int levels = o.levels >= 0 ? o.levels : automatic_levels(left.w, left.h, nd);
  const bool lr_full = o.lr_max_diff >= 0 && o.lr_full_match != 0;
  const int dmax = dmin + nd - 1;
  std::vector<ImgU16> lp, rp;
  std::vector<ImgU8> lvp, rvp;
  for (int l = 1; l <= levels; ++l) {
    lp.push_back(pyramid_down(l == 1 ? left : lp.back()));
    rp.push_back(pyramid_down(l == 1 ? right : rp.back()));
    if (!left_valid.empty()) lvp.push_back(reduce_mask(l == 1 ? left_valid : lvp.back()));
    if (!right_valid.empty()) rvp.push_back(reduce_mask(l == 1 ? right_valid : rvp.back()));
  }
  const ImgU8 none;
  LevelOut coarse;
  MatchOutput out;
  for (int l = levels; l >= 0; --l) {
    const std::size_t i = std::size_t(l - 1);
    const ImgU16& L = l == 0 ? left : lp[i];
    const ImgU16& R = l == 0 ? right : rp[i];
    const ImgU8& mL = l == 0 || left_valid.empty() ? left_valid : lvp[i];
    const ImgU8& mR = l == 0 || right_valid.empty() ? right_valid : rvp[i];
    const int ldmin = floor_shift(dmin, l);
    const int lnd = align16(ceil_shift(dmax, l) - ldmin + 1);
    const int speckle = o.speckle_size >> (2 * l);
    if (l == levels) {
      coarse = match_level(L, R, mL, mR, ldmin, lnd, nullptr, nullptr, lr_full && l > 0, speckle, o, l == 0, threads);
    } else {
      const Ranges lr = ranges_from_coarse(coarse.left.disparity, L.w, L.h, mL, mR, R.w, -1, ldmin, lnd, o, false);
      Ranges rr;
      if (lr_full) rr = ranges_from_coarse(coarse.right, R.w, R.h, mR, mL, L.w, 1, ldmin, lnd, o, true);
      coarse = match_level(L, R, mL, mR, ldmin, lnd, &lr, lr_full ? &rr : nullptr, lr_full && l > 0, speckle, o,
                           l == 0, threads);
    }
    out.cells += coarse.left.cells;
    out.tile_pixels += coarse.left.tile_pixels;
    if (l == 0) out.tiles_level0 = coarse.left.tiles;
  }
  out.disparity = std::move(coarse.left.disparity);
  out.confidence = std::move(coarse.left.confidence);
  out.labels = std::move(coarse.left.labels);
  out.lr_mask = std::move(coarse.lr_mask);
  out.disparity_min = dmin;
  out.num_disparities = align16(nd);
  out.levels = levels;
  return out;
}

}  // namespace oracle
