// Host driver of the GPU matcher and the extern "C" ABI of include/r3dg.h.
//
// match_stereo (stereo_matcher.cpp:1023-1095) runs entirely on the device:
// pyramids, Census, label ranges, tile layout, the 8 path scans, selection,
// blending, the left-right check and the post-filters of every level are
// kernels on the context's stream.  The host decides the tile layout (it needs
// one small read-back per level: the cell count of every tile, which sizes the
// aggregated volume and drives the reference's memory-budget tile halving).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "r3dg.h"
#include "runtime.hpp"

namespace r3dg {
namespace {

constexpr float kInf = __builtin_huge_valf();

int align16(int n) { return (n + 15) / 16 * 16; }
// stereo_matcher.cpp:889-890
int floor_shift(int v, int k) { return v >= 0 ? v >> k : -((-v + (1 << k) - 1) >> k); }
int ceil_shift(int v, int k) { return -floor_shift(-v, k); }

// stereo_matcher.cpp:1105-1122
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

// stereo_matcher.cpp:81-95
void check_options(const r3dg_options& o) {
  if (o.p1 < 0 || o.p2 < o.p1 || o.p2 > 8000)
    fail(R3DG_ERROR_INVALID_ARGUMENT, "stereo: invalid penalties P1 = " + std::to_string(o.p1) + ", P2 = " +
                                          std::to_string(o.p2) + " (0 <= P1 <= P2 <= 8000)");
  if (o.invalid_cost < 0 || o.invalid_cost > 255)
    fail(R3DG_ERROR_INVALID_ARGUMENT, "stereo: invalid_cost " + std::to_string(o.invalid_cost) + " outside [0, 255]");
  if (o.tile_size < 16 || o.tile_margin < 0 || o.tile_blend < 0)
    fail(R3DG_ERROR_INVALID_ARGUMENT, "stereo: invalid tiling");
  if (o.levels < -1 || o.levels > 12 || o.range_radius < 0 || o.range_margin < 0 || o.range_fill < 0)
    fail(R3DG_ERROR_INVALID_ARGUMENT, "stereo: invalid hierarchy levels / radius / margin / fill");
}

// stereo_matcher.cpp:69-79
std::array<std::uint16_t, 256> p2_table(const r3dg_options& o) {
  std::array<std::uint16_t, 256> lut{};
  for (int di = 0; di < 256; ++di) {
    int p2 = o.p2;
    if (o.p2_adapt > 0) p2 = std::max(o.p1 + 1, o.p2 * o.p2_adapt / (o.p2_adapt + di));
    lut[std::size_t(di)] = static_cast<std::uint16_t>(p2);
  }
  return lut;
}

bool view_empty(const r3dg_image& v) { return v.data == nullptr || v.width <= 0 || v.height <= 0; }

// ------------------------------------------------------------------ transfers
template <typename T>
DImg<T> alloc_img(Context& c, int w, int h) {
  DImg<T> d;
  d.w = w;
  d.h = h;
  d.s = w;
  d.own = DBuf<T>(c.arena, std::max<std::size_t>(std::size_t(w) * std::size_t(h), 1));
  d.p = d.own.get();
  return d;
}

template <typename T>
DImg<T> view_of(const DImg<T>& o) {
  DImg<T> d;
  d.p = o.p;
  d.w = o.w;
  d.h = o.h;
  d.s = o.s;
  return d;
}

// A caller image on the device: device views are used in place, host images
// are uploaded.  Empty views stay empty.
template <typename T>
DImg<T> input_img(Context& c, const r3dg_image& v, const char* name) {
  if (view_empty(v)) return DImg<T>{};
  if (v.stride < v.width) fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(name) + ": stride smaller than width");
  if (v.memory == R3DG_DEVICE) {
    DImg<T> d;
    d.p = static_cast<T*>(v.data);
    d.w = v.width;
    d.h = v.height;
    d.s = v.stride;
    return d;
  }
  DImg<T> d = alloc_img<T>(c, v.width, v.height);
  hip_check(hipMemcpy2DAsync(d.p, std::size_t(d.s) * sizeof(T), v.data, std::size_t(v.stride) * sizeof(T),
                             std::size_t(v.width) * sizeof(T), std::size_t(v.height), hipMemcpyHostToDevice, c.stream),
            "upload");
  return d;
}

template <typename T>
void output_img(Context& c, const r3dg_image& v, const T* src, int w, int h, std::int64_t ss, const char* name) {
  if (view_empty(v)) return;
  if (v.width != w || v.height != h)
    fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(name) + ": output size " + std::to_string(v.width) + "x" +
                                          std::to_string(v.height) + " differs from " + std::to_string(w) + "x" +
                                          std::to_string(h));
  if (v.stride < v.width) fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(name) + ": stride smaller than width");
  const hipMemcpyKind kind = v.memory == R3DG_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  hip_check(hipMemcpy2DAsync(v.data, std::size_t(v.stride) * sizeof(T), src, std::size_t(ss) * sizeof(T),
                             std::size_t(w) * sizeof(T), std::size_t(h), kind, c.stream),
            "download");
}

// Small host-to-device copy on the context's stream.  Waits for it, so the
// host vector may be released right after (pageable-memory copies are not
// guaranteed to have consumed the source when the async call returns).
template <typename T>
void upload(Context& c, DBuf<T>& dst, const std::vector<T>& src) {
  dst = DBuf<T>(c.arena, std::max<std::size_t>(src.size(), 1));
  if (!src.empty()) {
    hip_check(hipMemcpyAsync(dst.get(), src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, c.stream), "upload");
    hip_check(hipStreamSynchronize(c.stream), "hipStreamSynchronize");
  }
}

void sync(Context& c) { hip_check(hipStreamSynchronize(c.stream), "hipStreamSynchronize"); }
void check_launch() { hip_check(hipGetLastError(), "kernel launch"); }

// ------------------------------------------------------------------ stages
DImg<std::uint64_t> census(Context& c, const DImg<std::uint16_t>& img) {
  DImg<std::uint64_t> out = alloc_img<std::uint64_t>(c, img.w, img.h);
  launch_census(img.p, img.s, img.w, img.h, out.p, out.s, c.stream);
  check_launch();
  return out;
}

DImg<std::uint16_t> pyramid_u16(Context& c, const DImg<std::uint16_t>& img) {
  const int dw = (img.w + 1) / 2, dh = (img.h + 1) / 2;
  DBuf<std::uint32_t> tmp(c.arena, std::size_t(dw) * img.h);
  DImg<std::uint16_t> out = alloc_img<std::uint16_t>(c, dw, dh);
  launch_pyramid_u16(img.p, img.s, img.w, img.h, tmp.get(), out.p, out.s, c.stream);
  check_launch();
  return out;
}

DImg<std::uint8_t> pyramid_u8(Context& c, const DImg<std::uint8_t>& img, bool threshold) {
  const int dw = (img.w + 1) / 2, dh = (img.h + 1) / 2;
  DBuf<std::uint32_t> tmp(c.arena, std::size_t(dw) * img.h);
  DImg<std::uint8_t> out = alloc_img<std::uint8_t>(c, dw, dh);
  launch_pyramid_u8(img.p, img.s, img.w, img.h, tmp.get(), out.p, out.s, threshold, c.stream);
  check_launch();
  return out;
}

DImg<std::uint16_t> mirror_u16(Context& c, const DImg<std::uint16_t>& img) {
  DImg<std::uint16_t> out = alloc_img<std::uint16_t>(c, img.w, img.h);
  launch_mirror_u16(img.p, img.s, img.w, img.h, out.p, out.s, c.stream);
  check_launch();
  return out;
}

DImg<std::uint8_t> mirror_u8(Context& c, const DImg<std::uint8_t>& img) {
  if (img.empty()) return DImg<std::uint8_t>{};
  DImg<std::uint8_t> out = alloc_img<std::uint8_t>(c, img.w, img.h);
  launch_mirror_u8(img.p, img.s, img.w, img.h, out.p, out.s, c.stream);
  check_launch();
  return out;
}

DBuf<std::int32_t> intervals(Context& c, const DImg<std::uint8_t>& mask, int w, int h) {
  DBuf<std::int32_t> rows(c.arena, 2 * std::size_t(h));
  launch_valid_intervals(mask.empty() ? nullptr : mask.p, mask.s, w, h, rows.get(), c.stream);
  check_launch();
  return rows;
}

struct Ranges {
  DImg<std::uint16_t> lo, num;
};

Ranges ranges_from_coarse(Context& c, const DImg<float>& coarse, int width, int height,
                          const DImg<std::uint8_t>& fine_valid, const std::int32_t* partner_rows, int dmin, int nd,
                          const r3dg_options& o, int partner_sign, bool mirrored) {
  const int wc = coarse.w, hc = coarse.h;
  if (wc != (width + 1) / 2 || hc != (height + 1) / 2)
    fail(R3DG_ERROR_INVALID_ARGUMENT, "stereo: coarse map of " + std::to_string(wc) + "x" + std::to_string(hc) +
                                          " for a level of " + std::to_string(width) + "x" + std::to_string(height));
  if (coarse.s != coarse.w) fail(R3DG_ERROR_INTERNAL, "coarse map must be contiguous");
  Ranges r{alloc_img<std::uint16_t>(c, width, height), alloc_img<std::uint16_t>(c, width, height)};
  DBuf<float> scratch(c.arena, ranges_scratch_floats(wc, hc, o.range_fill));
  launch_ranges_from_coarse(coarse.p, wc, hc, width, height, fine_valid.empty() ? nullptr : fine_valid.p, fine_valid.s,
                            partner_rows, dmin, align16(nd), o.range_radius, o.range_margin, o.range_fill, partner_sign,
                            mirrored, scratch.get(), r.lo.p, r.num.p, r.lo.s, c.stream);
  check_launch();
  return r;
}

void speckles(Context& c, DImg<float>& d, int min_size, float max_diff) {
  if (min_size <= 1 || d.empty()) return;
  DBuf<std::int32_t> parent(c.arena, d.pixels()), size(c.arena, d.pixels());
  launch_speckles(d.p, d.w, d.h, min_size, max_diff, parent.get(), size.get(), c.stream);
  check_launch();
}

DImg<float> median(Context& c, const DImg<float>& d) {
  DImg<float> out = alloc_img<float>(c, d.w, d.h);
  launch_median3x3(d.p, d.w, d.h, out.p, c.stream);
  check_launch();
  return out;
}

void post_filter(Context& c, DImg<float>& d, const r3dg_options& o, int speckle) {
  speckles(c, d, speckle, o.speckle_diff);
  if (o.median != 0) d = median(c, d);
}

// ------------------------------------------------------------------ tiles
// One matching direction of one level (MatchContext, stereo_matcher.cpp:113-163).
struct Job {
  const DImg<std::uint16_t>* base = nullptr;
  const DImg<std::uint16_t>* partner = nullptr;
  DImg<std::uint64_t> cl, cr;
  const std::int32_t* lrows = nullptr;
  const std::int32_t* rrows = nullptr;
  DBuf<std::int32_t> rows4;
  const Ranges* ranges = nullptr;
  int dmin = 0;
  int D = 0;
  bool confidence = true;
  bool labels = false;
};

DBuf<std::int32_t> pack_rows(Context& c, const std::int32_t* base, const std::int32_t* partner, int h) {
  DBuf<std::int32_t> r(c.arena, 4 * std::size_t(h));
  launch_pack_rows(base, partner, h, r.get(), c.stream);
  check_launch();
  return r;
}

struct TilesOut {
  DImg<float> disp, conf;
  DImg<std::int16_t> lab;
  std::uint64_t cells = 0;
  std::uint64_t tile_pixels = 0;
  int tiles = 0;
};

// stereo_matcher.cpp:545-552
std::vector<int> split(int size, int tile) {
  const int n = std::max(1, (size + tile - 1) / tile);
  std::vector<int> b(std::size_t(n) + 1);
  for (int i = 0; i <= n; ++i) b[std::size_t(i)] = static_cast<int>(static_cast<long long>(size) * i / n);
  return b;
}

struct Grid {
  std::vector<std::array<int, 4>> cols, rows;  // extent [x0, x1), core [c0, c1)
};

// stereo_matcher.cpp:619-634
Grid layout(int width, int height, int tile_size, int ext) {
  Grid g;
  const auto bx = split(width, tile_size), by = split(height, tile_size);
  for (std::size_t i = 0; i + 1 < bx.size(); ++i)
    g.cols.push_back({std::max(0, bx[i] - ext), std::min(width, bx[i + 1] + ext), bx[i], bx[i + 1]});
  for (std::size_t j = 0; j + 1 < by.size(); ++j)
    g.rows.push_back({std::max(0, by[j] - ext), std::min(height, by[j + 1] + ext), by[j], by[j + 1]});
  return g;
}

// stereo_matcher.cpp:570-577 with the tile's exact cell count.
std::size_t estimated_tile_bytes(std::uint64_t cells64, int w_, int h_, int right_width) {
  const auto cells = static_cast<std::size_t>(cells64);
  const auto w = static_cast<std::size_t>(w_), h = static_cast<std::size_t>(h_);
  const std::size_t row_cells = cells / std::max<std::size_t>(h, 1);
  return cells * 2 + row_cells * 2 + 6 * (row_cells + w * 2 * 8) * 2 + w * h * 32 +
         (w + static_cast<std::size_t>(right_width) + 16) * (h + 6) * 8;
}

int path_scanlines(int p, int w, int h) { return p < 2 ? h : (p < 4 ? w : w + h - 1); }

// match_tiles (stereo_matcher.cpp:613-771) for several jobs at once: their
// tiles share the layout pass, the 8 path launches and the selection launch.
// Tile layout and aggregated volume S of a set of jobs (the tiles of all jobs
// share the layout pass and the 8 path launches).
struct Aggregation {
  DBuf<JobDesc> djobs;
  std::vector<TileDesc> td;
  DBuf<TileDesc> dtiles;
  std::vector<Grid> grids;
  std::vector<int> tile_base;
  std::vector<std::uint64_t> tcells;
  DBuf<std::uint64_t> drow_prefix, row_start, pix;
  DBuf<std::uint16_t> S;
  std::uint64_t total_rows = 0, total_px = 0, total_cells = 0;
  int max_w = 1, max_D = 16;
};

Aggregation aggregate_jobs(Context& c, std::vector<Job>& jobs, const r3dg_options& o) {
  Aggregation A;
  const int nj = static_cast<int>(jobs.size());
  const auto lut = p2_table(o);
  std::vector<JobDesc> jd(static_cast<std::size_t>(nj));
  for (int j = 0; j < nj; ++j) {
    const Job& J = jobs[std::size_t(j)];
    JobDesc& d = jd[std::size_t(j)];
    std::memset(&d, 0, sizeof(d));
    d.img = J.base->p;
    d.img_stride = J.base->s;
    d.cl = J.cl.p;
    d.cl_stride = J.cl.s;
    d.cr = J.cr.p;
    d.cr_stride = J.cr.s;
    d.rows4 = J.rows4.get();
    d.lo = J.ranges != nullptr ? J.ranges->lo.p : nullptr;
    d.num = J.ranges != nullptr ? J.ranges->num.p : nullptr;
    d.r_stride = J.ranges != nullptr ? J.ranges->lo.s : 0;
    d.width = J.base->w;
    d.height = J.base->h;
    d.pwidth = J.partner->w;
    d.dmin = J.dmin;
    d.D = J.D;
    d.p1 = o.p1;
    d.invalid_cost = o.invalid_cost;
    d.confidence = J.confidence ? 1 : 0;
    std::memcpy(d.lut, lut.data(), sizeof(d.lut));
  }
  DBuf<JobDesc>& djobs = A.djobs;
  upload(c, djobs, jd);

  const int ext = o.tile_margin + (o.tile_blend + 1) / 2;
  std::vector<int> tsize(static_cast<std::size_t>(nj), o.tile_size);
  std::vector<Grid>& grids = A.grids;
  grids.resize(static_cast<std::size_t>(nj));
  std::vector<TileDesc>& td = A.td;
  std::vector<int>& tile_base = A.tile_base;
  tile_base.resize(std::size_t(nj) + 1);
  std::vector<std::uint64_t> row_prefix;
  std::vector<std::uint64_t>& tcells = A.tcells;
  std::vector<std::uint32_t> tmax;
  DBuf<TileDesc>& dtiles = A.dtiles;
  DBuf<std::uint64_t>& drow_prefix = A.drow_prefix;
  DBuf<std::uint64_t>& row_start = A.row_start;
  DBuf<std::uint64_t>& pix = A.pix;
  DBuf<std::uint64_t> row_cells, tile_cells;
  DBuf<std::uint32_t> row_max, tile_max;
  std::uint64_t& total_rows = A.total_rows;
  std::uint64_t& total_px = A.total_px;
  std::vector<bool> fixed(static_cast<std::size_t>(nj), false);
  std::optional<StageTimer> layout_timer(std::in_place, c, R3DG_STAGE_LAYOUT);
  for (;;) {
    td.clear();
    row_prefix.clear();
    total_rows = 0;
    total_px = 0;
    for (int j = 0; j < nj; ++j) {
      const Job& J = jobs[std::size_t(j)];
      grids[std::size_t(j)] = layout(J.base->w, J.base->h, tsize[std::size_t(j)], ext);
      tile_base[std::size_t(j)] = static_cast<int>(td.size());
      for (const auto& r : grids[std::size_t(j)].rows) {
        for (const auto& cc : grids[std::size_t(j)].cols) {
          TileDesc t{};
          t.x0 = cc[0];
          t.y0 = r[0];
          t.w = cc[1] - cc[0];
          t.h = r[1] - r[0];
          t.job = j;
          t.px_base = total_px;
          t.row_base = total_rows;
          row_prefix.push_back(total_rows);
          total_px += std::uint64_t(t.w) * std::uint64_t(t.h);
          total_rows += std::uint64_t(t.h);
          td.push_back(t);
        }
      }
    }
    tile_base[std::size_t(nj)] = static_cast<int>(td.size());
    const int nt = static_cast<int>(td.size());
    upload(c, dtiles, td);
    upload(c, drow_prefix, row_prefix);
    pix = DBuf<std::uint64_t>(c.arena, total_px);
    row_cells = DBuf<std::uint64_t>(c.arena, total_rows);
    row_max = DBuf<std::uint32_t>(c.arena, total_rows);
    row_start = DBuf<std::uint64_t>(c.arena, total_rows);
    tile_cells = DBuf<std::uint64_t>(c.arena, std::size_t(nt));
    tile_max = DBuf<std::uint32_t>(c.arena, std::size_t(nt));
    launch_tile_row_scan(djobs.get(), dtiles.get(), drow_prefix.get(), nt, total_rows, pix.get(), row_cells.get(),
                         row_max.get(), c.stream);
    check_launch();
    launch_tile_cells(dtiles.get(), nt, row_cells.get(), row_max.get(), row_start.get(), tile_cells.get(),
                      tile_max.get(), c.stream);
    check_launch();
    tcells.assign(std::size_t(nt), 0);
    tmax.assign(std::size_t(nt), 0);
    hip_check(hipMemcpyAsync(tcells.data(), tile_cells.get(), sizeof(std::uint64_t) * nt, hipMemcpyDeviceToHost,
                             c.stream),
              "download");
    hip_check(hipMemcpyAsync(tmax.data(), tile_max.get(), sizeof(std::uint32_t) * nt, hipMemcpyDeviceToHost, c.stream),
              "download");
    sync(c);
    // Halve the tile size of a job while its largest tile exceeds the memory
    // budget (stereo_matcher.cpp:643-648).
    bool changed = false;
    for (int j = 0; j < nj; ++j) {
      if (fixed[std::size_t(j)]) continue;
      std::size_t mb = 0;
      for (int t = tile_base[std::size_t(j)]; t < tile_base[std::size_t(j) + 1]; ++t)
        mb = std::max(mb, estimated_tile_bytes(tcells[std::size_t(t)], td[std::size_t(t)].w, td[std::size_t(t)].h,
                                               jobs[std::size_t(j)].partner->w));
      if (tsize[std::size_t(j)] > 64 && mb > o.memory_budget) {
        tsize[std::size_t(j)] /= 2;
        changed = true;
      } else {
        fixed[std::size_t(j)] = true;
      }
    }
    if (!changed) break;
  }
  const int nt = static_cast<int>(td.size());
  std::uint64_t& total_cells = A.total_cells;
  int max_num = 16;
  int& max_w = A.max_w;
  int& max_D = A.max_D;
  for (int t = 0; t < nt; ++t) {
    td[std::size_t(t)].cell_base = total_cells;
    td[std::size_t(t)].max_num = static_cast<std::int32_t>(tmax[std::size_t(t)]);
    total_cells += tcells[std::size_t(t)];
    max_num = std::max<int>(max_num, static_cast<int>(tmax[std::size_t(t)]));
    max_w = std::max(max_w, td[std::size_t(t)].w);
  }
  for (const Job& J : jobs) max_D = std::max(max_D, J.D);
  if (max_num > kMaxLabels)
    fail(R3DG_ERROR_INVALID_ARGUMENT, "stereo: more than " + std::to_string(kMaxLabels) + " labels per pixel are not supported on the GPU");
  upload(c, dtiles, td);
  // Group prefixes of the 8 paths (4 scanlines per wave), one upload.
  std::vector<std::uint64_t> gp(8 * (std::size_t(nt) + 1), 0);
  for (int p = 0; p < 8; ++p) {
    std::uint64_t* g = gp.data() + std::size_t(p) * (std::size_t(nt) + 1);
    for (int t = 0; t < nt; ++t)
      g[t + 1] = g[t] + std::uint64_t((path_scanlines(p, td[std::size_t(t)].w, td[std::size_t(t)].h) +
                                       kScanPerWave - 1) / kScanPerWave);
  }
  DBuf<std::uint64_t> dgp;
  upload(c, dgp, gp);
  A.S = DBuf<std::uint16_t>(c.arena, total_cells + kSPad);
  DBuf<std::uint16_t>& S = A.S;
  layout_timer.reset();

  // Scratch rows for pixels wider than the kernel's LDS rows.
  std::uint64_t max_groups = 0;
  for (int p = 0; p < 8; ++p) max_groups = std::max(max_groups, gp[std::size_t(p) * (std::size_t(nt) + 1) + nt]);
  DBuf<std::uint16_t> scratch;
  if (const std::size_t n = path_scratch_elems(max_groups, max_num); n > 0) scratch = DBuf<std::uint16_t>(c.arena, n);

  // Matching costs of every cell (timed with the other cost preparation).
  DBuf<std::uint8_t> C(c.arena, total_cells + kSPad);
  {
    const StageTimer timer(c, R3DG_STAGE_PREPARE);
    launch_cost(djobs.get(), dtiles.get(), drow_prefix.get(), nt, total_rows, row_start.get(), pix.get(), C.get(),
                c.stream);
    check_launch();
  }

  // Aggregation: 8 path launches over all tiles.
  {
    const StageTimer timer(c, R3DG_STAGE_AGGREGATE);
    for (int p = 0; p < 8; ++p) {
      const std::size_t o8 = std::size_t(p) * (std::size_t(nt) + 1);
      launch_sgm_path(p, djobs.get(), dtiles.get(), dgp.get() + o8, nt, gp[o8 + std::size_t(nt)], row_start.get(),
                      pix.get(), C.get(), S.get(), scratch.get(), max_num, p > 0, c.stream);
      check_launch();
    }
  }
  c.agg_launches += 8;
  // Algorithmic bytes of the 8 launches: per cell C read 8 x 1, S written
  // 8 x 2 and read 7 x 2; per tile pixel and path its record (8) and intensity (2).
  c.agg_bytes += 38 * total_cells + 80 * total_px;

  return A;
}

std::vector<TilesOut> match_tiles(Context& c, std::vector<Job>& jobs, const r3dg_options& o, SelectParams sp) {
  const int nj = static_cast<int>(jobs.size());
  Aggregation A = aggregate_jobs(c, jobs, o);
  const int nt = static_cast<int>(A.td.size());
  const std::uint64_t total_px = A.total_px;
  // Selection into per-tile maps.
  bool any_labels = false;
  for (const Job& J : jobs) any_labels = any_labels || J.labels;
  DBuf<float> tdisp(c.arena, std::max<std::uint64_t>(total_px, 1)), tconf(c.arena, std::max<std::uint64_t>(total_px, 1));
  DBuf<std::int16_t> tlab;
  if (any_labels) tlab = DBuf<std::int16_t>(c.arena, std::max<std::uint64_t>(total_px, 1));
  {
    const StageTimer timer(c, R3DG_STAGE_SELECT);
    launch_select(A.djobs.get(), A.dtiles.get(), A.drow_prefix.get(), nt, A.total_rows, A.row_start.get(),
                  A.pix.get(), A.S.get(), sp, A.max_w, A.max_D, tdisp.get(), tconf.get(), tlab.get(), c.stream);
    check_launch();
  }
  A.S.reset();

  std::vector<TilesOut> out(static_cast<std::size_t>(nj));
  for (int j = 0; j < nj; ++j) {
    const Job& J = jobs[std::size_t(j)];
    const Grid& g = A.grids[std::size_t(j)];
    TilesOut& r = out[std::size_t(j)];
    r.disp = alloc_img<float>(c, J.base->w, J.base->h);
    if (J.confidence) r.conf = alloc_img<float>(c, J.base->w, J.base->h);
    if (J.labels) r.lab = alloc_img<std::int16_t>(c, J.base->w, J.base->h);
    std::vector<std::int32_t> cols, rows;
    for (const auto& e : g.cols) cols.insert(cols.end(), e.begin(), e.end());
    for (const auto& e : g.rows) rows.insert(rows.end(), e.begin(), e.end());
    DBuf<std::int32_t> dcols, drows;
    upload(c, dcols, cols);
    upload(c, drows, rows);
    GridDesc gd{};
    gd.cols = dcols.get();
    gd.rows = drows.get();
    gd.nx = static_cast<int>(g.cols.size());
    gd.ny = static_cast<int>(g.rows.size());
    gd.width = J.base->w;
    gd.height = J.base->h;
    gd.blend = o.tile_blend;
    gd.tile_base = A.tile_base[std::size_t(j)];
    {
      const StageTimer timer(c, R3DG_STAGE_BLEND);
      launch_blend(gd, A.dtiles.get(), tdisp.get(), tconf.get(), J.labels ? tlab.get() : nullptr, r.disp.p,
                   J.confidence ? r.conf.p : nullptr, J.labels ? r.lab.p : nullptr, c.stream);
      check_launch();
    }
    for (int t = A.tile_base[std::size_t(j)]; t < A.tile_base[std::size_t(j) + 1]; ++t) {
      r.cells += A.tcells[std::size_t(t)];
      r.tile_pixels += std::uint64_t(A.td[std::size_t(t)].w) * std::uint64_t(A.td[std::size_t(t)].h);
    }
    r.tiles = A.tile_base[std::size_t(j) + 1] - A.tile_base[std::size_t(j)];
    // dcols / drows return to the arena here; later uses of the same memory
    // are ordered after the blend on the context's stream.
  }
  return out;
}

// ------------------------------------------------------------------ levels
struct LevelIn {
  const DImg<std::uint16_t>* L;
  const DImg<std::uint16_t>* R;
  const DImg<std::uint8_t>* mL;
  const DImg<std::uint8_t>* mR;
  int dmin, nd;
};

struct LevelOut {
  TilesOut left;
  DImg<float> right;
  DImg<std::uint8_t> lr_mask;
};

void check_inputs(const DImg<std::uint16_t>& L, const DImg<std::uint16_t>& R, const DImg<std::uint8_t>& mL,
                  const DImg<std::uint8_t>& mR, int nd) {
  if (L.h != R.h || L.empty() || R.empty())
    fail(R3DG_ERROR_INVALID_ARGUMENT, "stereo: rectified images must be non-empty with equal height (" +
                                          std::to_string(L.w) + "x" + std::to_string(L.h) + ", " + std::to_string(R.w) +
                                          "x" + std::to_string(R.h) + ")");
  if (!mL.empty() && (mL.w != L.w || mL.h != L.h))
    fail(R3DG_ERROR_INVALID_ARGUMENT, "stereo: left mask size differs from the image");
  if (!mR.empty() && (mR.w != R.w || mR.h != R.h))
    fail(R3DG_ERROR_INVALID_ARGUMENT, "stereo: right mask size differs from the image");
  if (nd <= 0) fail(R3DG_ERROR_INVALID_ARGUMENT, "stereo: no disparities to search");
}

// match_level (stereo_matcher.cpp:831-886) with the coarser level's maps.
LevelOut match_level(Context& c, const LevelIn& in, const DImg<float>* coarse_left, const DImg<float>* coarse_right,
                     bool need_right, int speckle, const r3dg_options& o, bool want_debug) {
  check_inputs(*in.L, *in.R, *in.mL, *in.mR, in.nd);
  const int wl = in.L->w, wr = in.R->w, h = in.L->h;
  const int D = align16(in.nd);
  const bool check_lr = o.lr_max_diff >= 0;
  const bool full = check_lr && o.lr_full_match != 0;

  std::optional<StageTimer> prep_timer(std::in_place, c, R3DG_STAGE_PREPARE);
  DBuf<std::int32_t> lrows = intervals(c, *in.mL, wl, h);
  DBuf<std::int32_t> rrows = intervals(c, *in.mR, wr, h);
  Ranges lranges, rranges;
  if (coarse_left != nullptr) {
    const StageTimer timer(c, R3DG_STAGE_RANGES);
    lranges = ranges_from_coarse(c, *coarse_left, wl, h, *in.mL, rrows.get(), in.dmin, in.nd, o, -1, false);
  }

  std::vector<Job> jobs(1);
  jobs[0].base = in.L;
  jobs[0].partner = in.R;
  jobs[0].cl = census(c, *in.L);
  jobs[0].cr = census(c, *in.R);
  jobs[0].lrows = lrows.get();
  jobs[0].rrows = rrows.get();
  jobs[0].rows4 = pack_rows(c, lrows.get(), rrows.get(), h);
  jobs[0].ranges = coarse_left != nullptr ? &lranges : nullptr;
  jobs[0].dmin = in.dmin;
  jobs[0].D = D;
  jobs[0].confidence = true;
  jobs[0].labels = want_debug;

  DImg<std::uint16_t> mLi, mRi;
  DImg<std::uint8_t> mLv, mRv;
  DBuf<std::int32_t> mrows_base, mrows_partner;
  if (full) {
    // Right-to-left match on the mirrored pair (stereo_matcher.cpp:842-856):
    // base = mirror(R), partner = mirror(L), disparity offset wr - wl.
    mLi = mirror_u16(c, *in.L);
    mRi = mirror_u16(c, *in.R);
    mLv = mirror_u8(c, *in.mL);
    mRv = mirror_u8(c, *in.mR);
    mrows_base = intervals(c, mRv, wr, h);
    mrows_partner = intervals(c, mLv, wl, h);
    if (coarse_right != nullptr) {
      const StageTimer timer(c, R3DG_STAGE_RANGES);
      rranges = ranges_from_coarse(c, *coarse_right, wr, h, *in.mR, lrows.get(), in.dmin, in.nd, o, 1, true);
    }
    Job b;
    b.base = &mRi;
    b.partner = &mLi;
    b.cl = census(c, mRi);
    b.cr = census(c, mLi);
    b.lrows = mrows_base.get();
    b.rrows = mrows_partner.get();
    b.rows4 = pack_rows(c, mrows_base.get(), mrows_partner.get(), h);
    b.ranges = coarse_right != nullptr ? &rranges : nullptr;
    b.dmin = in.dmin + wr - wl;
    b.D = D;
    b.confidence = false;
    jobs.push_back(std::move(b));
  }
  SelectParams sp{};
  sp.subpixel = o.subpixel;
  sp.uniqueness = o.uniqueness;
  sp.lr_max_diff = o.lr_max_diff;
  sp.diagonal_lr = check_lr && !full ? 1 : 0;
  prep_timer.reset();
  std::vector<TilesOut> maps = match_tiles(c, jobs, o, sp);
  const StageTimer filter_timer(c, R3DG_STAGE_FILTER);

  LevelOut out;
  out.left = std::move(maps[0]);
  if (want_debug) {
    out.lr_mask = alloc_img<std::uint8_t>(c, wl, h);
    launch_fill_u8(out.lr_mask.p, out.lr_mask.pixels(), 0, c.stream);
    check_launch();
  }
  if (full) {
    out.left.cells += maps[1].cells;
    out.left.tile_pixels += maps[1].tile_pixels;
    out.right = alloc_img<float>(c, wr, h);
    launch_unmirror_right(maps[1].disp.p, wr, h, wl, out.right.p, c.stream);
    check_launch();
    DBuf<std::uint8_t> rbad;
    if (need_right) {
      // Both checks compare the unchecked maps (stereo_matcher.cpp:864-876).
      rbad = DBuf<std::uint8_t>(c.arena, out.right.pixels());
      launch_lr_check(out.right.p, wr, h, out.left.disp.p, wl, 1, o.lr_max_diff, rbad.get(), false, c.stream);
      check_launch();
    }
    launch_lr_check(out.left.disp.p, wl, h, out.right.p, wr, -1, o.lr_max_diff,
                    want_debug ? out.lr_mask.p : nullptr, true, c.stream);
    check_launch();
    if (need_right) {
      launch_apply_mask(out.right.p, out.right.pixels(), rbad.get(), c.stream);
      check_launch();
      post_filter(c, out.right, o, speckle);
    } else {
      out.right = DImg<float>{};
    }
  }
  post_filter(c, out.left.disp, o, speckle);
  return out;
}

struct MatchResult {
  LevelOut final;
  int levels = 0;
  std::uint64_t cells = 0;
  std::uint64_t tile_pixels = 0;
  int tiles0 = 0;
};

// match_stereo (stereo_matcher.cpp:1023-1095).
MatchResult match_stereo(Context& c, const DImg<std::uint16_t>& left, const DImg<std::uint16_t>& right,
                         const DImg<std::uint8_t>& lvalid, const DImg<std::uint8_t>& rvalid, int dmin, int nd,
                         const r3dg_options& o, bool want_debug) {
  check_options(o);
  if (nd <= 0) fail(R3DG_ERROR_INVALID_ARGUMENT, "stereo: no disparities to search");
  check_inputs(left, right, lvalid, rvalid, nd);
  const int levels = o.levels >= 0 ? o.levels : automatic_levels(left.w, left.h, nd);
  const bool lr_full = o.lr_max_diff >= 0 && o.lr_full_match != 0;
  const int dmax = dmin + nd - 1;
  std::vector<DImg<std::uint16_t>> lp, rp;
  std::vector<DImg<std::uint8_t>> lvp, rvp;
  std::optional<StageTimer> pyr_timer(std::in_place, c, R3DG_STAGE_PYRAMID);
  for (int l = 1; l <= levels; ++l) {
    lp.push_back(pyramid_u16(c, l == 1 ? left : lp.back()));
    rp.push_back(pyramid_u16(c, l == 1 ? right : rp.back()));
    if (!lvalid.empty()) lvp.push_back(pyramid_u8(c, l == 1 ? lvalid : lvp.back(), true));
    if (!rvalid.empty()) rvp.push_back(pyramid_u8(c, l == 1 ? rvalid : rvp.back(), true));
  }
  pyr_timer.reset();
  MatchResult res;
  res.levels = levels;
  LevelOut coarse;
  for (int l = levels; l >= 0; --l) {
    if (c.cancel.load() != 0) fail(R3DG_ERROR_CANCELLED, "cancelled");
    const std::size_t i = std::size_t(l - 1);
    LevelIn in{};
    in.L = l == 0 ? &left : &lp[i];
    in.R = l == 0 ? &right : &rp[i];
    in.mL = l == 0 || lvalid.empty() ? &lvalid : &lvp[i];
    in.mR = l == 0 || rvalid.empty() ? &rvalid : &rvp[i];
    in.dmin = floor_shift(dmin, l);
    in.nd = align16(ceil_shift(dmax, l) - in.dmin + 1);
    const int speckle = o.speckle_size >> (2 * l);
    if (l == levels) {
      coarse = match_level(c, in, nullptr, nullptr, lr_full && l > 0, speckle, o, l == 0 && want_debug);
    } else {
      LevelOut prev = std::move(coarse);
      coarse = match_level(c, in, &prev.left.disp, lr_full ? &prev.right : nullptr, lr_full && l > 0, speckle, o,
                           l == 0 && want_debug);
    }
    res.cells += coarse.left.cells;
    res.tile_pixels += coarse.left.tile_pixels;
    if (l == 0) res.tiles0 = coarse.left.tiles;
  }
  res.final = std::move(coarse);
  return res;
}

// ------------------------------------------------------------------ C ABI glue
thread_local std::string g_error;
thread_local int32_t g_status = R3DG_OK;

template <typename F>
int32_t guarded(Context* c, F&& f) {
  std::string msg;
  int32_t status = R3DG_OK;
  try {
    f();
    return R3DG_OK;
  } catch (const Failure& e) {
    status = e.status;
    msg = e.what();
  } catch (const std::bad_alloc&) {
    status = R3DG_ERROR_OUT_OF_MEMORY;
    msg = "host allocation failed";
  } catch (const std::exception& e) {
    status = R3DG_ERROR_INTERNAL;
    msg = e.what();
  } catch (...) {
    status = R3DG_ERROR_INTERNAL;
    msg = "unknown error";
  }
  if (c != nullptr) {
    c->error = msg;
    c->error_status = status;
    // Leave the stream idle so the context stays usable.
    (void)hipStreamSynchronize(c->stream);
    (void)hipGetLastError();
  }
  g_error = msg;
  g_status = status;
  return status;
}

int32_t copy_error(const std::string& m, int32_t status, char* buf, std::size_t size) {
  if (buf != nullptr && size > 0) {
    const std::size_t n = std::min(size - 1, m.size());
    std::memcpy(buf, m.data(), n);
    buf[n] = '\0';
  }
  return status;
}

template <typename F>
int32_t with_ctx(r3dg_ctx* h, F&& f) {
  if (h == nullptr) {
    g_error = "null context";
    g_status = R3DG_ERROR_INVALID_ARGUMENT;
    return g_status;
  }
  auto* c = reinterpret_cast<Context*>(h);
  const std::lock_guard<std::mutex> lock(c->mutex);
  return guarded(c, [&] {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    f(*c);
  });
}

const r3dg_options& opts_or_default(const r3dg_options* o, r3dg_options& tmp) {
  if (o != nullptr) return *o;
  r3dg_default_options(&tmp);
  return tmp;
}

}  // namespace
}  // namespace r3dg

using namespace r3dg;

extern "C" {

int32_t r3dg_abi_version(void) { return R3DG_ABI_VERSION; }

void r3dg_default_options(r3dg_options* o) {
  if (o == nullptr) return;
  o->p1 = 28;
  o->p2 = 200;
  o->p2_adapt = 4;
  o->invalid_cost = 31;
  o->subpixel = R3DG_SUBPIXEL_EQUIANGULAR;
  o->uniqueness = 0.0F;
  o->lr_max_diff = 1;
  o->lr_full_match = 1;
  o->speckle_size = 100;
  o->speckle_diff = 1.0F;
  o->median = 1;
  o->levels = -1;
  o->range_radius = 3;
  o->range_margin = 4;
  o->range_fill = 64;
  o->tile_size = 512;
  o->tile_margin = 32;
  o->tile_blend = 32;
  o->memory_budget = std::uint64_t{2} << 30;
}

int32_t r3dg_ctx_create(int32_t device, uint64_t hbm_budget, r3dg_ctx** out) {
  return guarded(nullptr, [&] {
    if (out == nullptr) fail(R3DG_ERROR_INVALID_ARGUMENT, "null output pointer");
    *out = nullptr;
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n)
      fail(R3DG_ERROR_INVALID_ARGUMENT, "no HIP device " + std::to_string(device) + " (" + std::to_string(n) + " found)");
    hip_check(hipSetDevice(device), "hipSetDevice");
    std::size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const std::uint64_t budget = hbm_budget != 0 ? hbm_budget : std::uint64_t(free_b) / 10 * 9;
    auto c = std::make_unique<Context>(budget);
    c->device = device;
    hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    *out = reinterpret_cast<r3dg_ctx*>(c.release());
  });
}

int32_t r3dg_ctx_destroy(r3dg_ctx* h) {
  if (h == nullptr) return R3DG_OK;
  auto* c = reinterpret_cast<Context*>(h);
  return guarded(nullptr, [&] {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    (void)hipStreamSynchronize(c->stream);
    c->arena.trim();
    (void)hipStreamDestroy(c->stream);
    delete c;
  });
}

int32_t r3dg_last_error(const r3dg_ctx* h, char* buf, size_t size) {
  if (h == nullptr) return copy_error(g_error, g_status, buf, size);
  const auto* c = reinterpret_cast<const Context*>(h);
  return copy_error(c->error, c->error_status, buf, size);
}

int32_t r3dg_global_error(char* buf, size_t size) { return copy_error(g_error, g_status, buf, size); }

int32_t r3dg_set_profiling(r3dg_ctx* h, int32_t enabled) {
  return with_ctx(h, [&](Context& c) { c.profiling = enabled != 0 ? 1 : 0; });
}

int32_t r3dg_cancel(r3dg_ctx* h, int32_t requested) {
  if (h == nullptr) return R3DG_ERROR_INVALID_ARGUMENT;
  reinterpret_cast<Context*>(h)->cancel.store(requested != 0 ? 1 : 0);
  return R3DG_OK;
}

int32_t r3dg_automatic_levels(int32_t width, int32_t height, int32_t num_disparities) {
  return automatic_levels(width, height, num_disparities);
}

int32_t r3dg_census(r3dg_ctx* h, r3dg_image img, r3dg_image out) {
  return with_ctx(h, [&](Context& c) {
    if (view_empty(img)) return;
    DImg<std::uint16_t> d = input_img<std::uint16_t>(c, img, "census input");
    DImg<std::uint64_t> r = census(c, d);
    output_img(c, out, r.p, r.w, r.h, r.s, "census output");
    sync(c);
  });
}

int32_t r3dg_pyramid_down_u16(r3dg_ctx* h, r3dg_image img, r3dg_image out) {
  return with_ctx(h, [&](Context& c) {
    if (view_empty(img)) fail(R3DG_ERROR_INVALID_ARGUMENT, "pyramid_down: empty image");
    DImg<std::uint16_t> d = input_img<std::uint16_t>(c, img, "pyramid input");
    DImg<std::uint16_t> r = pyramid_u16(c, d);
    output_img(c, out, r.p, r.w, r.h, r.s, "pyramid output");
    sync(c);
  });
}

int32_t r3dg_pyramid_down_u8(r3dg_ctx* h, r3dg_image img, r3dg_image out) {
  return with_ctx(h, [&](Context& c) {
    if (view_empty(img)) fail(R3DG_ERROR_INVALID_ARGUMENT, "pyramid_down: empty image");
    DImg<std::uint8_t> d = input_img<std::uint8_t>(c, img, "pyramid input");
    DImg<std::uint8_t> r = pyramid_u8(c, d, false);
    output_img(c, out, r.p, r.w, r.h, r.s, "pyramid output");
    sync(c);
  });
}

int32_t r3dg_cost_volume(r3dg_ctx* h, r3dg_image left, r3dg_image right, int32_t disparity_min,
                         int32_t num_disparities, const r3dg_options* options, r3dg_image lo, r3dg_image num,
                         uint16_t* out, uint64_t out_cells, int32_t out_memory) {
  return with_ctx(h, [&](Context& c) {
    r3dg_options tmp;
    const r3dg_options& o = opts_or_default(options, tmp);
    check_options(o);
    DImg<std::uint16_t> L = input_img<std::uint16_t>(c, left, "left");
    DImg<std::uint16_t> R = input_img<std::uint16_t>(c, right, "right");
    const DImg<std::uint8_t> none;
    check_inputs(L, R, none, none, num_disparities);
    const int D = align16(num_disparities);
    Ranges rg;
    const bool has = !view_empty(lo) || !view_empty(num);
    if (has) {
      if (view_empty(lo) || view_empty(num) || lo.width != L.w || lo.height != L.h || num.width != L.w ||
          num.height != L.h)
        fail(R3DG_ERROR_INVALID_ARGUMENT, "stereo: label ranges of the wrong size");
      rg.lo = input_img<std::uint16_t>(c, lo, "lo");
      rg.num = input_img<std::uint16_t>(c, num, "num");
      if (rg.lo.s != rg.num.s) {
        DImg<std::uint16_t> cn = alloc_img<std::uint16_t>(c, L.w, L.h);
        hip_check(hipMemcpy2DAsync(cn.p, cn.s * 2, rg.num.p, rg.num.s * 2, std::size_t(L.w) * 2, std::size_t(L.h),
                                   hipMemcpyDeviceToDevice, c.stream),
                  "copy");
        DImg<std::uint16_t> cl = alloc_img<std::uint16_t>(c, L.w, L.h);
        hip_check(hipMemcpy2DAsync(cl.p, cl.s * 2, rg.lo.p, rg.lo.s * 2, std::size_t(L.w) * 2, std::size_t(L.h),
                                   hipMemcpyDeviceToDevice, c.stream),
                  "copy");
        rg.num = std::move(cn);
        rg.lo = std::move(cl);
      }
      // check_ranges (stereo_matcher.cpp:97-111) on the host copy.
      std::vector<std::uint16_t> hl(L.pixels()), hn(L.pixels());
      hip_check(hipMemcpy2DAsync(hl.data(), std::size_t(L.w) * 2, rg.lo.p, rg.lo.s * 2, std::size_t(L.w) * 2,
                                 std::size_t(L.h), hipMemcpyDeviceToHost, c.stream),
                "download");
      hip_check(hipMemcpy2DAsync(hn.data(), std::size_t(L.w) * 2, rg.num.p, rg.num.s * 2, std::size_t(L.w) * 2,
                                 std::size_t(L.h), hipMemcpyDeviceToHost, c.stream),
                "download");
      sync(c);
      std::uint64_t total = 0;
      for (std::size_t i = 0; i < hl.size(); ++i) {
        const int a = hl[i], n = hn[i];
        if (n <= 0 || n % kLabelAlign != 0 || a + n > D)
          fail(R3DG_ERROR_INVALID_ARGUMENT, "stereo: invalid label range " + std::to_string(a) + " + " +
                                                std::to_string(n) + " for " + std::to_string(D) + " labels");
        total += std::uint64_t(n);
      }
      if (total != out_cells)
        fail(R3DG_ERROR_INVALID_ARGUMENT, "cost_volume: out_cells " + std::to_string(out_cells) + " != " +
                                              std::to_string(total));
    } else if (out_cells != std::uint64_t(L.w) * std::uint64_t(L.h) * std::uint64_t(D)) {
      fail(R3DG_ERROR_INVALID_ARGUMENT, "cost_volume: out_cells does not match w * h * labels");
    }
    if (out == nullptr) fail(R3DG_ERROR_INVALID_ARGUMENT, "cost_volume: null output");
    // One whole-image tile (stereo_matcher.cpp:1269-1278).
    r3dg_options one = o;
    one.tile_size = std::max({L.w, L.h, 16});
    one.tile_margin = 0;
    one.tile_blend = 0;
    one.memory_budget = ~std::uint64_t{0};
    DBuf<std::int32_t> lrows = intervals(c, none, L.w, L.h), rrows = intervals(c, none, R.w, R.h);
    std::vector<Job> jobs(1);
    jobs[0].base = &L;
    jobs[0].partner = &R;
    jobs[0].cl = census(c, L);
    jobs[0].cr = census(c, R);
    jobs[0].lrows = lrows.get();
    jobs[0].rrows = rrows.get();
    jobs[0].rows4 = pack_rows(c, lrows.get(), rrows.get(), L.h);
    jobs[0].ranges = has ? &rg : nullptr;
    jobs[0].dmin = disparity_min;
    jobs[0].D = D;
    Aggregation A = aggregate_jobs(c, jobs, one);
    if (A.total_cells != out_cells) fail(R3DG_ERROR_INTERNAL, "cost_volume: cell count mismatch");
    const hipMemcpyKind kind = out_memory == R3DG_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (out_cells > 0) hip_check(hipMemcpyAsync(out, A.S.get(), out_cells * 2, kind, c.stream), "download");
    sync(c);
  });
}

int32_t r3dg_match(r3dg_ctx* h, r3dg_image left, r3dg_image right, r3dg_image left_valid, r3dg_image right_valid,
                   int32_t disparity_min, int32_t num_disparities, const r3dg_options* options, r3dg_image disparity,
                   r3dg_image confidence, const r3dg_match_debug* debug, r3dg_match_info* info) {
  return with_ctx(h, [&](Context& c) {
    const auto t0 = std::chrono::steady_clock::now();
    r3dg_options tmp;
    const r3dg_options& o = opts_or_default(options, tmp);
    check_options(o);
    DImg<std::uint16_t> L = input_img<std::uint16_t>(c, left, "left");
    DImg<std::uint16_t> R = input_img<std::uint16_t>(c, right, "right");
    DImg<std::uint8_t> mL = input_img<std::uint8_t>(c, left_valid, "left_valid");
    DImg<std::uint8_t> mR = input_img<std::uint8_t>(c, right_valid, "right_valid");
    if (view_empty(disparity)) fail(R3DG_ERROR_INVALID_ARGUMENT, "match: empty disparity output");
    const bool want_debug = debug != nullptr && (!view_empty(debug->labels) || !view_empty(debug->lr_mask));
    c.agg_launches = 0;
    c.agg_bytes = 0;
    double stage_seconds[R3DG_STAGES] = {};
    c.collect(stage_seconds, R3DG_STAGES);  // drop marks of earlier calls
    std::fill(stage_seconds, stage_seconds + R3DG_STAGES, 0.0);
    MatchResult r = match_stereo(c, L, R, mL, mR, disparity_min, num_disparities, o, want_debug);
    const TilesOut& m = r.final.left;
    output_img(c, disparity, m.disp.p, m.disp.w, m.disp.h, m.disp.s, "disparity");
    if (!view_empty(confidence)) output_img(c, confidence, m.conf.p, m.conf.w, m.conf.h, m.conf.s, "confidence");
    if (want_debug) {
      if (!view_empty(debug->labels)) output_img(c, debug->labels, m.lab.p, m.lab.w, m.lab.h, m.lab.s, "labels");
      if (!view_empty(debug->lr_mask))
        output_img(c, debug->lr_mask, r.final.lr_mask.p, r.final.lr_mask.w, r.final.lr_mask.h, r.final.lr_mask.s,
                   "lr_mask");
    }
    sync(c);
    if (info != nullptr) {
      info->disparity_min = disparity_min;
      info->num_disparities = align16(num_disparities);
      info->levels = r.levels;
      info->tiles_level0 = r.tiles0;
      info->cells = r.cells;
      info->tile_pixels = r.tile_pixels;
      info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      info->aggregate_launches = c.agg_launches;
      info->aggregate_bytes = c.agg_bytes;
      c.collect(stage_seconds, R3DG_STAGES);
      for (int k = 0; k < R3DG_STAGES; ++k) info->stage_seconds[k] = stage_seconds[k];
    } else {
      c.collect(stage_seconds, R3DG_STAGES);
    }
  });
}

int32_t r3dg_remove_speckles(r3dg_ctx* h, r3dg_image disparity, int32_t min_size, float max_diff) {
  return with_ctx(h, [&](Context& c) {
    if (view_empty(disparity) || min_size <= 1) return;
    DImg<float> d = input_img<float>(c, disparity, "disparity");
    DImg<float> work = alloc_img<float>(c, d.w, d.h);
    hip_check(hipMemcpy2DAsync(work.p, work.s * 4, d.p, d.s * 4, std::size_t(d.w) * 4, std::size_t(d.h),
                               hipMemcpyDeviceToDevice, c.stream),
              "copy");
    speckles(c, work, min_size, max_diff);
    output_img(c, disparity, work.p, work.w, work.h, work.s, "disparity");
    sync(c);
  });
}

int32_t r3dg_median_filter_3x3(r3dg_ctx* h, r3dg_image disparity, r3dg_image out) {
  return with_ctx(h, [&](Context& c) {
    if (view_empty(disparity)) return;
    DImg<float> d = input_img<float>(c, disparity, "disparity");
    DImg<float> work = alloc_img<float>(c, d.w, d.h);
    hip_check(hipMemcpy2DAsync(work.p, work.s * 4, d.p, d.s * 4, std::size_t(d.w) * 4, std::size_t(d.h),
                               hipMemcpyDeviceToDevice, c.stream),
              "copy");
    DImg<float> m = median(c, work);
    output_img(c, out, m.p, m.w, m.h, m.s, "median output");
    sync(c);
  });
}

int32_t r3dg_ranges_from_coarse(r3dg_ctx* h, r3dg_image coarse, int32_t width, int32_t height, r3dg_image fine_valid,
                                r3dg_image partner_valid, int32_t partner_width, int32_t partner_sign,
                                int32_t disparity_min, int32_t num_disparities, const r3dg_options* options,
                                int32_t mirrored, r3dg_image lo, r3dg_image num) {
  return with_ctx(h, [&](Context& c) {
    r3dg_options tmp;
    const r3dg_options& o = opts_or_default(options, tmp);
    DImg<float> cin = input_img<float>(c, coarse, "coarse");
    DImg<float> cc = alloc_img<float>(c, std::max(cin.w, 1), std::max(cin.h, 1));
    if (!cin.empty())
      hip_check(hipMemcpy2DAsync(cc.p, cc.s * 4, cin.p, cin.s * 4, std::size_t(cin.w) * 4, std::size_t(cin.h),
                                 hipMemcpyDeviceToDevice, c.stream),
                "copy");
    cc.w = cin.w;
    cc.h = cin.h;
    cc.s = cin.w;
    DImg<std::uint8_t> fv = input_img<std::uint8_t>(c, fine_valid, "fine_valid");
    DImg<std::uint8_t> pv = input_img<std::uint8_t>(c, partner_valid, "partner_valid");
    if (!fv.empty() && (fv.w != width || fv.h != height))
      fail(R3DG_ERROR_INVALID_ARGUMENT, "ranges_from_coarse: fine mask size differs");
    if (!pv.empty() && (pv.w != partner_width || pv.h != height))
      fail(R3DG_ERROR_INVALID_ARGUMENT, "ranges_from_coarse: partner mask size differs");
    DBuf<std::int32_t> prow = intervals(c, pv, partner_width, height);
    Ranges r = ranges_from_coarse(c, cc, width, height, fv, prow.get(), disparity_min, num_disparities, o,
                                  partner_sign, mirrored != 0);
    output_img(c, lo, r.lo.p, r.lo.w, r.lo.h, r.lo.s, "lo");
    output_img(c, num, r.num.p, r.num.w, r.num.h, r.num.s, "num");
    sync(c);
  });
}

}  // extern "C"
