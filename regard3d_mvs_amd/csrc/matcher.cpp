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
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstring>
#include <deque>
#include <stdexcept>
#include <cctype>
#include <iterator>
#include <fstream>
#include <map>
#include <memory>
#include <optional>
#include <string>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

#include "kernels.hpp"
#include "r3dg.h"
#include "runtime.hpp"

namespace r3dg {
namespace {


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
// Host memory the library's runtime page-locked (r3dg_host_alloc): copied by
// DMA directly, without a staging copy.
bool host_pinned(const void* p) {
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

template <typename T>
DImg<T> input_img(Context& c, const r3dg_image& v, const char* name) {
  if (view_empty(v)) return DImg<T>{};
  if (v.stride < v.width) fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(name) + ": stride smaller than width");
  if (v.format != R3DG_FORMAT_NATIVE)
    fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(name) + ": this input takes no format conversion");
  if (v.memory == R3DG_DEVICE) {
    DImg<T> d;
    d.p = static_cast<T*>(v.data);
    d.w = v.width;
    d.h = v.height;
    d.s = v.stride;
    return d;
  }
  DImg<T> d = alloc_img<T>(c, v.width, v.height);
  const std::size_t row = std::size_t(v.width) * sizeof(T), bytes = row * std::size_t(v.height);
  const bool direct = bytes >= (std::size_t{1} << 16) && host_pinned(v.data);
  void* pin = !direct && bytes >= (std::size_t{1} << 20) ? c.pinned.get(bytes) : nullptr;
  if (pin != nullptr) {
    // Stage through pinned memory (released at the next stream sync).
    c.staged_blocks.push_back(pin);
    copy_rows(pin, row, v.data, std::size_t(v.stride) * sizeof(T), row, std::size_t(v.height));
    hip_check(hipMemcpyAsync(d.p, pin, bytes, hipMemcpyHostToDevice, c.stream), "upload");
    return d;
  }
  hip_check(hipMemcpy2DAsync(d.p, std::size_t(d.s) * sizeof(T), v.data, std::size_t(v.stride) * sizeof(T), row,
                             std::size_t(v.height), hipMemcpyHostToDevice, c.stream),
            "upload");
  return d;
}

// A uint16 image input: uint8 data (R3DG_FORMAT_U8X257) is transferred as
// bytes and widened to u8 * 257 on the device.
DImg<std::uint16_t> input_u16(Context& c, const r3dg_image& v, const char* name) {
  if (view_empty(v) || v.format == R3DG_FORMAT_NATIVE) return input_img<std::uint16_t>(c, v, name);
  if (v.format != R3DG_FORMAT_U8X257) fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(name) + ": unknown format");
  r3dg_image bytes = v;
  bytes.format = R3DG_FORMAT_NATIVE;
  const DImg<std::uint8_t> u = input_img<std::uint8_t>(c, bytes, name);
  DImg<std::uint16_t> d = alloc_img<std::uint16_t>(c, v.width, v.height);
  launch_widen_u8(u.p, u.s, u.w, u.h, d.p, d.s, c.stream);
  hip_check(hipGetLastError(), "kernel launch");
  return d;  // u's block returns to the arena; later users are ordered after the widening on the stream
}

template <typename T>
void output_img(Context& c, const r3dg_image& v, const T* src, int w, int h, std::int64_t ss, const char* name) {
  if (view_empty(v)) return;
  if (v.width != w || v.height != h)
    fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(name) + ": output size " + std::to_string(v.width) + "x" +
                                          std::to_string(v.height) + " differs from " + std::to_string(w) + "x" +
                                          std::to_string(h));
  if (v.stride < v.width) fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(name) + ": stride smaller than width");
  const std::size_t row = std::size_t(w) * sizeof(T), bytes = row * std::size_t(h);
  if (v.memory != R3DG_DEVICE && bytes >= (std::size_t{1} << 20) && !host_pinned(v.data)) {
    if (void* pin = c.pinned.get(bytes)) {
      // Into pinned memory; copied to the caller's buffer at the next sync.
      c.staged_blocks.push_back(pin);
      hip_check(hipMemcpy2DAsync(pin, row, src, std::size_t(ss) * sizeof(T), row, std::size_t(h),
                                 hipMemcpyDeviceToHost, c.stream),
                "download");
      c.staged_out.push_back({pin, v.data, row, std::size_t(h), std::size_t(v.stride) * sizeof(T)});
      return;
    }
  }
  const hipMemcpyKind kind = v.memory == R3DG_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  hip_check(hipMemcpy2DAsync(v.data, std::size_t(v.stride) * sizeof(T), src, std::size_t(ss) * sizeof(T), row,
                             std::size_t(h), kind, c.stream),
            "download");
}

// A float map to the caller's buffer on stream cs (a copy stream beside the
// context's stream): into a pinned block of the context's pool, recorded in
// outs / blocks for the host copy once cs is past it, or straight into the
// caller's device or page-locked buffer.
void download_map(Context& c, hipStream_t cs, const r3dg_image& v, const DImg<float>& m, const char* name,
                  std::vector<Context::StagedOut>& outs, std::vector<void*>& blocks) {
  if (view_empty(v)) return;
  if (v.width != m.w || v.height != m.h)
    fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(name) + ": output size " + std::to_string(v.width) + "x" +
                                          std::to_string(v.height) + " differs from " + std::to_string(m.w) + "x" +
                                          std::to_string(m.h));
  if (v.stride < v.width) fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(name) + ": stride smaller than width");
  const std::size_t row = std::size_t(m.w) * sizeof(float), rows = std::size_t(m.h), bytes = row * rows;
  if (v.memory != R3DG_DEVICE && bytes >= (std::size_t{1} << 20) && !host_pinned(v.data)) {
    if (void* pin = c.pinned.get(bytes)) {
      blocks.push_back(pin);
      hip_check(hipMemcpy2DAsync(pin, row, m.p, std::size_t(m.s) * sizeof(float), row, rows, hipMemcpyDeviceToHost,
                                 cs),
                "download");
      outs.push_back({pin, v.data, row, rows, std::size_t(v.stride) * sizeof(float)});
      return;
    }
  }
  const hipMemcpyKind kind = v.memory == R3DG_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  hip_check(hipMemcpy2DAsync(v.data, std::size_t(v.stride) * sizeof(float), m.p, std::size_t(m.s) * sizeof(float),
                             row, rows, kind, cs),
            "download");
}

// Side stream that carries the downloads of a batch (the path scans use the
// first two).
constexpr int kCopySide = 2;

// One batch group's maps on their way to the callers' buffers: the events
// before (context stream) and after (copy stream) the downloads, the pinned
// blocks to copy from, and the device maps the downloads read.
struct Outgoing {
  explicit Outgoing(Context& ctx) : c(&ctx) {}
  Outgoing(const Outgoing&) = delete;
  Outgoing& operator=(const Outgoing&) = delete;
  ~Outgoing() { release(); }
  // Waits for the downloads, then copies the staged maps out.
  void finish() {
    if (done != nullptr) hip_check(hipEventSynchronize(done), "hipEventSynchronize");
    for (const Context::StagedOut& o : outs) copy_rows(o.dst, o.dst_stride, o.pinned, o.row_bytes, o.row_bytes, o.rows);
    outs.clear();
    release();
  }
  // Blocks back to the pool, maps to the arena, events destroyed (after a
  // failure the context's streams are synchronised before either is reused).
  void release() {
    for (void* b : blocks) c->pinned.put(b);
    blocks.clear();
    maps.clear();
    if (done != nullptr) (void)hipEventDestroy(done);
    if (ready != nullptr) (void)hipEventDestroy(ready);
    done = ready = nullptr;
  }
  Context* c;
  hipEvent_t ready = nullptr, done = nullptr;
  std::vector<Context::StagedOut> outs;
  std::vector<void*> blocks;
  std::vector<DImg<float>> maps;
};
inline void swap(Outgoing& a, Outgoing& b) noexcept {
  std::swap(a.c, b.c);
  std::swap(a.ready, b.ready);
  std::swap(a.done, b.done);
  a.outs.swap(b.outs);
  a.blocks.swap(b.blocks);
  a.maps.swap(b.maps);
}

// Small host-to-device copy on the context's stream, staged through a pinned
// block (returned at the next sync), so the host vector may be released right
// away and the host does not wait for the device.
template <typename T>
void upload(Context& c, DBuf<T>& dst, const std::vector<T>& src) {
  dst = DBuf<T>(c.arena, std::max<std::size_t>(src.size(), 1));
  if (src.empty()) return;
  const std::size_t bytes = src.size() * sizeof(T);
  if (void* pin = c.pinned.get(bytes)) {
    c.staged_blocks.push_back(pin);
    std::memcpy(pin, src.data(), bytes);
    hip_check(hipMemcpyAsync(dst.get(), pin, bytes, hipMemcpyHostToDevice, c.stream), "upload");
    return;
  }
  hip_check(hipMemcpyAsync(dst.get(), src.data(), bytes, hipMemcpyHostToDevice, c.stream), "upload");
  hip_check(hipStreamSynchronize(c.stream), "hipStreamSynchronize");
}

// Several small host arrays uploaded with one copy: part k starts at a
// 256-byte aligned offset of one device block (one blit on the stream
// instead of one per array).
struct UploadPack {
  std::vector<unsigned char> host;
  template <typename T>
  std::size_t add(const std::vector<T>& v) {
    const std::size_t off = (host.size() + 255) / 256 * 256;
    host.resize(off + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(host.data() + off, v.data(), v.size() * sizeof(T));
    return off;
  }
  std::size_t zeros(std::size_t bytes) {
    const std::size_t off = (host.size() + 255) / 256 * 256;
    host.resize(off + bytes, 0);
    return off;
  }
};
// A typed pointer into a packed block (the block owns the memory).
template <typename T>
struct DView {
  T* p = nullptr;
  [[nodiscard]] T* get() const { return p; }
};
template <typename T>
DView<T> view(const DBuf<unsigned char>& b, std::size_t off) {
  return DView<T>{reinterpret_cast<T*>(b.get() + off)};
}

void sync(Context& c) { c.sync(); }
void check_launch() { hip_check(hipGetLastError(), "kernel launch"); }

// Cancellation point: the context's flag or the caller's poll function.
void poll_cancel(const Context& c) {
  if (c.cancel.load() != 0 || (c.cancel_poll != nullptr && c.cancel_poll(c.cancel_user) != 0))
    fail(R3DG_ERROR_CANCELLED, "cancelled");
}

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
// A device image with its pyramid, Census and mirrored copies per level,
// each computed on first use.  Single matches build one per input image; the
// serving loop keeps them across the pairs that share a view.  deque and map
// keep element addresses stable, so jobs may point into them.
struct View {
  DImg<std::uint16_t> img;
  DImg<std::uint8_t> mask;
  std::deque<DImg<std::uint16_t>> pyr;  // level l at pyr[l - 1]
  std::deque<DImg<std::uint8_t>> mpyr;
  std::map<int, DImg<std::uint64_t>> cen;

  const DImg<std::uint16_t>& image(Context& c, int l) {
    while (static_cast<int>(pyr.size()) < l) pyr.push_back(pyramid_u16(c, pyr.empty() ? img : pyr.back()));
    return l == 0 ? img : pyr[std::size_t(l - 1)];
  }
  const DImg<std::uint8_t>& valid(Context& c, int l) {
    if (mask.empty()) return mask;
    while (static_cast<int>(mpyr.size()) < l) mpyr.push_back(pyramid_u8(c, mpyr.empty() ? mask : mpyr.back(), true));
    return l == 0 ? mask : mpyr[std::size_t(l - 1)];
  }
  const DImg<std::uint64_t>& census_at(Context& c, int l) {
    auto it = cen.find(l);
    if (it == cen.end()) it = cen.emplace(l, census(c, image(c, l))).first;
    return it->second;
  }
  // Device bytes held (cache accounting).
  [[nodiscard]] std::uint64_t bytes() const {
    std::uint64_t b = 2 * img.pixels() + mask.pixels();
    for (const auto& x : pyr) b += 2 * x.pixels();
    for (const auto& x : mpyr) b += x.pixels();
    for (const auto& [l, x] : cen) b += 8 * x.pixels();
    return b;
  }
};

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
  bool mirrored = false;  // base / cl / cr are the unmirrored images (JobDesc::mirrored)
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

// Tile layout of a set of jobs: the tiles of all jobs share the layout pass,
// and chunks of consecutive tiles share the cost pass, the 8 path launches
// and the selection launch.
struct Aggregation {
  DBuf<unsigned char> meta;  // job descriptors (and the layout pass's tiles and row prefix)
  DView<JobDesc> djobs;
  std::vector<TileDesc> td;  // cell_base counts over all tiles
  DBuf<TileDesc> dtiles;
  std::vector<Grid> grids;
  std::vector<int> tile_base;
  std::vector<std::uint64_t> tcells;
  DBuf<std::uint64_t> bcorr, pix;  // bcorr: band corrections (TileDesc)
  std::uint64_t total_rows = 0, total_px = 0, total_cells = 0, max_tile_cells = 0;
  int max_w = 1, max_D = 16, max_num = 16;
};

// Aggregated volume S of tiles [t0, t1) of a layout; the chunk's tile
// descriptors count cells from its first tile.
struct Chunk {
  int t0 = 0, t1 = 0;
  DBuf<unsigned char> meta;  // tiles, row prefix, group prefixes, work-queue counters
  DView<TileDesc> dtiles;
  DView<std::uint64_t> drow_prefix;
  std::uint64_t rows = 0, cells = 0, px = 0;
  // Aggregated costs: S plus the slices of path scans that ran side by side
  // on other streams (the sum mod 2^16 of all of them).
  DBuf<std::uint16_t> S;
  std::vector<DBuf<std::uint16_t>> slices;
  std::vector<const std::uint16_t*> parts() const {
    std::vector<const std::uint16_t*> v{S.get()};
    for (const auto& b : slices) v.push_back(b.get());
    return v;
  }
};

// band_shift: log2 of the volume band width (TileDesc); 31 stores every tile
// row contiguously (the packed order of detail::sgm_cost_volume).
Aggregation layout_jobs(Context& c, std::vector<Job>& jobs, const r3dg_options& o, int band_shift = 3) {
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
    d.mirrored = J.mirrored ? 1 : 0;
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
  DView<TileDesc> dtiles;
  DView<std::uint64_t> drow_prefix;
  DBuf<std::uint64_t>& bcorr = A.bcorr;
  DBuf<std::uint64_t> bstart;
  DBuf<std::uint64_t>& pix = A.pix;
  DBuf<std::uint64_t> row_cells, tile_counts;  // tile_counts: cells [nt], then the max label counts (uint32)
  DBuf<std::uint32_t> row_max;
  std::uint64_t& total_rows = A.total_rows;
  std::uint64_t& total_px = A.total_px;
  std::vector<bool> fixed(static_cast<std::size_t>(nj), false);
  std::optional<StageTimer> layout_timer(std::in_place, c, R3DG_STAGE_LAYOUT);
  for (;;) {
    td.clear();
    row_prefix.clear();
    total_rows = 0;
    total_px = 0;
    std::uint64_t total_bands = 0;
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
          t.band_shift = band_shift;
          t.nbands = band_shift >= 30 ? 1 : (t.w + (1 << band_shift) - 1) >> band_shift;
          t.band_base = total_bands;
          row_prefix.push_back(total_rows);
          total_px += std::uint64_t(t.w) * std::uint64_t(t.h);
          total_rows += std::uint64_t(t.h);
          total_bands += std::uint64_t(t.nbands) * std::uint64_t(t.h);
          td.push_back(t);
        }
      }
    }
    tile_base[std::size_t(nj)] = static_cast<int>(td.size());
    const int nt = static_cast<int>(td.size());
    {
      UploadPack pk;
      const std::size_t oj = pk.add(jd), ot = pk.add(td), orp = pk.add(row_prefix);
      upload(c, A.meta, pk.host);
      A.djobs = view<JobDesc>(A.meta, oj);
      dtiles = view<TileDesc>(A.meta, ot);
      drow_prefix = view<std::uint64_t>(A.meta, orp);
    }
    pix = DBuf<std::uint64_t>(c.arena, total_px);
    row_cells = DBuf<std::uint64_t>(c.arena, total_rows);
    row_max = DBuf<std::uint32_t>(c.arena, total_rows);
    bstart = DBuf<std::uint64_t>(c.arena, total_bands);
    bcorr = DBuf<std::uint64_t>(c.arena, total_bands);
    tile_counts = DBuf<std::uint64_t>(c.arena, 2 * std::size_t(nt));
    launch_tile_row_scan(A.djobs.get(), dtiles.get(), drow_prefix.get(), nt, total_rows, pix.get(), row_cells.get(),
                         row_max.get(), bstart.get(), c.stream);
    check_launch();
    launch_tile_cells(dtiles.get(), nt, row_cells.get(), row_max.get(), bstart.get(), bcorr.get(), tile_counts.get(),
                      reinterpret_cast<std::uint32_t*>(tile_counts.get() + nt), c.stream);
    check_launch();
    // Cells and largest label count of every tile, in one download (into a
    // pinned block, so that the host waits in the stream synchronisation).
    std::vector<std::uint64_t> counts(2 * std::size_t(nt), 0);
    {
      const std::size_t cbytes = sizeof(std::uint64_t) * counts.size();
      struct Block {
        PinnedPool& pool;
        void* p;
        ~Block() {
          if (p != nullptr) pool.put(p);
        }
      } pin{c.pinned, c.pinned.get(cbytes)};
      hip_check(hipMemcpyAsync(pin.p != nullptr ? pin.p : counts.data(), tile_counts.get(), cbytes,
                               hipMemcpyDeviceToHost, c.stream),
                "download");
      sync(c);
      if (pin.p != nullptr) std::memcpy(counts.data(), pin.p, cbytes);
    }
    tcells.assign(counts.begin(), counts.begin() + nt);
    tmax.resize(std::size_t(nt));
    std::memcpy(tmax.data(), counts.data() + nt, sizeof(std::uint32_t) * nt);
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
  int& max_num = A.max_num;
  int& max_w = A.max_w;
  int& max_D = A.max_D;
  for (int t = 0; t < nt; ++t) {
    td[std::size_t(t)].cell_base = total_cells;
    A.max_tile_cells = std::max(A.max_tile_cells, tcells[std::size_t(t)]);
    td[std::size_t(t)].max_num = static_cast<std::int32_t>(tmax[std::size_t(t)]);
    total_cells += tcells[std::size_t(t)];
    max_num = std::max<int>(max_num, static_cast<int>(tmax[std::size_t(t)]));
    max_w = std::max(max_w, td[std::size_t(t)].w);
  }
  for (const Job& J : jobs) max_D = std::max(max_D, J.D);
  if (max_num > kMaxLabels)
    fail(R3DG_ERROR_INVALID_ARGUMENT, "stereo: more than " + std::to_string(kMaxLabels) + " labels per pixel are not supported on the GPU");
  // The path scans address a tile's volume with 32-bit byte offsets (2 B per
  // cell of S), so a tile holds fewer than 2^31 cells.
  if (A.max_tile_cells >= (std::uint64_t{1} << 31))
    fail(R3DG_ERROR_INVALID_ARGUMENT, "stereo: a tile of " + std::to_string(A.max_tile_cells) +
                                          " cells exceeds the 2^31 cells a tile volume may hold on the GPU; "
                                          "use a smaller tile_size or memory_budget");
  upload(c, A.dtiles, td);
  layout_timer.reset();
  return A;
}

// Costs and the 8 path scans of tiles [t0, t1) (their rows are consecutive).
// With `sum_slices` the slices of side-by-side scans are added into S.
Chunk aggregate_tiles(Context& c, Aggregation& A, int t0, int t1, bool sum_slices) {
  Chunk ch;
  ch.t0 = t0;
  ch.t1 = t1;
  const int nt = t1 - t0;
  std::vector<TileDesc> td(A.td.begin() + t0, A.td.begin() + t1);
  std::vector<std::uint64_t> row_prefix(static_cast<std::size_t>(nt));
  const std::uint64_t cell0 = td.front().cell_base, row0 = td.front().row_base;
  int max_num = 16;
  for (int t = 0; t < nt; ++t) {
    TileDesc& T = td[std::size_t(t)];
    T.cell_base -= cell0;
    row_prefix[std::size_t(t)] = T.row_base - row0;
    ch.cells += A.tcells[std::size_t(t0 + t)];
    ch.rows += std::uint64_t(T.h);
    ch.px += std::uint64_t(T.w) * std::uint64_t(T.h);
    max_num = std::max(max_num, T.max_num);
  }
  // Group prefixes of the 8 paths (kScanPerWave scanlines per wave).
  std::vector<std::uint64_t> gp(8 * (std::size_t(nt) + 1), 0);
  std::uint64_t max_groups = 0, sum_groups = 0;
  for (int p = 0; p < 8; ++p) {
    std::uint64_t* g = gp.data() + std::size_t(p) * (std::size_t(nt) + 1);
    for (int t = 0; t < nt; ++t)
      g[t + 1] = g[t] + std::uint64_t((path_scanlines(p, td[std::size_t(t)].w, td[std::size_t(t)].h) +
                                       kScanPerWave - 1) / kScanPerWave);
    max_groups = std::max(max_groups, g[nt]);
    sum_groups += g[nt];
  }
  // Tiles, row prefix, group prefixes and the 8 work-queue counters (zero)
  // in one upload.
  DView<std::uint64_t> dgp;
  DView<std::uint32_t> queues;
  {
    UploadPack pk;
    const std::size_t ot = pk.add(td), orp = pk.add(row_prefix), og = pk.add(gp), oq = pk.zeros(8 * sizeof(std::uint32_t));
    upload(c, ch.meta, pk.host);
    ch.dtiles = view<TileDesc>(ch.meta, ot);
    ch.drow_prefix = view<std::uint64_t>(ch.meta, orp);
    dgp = view<std::uint64_t>(ch.meta, og);
    queues = view<std::uint32_t>(ch.meta, oq);
  }
  ch.S = DBuf<std::uint16_t>(c.arena, ch.cells + kSPad);
  // Scratch rows of the persistent waves (pixels wider than the LDS rows) and
  // the work-queue counters of the 8 launches.
  // (a multiple of 4: launches use whole blocks of 4 waves, one slot each)
  const std::uint64_t resident = path_slots();
  const std::uint64_t slots = (std::min(resident, std::max<std::uint64_t>(max_groups, 1)) + 3) / 4 * 4;
  // The scans run side by side on 2 or 3 streams (3 where their groups
  // average less than the device holds), each stream summing its paths into a
  // slice of its own, when the slices and their scratch rows fit the budget:
  // a launch's last groups then overlap the next launch on another stream.
  // Not 4: a process gets 4 hardware queues and 4 such streams were traced
  // on 3 of them, so that one queue ran 4 launches in turn (measured: the
  // 1080p pair's aggregation 5.0 ms with 4 streams, 4.7 ms with 3).
  const std::uint64_t mean_groups = std::max<std::uint64_t>((sum_groups + 7) / 8, 1);
  int ways = static_cast<int>(std::clamp<std::uint64_t>((resident + mean_groups - 1) / mean_groups, 2, 3));
  const std::uint64_t slice_bytes =
      (ch.cells + kSPad) * sizeof(std::uint16_t) + path_scratch_elems(slots, max_num) * sizeof(std::uint16_t);
  const std::uint64_t free_bytes = c.arena.budget() > c.arena.live() ? c.arena.budget() - c.arena.live() : 0;
  while (ways > 1 && std::uint64_t(ways - 1) * slice_bytes + (ch.cells + kSPad) > free_bytes / 2) --ways;
  std::vector<DBuf<std::uint16_t>>& slices = ch.slices;
  for (int w = 1; w < ways; ++w) slices.emplace_back(c.arena, ch.cells + kSPad);
  DBuf<std::uint16_t> scratch(c.arena, std::size_t(ways) * path_scratch_elems(slots, max_num));

  // Matching costs of every cell (timed with the other cost preparation).
  DBuf<std::uint8_t> C(c.arena, ch.cells + kSPad);
  {
    const StageTimer timer(c, R3DG_STAGE_PREPARE);
    launch_cost(A.djobs.get(), ch.dtiles.get(), ch.drow_prefix.get(), nt, ch.rows, A.bcorr.get(), A.pix.get(),
                C.get(), c.stream);
    check_launch();
  }
  // Aggregation: 8 path launches over the chunk's tiles; path p runs on
  // stream p % ways and adds into that stream's slice.
  {
    const StageTimer timer(c, R3DG_STAGE_AGGREGATE);
    std::array<hipStream_t, 4> st{c.stream, nullptr, nullptr, nullptr};
    std::array<hipEvent_t, 4> ev{};
    if (ways > 1) {
      ev[0] = c.event();
      hip_check(hipEventRecord(ev[0], c.stream), "hipEventRecord");
      for (int w = 1; w < ways; ++w) {
        st[std::size_t(w)] = c.side_stream(w - 1);
        hip_check(hipStreamWaitEvent(st[std::size_t(w)], ev[0], 0), "hipStreamWaitEvent");
      }
    }
    for (int p = 0; p < 8; ++p) {
      const int w = p % ways;
      const std::size_t o8 = std::size_t(p) * (std::size_t(nt) + 1);
      std::uint16_t* S = w == 0 ? ch.S.get() : slices[std::size_t(w - 1)].get();
      launch_sgm_path(p, A.djobs.get(), ch.dtiles.get(), dgp.get() + o8, nt, gp[o8 + std::size_t(nt)],
                      A.bcorr.get(), A.pix.get(), C.get(), S, scratch.get() + std::size_t(w) * path_scratch_elems(slots, max_num),
                      slots, max_num, queues.get() + p, p >= ways, st[std::size_t(w)]);
      check_launch();
    }
    if (ways > 1) {
      for (int w = 1; w < ways; ++w) {
        ev[std::size_t(w)] = c.event();
        hip_check(hipEventRecord(ev[std::size_t(w)], st[std::size_t(w)]), "hipEventRecord");
        hip_check(hipStreamWaitEvent(c.stream, ev[std::size_t(w)], 0), "hipStreamWaitEvent");
      }
      if (sum_slices) {
        const std::vector<const std::uint16_t*> src = ch.parts();
        launch_sum_u16(ch.S.get(), src.data(), ways, ch.cells, c.stream);
        check_launch();
        slices.clear();
      }
      for (int w = 0; w < ways; ++w) c.event_pool.push_back(ev[std::size_t(w)]);
    }
  }
  c.agg_launches += 8;
  // Algorithmic bytes of the 8 launches: per cell C read 8 x 1, S written
  // 8 x 2 and read 8 - ways times x 2 (the first scan of each stream only
  // writes its slice); per tile pixel and path its record (8) and intensity (2).
  c.agg_bytes += std::uint64_t(40 - 2 * ways) * ch.cells + 80 * ch.px;
  return ch;
}

// Tile ranges whose volumes (S, costs, scratch) fit the device memory left
// in the context's budget, one chunk when everything fits.
std::vector<std::array<int, 2>> chunk_tiles(const Context& c, const Aggregation& A) {
  const std::uint64_t free_bytes = c.arena.budget() > c.arena.live() ? c.arena.budget() - c.arena.live() : 0;
  // 3 bytes per cell (S and C) with room for scratch rows and per-tile data;
  // at most 4 G cells, enough to fill the device, so that the volume blocks
  // are reused from chunk to chunk rather than freed and mapped again.
  const std::uint64_t cap = std::max(A.max_tile_cells, std::min(free_bytes / 4, std::uint64_t{4} << 30));
  std::vector<std::array<int, 2>> out;
  const int nt = static_cast<int>(A.td.size());
  int t0 = 0;
  std::uint64_t cells = 0;
  for (int t = 0; t < nt; ++t) {
    if (t > t0 && cells + A.tcells[std::size_t(t)] > cap) {
      out.push_back({t0, t});
      t0 = t;
      cells = 0;
    }
    cells += A.tcells[std::size_t(t)];
  }
  if (nt > 0) out.push_back({t0, nt});
  return out;
}

// match_tiles (stereo_matcher.cpp:613-771) for several jobs at once.
std::vector<TilesOut> match_tiles(Context& c, std::vector<Job>& jobs, const r3dg_options& o, SelectParams sp) {
  const int nj = static_cast<int>(jobs.size());
  Aggregation A = layout_jobs(c, jobs, o);
  const std::uint64_t total_px = A.total_px;
  bool any_labels = false;
  for (const Job& J : jobs) any_labels = any_labels || J.labels;
  DBuf<float> tdisp(c.arena, std::max<std::uint64_t>(total_px, 1)), tconf(c.arena, std::max<std::uint64_t>(total_px, 1));
  DBuf<std::int16_t> tlab;
  if (any_labels) tlab = DBuf<std::int16_t>(c.arena, std::max<std::uint64_t>(total_px, 1));
  for (const auto& [t0, t1] : chunk_tiles(c, A)) {
    poll_cancel(c);
    Chunk ch = aggregate_tiles(c, A, t0, t1, false);
    // Selection into per-tile maps.
    const StageTimer timer(c, R3DG_STAGE_SELECT);
    const std::vector<const std::uint16_t*> parts = ch.parts();
    launch_select(A.djobs.get(), ch.dtiles.get(), ch.drow_prefix.get(), t1 - t0, ch.rows, A.bcorr.get(),
                  A.pix.get(), parts.data(), static_cast<int>(parts.size()), sp, A.max_w, A.max_D, tdisp.get(),
                  tconf.get(), tlab.get(), c.stream);
    check_launch();
  }
  std::vector<TilesOut> out(static_cast<std::size_t>(nj));
  // Tile grids of every job (extents, cores) in one upload.
  std::vector<std::int32_t> grid_host;
  std::vector<std::size_t> grid_cols(static_cast<std::size_t>(nj)), grid_rows(static_cast<std::size_t>(nj));
  for (int j = 0; j < nj; ++j) {
    const Grid& g = A.grids[std::size_t(j)];
    grid_cols[std::size_t(j)] = grid_host.size();
    for (const auto& e : g.cols) grid_host.insert(grid_host.end(), e.begin(), e.end());
    grid_rows[std::size_t(j)] = grid_host.size();
    for (const auto& e : g.rows) grid_host.insert(grid_host.end(), e.begin(), e.end());
  }
  DBuf<std::int32_t> dgrids;
  upload(c, dgrids, grid_host);
  for (int j = 0; j < nj; ++j) {
    const Job& J = jobs[std::size_t(j)];
    const Grid& g = A.grids[std::size_t(j)];
    TilesOut& r = out[std::size_t(j)];
    r.disp = alloc_img<float>(c, J.base->w, J.base->h);
    if (J.confidence) r.conf = alloc_img<float>(c, J.base->w, J.base->h);
    if (J.labels) r.lab = alloc_img<std::int16_t>(c, J.base->w, J.base->h);
    GridDesc gd{};
    gd.cols = dgrids.get() + grid_cols[std::size_t(j)];
    gd.rows = dgrids.get() + grid_rows[std::size_t(j)];
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
  }
  // dgrids returns to the arena here; later uses of the same memory are
  // ordered after the blends on the context's stream.
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

// Jobs of one level of one pair and the buffers they point into
// (match_level, stereo_matcher.cpp:831-856).  Heap-allocated so that the
// jobs' pointers stay valid while the jobs of several pairs are matched
// together.
struct LevelPrep {
  LevelIn in{};
  bool full = false, need_right = false, want_debug = false;
  int speckle = 0;
  DBuf<std::int32_t> lrows, rrows, mrows_base, mrows_partner;
  Ranges lranges, rranges;
  std::vector<Job> jobs;
};

SelectParams select_params(const r3dg_options& o) {
  const bool check_lr = o.lr_max_diff >= 0;
  SelectParams sp{};
  sp.subpixel = o.subpixel;
  sp.uniqueness = o.uniqueness;
  sp.lr_max_diff = o.lr_max_diff;
  sp.diagonal_lr = check_lr && o.lr_full_match == 0 ? 1 : 0;
  return sp;
}

// Level l of the pair (Lv, Rv) with disparities [dmin, dmin + nd).
std::unique_ptr<LevelPrep> prepare_level(Context& c, View& Lv, View& Rv, int l, int dmin, int nd,
                                         const DImg<float>* coarse_left, const DImg<float>* coarse_right,
                                         bool need_right, int speckle, const r3dg_options& o, bool want_debug) {
  auto P = std::make_unique<LevelPrep>();
  LevelIn& in = P->in;
  in.L = &Lv.image(c, l);
  in.R = &Rv.image(c, l);
  in.mL = &Lv.valid(c, l);
  in.mR = &Rv.valid(c, l);
  in.dmin = dmin;
  in.nd = nd;
  check_inputs(*in.L, *in.R, *in.mL, *in.mR, in.nd);
  P->need_right = need_right;
  P->speckle = speckle;
  P->want_debug = want_debug;
  const int wl = in.L->w, wr = in.R->w, h = in.L->h;
  const int D = align16(in.nd);
  P->full = o.lr_max_diff >= 0 && o.lr_full_match != 0;

  const StageTimer prep_timer(c, R3DG_STAGE_PREPARE);
  // Row intervals of both images (and of the mirrored pair), one launch.
  P->lrows = DBuf<std::int32_t>(c.arena, 2 * std::size_t(h));
  P->rrows = DBuf<std::int32_t>(c.arena, 2 * std::size_t(h));
  DBuf<std::int32_t> rows4a(c.arena, 4 * std::size_t(h)), rows4b;
  if (P->full) {
    P->mrows_base = DBuf<std::int32_t>(c.arena, 2 * std::size_t(h));
    P->mrows_partner = DBuf<std::int32_t>(c.arena, 2 * std::size_t(h));
    rows4b = DBuf<std::int32_t>(c.arena, 4 * std::size_t(h));
  }
  launch_level_rows(in.mL->empty() ? nullptr : in.mL->p, in.mL->s, wl, in.mR->empty() ? nullptr : in.mR->p, in.mR->s,
                    wr, h, P->lrows.get(), P->rrows.get(), P->mrows_base.get(), P->mrows_partner.get(), rows4a.get(),
                    rows4b.get(), c.stream);
  check_launch();
  if (coarse_left != nullptr) {
    const StageTimer timer(c, R3DG_STAGE_RANGES);
    P->lranges = ranges_from_coarse(c, *coarse_left, wl, h, *in.mL, P->rrows.get(), in.dmin, in.nd, o, -1, false);
  }
  P->jobs.resize(1);
  Job& a = P->jobs[0];
  a.base = in.L;
  a.partner = in.R;
  a.cl = view_of(Lv.census_at(c, l));
  a.cr = view_of(Rv.census_at(c, l));
  a.lrows = P->lrows.get();
  a.rrows = P->rrows.get();
  a.rows4 = std::move(rows4a);
  a.ranges = coarse_left != nullptr ? &P->lranges : nullptr;
  a.dmin = in.dmin;
  a.D = D;
  a.confidence = true;
  a.labels = want_debug;
  if (P->full) {
    // Right-to-left match on the mirrored pair (stereo_matcher.cpp:842-856):
    // base = mirror(R), partner = mirror(L), disparity offset wr - wl.
    if (coarse_right != nullptr) {
      const StageTimer timer(c, R3DG_STAGE_RANGES);
      P->rranges = ranges_from_coarse(c, *coarse_right, wr, h, *in.mR, P->lrows.get(), in.dmin, in.nd, o, 1, true);
    }
    Job b;
    // The unmirrored images and their Census, read backwards
    // (JobDesc::mirrored); the partner image only gives its width.
    b.base = &Rv.image(c, l);
    b.partner = &Lv.image(c, l);
    b.cl = view_of(Rv.census_at(c, l));
    b.cr = view_of(Lv.census_at(c, l));
    b.mirrored = true;
    b.lrows = P->mrows_base.get();
    b.rrows = P->mrows_partner.get();
    b.rows4 = std::move(rows4b);
    b.ranges = coarse_right != nullptr ? &P->rranges : nullptr;
    b.dmin = in.dmin + wr - wl;
    b.D = D;
    b.confidence = false;
    P->jobs.push_back(std::move(b));
  }
  return P;
}

// LR checks and post filters of a matched level (stereo_matcher.cpp:857-886);
// maps holds the tile maps of P's jobs.
LevelOut finish_level(Context& c, LevelPrep& P, std::vector<TilesOut>& maps, const r3dg_options& o) {
  const StageTimer filter_timer(c, R3DG_STAGE_FILTER);
  const int wl = P.in.L->w, wr = P.in.R->w, h = P.in.L->h;
  LevelOut out;
  out.left = std::move(maps[0]);
  if (P.want_debug) {
    out.lr_mask = alloc_img<std::uint8_t>(c, wl, h);
    launch_fill_u8(out.lr_mask.p, out.lr_mask.pixels(), 0, c.stream);
    check_launch();
  }
  if (P.full) {
    out.left.cells += maps[1].cells;
    out.left.tile_pixels += maps[1].tile_pixels;
    out.right = alloc_img<float>(c, wr, h);
    launch_unmirror_right(maps[1].disp.p, wr, h, wl, out.right.p, c.stream);
    check_launch();
    DBuf<std::uint8_t> rbad;
    if (P.need_right) {
      // Both checks compare the unchecked maps (stereo_matcher.cpp:864-876).
      rbad = DBuf<std::uint8_t>(c.arena, out.right.pixels());
      launch_lr_check(out.right.p, wr, h, out.left.disp.p, wl, 1, o.lr_max_diff, rbad.get(), false, c.stream);
      check_launch();
    }
    launch_lr_check(out.left.disp.p, wl, h, out.right.p, wr, -1, o.lr_max_diff,
                    P.want_debug ? out.lr_mask.p : nullptr, true, c.stream);
    check_launch();
    if (P.need_right) {
      launch_apply_mask(out.right.p, out.right.pixels(), rbad.get(), c.stream);
      check_launch();
      post_filter(c, out.right, o, P.speckle);
    } else {
      out.right = DImg<float>{};
    }
  }
  post_filter(c, out.left.disp, o, P.speckle);
  return out;
}

struct MatchResult {
  LevelOut final;
  int levels = 0;
  std::uint64_t cells = 0;
  std::uint64_t tile_pixels = 0;
  int tiles0 = 0;
};

// One pair through the levels of match_stereo (stereo_matcher.cpp:1023-1095).
struct PairRun {
  View* lv = nullptr;
  View* rv = nullptr;
  int dmin = 0, nd = 0, levels = 0;
  bool lr_full = false, want_debug = false;
  int next = 0;  // next level to match (levels down to 0)
  LevelOut coarse;
  MatchResult res;
};

// Checks and the image pyramids of a pair.
void start_pair(Context& c, PairRun& P, const r3dg_options& o) {
  if (P.nd <= 0) fail(R3DG_ERROR_INVALID_ARGUMENT, "stereo: no disparities to search");
  check_inputs(P.lv->img, P.rv->img, P.lv->mask, P.rv->mask, P.nd);
  P.levels = o.levels >= 0 ? o.levels : automatic_levels(P.lv->img.w, P.lv->img.h, P.nd);
  P.lr_full = o.lr_max_diff >= 0 && o.lr_full_match != 0;
  P.res.levels = P.levels;
  P.next = P.levels;
  const StageTimer pyr_timer(c, R3DG_STAGE_PYRAMID);
  P.lv->image(c, P.levels);
  P.rv->image(c, P.levels);
  P.lv->valid(c, P.levels);
  P.rv->valid(c, P.levels);
}

std::unique_ptr<LevelPrep> prepare_pair_level(Context& c, PairRun& P, int l, const r3dg_options& o) {
  const int dmin = floor_shift(P.dmin, l);
  const int nd = align16(ceil_shift(P.dmin + P.nd - 1, l) - dmin + 1);
  const int speckle = o.speckle_size >> (2 * l);
  const bool top = l == P.levels;
  return prepare_level(c, *P.lv, *P.rv, l, dmin, nd, top ? nullptr : &P.coarse.left.disp,
                       top || !P.lr_full ? nullptr : &P.coarse.right, P.lr_full && l > 0, speckle, o,
                       l == 0 && P.want_debug);
}

void finish_pair_level(Context& c, PairRun& P, int l, LevelPrep& prep, std::vector<TilesOut>& maps,
                       const r3dg_options& o) {
  P.coarse = finish_level(c, prep, maps, o);
  P.res.cells += P.coarse.left.cells;
  P.res.tile_pixels += P.coarse.left.tile_pixels;
  P.next = l - 1;
  if (l == 0) {
    P.res.tiles0 = P.coarse.left.tiles;
    P.res.final = std::move(P.coarse);
  }
}

// One step of several pairs: the next level of each, their jobs matched
// together (one layout, shared cost, path and selection launches).
void match_step(Context& c, const std::vector<PairRun*>& part, const r3dg_options& o) {
  std::vector<std::unique_ptr<LevelPrep>> preps;
  std::vector<Job> jobs;
  for (PairRun* P : part) {
    preps.push_back(prepare_pair_level(c, *P, P->next, o));
    for (Job& J : preps.back()->jobs) jobs.push_back(std::move(J));
  }
  std::vector<TilesOut> maps = match_tiles(c, jobs, o, select_params(o));
  std::size_t k = 0;
  for (std::size_t q = 0; q < part.size(); ++q) {
    const std::size_t n = preps[q]->full ? 2 : 1;
    std::vector<TilesOut> mine;
    for (std::size_t j = 0; j < n; ++j) mine.push_back(std::move(maps[k + j]));
    k += n;
    finish_pair_level(c, *part[q], part[q]->next, *preps[q], mine, o);
  }
}

// match_stereo (stereo_matcher.cpp:1023-1095).
MatchResult match_stereo(Context& c, View& Lv, View& Rv, int dmin, int nd, const r3dg_options& o,
                         bool want_debug) {
  check_options(o);
  PairRun P;
  P.lv = &Lv;
  P.rv = &Rv;
  P.dmin = dmin;
  P.nd = nd;
  P.want_debug = want_debug;
  start_pair(c, P, o);
  while (P.next >= 0) {
    poll_cancel(c);
    match_step(c, {&P}, o);
  }
  return std::move(P.res);
}

// match_stereo of several pairs with their levels in lockstep, finest levels
// together: at each step the jobs of every pair that has the level share one
// layout, the cost passes, the path launches and the selection launches.
// Each pair's maps are the same as when it is matched alone (tiles, ranges and
// every per-pixel step depend only on the pair).
void match_batch(Context& c, const std::vector<PairRun*>& pairs, const r3dg_options& o) {
  int top = 0;
  for (PairRun* P : pairs) {
    start_pair(c, *P, o);
    top = std::max(top, P->levels);
  }
  for (int l = top; l >= 0; --l) {
    poll_cancel(c);
    std::vector<PairRun*> part;
    for (PairRun* P : pairs)
      if (P->next == l) part.push_back(P);
    if (!part.empty()) match_step(c, part, o);
  }
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
    // Leave the streams idle so the context stays usable (blocks the failed
    // call released go back to the arena only once no kernel uses them);
    // staged outputs of the failed call are dropped.
    c->sync_all();
    (void)hipGetLastError();
    c->drop_staged();
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

// ------------------------------------------------------------------ serving
using ViewCache = std::unordered_map<std::uint64_t, std::shared_ptr<View>>;

ViewCache& view_cache(Context& c) {
  if (!c.views) c.views = std::make_shared<ViewCache>();
  return *static_cast<ViewCache*>(c.views.get());
}

// Device bytes a pair is expected to need: per-pixel buffers (pyramids,
// Census, mirrored copies, maps) and its largest level's volumes (S and
// costs, 3 bytes per cell) with about 48 labels per pixel when ranges come
// from a coarser level.  Volumes beyond the memory left are chunked anyway.
// The part that cannot be chunked: images, pyramids, Census and maps.
// Sizes of a pair for its memory estimates.  The disparity output has the
// left image's size also where the left image comes from a cached view (the
// pair's own left image is then empty); a right image from a view is taken
// as wide as the left.
struct PairSize {
  std::uint64_t wl, wr, h;
};
PairSize pair_size(const r3dg_pair& q) {
  const std::uint64_t wl = std::uint64_t(std::max(q.disparity.width, 0));
  const std::uint64_t h = std::uint64_t(std::max(q.disparity.height, 0));
  const std::uint64_t wr = view_empty(q.right) ? wl : std::uint64_t(std::max(q.right.width, 0));
  return {wl, wr, h};
}

std::uint64_t pair_fixed_bytes(const r3dg_pair& q) {
  const PairSize z = pair_size(q);
  return 96 * (z.wl + z.wr) * z.h;
}

std::uint64_t pair_footprint(const r3dg_pair& q, const r3dg_options& o) {
  const PairSize z = pair_size(q);
  const std::uint64_t wl = z.wl, h = z.h;
  const int levels = o.levels >= 0 ? o.levels
                                   : automatic_levels(static_cast<int>(wl), static_cast<int>(h), q.num_disparities);
  const std::uint64_t labels = levels > 0 ? 48 : std::uint64_t(align16(std::max(q.num_disparities, 1)));
  return pair_fixed_bytes(q) + 3 * 12 * wl * h * labels / 5;
}

struct Request {
  r3dg_pair* pair = nullptr;
  std::uint64_t left_view = 0, right_view = 0;
  std::uint64_t footprint = 0;
  std::chrono::steady_clock::time_point submitted;
  int32_t status = R3DG_PENDING;
  // Worker-side state while admitted.
  std::shared_ptr<View> lv, rv;
  PairRun run;
  void* pre[2] = {nullptr, nullptr};  // left / right host image copied into a pinned block at admission
};

// Request loop of a context (r3dg_server_start).
class Server {
 public:
  Server(Context& c, const r3dg_options& o, std::uint64_t budget) : c_(c), o_(o), budget_(budget) {
    worker_ = std::thread([this] { loop(); });
  }
  ~Server() { stop(); }

  void stop() {
    {
      const std::lock_guard<std::mutex> lock(m_);
      stop_ = true;
    }
    cv_.notify_all();
    if (worker_.joinable()) worker_.join();
  }

  // A view put (host copies of the image and mask) or evict, applied by the
  // loop before it admits the requests submitted after it.
  struct ViewOp {
    std::uint64_t id = 0;
    bool evict = false;
    int w = 0, h = 0, mw = 0, mh = 0;
    std::vector<std::uint16_t> img;
    std::vector<std::uint8_t> img8;  // uint8 image (R3DG_FORMAT_U8X257), widened on the device
    std::vector<std::uint8_t> mask;
  };

  void view_op(ViewOp op) {
    {
      const std::lock_guard<std::mutex> lock(m_);
      if (stop_) fail(R3DG_ERROR_INVALID_ARGUMENT, "serve: the request loop is stopping");
      view_ops_.push_back(std::move(op));
    }
    cv_.notify_all();
  }

  std::uint64_t submit(r3dg_pair* pair, std::uint64_t lv, std::uint64_t rv) {
    auto r = std::make_unique<Request>();
    r->pair = pair;
    r->left_view = lv;
    r->right_view = rv;
    r->footprint = pair_footprint(*pair, o_);
    // Refused at admission: a pair whose images, pyramids, Census and maps
    // alone exceed the context's device budget can never run (its volumes
    // could be chunked, those cannot).
    const std::uint64_t fixed = pair_fixed_bytes(*pair);
    if (fixed > c_.arena.budget())
      fail(R3DG_ERROR_BUDGET, "serve: the pair needs at least " + std::to_string(fixed >> 20) +
                                  " MiB of device memory, more than the context's budget of " +
                                  std::to_string(c_.arena.budget() >> 20) + " MiB");
    r->submitted = std::chrono::steady_clock::now();
    pair->status = R3DG_PENDING;
    std::uint64_t t;
    {
      const std::lock_guard<std::mutex> lock(m_);
      if (stop_) fail(R3DG_ERROR_INVALID_ARGUMENT, "serve: the request loop is stopping");
      t = next_++;
      pending_.push_back(t);
      reqs_.emplace(t, std::move(r));
    }
    cv_.notify_all();
    return t;
  }

  void inject_step_failure(std::int64_t after) { step_fail_after_.store(after < 0 ? -1 : after); }

  r3dg_server_stats stats() {
    const std::lock_guard<std::mutex> lock(m_);
    r3dg_server_stats s{};
    s.submitted = next_ - 1;
    s.completed = completed_;
    s.failed = failed_;
    s.steps = steps_.load(std::memory_order_relaxed);
    s.requeued = requeued_;
    s.pending = pending_.size();
    s.active = active_n_.load(std::memory_order_relaxed);
    s.in_use = in_use_;
    s.view_bytes = view_bytes_.load(std::memory_order_relaxed);
    s.admission_scale = scale_.load(std::memory_order_relaxed);
    s.loop_seconds = loop_s_.load(std::memory_order_relaxed);
    s.wait_seconds = wait_s_.load(std::memory_order_relaxed);
    return s;
  }

  // R3DG_PENDING, or the finished pair's status (the ticket is released).
  int32_t wait(std::uint64_t t, double timeout) {
    std::unique_lock<std::mutex> lock(m_);
    auto it = reqs_.find(t);
    if (it == reqs_.end()) fail(R3DG_ERROR_NOT_FOUND, "serve: unknown ticket " + std::to_string(t));
    auto ready = [&] { return it->second->status != R3DG_PENDING; };
    if (timeout < 0) {
      done_cv_.wait(lock, ready);
    } else if (timeout > 0) {
      done_cv_.wait_for(lock, std::chrono::duration<double>(timeout), ready);
    }
    if (!ready()) return R3DG_PENDING;
    const int32_t st = it->second->status;
    reqs_.erase(it);
    return st;
  }

 private:
  // Finished pairs handed to the copier: the events before (on the stream)
  // and after (on the copy stream) their downloads, the staged maps and
  // blocks, and the device maps the downloads read.
  struct Publish {
    Request* r = nullptr;
    hipEvent_t ready = nullptr, done = nullptr;
    std::vector<Context::StagedOut> outs;
    std::vector<void*> blocks;
    DImg<float> disp, conf;
  };

  // A finished pair's device state and admission share released (loop thread).
  void release(Request* r) {
    drop_prestaged(r);
    r->run = PairRun{};
    r->lv.reset();
    r->rv.reset();
    {
      const std::lock_guard<std::mutex> lock(m_);
      in_use_ -= r->footprint;
    }
    hold_ = false;  // memory was returned: admission may resume
  }
  // The pair's status made visible to its waiter, once its maps are in place.
  void publish(Request* r, int32_t status) {
    {
      const std::lock_guard<std::mutex> lock(m_);
      r->pair->info.seconds =
          std::chrono::duration<double>(std::chrono::steady_clock::now() - r->submitted).count();
      r->pair->status = status;
      r->status = status;
      ++completed_;
      if (status != R3DG_OK) ++failed_;
    }
    done_cv_.notify_all();
  }
  void complete(Request* r, int32_t status) {
    release(r);
    publish(r, status);
  }

  // Finished pairs' maps leave on the copy stream while the loop runs its
  // next steps.  The copier waits for a pair's downloads, copies the maps it
  // staged in pinned blocks to the caller's buffers, publishes the pair, and
  // hands the blocks and the device maps back to the loop thread (returned_,
  // retired_), which owns the pool and the arena.
  void copier() {
    for (;;) {
      Publish p;
      {
        std::unique_lock<std::mutex> lock(pm_);
        pcv_.wait(lock, [&] { return copier_stop_ || !publish_.empty(); });
        if (publish_.empty()) return;
        p = std::move(publish_.front());
        publish_.pop_front();
      }
      int32_t status = R3DG_OK;
      if (p.done != nullptr && hipEventSynchronize(p.done) != hipSuccess) status = R3DG_ERROR_DEVICE;
      destroy_events(p);
      if (status == R3DG_OK)
        for (const Context::StagedOut& o : p.outs)
          copy_rows(o.dst, o.dst_stride, o.pinned, o.row_bytes, o.row_bytes, o.rows);
      Request* r = p.r;
      {
        const std::lock_guard<std::mutex> lock(pm_);
        returned_.insert(returned_.end(), p.blocks.begin(), p.blocks.end());
        p.blocks.clear();
        retired_.push_back(std::move(p));
      }
      retired_n_.fetch_add(1, std::memory_order_release);
      cv_.notify_all();  // the loop frees the maps
      publish(r, status);
    }
  }
  static void destroy_events(Publish& p) {
    if (p.done != nullptr) (void)hipEventDestroy(p.done);
    if (p.ready != nullptr) (void)hipEventDestroy(p.ready);
    p.done = p.ready = nullptr;
  }
  // Downloads of a finished pair's map on the copy stream (pinned blocks are
  // copied to the caller's buffer by the copier).
  void download(Publish& p, const r3dg_image& v, const DImg<float>& m, const char* name) {
    download_map(c_, cs_, v, m, name, p.outs, p.blocks);
  }
  // Host rows to device memory through a pinned block of the context's pool
  // (one transfer at full speed instead of the runtime's pageable chunks, and
  // the host rows are free on return); the block returns to the pool at the
  // next wait on the stream.
  void staged_upload(void* dst, const void* src, std::size_t row_bytes, std::size_t rows) {
    const std::size_t bytes = row_bytes * rows;
    void* pin = c_.pinned.get(bytes);
    if (pin == nullptr) fail(R3DG_ERROR_OUT_OF_MEMORY, "view: no page-locked staging memory");
    c_.staged_blocks.push_back(pin);
    copy_rows(pin, row_bytes, src, row_bytes, row_bytes, rows);
    hip_check(hipMemcpyAsync(dst, pin, bytes, hipMemcpyHostToDevice, c_.stream), "upload");
  }
  // Pinned blocks the copier is done with, back into the context's pool
  // (the loop thread, holding the context lock).
  void take_back_blocks() {
    std::vector<void*> back;
    std::vector<Publish> done;
    {
      const std::lock_guard<std::mutex> lock(pm_);
      back.swap(returned_);
      done.swap(retired_);
      retired_n_.store(0, std::memory_order_relaxed);
    }
    for (void* b : back) c_.pinned.put(b);
    done.clear();  // device maps back to the arena
  }

  void loop() {
    (void)hipSetDevice(c_.device);
    if (hipStreamCreateWithFlags(&cs_, hipStreamNonBlocking) != hipSuccess) cs_ = nullptr;
    std::thread cp([this] { copier(); });
    run_steps();
    {
      const std::lock_guard<std::mutex> lock(pm_);
      copier_stop_ = true;
    }
    pcv_.notify_all();
    cp.join();
    const std::lock_guard<std::mutex> ctx_lock(c_.mutex);
    take_back_blocks();
    (void)guarded(&c_, [&] { sync(c_); });  // the last steps' staging blocks back to the pool
    if (cs_ != nullptr) (void)hipStreamDestroy(cs_);
    cs_ = nullptr;
  }

  std::shared_ptr<View> view_for(std::uint64_t id, const r3dg_image& img, const r3dg_image& mask, const char* name,
                                 void* pre) {
    if (id != 0) {
      ViewCache& vc = view_cache(c_);
      auto it = vc.find(id);
      if (it == vc.end()) fail(R3DG_ERROR_NOT_FOUND, std::string("serve: unknown ") + name + " view " + std::to_string(id));
      return it->second;
    }
    auto v = std::make_shared<View>();
    if (pre != nullptr) {
      // Already in a pinned block (prestage): one transfer, widened on the device for uint8.
      c_.staged_blocks.push_back(pre);
      if (img.format == R3DG_FORMAT_U8X257) {
        const DImg<std::uint8_t> u = alloc_img<std::uint8_t>(c_, img.width, img.height);
        hip_check(hipMemcpyAsync(u.p, pre, u.pixels(), hipMemcpyHostToDevice, c_.stream), "upload");
        v->img = alloc_img<std::uint16_t>(c_, img.width, img.height);
        launch_widen_u8(u.p, u.s, u.w, u.h, v->img.p, v->img.s, c_.stream);
        check_launch();
      } else {
        v->img = alloc_img<std::uint16_t>(c_, img.width, img.height);
        hip_check(hipMemcpyAsync(v->img.p, pre, v->img.pixels() * 2, hipMemcpyHostToDevice, c_.stream), "upload");
      }
    } else {
      v->img = input_u16(c_, img, name);
    }
    v->mask = input_img<std::uint8_t>(c_, mask, "valid");
    return v;
  }
  // The host images of the pairs about to be admitted, copied into pinned
  // blocks by several threads at once (one pair after the other, the copies
  // would leave the device idle between the pairs' uploads).
  void prestage(const std::vector<Request*>& admit) {
    struct Copy {
      void* pin;
      const void* src;
      std::size_t row, stride, rows;
    };
    std::vector<Copy> jobs;
    for (Request* r : admit) {
      const r3dg_pair& q = *r->pair;
      const r3dg_image* imgs[2] = {r->left_view == 0 ? &q.left : nullptr, r->right_view == 0 ? &q.right : nullptr};
      for (int k = 0; k < 2; ++k) {
        const r3dg_image* v = imgs[k];
        if (v == nullptr || view_empty(*v) || v->memory != R3DG_HOST || v->stride < v->width || v->data == nullptr)
          continue;
        if (v->format != R3DG_FORMAT_NATIVE && v->format != R3DG_FORMAT_U8X257) continue;
        const std::size_t es = v->format == R3DG_FORMAT_U8X257 ? 1 : 2;
        const std::size_t row = std::size_t(v->width) * es, rows = std::size_t(v->height);
        if (row * rows < (std::size_t{1} << 20) || host_pinned(v->data)) continue;
        void* pin = c_.pinned.get(row * rows);
        if (pin == nullptr) continue;
        r->pre[k] = pin;
        jobs.push_back({pin, v->data, row, std::size_t(v->stride) * es, rows});
      }
    }
    if (jobs.empty()) return;
    std::atomic<std::size_t> next{0};
    auto work = [&] {
      for (std::size_t j; (j = next.fetch_add(1)) < jobs.size();) {
        const Copy& J = jobs[j];
        for (std::size_t y = 0; y < J.rows; ++y)
          std::memcpy(static_cast<char*>(J.pin) + y * J.row, static_cast<const char*>(J.src) + y * J.stride, J.row);
      }
    };
    std::vector<std::thread> th;
    for (std::size_t k = 1; k < std::min<std::size_t>(8, jobs.size()); ++k) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
  }
  // Pinned blocks of a request's prestaged images that no upload took.
  void drop_prestaged(Request* r) {
    for (void*& b : r->pre) {
      if (b != nullptr) c_.pinned.put(b);
      b = nullptr;
    }
  }

  // Steps until the loop is stopped and has nothing left to run.
  void run_steps() {
    std::vector<Request*> active;
    for (;;) {
      std::vector<Request*> admit;
      std::deque<ViewOp> ops;
      {
        std::unique_lock<std::mutex> lock(m_);
        cv_.wait(lock, [&] {
          return stop_ || !pending_.empty() || !active.empty() || !view_ops_.empty() ||
                 retired_n_.load(std::memory_order_acquire) > 0;
        });
        if (stop_ && pending_.empty() && active.empty() && view_ops_.empty()) return;
        ops.swap(view_ops_);
        // Admission by estimated footprint against the budget left beside
        // the cached views (and below the context's own device budget); a
        // pair that fits nowhere runs alone.
        // After a memory shortage nothing new is admitted until a pair in
        // flight completes: admitting again at once would fail again, after
        // freeing the arena's cached blocks, at every step.
        const std::uint64_t cap = admission_cap();
        while (!pending_.empty() && !(hold_ && !active.empty())) {
          Request* r = reqs_.at(pending_.front()).get();
          if ((!active.empty() || !admit.empty()) && in_use_ + r->footprint > cap) break;
          in_use_ += r->footprint;
          admit.push_back(r);
          pending_.pop_front();
        }
      }
      const std::lock_guard<std::mutex> ctx_lock(c_.mutex);
      const auto busy0 = std::chrono::steady_clock::now();
      const double wait0 = c_.wait_seconds;
      // Time of this iteration, added when it ends (any of its exits).
      struct Clock {
        Server& s;
        std::chrono::steady_clock::time_point t0;
        double w0;
        ~Clock() {
          s.loop_s_.store(s.loop_s_.load(std::memory_order_relaxed) +
                              std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(),
                          std::memory_order_relaxed);
          s.wait_s_.store(s.wait_s_.load(std::memory_order_relaxed) + (s.c_.wait_seconds - w0),
                          std::memory_order_relaxed);
        }
      } clock{*this, busy0, wait0};
      // No stage timing inside the loop (nothing would collect the events);
      // direct calls, under the same lock, keep the context's setting.
      struct NoStageTiming {
        Context& c;
        int saved;
        ~NoStageTiming() { c.profiling = saved; }
      } no_timing{c_, std::exchange(c_.profiling, 0)};
      (void)hipSetDevice(c_.device);
      take_back_blocks();
      // View puts: their uploads queued on the stream from pinned copies (no
      // wait: the cache holds a view from the put on, and later kernels are
      // ordered after its upload).
      for (ViewOp& op : ops) {
        (void)guarded(&c_, [&] {
          ViewCache& vc = view_cache(c_);
          if (op.evict) {
            vc.erase(op.id);
            return;
          }
          auto v = std::make_shared<View>();
          if (!op.img8.empty()) {
            const DImg<std::uint8_t> u = alloc_img<std::uint8_t>(c_, op.w, op.h);
            staged_upload(u.p, op.img8.data(), std::size_t(op.w), std::size_t(op.h));
            v->img = alloc_img<std::uint16_t>(c_, op.w, op.h);
            launch_widen_u8(u.p, u.s, u.w, u.h, v->img.p, v->img.s, c_.stream);
            check_launch();
          } else {
            v->img = alloc_img<std::uint16_t>(c_, op.w, op.h);
            staged_upload(v->img.p, op.img.data(), std::size_t(op.w) * 2, std::size_t(op.h));
          }
          if (!op.mask.empty()) {
            v->mask = alloc_img<std::uint8_t>(c_, op.mw, op.mh);
            staged_upload(v->mask.p, op.mask.data(), std::size_t(op.mw), std::size_t(op.mh));
          }
          vc[op.id] = std::move(v);
        });
      }
      if (!ops.empty()) update_view_bytes();
      prestage(admit);
      std::vector<Request*> retry;  // admitted while memory was short: back to the queue
      for (Request* r : admit) {
        if (!retry.empty()) {
          retry.push_back(r);
          continue;
        }
        const int32_t st = guarded(&c_, [&] {
          r3dg_pair& q = *r->pair;
          std::memset(&q.info, 0, sizeof(q.info));
          if (view_empty(q.disparity)) fail(R3DG_ERROR_INVALID_ARGUMENT, "match: empty disparity output");
          // A prestaged block passes to its upload (back to the pool at the next wait).
          r->lv = view_for(r->left_view, q.left, q.left_valid, "left", std::exchange(r->pre[0], nullptr));
          r->rv = view_for(r->right_view, q.right, q.right_valid, "right", std::exchange(r->pre[1], nullptr));
          r->run = PairRun{};
          r->run.lv = r->lv.get();
          r->run.rv = r->rv.get();
          r->run.dmin = q.disparity_min;
          r->run.nd = q.num_disparities;
          start_pair(c_, r->run, o_);
        });
        if (st == R3DG_OK) {
          active.push_back(r);
        } else if (shortage(st) && !active.empty()) {
          retry.push_back(r);
        } else {
          complete(r, st);
        }
      }
      if (!retry.empty()) hold_ = true;
      requeue(retry);
      active_n_.store(active.size(), std::memory_order_relaxed);
      if (active.empty()) continue;
      std::vector<PairRun*> runs;
      for (Request* r : active) runs.push_back(&r->run);
      bool armed = false;  // test hook: a failure inside a step of several pairs
      if (active.size() >= 2) {
        const std::int64_t k = step_fail_after_.exchange(-1);
        if (k >= 0) {
          c_.arena.inject_failure(k);
          armed = true;
        }
      }
      // Finished pairs' maps are downloaded on the copy stream after the
      // step's kernels, overlapping the next steps; the copier completes them.
      std::vector<Publish> pubs;
      const int32_t st = guarded(&c_, [&] {
        match_step(c_, runs, o_);
        for (Request* r : active) {
          if (r->run.next >= 0) continue;
          if (cs_ == nullptr) fail(R3DG_ERROR_DEVICE, "serve: no copy stream");
          r3dg_pair& q = *r->pair;
          TilesOut& m = r->run.res.final.left;
          pubs.emplace_back();
          Publish& last = pubs.back();
          last.r = r;
          hip_check(hipEventCreateWithFlags(&last.ready, hipEventDisableTiming), "hipEventCreate");
          hip_check(hipEventRecord(last.ready, c_.stream), "hipEventRecord");
          hip_check(hipStreamWaitEvent(cs_, last.ready, 0), "hipStreamWaitEvent");
          download(last, q.disparity, m.disp, "disparity");
          if (!view_empty(q.confidence)) download(last, q.confidence, m.conf, "confidence");
          hip_check(hipEventCreateWithFlags(&last.done, hipEventDisableTiming), "hipEventCreate");
          hip_check(hipEventRecord(last.done, cs_), "hipEventRecord");
          last.disp = std::move(m.disp);  // alive until the copy stream is done with it
          last.conf = std::move(m.conf);
          q.info.disparity_min = q.disparity_min;
          q.info.num_disparities = align16(q.num_disparities);
          q.info.levels = r->run.res.levels;
          q.info.tiles_level0 = r->run.res.tiles0;
          q.info.cells = r->run.res.cells;
          q.info.tile_pixels = r->run.res.tile_pixels;
        }
        // No wait here: the next iteration's admission (host staging of new
        // pairs' images) overlaps this step's last kernels, and the next
        // step's layout waits for them.  Blocks released meanwhile are reused
        // in stream order.
      });
      if (armed) c_.arena.inject_failure(-1);  // one the step did not reach is dropped
      if (st != R3DG_OK) {
        // Nothing to publish: once the copy stream is idle the blocks and
        // maps go back, and the pairs complete (or run again) with the step's
        // status below.
        if (cs_ != nullptr) (void)hipStreamSynchronize(cs_);
        for (Publish& p : pubs) {
          for (void* b : p.blocks) c_.pinned.put(b);
          destroy_events(p);
        }
        pubs.clear();
      }
      steps_.fetch_add(1, std::memory_order_relaxed);
      if (shortage(st) && active.size() > 1) {
        // Device memory ran out with several pairs in flight: start them
        // again later, in their order, and admit less from now on.
        budget_scale_ *= 0.75;
        scale_.store(budget_scale_, std::memory_order_relaxed);
        requeue(active);
        active.clear();
        active_n_.store(0, std::memory_order_relaxed);
        update_view_bytes();
        continue;
      }
      std::vector<Request*> still;
      for (Request* r : active) {
        if (st != R3DG_OK) {
          complete(r, st);
        } else if (r->run.next < 0) {
          release(r);  // published by the copier
        } else {
          still.push_back(r);
        }
      }
      if (!pubs.empty()) {
        {
          const std::lock_guard<std::mutex> lock(pm_);
          for (Publish& p : pubs) publish_.push_back(std::move(p));
        }
        pcv_.notify_one();
      }
      active.swap(still);
      active_n_.store(active.size(), std::memory_order_relaxed);
      update_view_bytes();
    }
  }

  // Out of device memory: the arena's budget or the device itself.
  static bool shortage(int32_t st) { return st == R3DG_ERROR_BUDGET || st == R3DG_ERROR_OUT_OF_MEMORY; }

  // Device bytes the admission may promise to requests: the loop's budget,
  // at most the context's own, times the scale that memory shortages
  // lowered, minus the cached views.
  std::uint64_t admission_cap() const {
    const double cap = static_cast<double>(std::min(budget_, c_.arena.budget())) * budget_scale_;
    const double views = static_cast<double>(view_bytes_.load(std::memory_order_relaxed));
    return cap > views ? static_cast<std::uint64_t>(cap - views) : 0;
  }
  void update_view_bytes() {
    std::uint64_t b = 0;
    if (c_.views)
      for (const auto& [id, v] : view_cache(c_)) b += v->bytes();
    view_bytes_.store(b, std::memory_order_relaxed);
  }
  // Requests back to the front of the queue, their state released.
  void requeue(std::vector<Request*>& rs) {
    if (rs.empty()) return;
    const std::lock_guard<std::mutex> lock(m_);
    for (auto it = rs.rbegin(); it != rs.rend(); ++it) {
      Request* r = *it;
      drop_prestaged(r);
      r->run = PairRun{};
      r->lv.reset();
      r->rv.reset();
      in_use_ -= r->footprint;
      ++requeued_;
      for (const auto& [t, q] : reqs_) {
        if (q.get() == r) {
          pending_.push_front(t);
          break;
        }
      }
    }
    rs.clear();
  }

  Context& c_;
  const r3dg_options o_;
  const std::uint64_t budget_;
  double budget_scale_ = 1.0;                      // loop thread only
  bool hold_ = false;                              // loop thread only: no admission until a completion
  std::atomic<std::uint64_t> view_bytes_{0};       // cached views, updated by the loop
  std::atomic<double> scale_{1.0};                 // budget_scale_ for the counters
  std::atomic<double> loop_s_{0.0}, wait_s_{0.0};   // loop thread: time in steps, of it waiting for the device
  std::atomic<std::uint64_t> steps_{0}, active_n_{0};
  std::atomic<std::int64_t> step_fail_after_{-1};
  std::uint64_t completed_ = 0, failed_ = 0, requeued_ = 0;  // under m_
  std::mutex m_;
  std::condition_variable cv_, done_cv_;
  std::deque<std::uint64_t> pending_;
  std::deque<ViewOp> view_ops_;
  std::unordered_map<std::uint64_t, std::unique_ptr<Request>> reqs_;
  std::uint64_t next_ = 1, in_use_ = 0;
  bool stop_ = false;
  hipStream_t cs_ = nullptr;     // copy stream (created and destroyed by the loop thread)
  std::mutex pm_;
  std::condition_variable pcv_;
  std::deque<Publish> publish_;  // under pm_
  std::vector<void*> returned_;  // under pm_: blocks the copier is done with
  std::vector<Publish> retired_; // under pm_: published pairs whose device maps the loop frees
  std::atomic<int> retired_n_{0};
  bool copier_stop_ = false;     // under pm_
  std::thread worker_;
};

Server* server_of(Context& c) {
  if (!c.server) fail(R3DG_ERROR_INVALID_ARGUMENT, "serve: no request loop (r3dg_server_start)");
  return static_cast<Server*>(c.server.get());
}

}  // namespace
// ------------------------------------------------------------------ rectification
// rectify_image (rectify/rectification.cpp:288-350) for one channel of T.
template <typename T>
void rectify(Context& c, const r3dg_image& src, const r3dg_camera& cam, const r3dg_camera& tgt, const r3dg_image& out,
             const r3dg_image& valid) {
  if (src.width != cam.width || src.height != cam.height)
    fail(R3DG_ERROR_INVALID_ARGUMENT, "rectify_image: image is " + std::to_string(src.width) + "x" +
                                          std::to_string(src.height) + ", camera is " + std::to_string(cam.width) +
                                          "x" + std::to_string(cam.height));
  if (cam.model < R3DG_CAMERA_PINHOLE || cam.model > R3DG_CAMERA_FISHEYE)
    fail(R3DG_ERROR_INVALID_ARGUMENT, "rectify_image: unknown camera model " + std::to_string(cam.model));
  // Same centre (rectification.cpp:293-295).
  auto norm3 = [](double a, double b, double cc) { return std::sqrt(a * a + b * b + cc * cc); };
  const double dc = norm3(cam.center[0] - tgt.center[0], cam.center[1] - tgt.center[1], cam.center[2] - tgt.center[2]);
  if (dc > 1e-9 * (1.0 + norm3(cam.center[0], cam.center[1], cam.center[2])))
    fail(R3DG_ERROR_INVALID_ARGUMENT, "rectify_image: the target camera must have the same centre as the source");
  if (tgt.width <= 0 || tgt.height <= 0) fail(R3DG_ERROR_INVALID_ARGUMENT, "rectify_image: empty target camera");
  // Anti-aliasing level (rectification.cpp:298-312): the pyramid level that
  // keeps at least 1/1.5 of the target's resolution.
  int level = 0;
  while (cam.fx / tgt.fx > 1.5 * std::ldexp(1.0, level) && (cam.width >> (level + 1)) >= 16 &&
         (cam.height >> (level + 1)) >= 16)
    ++level;
  DImg<T> lv = input_img<T>(c, src, "rectify input");  // empty source: every target pixel is outside
  for (int l = 0; l < level && !lv.empty(); ++l) {
    const int dw = (lv.w + 1) / 2, dh = (lv.h + 1) / 2;
    DImg<T> next = alloc_img<T>(c, dw, dh);
    if constexpr (std::is_same_v<T, float>) {
      DBuf<float> tmp(c.arena, std::size_t(dw) * lv.h);
      launch_pyramid_f32(lv.p, lv.s, lv.w, lv.h, tmp.get(), next.p, next.s, c.stream);
    } else {
      DBuf<std::uint32_t> tmp(c.arena, std::size_t(dw) * lv.h);
      if constexpr (std::is_same_v<T, std::uint16_t>) {
        launch_pyramid_u16(lv.p, lv.s, lv.w, lv.h, tmp.get(), next.p, next.s, c.stream);
      } else {
        launch_pyramid_u8(lv.p, lv.s, lv.w, lv.h, tmp.get(), next.p, next.s, false, c.stream);
      }
    }
    check_launch();
    lv = std::move(next);
  }
  // pyramid_camera (rectification.cpp:152-164) and m = R_src * R_tgt^T.
  RectifyGeom g{};
  const double sc = std::ldexp(1.0, -level);
  g.fx = cam.fx * sc;
  g.fy = cam.fy * sc;
  g.cx = cam.cx * sc;
  g.cy = cam.cy * sc;
  for (int k = 0; k < 5; ++k) g.params[k] = cam.params[k];
  g.model = cam.model;
  g.sw = lv.w;
  g.sh = lv.h;
  g.tw = tgt.width;
  g.th = tgt.height;
  g.tfx = tgt.fx;
  g.tfy = tgt.fy;
  g.tcx = tgt.cx;
  g.tcy = tgt.cy;
  const double* R = cam.rotation;
  const double* Q = tgt.rotation;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) g.m[3 * i + j] = R[3 * i] * Q[3 * j] + R[3 * i + 1] * Q[3 * j + 1] + R[3 * i + 2] * Q[3 * j + 2];
  // Device outputs of the target's size are written in place; host ones go
  // through a device image and the staged download.
  auto direct = [&](const r3dg_image& im) {
    return im.memory == R3DG_DEVICE && im.width == tgt.width && im.height == tgt.height && im.stride >= im.width;
  };
  DImg<T> o;
  if (direct(out)) {
    o.p = static_cast<T*>(out.data);
    o.w = out.width;
    o.h = out.height;
    o.s = out.stride;
  } else {
    o = alloc_img<T>(c, tgt.width, tgt.height);
  }
  DImg<std::uint8_t> v;
  if (direct(valid)) {
    v.p = static_cast<std::uint8_t*>(valid.data);
    v.w = valid.width;
    v.h = valid.height;
    v.s = valid.stride;
  } else if (!view_empty(valid)) {
    v = alloc_img<std::uint8_t>(c, tgt.width, tgt.height);
  }
  if constexpr (std::is_same_v<T, float>) {
    launch_rectify_f32(lv.p, lv.s, g, o.p, o.s, v.p, v.s, c.stream);
  } else if constexpr (std::is_same_v<T, std::uint16_t>) {
    launch_rectify_u16(lv.p, lv.s, g, o.p, o.s, v.p, v.s, c.stream);
  } else {
    launch_rectify_u8(lv.p, lv.s, g, o.p, o.s, v.p, v.s, c.stream);
  }
  check_launch();
  if (!direct(out)) output_img(c, out, o.p, o.w, o.h, o.s, "rectify output");
  if (!view_empty(valid) && !direct(valid)) output_img(c, valid, v.p, v.w, v.h, v.s, "rectify valid");
  sync(c);
}

// ---- file formats (io/pfm.cpp, mvs/depth_map.cpp) -----------------------
std::vector<std::uint8_t> read_bytes(const char* path) {
  if (path == nullptr) fail(R3DG_ERROR_INVALID_ARGUMENT, "null path");
  std::ifstream f(path, std::ios::binary);
  if (!f) fail(R3DG_ERROR_NOT_FOUND, std::string("cannot open ") + path);
  return std::vector<std::uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

void write_bytes(const char* path, const std::vector<std::uint8_t>& data) {
  if (path == nullptr) fail(R3DG_ERROR_INVALID_ARGUMENT, "null path");
  std::ofstream f(path, std::ios::binary);
  if (!f) fail(R3DG_ERROR_INVALID_ARGUMENT, std::string("cannot write ") + path);
  f.write(reinterpret_cast<const char*>(data.data()), static_cast<std::streamsize>(data.size()));
  if (!f) fail(R3DG_ERROR_INTERNAL, std::string("error writing ") + path);
}

// PFM header (pfm.cpp:30-62): magic, width, height, scale; the data start
// one whitespace character after the scale.
struct PfmHeader {
  int width = 0, height = 0, channels = 0;
  bool little = true;
  std::size_t data = 0;
};

PfmHeader pfm_header(const std::vector<std::uint8_t>& d, const char* path) {
  std::size_t pos = 0;
  auto token = [&]() {
    while (pos < d.size() && std::isspace(d[pos]) != 0) ++pos;
    std::string t;
    while (pos < d.size() && std::isspace(d[pos]) == 0) t += static_cast<char>(d[pos++]);
    return t;
  };
  PfmHeader h;
  const std::string magic = token();
  if (magic == "Pf") {
    h.channels = 1;
  } else if (magic == "PF") {
    h.channels = 3;
  } else {
    fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(path) + ": not a PFM file");
  }
  double scale = 0.0;
  try {
    h.width = std::stoi(token());
    h.height = std::stoi(token());
    scale = std::stod(token());
  } catch (const std::exception&) {
    fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(path) + ": invalid PFM header");
  }
  ++pos;
  if (h.width <= 0 || h.height <= 0 || scale == 0.0) fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(path) + ": invalid PFM header");
  h.little = scale < 0.0;
  h.data = pos;
  const std::size_t bytes = std::size_t(h.width) * std::size_t(h.channels) * std::size_t(h.height) * 4;
  if (d.size() < pos + bytes) fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(path) + ": truncated PFM data");
  return h;
}

// .r3dd (depth_map.cpp:94-221), native little-endian PODs.
constexpr char kDepthMagic[4] = {'R', '3', 'D', 'D'};
constexpr std::uint32_t kDepthVersion = 2;

struct ByteReader {
  const std::vector<std::uint8_t>& d;
  std::size_t pos = 0;
  template <typename T>
  T pod() {
    if (pos + sizeof(T) > d.size()) throw std::out_of_range("truncated");
    T v;
    std::memcpy(&v, d.data() + pos, sizeof(T));
    pos += sizeof(T);
    return v;
  }
  void bytes(void* out, std::size_t n) {
    if (pos + n > d.size()) throw std::out_of_range("truncated");
    std::memcpy(out, d.data() + pos, n);
    pos += n;
  }
};

void depth_header(ByteReader& r, r3dg_depth_map& m, const char* path) {
  char magic[4];
  r.bytes(magic, 4);
  const auto version = r.pod<std::uint32_t>();
  if (std::memcmp(magic, kDepthMagic, 4) != 0 || version != kDepthVersion)
    fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(path) + " is not an r3d depth map (version 2)");
  m.view = r.pod<std::uint32_t>();
  m.camera_width = r.pod<std::int32_t>();
  m.camera_height = r.pod<std::int32_t>();
  m.fx = r.pod<double>();
  m.fy = r.pod<double>();
  m.cx = r.pod<double>();
  m.cy = r.pod<double>();
  for (double& v : m.rotation) v = r.pod<double>();
  for (double& v : m.center) v = r.pod<double>();
  m.num_partners = r.pod<std::uint32_t>();
  if (m.num_partners > 16) fail(R3DG_ERROR_INVALID_ARGUMENT, std::string("corrupt depth map ") + path);
  for (std::uint32_t i = 0; i < m.num_partners; ++i) m.partners[i] = r.pod<std::uint32_t>();
  m.width = r.pod<std::int32_t>();
  m.height = r.pod<std::int32_t>();
  if (m.width <= 0 || m.height <= 0 || m.width > (1 << 16) || m.height > (1 << 16))
    fail(R3DG_ERROR_INVALID_ARGUMENT, std::string("corrupt depth map ") + path);
}

}  // namespace r3dg

// ---- helpers of the row-kernel entry points
namespace r3dg {
namespace {

template <typename T>
DBuf<T> to_device(Context& c, const T* src, std::size_t n) {
  DBuf<T> d(c.arena, std::max<std::size_t>(n, 1));
  if (n > 0) hip_check(hipMemcpyAsync(d.get(), src, n * sizeof(T), hipMemcpyHostToDevice, c.stream), "upload");
  return d;
}

template <typename T>
void to_host(Context& c, T* dst, const DBuf<T>& src, std::size_t n) {
  if (n > 0) hip_check(hipMemcpyAsync(dst, src.get(), n * sizeof(T), hipMemcpyDeviceToHost, c.stream), "download");
}

// Checks a RowLabels of `width` pixels so that every kernel access stays in
// its buffers: offset[x] + num[x] <= offset[width], num[x] <= max_num (0: no
// bound).  Returns offset[width].
std::size_t check_labels(const r3dg_row_labels* L, std::size_t width, std::size_t max_num, const char* name) {
  if (width == 0) return 0;
  if (L == nullptr || L->lo == nullptr || L->num == nullptr || L->offset == nullptr)
    fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(name) + ": null labels");
  const std::size_t total = L->offset[width];
  for (std::size_t x = 0; x < width; ++x) {
    if (L->offset[x] > total || total - L->offset[x] < L->num[x])
      fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(name) + ": labels of pixel " + std::to_string(x) +
                                            " run past offset[width]");
    if (max_num != 0 && L->num[x] > max_num)
      fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(name) + ": pixel " + std::to_string(x) + " has more than max_num labels");
  }
  if (total > (std::size_t{1} << 40)) fail(R3DG_ERROR_INVALID_ARGUMENT, std::string(name) + ": row too large");
  return total;
}

DBuf<std::uint64_t> offsets_to_device(Context& c, const std::size_t* off, std::size_t width) {
  std::vector<std::uint64_t> v(off, off + width + 1);
  DBuf<std::uint64_t> d;
  upload(c, d, v);
  return d;
}

}  // namespace
}  // namespace r3dg

using namespace r3dg;

// The context's request loop, or nullptr (without waiting for a loop step).
static std::shared_ptr<Server> running_server(Context& c) {
  const std::lock_guard<std::mutex> lock(c.server_mutex);
  return std::static_pointer_cast<Server>(c.server);
}

extern "C" {

int32_t r3dg_abi_version(void) { return R3DG_ABI_VERSION; }

int32_t r3dg_host_alloc(size_t bytes, void** out) {
  return guarded(nullptr, [&] {
    if (out == nullptr) fail(R3DG_ERROR_INVALID_ARGUMENT, "host_alloc: null output");
    *out = nullptr;
    void* p = nullptr;
    const hipError_t e = hipHostMalloc(&p, std::max<size_t>(bytes, 1), hipHostMallocDefault);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      fail(R3DG_ERROR_OUT_OF_MEMORY, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    }
    *out = p;
  });
}

int32_t r3dg_host_free(void* p) {
  return guarded(nullptr, [&] {
    if (p != nullptr) hip_check(hipHostFree(p), "hipHostFree");
  });
}

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
    // Device address of this GPU's fixed scratch array (entry points that
    // allocate nothing), looked up once per device with the device current.
    // A failed lookup is reported by those entry points, not here.
    (void)static_scratch_init(device);
    *out = reinterpret_cast<r3dg_ctx*>(c.release());
  });
}

int32_t r3dg_ctx_destroy(r3dg_ctx* h) {
  if (h == nullptr) return R3DG_OK;
  auto* c = reinterpret_cast<Context*>(h);
  (void)r3dg_server_stop(h);
  return guarded(nullptr, [&] {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    (void)hipStreamSynchronize(c->stream);
    c->views.reset();
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

int32_t r3dg_inject_alloc_failure(r3dg_ctx* h, int64_t after) {
  return with_ctx(h, [&](Context& c) { c.arena.inject_failure(after < 0 ? -1 : after); });
}

int32_t r3dg_cancel(r3dg_ctx* h, int32_t requested) {
  if (h == nullptr) return R3DG_ERROR_INVALID_ARGUMENT;
  reinterpret_cast<Context*>(h)->cancel.store(requested != 0 ? 1 : 0);
  return R3DG_OK;
}

int32_t r3dg_set_cancel_poll(r3dg_ctx* h, int32_t (*poll)(void*), void* user) {
  return with_ctx(h, [&](Context& c) {
    c.cancel_poll = poll;
    c.cancel_user = user;
  });
}


int32_t r3dg_census_row(r3dg_ctx* h, const uint16_t* const rows[7], uint64_t* out, size_t n) {
  return with_ctx(h, [&](Context& c) {
    if (n == 0) return;
    if (rows == nullptr || out == nullptr) fail(R3DG_ERROR_INVALID_ARGUMENT, "census_row: null pointer");
    const std::size_t stride = n + 8;
    std::vector<std::uint16_t> packed(7 * stride);
    for (int j = 0; j < 7; ++j) {
      if (rows[j] == nullptr) fail(R3DG_ERROR_INVALID_ARGUMENT, "census_row: null row");
      std::memcpy(packed.data() + std::size_t(j) * stride, rows[j], stride * sizeof(std::uint16_t));
    }
    DBuf<std::uint16_t> drows;
    upload(c, drows, packed);
    DBuf<std::uint64_t> dout(c.arena, n);
    launch_census_row(drows.get(), n, dout.get(), c.stream);
    check_launch();
    to_host(c, out, dout, n);
    sync(c);
  });
}

int32_t r3dg_census_costs(r3dg_ctx* h, const uint64_t* left, const uint64_t* rev, size_t width,
                          const r3dg_row_labels* labels, int32_t rev_origin, uint16_t* cost) {
  return with_ctx(h, [&](Context& c) {
    if (width == 0) return;
    if (left == nullptr || rev == nullptr || cost == nullptr) fail(R3DG_ERROR_INVALID_ARGUMENT, "census_costs: null pointer");
    const std::size_t cells = check_labels(labels, width, 0, "census_costs");
    std::int64_t rmin = INT64_MAX, rmax = INT64_MIN;
    for (std::size_t x = 0; x < width; ++x) {
      if (labels->num[x] == 0) continue;
      const std::int64_t b = std::int64_t(width) - 1 - std::int64_t(x) + labels->lo[x] - rev_origin;
      rmin = std::min(rmin, b);
      rmax = std::max(rmax, b + labels->num[x]);
    }
    if (rmin > rmax) return;  // no labels
    const DBuf<std::uint64_t> dl = to_device(c, left, width);
    const DBuf<std::uint64_t> dr = to_device(c, rev + rmin, std::size_t(rmax - rmin));
    const DBuf<std::int32_t> dlo = to_device(c, labels->lo, width);
    const DBuf<std::uint16_t> dnum = to_device(c, labels->num, width);
    const DBuf<std::uint64_t> doff = offsets_to_device(c, labels->offset, width);
    DBuf<std::uint16_t> dcost(c.arena, std::max<std::size_t>(cells, 1));
    launch_census_costs(dl.get(), dr.get(), rmin, width, dlo.get(), dnum.get(), doff.get(), rev_origin, dcost.get(),
                        c.stream);
    check_launch();
    to_host(c, cost, dcost, cells);
    sync(c);
  });
}

int32_t r3dg_sgm_row(r3dg_ctx* h, const r3dg_sgm_row_args* a) {
  return with_ctx(h, [&](Context& c) {
    if (a == nullptr) fail(R3DG_ERROR_INVALID_ARGUMENT, "sgm_row: null arguments");
    const std::size_t w = a->width;
    if (w == 0) return;
    if (a->cost == nullptr || a->p2 == nullptr || a->cur == nullptr || a->cur_min == nullptr ||
        a->scratch == nullptr || a->sum == nullptr || (a->has_prev && (a->prev == nullptr || a->prev_min == nullptr)))
      fail(R3DG_ERROR_INVALID_ARGUMENT, "sgm_row: null buffer");
    if (a->max_num == 0 || a->max_num > 65535) fail(R3DG_ERROR_INVALID_ARGUMENT, "sgm_row: max_num out of range");
    const std::size_t cells = check_labels(&a->labels, w, a->max_num, "sgm_row");
    const std::size_t path = cells + w * 16;
    std::size_t pcells = 0, ppath = 0;
    if (a->has_prev) {
      pcells = check_labels(&a->prev_labels, w, a->max_num, "sgm_row (preceding row)");
      ppath = pcells + w * 16;
    }
    const std::size_t stride = a->max_num + 16;
    SgmRowDev d{};
    const DBuf<std::uint16_t> dcost = to_device(c, a->cost, cells), dp2 = to_device(c, a->p2, 4 * w),
                              dstart = to_device(c, static_cast<const std::uint16_t*>(a->scratch), stride);
    DBuf<std::uint16_t> dcur = to_device(c, static_cast<const std::uint16_t*>(a->cur), 3 * path);
    DBuf<std::uint16_t> dsum = a->accumulate ? to_device(c, static_cast<const std::uint16_t*>(a->sum), cells)
                                             : DBuf<std::uint16_t>(c.arena, std::max<std::size_t>(cells, 1));
    DBuf<std::uint16_t> dcur_min(c.arena, 3 * w), part(c.arena, std::max<std::size_t>(cells, 1)),
        hbuf(c.arena, 2 * (a->max_num + 2));
    const DBuf<std::int32_t> dlo = to_device(c, a->labels.lo, w);
    const DBuf<std::uint16_t> dnum = to_device(c, a->labels.num, w);
    const DBuf<std::uint64_t> doff = offsets_to_device(c, a->labels.offset, w);
    DBuf<std::uint16_t> dprev, dprev_min, dpnum;
    DBuf<std::int32_t> dplo;
    DBuf<std::uint64_t> dpoff;
    if (a->has_prev) {
      dprev = to_device(c, a->prev, 3 * ppath);
      dprev_min = to_device(c, a->prev_min, 3 * w);
      dplo = to_device(c, a->prev_labels.lo, w);
      dpnum = to_device(c, a->prev_labels.num, w);
      dpoff = offsets_to_device(c, a->prev_labels.offset, w);
    }
    d.cost = dcost.get();
    d.p2 = dp2.get();
    d.prev = dprev.get();
    d.prev_min = dprev_min.get();
    d.cur = dcur.get();
    d.cur_min = dcur_min.get();
    d.sum = dsum.get();
    d.start = dstart.get();
    d.lo = dlo.get();
    d.num = dnum.get();
    d.offset = doff.get();
    d.prev_lo = dplo.get();
    d.prev_num = dpnum.get();
    d.prev_offset = dpoff.get();
    d.width = w;
    d.max_num = a->max_num;
    d.p1 = a->p1;
    d.direction = a->direction > 0 ? 1 : -1;
    d.has_prev = a->has_prev ? 1 : 0;
    d.accumulate = a->accumulate ? 1 : 0;
    launch_sgm_row(d, part.get(), hbuf.get(), c.stream);
    check_launch();
    to_host(c, a->cur, dcur, 3 * path);
    to_host(c, a->cur_min, dcur_min, 3 * w);
    to_host(c, a->sum, dsum, cells);
    sync(c);
  });
}

int32_t r3dg_wta_row(r3dg_ctx* h, const uint16_t* sum, size_t width, const r3dg_row_labels* labels, uint16_t* best,
                     uint16_t* best_cost, uint16_t* second, uint16_t* right_cost, uint16_t* right_best,
                     ptrdiff_t right_shift) {
  return with_ctx(h, [&](Context& c) {
    if (width == 0) return;
    if (sum == nullptr || best == nullptr || best_cost == nullptr || second == nullptr)
      fail(R3DG_ERROR_INVALID_ARGUMENT, "wta_row: null pointer");
    if ((right_cost == nullptr) != (right_best == nullptr))
      fail(R3DG_ERROR_INVALID_ARGUMENT, "wta_row: right_cost and right_best go together");
    const std::size_t cells = check_labels(labels, width, 65535, "wta_row");
    for (std::size_t x = 0; x < width; ++x)
      if (labels->num[x] == 0) fail(R3DG_ERROR_INVALID_ARGUMENT, "wta_row: a pixel without labels");
    std::int64_t imin = INT64_MAX, imax = INT64_MIN;
    if (right_cost != nullptr) {
      for (std::size_t x = 0; x < width; ++x) {
        const std::int64_t top = std::int64_t(x) - labels->lo[x] + right_shift;
        imin = std::min(imin, top - (labels->num[x] - 1));
        imax = std::max(imax, top);
      }
      if (imin < 0) fail(R3DG_ERROR_INVALID_ARGUMENT, "wta_row: right index below 0 (right_shift too small)");
    }
    const DBuf<std::uint16_t> dsum = to_device(c, sum, cells);
    const DBuf<std::int32_t> dlo = to_device(c, labels->lo, width);
    const DBuf<std::uint16_t> dnum = to_device(c, labels->num, width);
    const DBuf<std::uint64_t> doff = offsets_to_device(c, labels->offset, width);
    DBuf<std::uint16_t> db(c.arena, width), dc(c.arena, width), ds(c.arena, width), drc, drb;
    std::size_t count = 0;
    if (right_cost != nullptr) {
      count = std::size_t(imax - imin + 1);
      drc = to_device(c, static_cast<const std::uint16_t*>(right_cost) + imin, count);
      drb = to_device(c, static_cast<const std::uint16_t*>(right_best) + imin, count);
    }
    launch_wta_row(dsum.get(), width, dlo.get(), dnum.get(), doff.get(), db.get(), dc.get(), ds.get(), right_shift,
                   imin, count, drc.get(), drb.get(), c.stream);
    check_launch();
    to_host(c, best, db, width);
    to_host(c, best_cost, dc, width);
    to_host(c, second, ds, width);
    if (count > 0) {
      to_host(c, right_cost + imin, drc, count);
      to_host(c, right_best + imin, drb, count);
    }
    sync(c);
  });
}

int32_t r3dg_automatic_levels(int32_t width, int32_t height, int32_t num_disparities) {
  return automatic_levels(width, height, num_disparities);
}

int32_t r3dg_pfm_header(const char* path, int32_t* width, int32_t* height, int32_t* channels) {
  return guarded(nullptr, [&] {
    const PfmHeader h = pfm_header(read_bytes(path), path);
    if (width != nullptr) *width = h.width;
    if (height != nullptr) *height = h.height;
    if (channels != nullptr) *channels = h.channels;
  });
}

int32_t r3dg_read_pfm(const char* path, float* out, int64_t stride) {
  return guarded(nullptr, [&] {
    const std::vector<std::uint8_t> d = read_bytes(path);
    const PfmHeader h = pfm_header(d, path);
    const std::size_t row = std::size_t(h.width) * std::size_t(h.channels);
    if (out == nullptr || stride < static_cast<int64_t>(row)) fail(R3DG_ERROR_INVALID_ARGUMENT, "read_pfm: output too small");
    const bool swap = !h.little;  // this library runs on little-endian hosts
    for (int y = 0; y < h.height; ++y) {
      const std::uint8_t* src = d.data() + h.data + std::size_t(h.height - 1 - y) * row * 4;
      float* dst = out + std::int64_t(y) * stride;
      for (std::size_t i = 0; i < row; ++i) {
        std::uint32_t u;
        std::memcpy(&u, src + 4 * i, 4);
        if (swap) u = ((u & 0xFFU) << 24) | ((u & 0xFF00U) << 8) | ((u >> 8) & 0xFF00U) | (u >> 24);
        std::memcpy(dst + i, &u, 4);
      }
    }
  });
}

int32_t r3dg_write_pfm(const char* path, const float* data, int32_t width, int32_t height, int32_t channels,
                       int64_t stride) {
  return guarded(nullptr, [&] {
    if (channels != 1 && channels != 3)
      fail(R3DG_ERROR_INVALID_ARGUMENT, "write_pfm: " + std::to_string(channels) + " channels (1 or 3 supported)");
    if (data == nullptr || width <= 0 || height <= 0 || stride < int64_t(width) * channels)
      fail(R3DG_ERROR_INVALID_ARGUMENT, "write_pfm: invalid image");
    const std::string header =
        std::string(channels == 1 ? "Pf" : "PF") + "\n" + std::to_string(width) + " " + std::to_string(height) + "\n-1\n";
    std::vector<std::uint8_t> out(header.begin(), header.end());
    const std::size_t row = std::size_t(width) * std::size_t(channels) * 4;
    out.reserve(out.size() + row * std::size_t(height));
    for (int y = height - 1; y >= 0; --y) {
      const auto* r = reinterpret_cast<const std::uint8_t*>(data + std::int64_t(y) * stride);
      out.insert(out.end(), r, r + row);
    }
    write_bytes(path, out);
  });
}

int32_t r3dg_read_depth_map_header(const char* path, r3dg_depth_map* map) {
  return guarded(nullptr, [&] {
    if (map == nullptr) fail(R3DG_ERROR_INVALID_ARGUMENT, "null depth map");
    const std::vector<std::uint8_t> d = read_bytes(path);
    ByteReader r{d};
    try {
      depth_header(r, *map, path);
    } catch (const std::out_of_range&) {
      fail(R3DG_ERROR_INVALID_ARGUMENT, std::string("truncated depth map ") + path);
    }
  });
}

int32_t r3dg_read_depth_map(const char* path, r3dg_depth_map* map) {
  return guarded(nullptr, [&] {
    if (map == nullptr || map->depth == nullptr || map->confidence == nullptr || map->pair_mask == nullptr ||
        map->support == nullptr || map->color == nullptr)
      fail(R3DG_ERROR_INVALID_ARGUMENT, "read_depth_map: missing image buffers");
    const std::vector<std::uint8_t> d = read_bytes(path);
    ByteReader r{d};
    r3dg_depth_map h = *map;
    try {
      depth_header(r, h, path);
      if (h.width != map->width || h.height != map->height)
        fail(R3DG_ERROR_INVALID_ARGUMENT, "read_depth_map: buffers sized for another grid");
      const std::size_t n = std::size_t(h.width) * std::size_t(h.height);
      r.bytes(map->depth, 4 * n);
      r.bytes(map->confidence, 4 * n);
      r.bytes(map->pair_mask, 2 * n);
      r.bytes(map->support, n);
      r.bytes(map->color, 3 * n);
    } catch (const std::out_of_range&) {
      fail(R3DG_ERROR_INVALID_ARGUMENT, std::string("truncated depth map ") + path);
    }
    h.depth = map->depth;
    h.confidence = map->confidence;
    h.pair_mask = map->pair_mask;
    h.support = map->support;
    h.color = map->color;
    *map = h;
  });
}

int32_t r3dg_write_depth_map(const char* path, const r3dg_depth_map* map) {
  return guarded(nullptr, [&] {
    if (map == nullptr || map->num_partners > 16 || map->width <= 0 || map->height <= 0 || map->depth == nullptr ||
        map->confidence == nullptr || map->pair_mask == nullptr || map->support == nullptr || map->color == nullptr)
      fail(R3DG_ERROR_INVALID_ARGUMENT, "write_depth_map: invalid depth map");
    std::vector<std::uint8_t> out;
    auto pod = [&](const auto& v) {
      const auto* b = reinterpret_cast<const std::uint8_t*>(&v);
      out.insert(out.end(), b, b + sizeof(v));
    };
    auto raw = [&](const void* p, std::size_t n) {
      const auto* b = static_cast<const std::uint8_t*>(p);
      out.insert(out.end(), b, b + n);
    };
    raw(kDepthMagic, 4);
    pod(kDepthVersion);
    pod(map->view);
    pod(map->camera_width);
    pod(map->camera_height);
    for (const double v : {map->fx, map->fy, map->cx, map->cy}) pod(v);
    for (const double v : map->rotation) pod(v);
    for (const double v : map->center) pod(v);
    pod(map->num_partners);
    for (std::uint32_t i = 0; i < map->num_partners; ++i) pod(map->partners[i]);
    pod(map->width);
    pod(map->height);
    const std::size_t n = std::size_t(map->width) * std::size_t(map->height);
    raw(map->depth, 4 * n);
    raw(map->confidence, 4 * n);
    raw(map->pair_mask, 2 * n);
    raw(map->support, n);
    raw(map->color, 3 * n);
    write_bytes(path, out);
  });
}

int32_t r3dg_combine_pairs(r3dg_ctx* h, const r3dg_grid* grid, const r3dg_pair_geometry* pairs, int32_t count,
                           const r3dg_multi_options* options, r3dg_image depth, r3dg_image confidence,
                           r3dg_image pair_mask) {
  return with_ctx(h, [&](Context& c) {
    if (grid == nullptr || options == nullptr) fail(R3DG_ERROR_INVALID_ARGUMENT, "combine_pairs: null grid or options");
    if (count < 0 || count > kMaxCombinePairs)
      fail(R3DG_ERROR_INVALID_ARGUMENT, "combine_pairs: at most " + std::to_string(kMaxCombinePairs) + " pairs per view (" +
                                            std::to_string(count) + " given)");
    if (count > 0 && pairs == nullptr) fail(R3DG_ERROR_INVALID_ARGUMENT, "combine_pairs: no pairs");
    const int w = grid->width, hh = grid->height;
    if (w < 0 || hh < 0) fail(R3DG_ERROR_INVALID_ARGUMENT, "combine_pairs: negative grid size");
    // The pairs' maps on the device (host maps are uploaded once).
    std::vector<DImg<float>> maps;
    maps.reserve(2 * std::size_t(count));
    std::vector<CombinePair> cp(static_cast<std::size_t>(count));
    for (int i = 0; i < count; ++i) {
      const r3dg_pair_geometry& q = pairs[i];
      if (view_empty(q.disparity)) fail(R3DG_ERROR_INVALID_ARGUMENT, "combine_pairs: empty disparity map");
      maps.push_back(input_img<float>(c, q.disparity, "combine disparity"));
      const DImg<float>& d = maps.back();
      CombinePair& P = cp[std::size_t(i)];
      for (int k = 0; k < 9; ++k) P.g[k] = q.grid_to_rect[k];
      P.fb = q.fb;
      P.offset = q.offset;
      for (int k = 0; k < 3; ++k) P.pc[k] = q.partner_center[k];
      P.disp = d.p;
      P.ds = d.s;
      P.dw = d.w;
      P.dh = d.h;
      if (!view_empty(q.confidence)) {
        maps.push_back(input_img<float>(c, q.confidence, "combine confidence"));
        const DImg<float>& cf = maps.back();
        P.conf = cf.p;
        P.cs = cf.s;
        P.cw = cf.w;
        P.ch = cf.h;
      }
    }
    DBuf<CombinePair> dpairs;
    upload(c, dpairs, cp);
    CombineGrid g{};
    for (int k = 0; k < 9; ++k) {
      g.inv_k[k] = grid->inv_k[k];
      g.rot_t[k] = grid->rot_t[k];
    }
    for (int k = 0; k < 3; ++k) g.center[k] = grid->center[k];
    g.inlier = options->inlier_px;
    g.max_res = options->max_reprojection_px;
    g.cos_min = options->cos_min_angle;
    g.min_pairs = std::max(1, options->min_pairs);  // multi_stereo.cpp:99
    g.npairs = count;
    g.w = w;
    g.h = hh;
    // Device outputs of the grid's size are written in place.
    auto direct = [&](const r3dg_image& im) {
      return im.memory == R3DG_DEVICE && im.width == w && im.height == hh && im.stride >= im.width;
    };
    auto target = [&](const r3dg_image& im, auto* type_tag) {
      using T = std::remove_pointer_t<decltype(type_tag)>;
      DImg<T> d;
      if (direct(im)) {
        d.p = static_cast<T*>(im.data);
        d.w = w;
        d.h = hh;
        d.s = im.stride;
      } else {
        d = alloc_img<T>(c, w, hh);
      }
      return d;
    };
    DImg<float> dep = target(depth, static_cast<float*>(nullptr)), cnf = target(confidence, static_cast<float*>(nullptr));
    DImg<std::uint16_t> msk = target(pair_mask, static_cast<std::uint16_t*>(nullptr));
    launch_combine(dpairs.get(), g, dep.p, dep.s, cnf.p, cnf.s, msk.p, msk.s, c.stream);
    check_launch();
    if (!direct(depth)) output_img(c, depth, dep.p, dep.w, dep.h, dep.s, "combine depth");
    if (!direct(confidence)) output_img(c, confidence, cnf.p, cnf.w, cnf.h, cnf.s, "combine confidence");
    if (!direct(pair_mask)) output_img(c, pair_mask, msk.p, msk.w, msk.h, msk.s, "combine pair mask");
    sync(c);
  });
}

int32_t r3dg_rectify(r3dg_ctx* h, int32_t dtype, r3dg_image src, const r3dg_camera* camera,
                     const r3dg_camera* target, r3dg_image out, r3dg_image valid) {
  return with_ctx(h, [&](Context& c) {
    if (camera == nullptr || target == nullptr) fail(R3DG_ERROR_INVALID_ARGUMENT, "rectify_image: null camera");
    switch (dtype) {
      case R3DG_U8: rectify<std::uint8_t>(c, src, *camera, *target, out, valid); break;
      case R3DG_U16: rectify<std::uint16_t>(c, src, *camera, *target, out, valid); break;
      case R3DG_F32: rectify<float>(c, src, *camera, *target, out, valid); break;
      default: fail(R3DG_ERROR_INVALID_ARGUMENT, "rectify_image: unknown element type " + std::to_string(dtype));
    }
  });
}

int32_t r3dg_census(r3dg_ctx* h, r3dg_image img, r3dg_image out) {
  return with_ctx(h, [&](Context& c) {
    if (view_empty(img)) return;
    DImg<std::uint16_t> d = input_u16(c, img, "census input");
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
    DImg<std::uint16_t> L = input_u16(c, left, "left");
    DImg<std::uint16_t> R = input_u16(c, right, "right");
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
    Aggregation A = layout_jobs(c, jobs, one, 31);  // rows contiguous: the reference's packed order
    if (A.total_cells != out_cells) fail(R3DG_ERROR_INTERNAL, "cost_volume: cell count mismatch");
    const Chunk ch = aggregate_tiles(c, A, 0, static_cast<int>(A.td.size()), true);
    const hipMemcpyKind kind = out_memory == R3DG_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (out_cells > 0) hip_check(hipMemcpyAsync(out, ch.S.get(), out_cells * 2, kind, c.stream), "download");
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
    View Lv, Rv;
    Lv.img = input_u16(c, left, "left");
    Rv.img = input_u16(c, right, "right");
    Lv.mask = input_img<std::uint8_t>(c, left_valid, "left_valid");
    Rv.mask = input_img<std::uint8_t>(c, right_valid, "right_valid");
    if (view_empty(disparity)) fail(R3DG_ERROR_INVALID_ARGUMENT, "match: empty disparity output");
    const bool want_debug = debug != nullptr && (!view_empty(debug->labels) || !view_empty(debug->lr_mask));
    c.agg_launches = 0;
    c.agg_bytes = 0;
    double stage_seconds[R3DG_STAGES] = {};
    c.collect(stage_seconds, R3DG_STAGES);  // drop marks of earlier calls
    std::fill(stage_seconds, stage_seconds + R3DG_STAGES, 0.0);
    MatchResult r = match_stereo(c, Lv, Rv, disparity_min, num_disparities, o, want_debug);
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

int32_t r3dg_match_batch(r3dg_ctx* h, r3dg_pair* pairs, int32_t count, const r3dg_options* options) {
  int32_t first_error = R3DG_OK;
  std::string first_message;
  const int32_t rc = with_ctx(h, [&](Context& c) {
    const auto t0 = std::chrono::steady_clock::now();
    r3dg_options tmp;
    const r3dg_options& o = opts_or_default(options, tmp);
    check_options(o);
    if (count < 0 || (count > 0 && pairs == nullptr)) fail(R3DG_ERROR_INVALID_ARGUMENT, "match_batch: no pairs");
    struct Inputs {
      View L, R;
      PairRun run;
    };
    std::vector<std::unique_ptr<Inputs>> in(static_cast<std::size_t>(count));
    std::vector<int> ok;
    for (int i = 0; i < count; ++i) {
      pairs[i].status = R3DG_OK;
      std::memset(&pairs[i].info, 0, sizeof(pairs[i].info));
      ok.push_back(i);
    }
    // A pair's images are uploaded (and checked) when its group is matched,
    // so that host staging memory is bounded by one group.
    auto load = [&](int i) -> bool {
      r3dg_pair& q = pairs[i];
      try {
        auto x = std::make_unique<Inputs>();
        x->L.img = input_u16(c, q.left, "left");
        x->R.img = input_u16(c, q.right, "right");
        x->L.mask = input_img<std::uint8_t>(c, q.left_valid, "left_valid");
        x->R.mask = input_img<std::uint8_t>(c, q.right_valid, "right_valid");
        if (view_empty(q.disparity)) fail(R3DG_ERROR_INVALID_ARGUMENT, "match: empty disparity output");
        if (q.num_disparities <= 0) fail(R3DG_ERROR_INVALID_ARGUMENT, "stereo: no disparities to search");
        check_inputs(x->L.img, x->R.img, x->L.mask, x->R.mask, q.num_disparities);
        PairRun& P = x->run;
        P.lv = &x->L;
        P.rv = &x->R;
        P.dmin = q.disparity_min;
        P.nd = q.num_disparities;
        in[std::size_t(i)] = std::move(x);
        return true;
      } catch (const Failure& e) {
        q.status = e.status;
        if (first_error == R3DG_OK) {
          first_error = e.status;
          first_message = e.what();
        }
        return false;
      }
    };
    c.agg_launches = 0;
    c.agg_bytes = 0;
    double stage_seconds[R3DG_STAGES] = {};
    c.collect(stage_seconds, R3DG_STAGES);  // drop marks of earlier calls
    std::fill(stage_seconds, stage_seconds + R3DG_STAGES, 0.0);
    // Groups whose per-pixel buffers (pyramids, Census, maps: about 96 bytes
    // per input pixel) take at most half of the memory left, and of at most
    // 256 M left pixels (enough to fill the device at every level but the
    // coarsest); the volumes are chunked to the rest by match_tiles.
    const std::uint64_t room = c.arena.budget() > c.arena.live() ? (c.arena.budget() - c.arena.live()) / 2 : 0;
    std::size_t g0 = 0;
    Outgoing prev(c);
    while (g0 < ok.size()) {
      std::size_t g1 = g0;
      std::uint64_t need = 0, px = 0;
      while (g1 < ok.size()) {
        const r3dg_pair& q = pairs[ok[g1]];
        const std::uint64_t est = 96 * (std::uint64_t(q.left.width) + std::uint64_t(q.right.width)) *
                                  std::uint64_t(q.left.height);
        const std::uint64_t p = std::uint64_t(q.left.width) * std::uint64_t(q.left.height);
        if (g1 > g0 && (need + est > room || px + p > (std::uint64_t{256} << 20))) break;
        need += est;
        px += p;
        ++g1;
      }
      std::vector<PairRun*> group;
      std::vector<int> members;
      for (std::size_t k = g0; k < g1; ++k) {
        if (!load(ok[k])) continue;
        group.push_back(&in[std::size_t(ok[k])]->run);
        members.push_back(ok[k]);
      }
      if (!group.empty()) match_batch(c, group, o);
      // The group's maps go down on a copy stream while the next group is
      // staged and matched; their host copies follow the next group's
      // launches (or the last group's).
      Outgoing cur(c);
      if (!members.empty()) {
        const hipStream_t cs = c.side_stream(kCopySide);
        hip_check(hipEventCreateWithFlags(&cur.ready, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventRecord(cur.ready, c.stream), "hipEventRecord");
        hip_check(hipStreamWaitEvent(cs, cur.ready, 0), "hipStreamWaitEvent");
        for (const int i : members) {
          r3dg_pair& q = pairs[i];
          Inputs& x = *in[std::size_t(i)];
          TilesOut& m = x.run.res.final.left;
          download_map(c, cs, q.disparity, m.disp, "disparity", cur.outs, cur.blocks);
          if (!view_empty(q.confidence)) download_map(c, cs, q.confidence, m.conf, "confidence", cur.outs, cur.blocks);
          cur.maps.push_back(std::move(m.disp));
          cur.maps.push_back(std::move(m.conf));
          q.info.disparity_min = q.disparity_min;
          q.info.num_disparities = align16(q.num_disparities);
          q.info.levels = x.run.res.levels;
          q.info.tiles_level0 = x.run.res.tiles0;
          q.info.cells = x.run.res.cells;
          q.info.tile_pixels = x.run.res.tile_pixels;
          in[std::size_t(i)].reset();
        }
        hip_check(hipEventCreateWithFlags(&cur.done, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventRecord(cur.done, cs), "hipEventRecord");
      }
      prev.finish();
      swap(prev, cur);
      g0 = g1;
    }
    prev.finish();
    sync(c);
    c.collect(stage_seconds, R3DG_STAGES);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (const int i : ok) {
      if (pairs[i].status != R3DG_OK) continue;
      pairs[i].info.seconds = secs;
      pairs[i].info.aggregate_launches = c.agg_launches;
      pairs[i].info.aggregate_bytes = c.agg_bytes;
      for (int k = 0; k < R3DG_STAGES; ++k) pairs[i].info.stage_seconds[k] = stage_seconds[k];
    }
    if (first_error != R3DG_OK) fail(first_error, first_message);
  });
  if (rc != R3DG_OK && first_error == R3DG_OK && pairs != nullptr) {
    for (int i = 0; i < count; ++i)
      if (pairs[i].status == R3DG_OK) pairs[i].status = rc;
  }
  return rc;
}

int32_t r3dg_server_start(r3dg_ctx* h, const r3dg_options* options, uint64_t hbm_budget) {
  return with_ctx(h, [&](Context& c) {
    r3dg_options tmp;
    const r3dg_options& o = opts_or_default(options, tmp);
    check_options(o);
    const std::lock_guard<std::mutex> lock(c.server_mutex);
    if (c.server) fail(R3DG_ERROR_INVALID_ARGUMENT, "serve: the request loop is already running");
    const std::uint64_t budget = hbm_budget != 0 ? hbm_budget : c.arena.budget() / 10 * 8;
    c.server = std::make_shared<Server>(c, o, budget);
  });
}

int32_t r3dg_server_stop(r3dg_ctx* h) {
  if (h == nullptr) return with_ctx(h, [](Context&) {});
  auto* c = reinterpret_cast<Context*>(h);
  std::shared_ptr<void> srv;
  {
    const std::lock_guard<std::mutex> lock(c->server_mutex);
    srv = std::move(c->server);
  }
  // The loop takes the context lock for its steps: stop it without holding it.
  if (srv) static_cast<Server*>(srv.get())->stop();
  return R3DG_OK;
}

int32_t r3dg_server_inject_step_failure(r3dg_ctx* h, int64_t after) {
  if (h == nullptr) return with_ctx(nullptr, [](Context&) {});
  auto* c = reinterpret_cast<Context*>(h);
  return guarded(nullptr, [&] {
    const std::shared_ptr<Server> srv = running_server(*c);
    if (!srv) fail(R3DG_ERROR_INVALID_ARGUMENT, "serve: no request loop (r3dg_server_start)");
    srv->inject_step_failure(after);
  });
}

int32_t r3dg_server_stats_get(r3dg_ctx* h, r3dg_server_stats* out) {
  if (h == nullptr || out == nullptr) return with_ctx(nullptr, [](Context&) {});
  auto* c = reinterpret_cast<Context*>(h);
  return guarded(nullptr, [&] {
    const std::shared_ptr<Server> srv = running_server(*c);
    if (!srv) fail(R3DG_ERROR_INVALID_ARGUMENT, "serve: no request loop (r3dg_server_start)");
    *out = srv->stats();
  });
}

int32_t r3dg_submit(r3dg_ctx* h, r3dg_pair* pair, uint64_t left_view, uint64_t right_view, uint64_t* ticket) {
  if (h == nullptr || pair == nullptr || ticket == nullptr) return with_ctx(nullptr, [](Context&) {});
  auto* c = reinterpret_cast<Context*>(h);
  return guarded(nullptr, [&] {
    const std::shared_ptr<Server> srv = running_server(*c);
    if (!srv) fail(R3DG_ERROR_INVALID_ARGUMENT, "serve: no request loop (r3dg_server_start)");
    *ticket = srv->submit(pair, left_view, right_view);
  });
}

namespace {
int32_t wait_ticket(r3dg_ctx* h, uint64_t ticket, double timeout) {
  if (h == nullptr) return with_ctx(nullptr, [](Context&) {});
  auto* c = reinterpret_cast<Context*>(h);
  int32_t st = R3DG_OK;
  const int32_t rc = guarded(nullptr, [&] {
    const std::shared_ptr<Server> srv = running_server(*c);
    if (!srv) fail(R3DG_ERROR_INVALID_ARGUMENT, "serve: no request loop (r3dg_server_start)");
    st = srv->wait(ticket, timeout);
  });
  return rc != R3DG_OK ? rc : st;
}
}  // namespace

int32_t r3dg_poll(r3dg_ctx* h, uint64_t ticket) { return wait_ticket(h, ticket, 0.0); }

int32_t r3dg_wait(r3dg_ctx* h, uint64_t ticket, double timeout_seconds) {
  return wait_ticket(h, ticket, timeout_seconds);
}

int32_t r3dg_view_put(r3dg_ctx* h, uint64_t view, r3dg_image image, r3dg_image valid) {
  if (h != nullptr && image.memory == R3DG_HOST && (view_empty(valid) || valid.memory == R3DG_HOST)) {
    auto* c = reinterpret_cast<Context*>(h);
    if (std::shared_ptr<Server> srv = running_server(*c)) {
      // Host images while the loop runs: copy them and let the loop upload
      // them between its steps, so the caller does not wait for a step.
      return guarded(nullptr, [&] {
        if (view == 0) fail(R3DG_ERROR_INVALID_ARGUMENT, "view: id 0 is reserved");
        if (view_empty(image) || image.stride < image.width) fail(R3DG_ERROR_INVALID_ARGUMENT, "view: bad image");
        if (image.format != R3DG_FORMAT_NATIVE && image.format != R3DG_FORMAT_U8X257)
          fail(R3DG_ERROR_INVALID_ARGUMENT, "view: unknown format");
        Server::ViewOp op;
        op.id = view;
        op.w = image.width;
        op.h = image.height;
        const std::size_t w = std::size_t(image.width), hh = std::size_t(image.height);
        if (image.format == R3DG_FORMAT_U8X257) {
          // uint8 rows, uploaded as bytes and widened to u8 * 257 on the device.
          op.img8.resize(w * hh);
          copy_rows(op.img8.data(), w, image.data, std::size_t(image.stride), w, hh);
        } else {
          op.img.resize(w * hh);
          copy_rows(op.img.data(), w * 2, image.data, std::size_t(image.stride) * 2, w * 2, hh);
        }
        if (!view_empty(valid)) {
          if (valid.width != image.width || valid.height != image.height || valid.stride < valid.width)
            fail(R3DG_ERROR_INVALID_ARGUMENT, "view: mask size differs from the image");
          op.mw = valid.width;
          op.mh = valid.height;
          op.mask.resize(std::size_t(valid.width) * std::size_t(valid.height));
          for (int y = 0; y < valid.height; ++y)
            std::memcpy(op.mask.data() + std::size_t(y) * std::size_t(valid.width),
                        static_cast<const std::uint8_t*>(valid.data) + std::size_t(y) * std::size_t(valid.stride),
                        std::size_t(valid.width));
        }
        srv->view_op(std::move(op));
      });
    }
  }
  return with_ctx(h, [&](Context& c) {
    if (view == 0) fail(R3DG_ERROR_INVALID_ARGUMENT, "view: id 0 is reserved");
    if (view_empty(image)) fail(R3DG_ERROR_INVALID_ARGUMENT, "view: empty image");
    auto v = std::make_shared<View>();
    v->img = input_u16(c, image, "view");
    v->mask = input_img<std::uint8_t>(c, valid, "view valid");
    if (!v->mask.empty() && (v->mask.w != v->img.w || v->mask.h != v->img.h))
      fail(R3DG_ERROR_INVALID_ARGUMENT, "view: mask size differs from the image");
    if (image.memory == R3DG_DEVICE) {
      // Own a copy: the caller's device memory may change after the call.
      DImg<std::uint16_t> own = alloc_img<std::uint16_t>(c, v->img.w, v->img.h);
      hip_check(hipMemcpy2DAsync(own.p, std::size_t(own.s) * 2, v->img.p, std::size_t(v->img.s) * 2,
                                 std::size_t(v->img.w) * 2, std::size_t(v->img.h), hipMemcpyDeviceToDevice, c.stream),
                "copy");
      v->img = std::move(own);
    }
    if (valid.memory == R3DG_DEVICE && !v->mask.empty()) {
      DImg<std::uint8_t> own = alloc_img<std::uint8_t>(c, v->mask.w, v->mask.h);
      hip_check(hipMemcpy2DAsync(own.p, std::size_t(own.s), v->mask.p, std::size_t(v->mask.s), std::size_t(v->mask.w),
                                 std::size_t(v->mask.h), hipMemcpyDeviceToDevice, c.stream),
                "copy");
      v->mask = std::move(own);
    }
    sync(c);
    view_cache(c)[view] = std::move(v);
  });
}

int32_t r3dg_view_evict(r3dg_ctx* h, uint64_t view) {
  if (h != nullptr) {
    auto* c = reinterpret_cast<Context*>(h);
    if (std::shared_ptr<Server> srv = running_server(*c)) {
      // In order with the puts queued before it.
      return guarded(nullptr, [&] {
        Server::ViewOp op;
        op.id = view;
        op.evict = true;
        srv->view_op(std::move(op));
      });
    }
  }
  return with_ctx(h, [&](Context& c) {
    if (view_cache(c).erase(view) == 0) fail(R3DG_ERROR_NOT_FOUND, "view: unknown id " + std::to_string(view));
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

static_assert(R3DG_STATIC_SPECKLE_PIXELS == kStaticSpecklePixels, "r3dg.h and the scratch array agree");

int32_t r3dg_remove_speckles_static(r3dg_ctx* h, r3dg_image disparity, int32_t min_size, float max_diff) {
  return with_ctx(h, [&](Context& c) {
    if (view_empty(disparity) || min_size <= 1) return;
    if (disparity.memory != R3DG_DEVICE)
      fail(R3DG_ERROR_INVALID_ARGUMENT, "remove_speckles_static: the map must be in device memory");
    if (disparity.format != R3DG_FORMAT_NATIVE)
      fail(R3DG_ERROR_INVALID_ARGUMENT, "remove_speckles_static: this input takes no format conversion");
    if (disparity.stride != disparity.width)
      fail(R3DG_ERROR_INVALID_ARGUMENT, "remove_speckles_static: the map must be contiguous (stride == width)");
    const std::int64_t n = std::int64_t(disparity.width) * std::int64_t(disparity.height);
    if (disparity.width <= 0 || disparity.height <= 0 || n > kStaticSpecklePixels)
      fail(R3DG_ERROR_INVALID_ARGUMENT, "remove_speckles_static: " + std::to_string(disparity.width) + "x" +
                                            std::to_string(disparity.height) + " exceeds the fixed scratch of " +
                                            std::to_string(kStaticSpecklePixels) + " pixels");
    std::int32_t* scr = static_scratch(c.device);
    if (scr == nullptr)
      fail(R3DG_ERROR_DEVICE, std::string("remove_speckles_static: hipGetSymbolAddress(scratch) failed: ") +
                                  hipGetErrorString(static_scratch_init(c.device)));
    // Every call on this GPU shares the array: one at a time, finished
    // before the next starts.  The kernels write every parent and size they
    // read, so nothing is cleared first.
    const std::lock_guard<std::mutex> use(static_scratch_mutex(c.device));
    launch_speckles(static_cast<float*>(disparity.data), disparity.width, disparity.height, min_size, max_diff, scr,
                    scr + kStaticSpecklePixels, c.stream);
    check_launch();
    hip_check(hipStreamSynchronize(c.stream), "hipStreamSynchronize");
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
