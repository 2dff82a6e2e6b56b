// Host-side launchers of the gfx950 kernels (kernels_image.hip, kernels_sgm.hip).
// Every launcher enqueues on `s` and returns without synchronising.  Pointers
// are device pointers; strides are in elements.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "sgm_types.hpp"

namespace r3dg {

// ---- images (kernels_image.hip)
void launch_census(const std::uint16_t* img, std::int64_t is, int w, int h, std::uint64_t* out, std::int64_t os,
                   hipStream_t s);
void launch_pyramid_u16(const std::uint16_t* in, std::int64_t is, int w, int h, std::uint32_t* tmp,
                        std::uint16_t* out, std::int64_t os, hipStream_t s);
void launch_pyramid_u8(const std::uint8_t* in, std::int64_t is, int w, int h, std::uint32_t* tmp, std::uint8_t* out,
                       std::int64_t os, bool threshold, hipStream_t s);
// uint8 image to uint16 = u8 * 257 (io/image_io.hpp:25-26).
void launch_widen_u8(const std::uint8_t* in, std::int64_t is, int w, int h, std::uint16_t* out, std::int64_t os,
                     hipStream_t s);
// rows[y] = {first, last} valid column (mask >= 128); mask == nullptr: {0, w - 1}.
// The row intervals of a level's left and right masks (nullptr: all valid),
// with those of the mirrored pair (mbase / mpartner, nullptr for none) and
// the packed rows [h][4] of both directions (rows4b with the mirrored pair).
void launch_level_rows(const std::uint8_t* ml, std::int64_t msl, int wl, const std::uint8_t* mr, std::int64_t msr,
                       int wr, int h, std::int32_t* lrows, std::int32_t* rrows, std::int32_t* mbase,
                       std::int32_t* mpartner, std::int32_t* rows4a, std::int32_t* rows4b, hipStream_t s);
void launch_valid_intervals(const std::uint8_t* mask, std::int64_t ms, int w, int h, std::int32_t* rows,
                            hipStream_t s);
// pyramid_down<float> (pyramid.cpp:34-85): float sums, * (1 / 256).
void launch_pyramid_f32(const float* in, std::int64_t is, int w, int h, float* tmp, float* out, std::int64_t os,
                        hipStream_t s);
// Geometry of one rectify_image call (rectification.cpp:288-350): the target
// pinhole intrinsics, the target-to-source rotation m (row-major) and the
// source camera at the sampled pyramid level.
struct RectifyGeom {
  double m[9];
  double tfx, tfy, tcx, tcy;
  double fx, fy, cx, cy;
  double params[5];
  int tw, th;
  int sw, sh;
  int model;
};
// out (tw x th) and valid (may be null) from the source level image src.
void launch_rectify_u8(const std::uint8_t* src, std::int64_t ss, const RectifyGeom& g, std::uint8_t* out,
                       std::int64_t os, std::uint8_t* valid, std::int64_t vs, hipStream_t s);
void launch_rectify_u16(const std::uint16_t* src, std::int64_t ss, const RectifyGeom& g, std::uint16_t* out,
                        std::int64_t os, std::uint8_t* valid, std::int64_t vs, hipStream_t s);
void launch_rectify_f32(const float* src, std::int64_t ss, const RectifyGeom& g, float* out, std::int64_t os,
                        std::uint8_t* valid, std::int64_t vs, hipStream_t s);
// combine_pairs (mvs/multi_stereo.cpp:74-216): one pair of a view (device
// pointers) and the grid with the options.
struct CombinePair {
  double g[9];       // grid pixel -> rectified pixel, row-major
  double fb, offset;
  double pc[3];      // partner centre
  const float* disp;
  std::int64_t ds;
  const float* conf;  // may be null
  std::int64_t cs;
  int dw, dh, cw, ch;
};
struct CombineGrid {
  double inv_k[9], rot_t[9], center[3];
  double inlier, max_res, cos_min;
  int w, h, min_pairs, npairs;
};
constexpr int kMaxCombinePairs = 16;  // multi_stereo.cpp:23
void launch_combine(const CombinePair* pairs, const CombineGrid& g, float* depth, std::int64_t ds, float* conf,
                    std::int64_t cs, std::uint16_t* mask, std::int64_t ms, hipStream_t s);
void launch_fill_f32(float* p, std::size_t n, float v, hipStream_t s);
void launch_fill_u8(std::uint8_t* p, std::size_t n, std::uint8_t v, hipStream_t s);

// Left-right check (stereo_matcher.cpp:777-805, 842-880).
// right(x) = mirrored(wr - 1 - x) - (wr - wl)
void launch_unmirror_right(const float* mirrored, int wr, int h, int wl, float* right, hipStream_t s);
// bad[x] = map valid and left_right_differ(map, other, sign) (1 / 0); with
// `apply`, invalidates map[x] too.  Reads `other` only.
void launch_lr_check(float* map, int w, int h, const float* other, int wo, int sign, int max_diff,
                     std::uint8_t* bad, bool apply, hipStream_t s);
void launch_apply_mask(float* map, std::size_t n, const std::uint8_t* bad, hipStream_t s);

// Post-filters (stereo_matcher.cpp:919-1021).  d is contiguous w x h.
void launch_speckles(float* d, int w, int h, int min_size, float max_diff, std::int32_t* parent, std::int32_t* size,
                     hipStream_t s);
void launch_median3x3(const float* in, int w, int h, float* out, hipStream_t s);

// Fixed scratch of the entry points that allocate nothing (static_scratch.hip):
// a device array of 2 * kStaticSpecklePixels int32 per GPU (32 MiB), the
// speckle filter's parents and component sizes.  static_scratch_init looks up
// the array's device address of `device` (which must be current) once;
// static_scratch returns it (nullptr before a successful init).  Calls that
// use the array on one GPU take static_scratch_mutex(device).
constexpr std::int64_t kStaticSpecklePixels = std::int64_t{1} << 22;
hipError_t static_scratch_init(int device);
std::int32_t* static_scratch(int device);
std::mutex& static_scratch_mutex(int device);

// ranges_from_coarse (stereo_matcher.cpp:1124-1267).  scratch needs
// ranges_scratch_floats(wc, hc) floats.
std::size_t ranges_scratch_floats(int wc, int hc, int range_fill);
void launch_ranges_from_coarse(const float* coarse, int wc, int hc, int width, int height, const std::uint8_t* fine_valid,
                               std::int64_t fvs, const std::int32_t* partner_rows, int dmin, int n_all, int radius,
                               int margin, int fill, int partner_sign, bool mirrored, float* scratch,
                               std::uint16_t* lo, std::uint16_t* num, std::int64_t rs, hipStream_t s);

// ---- aggregation (kernels_sgm.hip)
// Cells of every tile row (row_cells[row_base + y]) and the pixel records
// (cell offset inside the row, lo, num); one block per tile row.
// row_max = largest label count of the row; bstart[band_base + b * h + y] =
// row offset of band b's first pixel.
void launch_tile_row_scan(const JobDesc* jobs, const TileDesc* tiles, const std::uint64_t* row_prefix, int ntiles,
                          std::uint64_t total_rows, std::uint64_t* pix, std::uint64_t* row_cells,
                          std::uint32_t* row_max, std::uint64_t* bstart, hipStream_t s);
// Band corrections of every tile (see TileDesc): bcorr[band_base + b * h + y];
// tile_cells[t] = cells of the tile, tile_max[t] = its largest label count.
void launch_tile_cells(const TileDesc* tiles, int ntiles, const std::uint64_t* row_cells, const std::uint32_t* row_max,
                       const std::uint64_t* bstart, std::uint64_t* bcorr, std::uint64_t* tile_cells,
                       std::uint32_t* tile_max, hipStream_t s);
// Path kernel geometry: a wave64 scans kScanPerWave scanlines with
// kLanesPerScan lanes each (group prefixes count waves).
constexpr int kLanesPerScan = 8;
constexpr int kScanPerWave = 64 / kLanesPerScan;
// Matching costs (u8) of every tile cell into C (same cell indexing as S).
void launch_cost(const JobDesc* jobs, const TileDesc* tiles, const std::uint64_t* row_prefix, int ntiles,
                 std::uint64_t total_rows, const std::uint64_t* bcorr, const std::uint64_t* pix, std::uint8_t* C,
                 hipStream_t s);
// One SGM path over every tile (path index 0..7, see kernels_sgm.hip): a
// persistent launch of at most `slots` waves (path_slots(): the waves that fit
// on the device) taking scanline groups from `queue` (a zeroed counter).
// `scratch` holds path_scratch_elems(slots, max_num) uint16: per wave and
// scanline two rows for the path costs of pixels wider than the LDS rows.
std::uint64_t path_slots();
std::size_t path_scratch_elems(std::uint64_t slots, int max_num);
int path_scratch_row(int max_num);
void launch_sgm_path(int path, const JobDesc* jobs, const TileDesc* tiles, const std::uint64_t* group_prefix, int ntiles,
                     std::uint64_t groups, const std::uint64_t* bcorr, const std::uint64_t* pix,
                     const std::uint8_t* C, std::uint16_t* S, std::uint16_t* scratch, std::uint64_t slots, int max_num,
                     std::uint32_t* queue, bool accumulate, hipStream_t s);
// Extra elements S and C hold past their last cell.
constexpr std::size_t kSPad = 64;
// dst[i] = (src[0][i] + ... + src[nsrc - 1][i]) mod 2^16 for i < n rounded up
// to a multiple of 8 (buffers padded by kSPad); dst may be src[0].
void launch_sum_u16(std::uint16_t* dst, const std::uint16_t* const* src, int nsrc, std::uint64_t n, hipStream_t s);
// WTA + sub-pixel + checks of every tile row into per-tile-pixel results.
void launch_select(const JobDesc* jobs, const TileDesc* tiles, const std::uint64_t* row_prefix, int ntiles,
                   std::uint64_t total_rows, const std::uint64_t* row_start, const std::uint64_t* pix,
                   const std::uint16_t* const* S, int nslices, SelectParams p, int max_w, int max_D, float* tdisp,
                   float* tconf, std::int16_t* tlab, hipStream_t s);
// Packs two valid-interval arrays ([h][2] each) into rows4 ([h][4]).
void launch_pack_rows(const std::int32_t* base, const std::int32_t* partner, int h, std::int32_t* rows4,
                      hipStream_t s);
// Blends the tiles of one job into a full-size map.
void launch_blend(GridDesc g, const TileDesc* tiles, const float* tdisp, const float* tconf, const std::int16_t* tlab,
                  float* disp, float* conf, std::int16_t* lab, hipStream_t s);

// ---- row kernels of the reference's interface (census.hpp:25-29, sgm_kernels.hpp:41-97)
// rows: 7 rows of n + 8 elements (image rows y - 3 .. y + 3, padded by 4 replicated columns).
void launch_census_row(const std::uint16_t* rows, std::uint64_t n, std::uint64_t* out, hipStream_t s);
// rev holds the caller's rev[rmin ..]; offsets are std::size_t of the caller's RowLabels.
void launch_census_costs(const std::uint64_t* left, const std::uint64_t* rev, std::int64_t rmin, std::uint64_t width,
                         const std::int32_t* lo, const std::uint16_t* num, const std::uint64_t* offset,
                         std::int32_t origin, std::uint16_t* cost, hipStream_t s);
// SgmRowArgs with device copies of every buffer; start is the zero start
// buffer (scratch[0 .. path_stride(max_num))).
struct SgmRowDev {
  const std::uint16_t* cost;
  const std::uint16_t* p2;        // [4][width]
  const std::uint16_t* prev;      // [3][prev path size]
  const std::uint16_t* prev_min;  // [3][width]
  std::uint16_t* cur;             // [3][path size]
  std::uint16_t* cur_min;         // [3][width]
  std::uint16_t* sum;
  const std::uint16_t* start;
  const std::int32_t* lo;
  const std::uint16_t* num;
  const std::uint64_t* offset;    // [width + 1]
  const std::int32_t* prev_lo;
  const std::uint16_t* prev_num;
  const std::uint64_t* prev_offset;
  std::uint64_t width;
  std::uint64_t max_num;
  std::uint32_t p1;
  std::int32_t direction;
  std::int32_t has_prev;
  std::int32_t accumulate;
};
// part: one uint16 per cell (paths 1..3); hbuf: 2 * (max_num + 2) uint16.
void launch_sgm_row(const SgmRowDev& a, std::uint16_t* part, std::uint16_t* hbuf, hipStream_t s);
// rc / rb: the caller's right arrays at index imin, count entries (may be null).
void launch_wta_row(const std::uint16_t* sum, std::uint64_t width, const std::int32_t* lo, const std::uint16_t* num,
                    const std::uint64_t* offset, std::uint16_t* best, std::uint16_t* cost, std::uint16_t* second,
                    std::int64_t shift, std::int64_t imin, std::uint64_t count, std::uint16_t* rc, std::uint16_t* rb,
                    hipStream_t s);

}  // namespace r3dg
