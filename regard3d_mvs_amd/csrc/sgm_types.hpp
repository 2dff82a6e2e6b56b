// Descriptors shared by the host driver and the gfx950 kernels.
//
// A *job* is one matching direction of one pyramid level of one pair: a base
// image matched against a partner image (the left-to-right match, or the
// right-to-left match on the mirrored pair, stereo_matcher.cpp:842-863).  A
// *tile* is one of the job's overlapping tiles (stereo_matcher.cpp:613-653);
// aggregation restarts at tile borders, so tiles are independent work units
// and one launch can cover the tiles of many jobs (both directions of a level,
// or every pair of a batch).
//
// HBM layout of the aggregated costs S (uint16) and the costs C (uint8): every
// tile owns a contiguous volume at cell_base, stored band by band (TileDesc
// below); a pixel's labels follow each other, at the cell offset held in its
// pixel record pix[px_base + y*w + x] (the TileLabels ranges of
// stereo_matcher.cpp:172-224).  Label counts are multiples of 16, so every
// pixel's labels start on a 32-byte boundary of S.
//
// Pixel record (uint64): bits 0-31 cell offset inside the tile row, bits 32-47
// first label `lo`, bits 48-63 label count `num`.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace r3dg {

constexpr int kLabelAlign = 16;
constexpr int kMaxLabels = 4000;  // per pixel and level: the path kernel keeps label count / 16 in 8 bits

__host__ __device__ constexpr std::uint64_t pix_record(std::uint32_t off, int lo, int num) {
  return std::uint64_t(off) | (std::uint64_t(std::uint16_t(lo)) << 32) | (std::uint64_t(std::uint16_t(num)) << 48);
}
__host__ __device__ constexpr std::uint32_t pix_off(std::uint64_t r) { return std::uint32_t(r); }
__host__ __device__ constexpr int pix_lo(std::uint64_t r) { return int((r >> 32) & 0xFFFFU); }
__host__ __device__ constexpr int pix_num(std::uint64_t r) { return int((r >> 48) & 0xFFFFU); }
constexpr int kMaxLevels = 13;

struct JobDesc {
  const std::uint16_t* img;  // base image (P2 intensities), width x height
  const std::uint64_t* cl;   // base Census
  const std::uint64_t* cr;   // partner Census, pwidth x height
  const std::int32_t* rows4; // [height][4] valid columns of the row: base first, last; partner first, last
  const std::uint16_t* lo;   // per-pixel label ranges (nullptr: full range)
  const std::uint16_t* num;
  std::int64_t img_stride;   // elements
  std::int64_t cl_stride;
  std::int64_t cr_stride;
  std::int64_t r_stride;
  std::int32_t width;
  std::int32_t height;
  std::int32_t pwidth;
  std::int32_t dmin;
  std::int32_t D;            // aligned label count of the level
  std::int32_t p1;
  std::int32_t invalid_cost;
  std::int32_t confidence;   // select writes confidence
  // 1: a mirrored job whose img, cl and cr are those of the unmirrored
  // images: base column X reads img[width - 1 - X] and cl[width - 1 - X],
  // partner column c reads cr[pwidth - 1 - c].  The Census of a mirrored
  // image is that of the image read backwards with its bits permuted
  // (neighbour (dy, dx) <-> (dy, -dx)), and a permutation applied to both
  // sides leaves Hamming distances unchanged.
  std::int32_t mirrored;
  std::uint16_t lut[256];    // P2 by 8-bit intensity step (stereo_matcher.cpp:69-79)
};

// Cells of a tile are stored band by band: columns [b << band_shift,
// (b + 1) << band_shift) of every row, rows in order, then the next band.
// Pixel (x, y)'s labels start at cell_base + bcorr[band_base + (x >> band_shift)
// * h + y] + off(x, y), off being the pixel record's offset in its row.
struct TileDesc {
  std::int32_t x0, y0, w, h;  // in the job's base image
  std::int32_t job;
  std::int32_t max_num;       // largest label count of the tile
  std::uint64_t cell_base;    // first cell of the tile in S
  std::uint64_t px_base;      // first tile pixel in per-pixel arrays
  std::uint64_t row_base;     // first tile row in per-row arrays
  std::int32_t band_shift;    // log2 of the band width (31: one band per row)
  std::int32_t nbands;
  std::uint64_t band_base;    // first (band, row) entry in bcorr
};

// Tile grid of one job for blending (stereo_matcher.cpp:688-766): tile (i, j)
// of the grid is tile tile_base + j * nx + i; cols[i] / rows[j] hold the tile
// extent [x0, x1) and its core [c0, c1).
struct GridDesc {
  const std::int32_t* cols;  // [nx][4]
  const std::int32_t* rows;  // [ny][4]
  std::int32_t nx, ny;
  std::int32_t width, height;
  std::int32_t blend;
  std::int32_t tile_base;
};

struct SelectParams {
  std::int32_t subpixel;     // r3dg_subpixel
  float uniqueness;
  std::int32_t lr_max_diff;
  std::int32_t diagonal_lr;  // check against the right WTA of the same volume
};

}  // namespace r3dg
