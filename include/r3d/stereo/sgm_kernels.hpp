// Row kernels of the r3d matcher API (reference: src/r3d/stereo/sgm_kernels.hpp:10-97),
// served on the GPU through the C ABI (r3dg_census_costs, r3dg_sgm_row,
// r3dg_wta_row).  Same types, constants and semantics as the reference.
// The reference compiles each kernel for several CPU SIMD targets plus a
// scalar twin; here both names run the same gfx950 kernels, which give the
// scalar results (the reference's contract for every target).
#pragma once

#include <cstddef>
#include <cstdint>

namespace r3d::sgm {

// Path buffers pad each pixel's labels by kLabelOffset elements on both sides;
// the elements just before label 0 and after the last label are kSentinel.
inline constexpr std::size_t kLabelOffset = 8;
inline constexpr std::uint16_t kSentinel = 0xFFFF;
// Label counts are multiples of this.
inline constexpr std::size_t kLabelAlign = 16;

[[nodiscard]] constexpr std::size_t path_stride(std::size_t num_disp) noexcept { return num_disp + 2 * kLabelOffset; }

// Labels of one row: pixel x has num[x] labels lo[x] .. lo[x] + num[x] - 1,
// packed from offset[x] (offset[0] = 0, offset[x + 1] = offset[x] + num[x]).
struct RowLabels {
  const std::int32_t* lo = nullptr;
  const std::uint16_t* num = nullptr;
  const std::size_t* offset = nullptr;  // [width + 1]

  [[nodiscard]] std::size_t path_offset(std::size_t x) const noexcept {
    return offset[x] + x * 2 * kLabelOffset + kLabelOffset;
  }
  [[nodiscard]] std::size_t path_size(std::size_t width) const noexcept {
    return offset[width] + width * 2 * kLabelOffset;
  }
};

// cost[offset[x] + j] = popcount(left[x] ^ rev[width - 1 - x + lo[x] + j - rev_origin]).
void census_costs(const std::uint64_t* left, const std::uint64_t* rev, std::size_t width, const RowLabels& labels,
                  std::int32_t rev_origin, std::uint16_t* cost);

// One row of one 4-path SGM pass (Hirschmueller 2008, eq. 13); direction +1:
// predecessors (x-1, y), (x, y-1), (x-1, y-1), (x+1, y-1); -1: mirrored.
struct SgmRowArgs {
  const std::uint16_t* cost = nullptr;      // packed per `labels`
  const std::uint16_t* p2 = nullptr;        // [4][width]
  const std::uint16_t* prev = nullptr;      // [3][prev_labels.path_size]
  const std::uint16_t* prev_min = nullptr;  // [3][width]
  std::uint16_t* cur = nullptr;             // [3][labels.path_size]
  std::uint16_t* cur_min = nullptr;         // [3][width]
  std::uint16_t* scratch = nullptr;         // [kScratchBuffers][path_stride(max_num)]; the first is the zero start
  std::uint16_t* sum = nullptr;             // packed per `labels`
  RowLabels labels;
  RowLabels prev_labels;
  std::size_t width = 0;
  std::size_t max_num = 0;
  std::uint16_t p1 = 0;
  int direction = 1;
  bool has_prev = false;
  bool accumulate = false;
};

inline constexpr std::size_t kScratchBuffers = 7;

void sgm_row(const SgmRowArgs& args);

// Winner-takes-all over a row of aggregated costs; see the reference header.
void wta_row(const std::uint16_t* sum, std::size_t width, const RowLabels& labels, std::uint16_t* best,
             std::uint16_t* best_cost, std::uint16_t* second, std::uint16_t* right_cost, std::uint16_t* right_best,
             std::ptrdiff_t right_shift);

namespace scalar {
void census_costs(const std::uint64_t* left, const std::uint64_t* rev, std::size_t width, const RowLabels& labels,
                  std::int32_t rev_origin, std::uint16_t* cost);
void sgm_row(const SgmRowArgs& args);
void wta_row(const std::uint16_t* sum, std::size_t width, const RowLabels& labels, std::uint16_t* best,
             std::uint16_t* best_cost, std::uint16_t* second, std::uint16_t* right_cost, std::uint16_t* right_best,
             std::ptrdiff_t right_shift);
}  // namespace scalar

}  // namespace r3d::sgm
