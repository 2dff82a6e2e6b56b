// gfx950 kernels of the SGM core: tile label layout, matching costs, the 8
// path scans, winner-takes-all selection and tile blending.
//
// Aggregation dataflow (HBM-bound, no MFMA: nothing here is a contraction).
// k_cost writes the matching cost of every tile cell once per level as one
// byte (popcount of the Census XOR, or invalid_cost).  Each path direction is
// one persistent launch over every tile of a chunk: waves take groups of 8
// adjacent scanlines from a work queue; a scanline is 8 lanes, each holding 4
// or 8 consecutive labels of the current pixel as packed uint16 pairs, and a
// pixel's labels are processed in segments (32 labels, then 64 at a time).
// The predecessor's path costs sit in an LDS row with 0xFFFF sentinels around
// its labels (a global scratch row for pixels of more than kNarrow labels),
// the min over a pixel's labels is an 8-lane DPP reduction, and costs, S and
// the pixel records are prefetched a few steps ahead.  The first path of a
// stream writes its slice of S, the others add to it (uint16 wrap-around, so
// the order of the paths does not matter); the selection sums the slices.
//
// Recurrence (sgm_kernels.cpp:176-228, 387-403; stereo_matcher.cpp:1280-1362):
//   L(p,d) = C(p,d) + min(L(q,d), sat16(min(L(q,d-1), L(q,d+1)) + P1),
//                         sat16(min_q + P2(p,q))) - min_q            (mod 2^16)
// with L = C at the first pixel of a path inside the tile.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

#include "kernels.hpp"

namespace r3dg {
namespace {

constexpr float kInf = __builtin_huge_valf();

__device__ __forceinline__ int find_tile(const std::uint64_t* __restrict__ prefix, int ntiles, std::uint64_t id) {
  int lo = 0, hi = ntiles - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= id) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  return lo;
}

// The same search by a whole wave (every lane active, the same id in every
// lane): each round samples 64 evenly spaced prefixes and keeps the interval
// between the last sample <= id and the next, so 4096 tiles take two rounds
// of loads instead of twelve.
__device__ __forceinline__ int find_tile_wave(const std::uint64_t* __restrict__ prefix, int ntiles, std::uint64_t id) {
  const int lane = static_cast<int>(__lane_id());
  int lo = 0, span = ntiles;  // prefix[lo] <= id; the tile lies in [lo, lo + span)
  while (span > 1) {
    const int step = (span + 63) >> 6;
    const int i = lane * step;
    const bool le = i < span && prefix[lo + i] <= id;
    const int c = __popcll(__ballot(le));
    lo += (c - 1) * step;
    span = min(step, span - (c - 1) * step);
  }
  return lo;
}

// Pointers read out of JobDesc / TileDesc are generic to the compiler; these
// views make their loads and stores global_* instructions (flat ones also count
// against the LDS wait counter, which would serialise the prefetch below).
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gptr(const T* p) {
  return (const __attribute__((address_space(1))) T*)(p);
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gptr_mut(T* p) {
  return (__attribute__((address_space(1))) T*)(p);
}

__device__ __forceinline__ void label_range(const JobDesc& J, int X, int Y, int& lo, int& num) {
  if (J.lo != nullptr) {
    lo = J.lo[Y * J.r_stride + X];
    num = J.num[Y * J.r_stride + X];
  } else {
    lo = 0;
    num = J.D;
  }
}

// ------------------------------------------------------------------ layout
// Inclusive scan of v over a block of W waves: shuffles inside each wave, the
// wave totals through LDS (wsum[W]).  Returns this thread's inclusive prefix
// and the block total; every thread of the block must call it.
template <int W, typename U>
__device__ __forceinline__ U block_scan_incl(U v, U* wsum, U& total) {
  const int lane = static_cast<int>(threadIdx.x & 63), wave = static_cast<int>(threadIdx.x >> 6);
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const U t = __shfl_up(v, d);
    if (lane >= d) v += t;
  }
  if (lane == 63) wsum[wave] = v;
  __syncthreads();
  U off = 0, all = 0;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    off += k < wave ? wsum[k] : U(0);
    all += wsum[k];
  }
  total = all;
  __syncthreads();  // wsum is reused by the next call
  return v + off;
}

// Inclusive scan of v over the 64 lanes of a wave.
__device__ __forceinline__ std::uint32_t wave_scan_incl(std::uint32_t v) {
  const int lane = static_cast<int>(threadIdx.x & 63);
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const std::uint32_t t = __shfl_up(v, static_cast<unsigned>(d));
    if (lane >= d) v += t;
  }
  return v;
}

// Exclusive scan of the label counts of one tile row into the pixel records:
// cell offset inside the row, lo, num.  One wave per row (4 rows per block),
// 64 x kItems pixels per pass, the pass's label ranges loaded before its
// scans; no block barriers, so a row waits only on its own loads.
__global__ __launch_bounds__(256) void k_tile_row_scan(const JobDesc* __restrict__ jobs,
                                                       const TileDesc* __restrict__ tiles,
                                                       const std::uint64_t* __restrict__ row_prefix, int ntiles,
                                                       std::uint64_t total_rows,
                                                       std::uint64_t* __restrict__ pix,
                                                       std::uint64_t* __restrict__ row_cells,
                                                       std::uint32_t* __restrict__ row_max,
                                                       std::uint64_t* __restrict__ bstart) {
  constexpr int kItems = 8;
  const int lane = static_cast<int>(threadIdx.x & 63);
  const std::uint64_t r = std::uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (r >= total_rows) return;  // whole waves
  const int t = find_tile_wave(row_prefix, ntiles, r);
  const TileDesc T = tiles[t];
  const JobDesc& J = jobs[T.job];
  const int y = static_cast<int>(r - row_prefix[t]);
  const int Y = T.y0 + y;
  const auto jlo = gptr(J.lo), jnum = gptr(J.num);
  const std::int64_t rbase = std::int64_t(Y) * J.r_stride + T.x0;
  const int bmask = (1 << min(T.band_shift, 30)) - 1;
  std::uint64_t* prow = pix + T.px_base + std::uint64_t(y) * T.w;
  std::uint32_t carry = 0, vmax = 0;
  for (int base = 0; base < T.w; base += 64 * kItems) {
    int lo[kItems], num[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
      const int x = base + 64 * k + lane;
      lo[k] = 0;
      num[k] = 0;
      if (x < T.w) {
        if (jlo != nullptr) {
          lo[k] = jlo[rbase + x];
          num[k] = jnum[rbase + x];
        } else {
          num[k] = J.D;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
      const int x = base + 64 * k + lane;
      const std::uint32_t v = static_cast<std::uint32_t>(num[k]);
      const std::uint32_t incl = wave_scan_incl(v);
      if (x < T.w) {
        const std::uint32_t off = carry + incl - v;
        prow[x] = pix_record(off, lo[k], num[k]);
        // Row offset of the first pixel of each band.
        if ((x & bmask) == 0 && (x >> T.band_shift) < T.nbands)
          bstart[T.band_base + std::uint64_t(x >> T.band_shift) * T.h + y] = off;
      }
      carry += static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63));
      vmax = max(vmax, v);
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) vmax = max(vmax, static_cast<std::uint32_t>(__shfl_xor(vmax, d)));
  if (lane == 0) {
    row_cells[T.row_base + y] = carry;
    row_max[T.row_base + y] = vmax;
  }
}

// Totals of one tile (tile_cells, tile_max) and the band corrections: an
// exclusive scan of the band-row cell counts in band-major order, minus the
// row offset of the band's first pixel.
constexpr int kCellsThreads = 1024;  // k_tile_cells: one block of 16 waves per tile
__global__ __launch_bounds__(kCellsThreads) void k_tile_cells(const TileDesc* __restrict__ tiles,
                                                    const std::uint64_t* __restrict__ row_cells,
                                                    const std::uint32_t* __restrict__ row_max,
                                                    const std::uint64_t* __restrict__ bstart,
                                                    std::uint64_t* __restrict__ bcorr,
                                                    std::uint64_t* __restrict__ tile_cells,
                                                    std::uint32_t* __restrict__ tile_max) {
  constexpr int kItems = 4;  // consecutive band rows per thread and pass
  const TileDesc T = tiles[blockIdx.x];
  __shared__ std::uint64_t wsum[kCellsThreads / 64];
  __shared__ std::uint32_t smax;
  if (threadIdx.x == 0) smax = 0;
  __syncthreads();
  for (int y = threadIdx.x; y < T.h; y += kCellsThreads) atomicMax(&smax, row_max[T.row_base + y]);
  const int n = T.nbands * T.h;
  std::uint64_t carry = 0;
  for (int base = 0; base < n; base += kCellsThreads * kItems) {
    const int e0 = base + kItems * static_cast<int>(threadIdx.x);
    std::uint64_t v[kItems], b0[kItems], sum = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const int e = e0 + i;
      v[i] = 0;
      b0[i] = 0;
      if (e < n) {
        const int b = e / T.h, y = e - b * T.h;
        b0[i] = bstart[T.band_base + e];
        const std::uint64_t b1 = b + 1 < T.nbands ? bstart[T.band_base + e + T.h] : row_cells[T.row_base + y];
        v[i] = b1 - b0[i];
      }
      sum += v[i];
    }
    std::uint64_t total;
    std::uint64_t before = carry + block_scan_incl<kCellsThreads / 64>(sum, wsum, total) - sum;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      if (e0 + i < n) bcorr[T.band_base + e0 + i] = before - b0[i];
      before += v[i];
    }
    carry += total;
  }
  if (threadIdx.x == 0) {
    tile_cells[blockIdx.x] = carry;
    tile_max[blockIdx.x] = smax;
  }
}

// ------------------------------------------------------------------ paths
// Path directions (dx, dy): the predecessor of p is p - (dx, dy).
//   0 (1,0)  1 (-1,0)  2 (0,1)  3 (0,-1)  4 (1,1)  5 (-1,-1)  6 (-1,1)  7 (1,-1)
// Scanlines of a w x h tile: rows (dy == 0), columns (dx == 0), diagonals
// x - y = s - (h - 1) (dx == dy) or anti-diagonals x + y = s (dx == -dy).
template <int DX, int DY>
__device__ __forceinline__ int scanlines(int w, int h) {
  return DY == 0 ? h : (DX == 0 ? w : w + h - 1);
}

template <int DX, int DY>
__device__ __forceinline__ void scan_start(int s, int w, int h, int& x, int& y, int& len) {
  if (DY == 0) {
    y = s;
    x = DX > 0 ? 0 : w - 1;
    len = w;
  } else if (DX == 0) {
    x = s;
    y = DY > 0 ? 0 : h - 1;
    len = h;
  } else if (DX == DY) {
    const int c = s - (h - 1);
    const int x0 = c >= 0 ? c : 0, y0 = c >= 0 ? 0 : -c;
    len = min(w - x0, h - y0);
    x = DX > 0 ? x0 : x0 + len - 1;
    y = DX > 0 ? y0 : y0 + len - 1;
  } else {
    const int x0 = min(s, w - 1), y0 = s - x0;  // end on the top / right border
    len = min(x0 + 1, h - y0);
    x = DX < 0 ? x0 : x0 - (len - 1);
    y = DX < 0 ? y0 : y0 + (len - 1);
  }
}

// ------------------------------------------------------------------ costs
// Matching costs of every tile cell (stereo_matcher.cpp:233-247, census_costs):
// the Hamming distance of the base Census and the partner Census at column
// X - dmin - lo - j when that column lies in the partner row's valid interval,
// invalid_cost when it does not, 0 for base pixels outside the base row's
// valid interval.  One block per tile row; a thread per 16-label chunk writes
// 16 bytes.  Rows of at most kPx pixels stage everything a chunk reads in
// LDS (pixel cell offsets and lo, base Census, band corrections, a
// chunk-to-pixel map, the partner Census window), so the chunk loop touches
// global memory only to store; wider rows search their pixel records in
// global memory.
constexpr int kCostThreads = 256;
__global__ __launch_bounds__(kCostThreads) void k_cost(const JobDesc* __restrict__ jobs, const TileDesc* __restrict__ tiles,
                                              const std::uint64_t* __restrict__ row_prefix, int ntiles,
                                              const std::uint64_t* __restrict__ bcorr,
                                              const std::uint64_t* __restrict__ pix, std::uint8_t* __restrict__ C) {
  // Sized for the tiles of the default options (at most 600 columns, 512
  // labels): 24 KB of LDS, six blocks per CU.
  constexpr int kPx = 640;          // row pixels staged in LDS
  constexpr int kBands = kPx / 8;   // band corrections staged (8-column bands)
  constexpr int kMap = 2048;        // chunks mapped to their pixel
  constexpr int kWin = 1280;        // partner Census columns staged
  __shared__ std::uint32_t soff[kPx];
  __shared__ std::uint16_t slo[kPx];
  __shared__ std::uint64_t scl[kPx];
  __shared__ std::uint64_t sbc[kBands];
  __shared__ std::uint16_t smap[kMap];
  __shared__ std::uint64_t swin[kWin];
  const std::uint64_t r = blockIdx.x;
  const int t = find_tile_wave(row_prefix, ntiles, r);
  const TileDesc T = tiles[t];
  const JobDesc& J = jobs[T.job];
  const int y = static_cast<int>(r - row_prefix[t]);
  const int Y = T.y0 + y;
  const auto rec = gptr(pix + T.px_base + std::uint64_t(y) * T.w);
  const auto rows = gptr(J.rows4) + 4 * Y;
  const auto cl = gptr(J.cl) + std::size_t(Y) * J.cl_stride;
  const auto cr = gptr(J.cr) + std::size_t(Y) * J.cr_stride;
  const std::uint64_t* bc = bcorr + T.band_base + y;
  const int pw1 = J.pwidth - 1;
  const bool mir = J.mirrored != 0;
  auto bcol = [&](int X) { return mir ? J.width - 1 - X : X; };  // column of cl
  auto pcol = [&](int c) { return mir ? pw1 - c : c; };         // column of cr
  // Partner columns a cost of this row can use: X - dmin - label for X in the
  // tile and label in [0, D); every valid column lies in [c0, c1].
  const int c0 = max(0, T.x0 - J.dmin - (J.D - 1)), c1 = min(pw1, T.x0 + T.w - 1 - J.dmin);
  const bool windowed = c1 >= c0 && c1 - c0 < kWin;
  const bool staged = T.w <= kPx && T.nbands <= kBands;
  const std::uint64_t last = rec[T.w - 1];
  // Staging: every global load of a thread is issued before the first LDS
  // store, so that the block waits for one memory latency, not one per
  // element (the launch has kCostThreads threads).
  constexpr int kWinIt = (kWin + kCostThreads - 1) / kCostThreads;
  constexpr int kPxIt = (kPx + kCostThreads - 1) / kCostThreads;
  static_assert(kBands <= kCostThreads, "one band correction per thread");
  const int tid = static_cast<int>(threadIdx.x);
  std::uint64_t wv[kWinIt], pv[kPxIt], cv[kPxIt], bv = 0;
#pragma unroll
  for (int i = 0; i < kWinIt; ++i) {
    const int k = tid + kCostThreads * i;
    wv[i] = windowed && k <= c1 - c0 ? cr[pcol(c0 + k)] : 0;
  }
  if (staged) {
#pragma unroll
    for (int i = 0; i < kPxIt; ++i) {
      const int x = tid + kCostThreads * i;
      pv[i] = x < T.w ? rec[x] : 0;
      cv[i] = x < T.w ? cl[bcol(T.x0 + x)] : 0;
    }
    if (tid < T.nbands) bv = bc[std::size_t(tid) * T.h];
  }
#pragma unroll
  for (int i = 0; i < kWinIt; ++i) {
    const int k = tid + kCostThreads * i;
    if (windowed && k <= c1 - c0) swin[k] = wv[i];
  }
  if (staged) {
#pragma unroll
    for (int i = 0; i < kPxIt; ++i) {
      const int x = tid + kCostThreads * i;
      if (x < T.w) {
        const std::uint64_t pr = pv[i];
        const std::uint32_t off = pix_off(pr);
        soff[x] = off;
        slo[x] = static_cast<std::uint16_t>(pix_lo(pr));
        scl[x] = cv[i];
        const std::uint32_t e =
            min((off + static_cast<std::uint32_t>(pix_num(pr))) >> 4, static_cast<std::uint32_t>(kMap));
        for (std::uint32_t k = off >> 4; k < e; ++k) smap[k] = static_cast<std::uint16_t>(x);
      }
    }
    if (tid < T.nbands) sbc[tid] = bv;
  }
  __syncthreads();
  const std::uint32_t chunks = (pix_off(last) + static_cast<std::uint32_t>(pix_num(last))) >> 4;
  const int bfirst = rows[0], blast = rows[1], pfirst = rows[2], plast = rows[3];
  const std::uint32_t invalid = static_cast<std::uint32_t>(J.invalid_cost);
  std::uint8_t* out = C + T.cell_base;
  auto partner = [&](int xr) {
    return windowed ? swin[min(max(xr, c0), c1) - c0] : cr[pcol(min(max(xr, 0), pw1))];
  };
  // Costs of chunk k of pixel xa (first cell off, first label lo, Census c),
  // stored at the band row offset bco.
  auto chunk = [&](std::uint32_t k, int xa, std::uint32_t off, int lo, std::uint64_t c, std::uint64_t bco) {
    const int X = T.x0 + xa;
    const int jb = static_cast<int>(16 * k - off);
    const int xr0 = X - J.dmin - lo - jb;
    const bool lvalid = X >= bfirst && X <= blast;
    std::uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      std::uint32_t v = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int xr = xr0 - 4 * q - b;
        std::uint32_t cost = xr >= pfirst && xr <= plast ? static_cast<std::uint32_t>(__popcll(c ^ partner(xr)))
                                                         : invalid;
        cost = lvalid ? cost : 0U;
        v |= cost << (8 * b);
      }
      w[q] = v;
    }
    *reinterpret_cast<uint4*>(out + bco + 16 * std::size_t(k)) = make_uint4(w[0], w[1], w[2], w[3]);
  };
  if (staged) {
    const bool mapped = chunks <= static_cast<std::uint32_t>(kMap);
    for (std::uint32_t k = threadIdx.x; k < chunks; k += blockDim.x) {
      int xa;
      if (mapped) {
        xa = smap[k];
      } else {
        // Every pixel holds at least one chunk, so the pixel of chunk k is at most k.
        xa = 0;
        int xb = min(static_cast<int>(k), T.w - 1);
        while (xa < xb) {
          const int mid = (xa + xb + 1) >> 1;
          if (soff[mid] <= 16 * k) {
            xa = mid;
          } else {
            xb = mid - 1;
          }
        }
      }
      chunk(k, xa, soff[xa], slo[xa], scl[xa], sbc[xa >> T.band_shift]);
    }
  } else {
    for (std::uint32_t k = threadIdx.x; k < chunks; k += blockDim.x) {
      int xa = 0, xb = min(static_cast<int>(k), T.w - 1);
      while (xa < xb) {
        const int mid = (xa + xb + 1) >> 1;
        if (pix_off(rec[mid]) <= 16 * k) {
          xa = mid;
        } else {
          xb = mid - 1;
        }
      }
      const std::uint64_t pr = rec[xa];
      chunk(k, xa, pix_off(pr), pix_lo(pr), cl[bcol(T.x0 + xa)], bc[std::size_t(xa >> T.band_shift) * T.h]);
    }
  }
}

// ------------------------------------------------------------------ path scans
// Packed pairs of uint16 path costs: v_pk_* instructions process two labels.
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ u16x2 as_pair(std::uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ std::uint32_t as_u32(u16x2 v) { return __builtin_bit_cast(std::uint32_t, v); }

// Min over the 8 lanes of a scanline (quad swaps, then the half-row mirror).
__device__ __forceinline__ std::uint32_t min8(std::uint32_t v) {
  v = min(v, static_cast<std::uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0xB1, 0xF, 0xF, false)));
  v = min(v, static_cast<std::uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x4E, 0xF, 0xF, false)));
  v = min(v, static_cast<std::uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x141, 0xF, 0xF, false)));
  return v;
}

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Loads and stores at a byte offset from a wave-uniform base (SGPR base +
// 32-bit VGPR offset addressing).
template <typename T>
__device__ __forceinline__ T ldb(const void* base, std::uint32_t off) {
  const auto p = (const __attribute__((address_space(1))) char*)(base);
  return *reinterpret_cast<const __attribute__((address_space(1))) T*>(p + off);
}
template <typename T>
__device__ __forceinline__ void stb(void* base, std::uint32_t off, T v) {
  const auto p = (__attribute__((address_space(1))) char*)(base);
  *reinterpret_cast<__attribute__((address_space(1))) T*>(p + off) = v;
}

// s_waitcnt immediate (gfx9 encoding) that waits for every vector memory
// operation and nothing else: vmcnt 0, expcnt 7, lgkmcnt 15.
constexpr int kVmcnt0 = 0x0F70;
constexpr int kWaves = 4;                     // waves per block
constexpr int kDist = 3;                      // steps between a segment prefetch and its use
constexpr int kNarrow = 368;                  // labels whose path costs stay in LDS
constexpr int kPad = 12;                      // sentinel slots before label 0 and after the last label
constexpr int kRow = kPad + kNarrow + kPad;   // uint16 slots of an LDS row

// Label segments of a pixel: segment 0 holds labels [0, 32), lane l holds
// four of them from 4l; segment s >= 1 holds [32 + 64 (s - 1), 32 + 64 s),
// lane l eight of them from 32 + 64 (s - 1) + 8l.  A lane's labels lie all
// inside or all outside a pixel's range (label counts are multiples of 16).
template <int P>
__device__ __forceinline__ int seg_first(int s, int l8) {
  return P == 4 ? 4 * l8 : 32 + 64 * (s - 1) + 8 * l8;
}
__device__ __forceinline__ int segments(int num) { return num <= 32 ? 1 : 1 + (num - 32 + 63) / 64; }

// Costs (P bytes) and S (P uint16) of one lane in one segment.
template <int P>
struct SegIn {
  std::uint32_t c[P / 4];
  std::uint32_t s[P / 2];
};

// Per-step values of one scanline pixel, read from the wave's step ring:
// the pixel's first cell in the tile volume and a packed word holding its
// first label (bits 0-15), the P2 table index |I - I_prev| >> 8 of the step
// (bits 16-23) and its label count / 16 (bits 24-31); p2 is the table
// entry itself, read kDist steps before the step.
struct StepIn {
  std::uint32_t cell;
  std::uint32_t meta;
  std::uint32_t p2;
};
static_assert(kMaxLabels / kLabelAlign <= 255, "label count / 16 must fit 8 bits");
__device__ __forceinline__ int step_lo(const StepIn& u) { return static_cast<int>(u.meta & 0xFFFFU); }
__device__ __forceinline__ std::uint32_t step_p2_index(const StepIn& u) { return (u.meta >> 16) & 0xFFU; }
__device__ __forceinline__ int step_num(const StepIn& u) { return static_cast<int>(u.meta >> 24) * kLabelAlign; }
// Pixel record, band-row offset and intensity of step 8 b + l of a scanline,
// held by lane l of its group: the loads of 8 steps are one batch.
struct StepBatch {
  std::uint64_t rec;
  std::uint32_t rs;
  std::uint16_t ip;  // widened where it is used, so no wait follows the load
};
// Segments 0 and 1 of a pixel, loaded one step ahead.
struct StepSegs {
  SegIn<4> a;
  SegIn<8> b;
};

// One path direction over every tile.  A wave scans 8 adjacent scanlines of a
// tile, 8 lanes per scanline, a lane holding 4 or 8 consecutive labels of a
// segment as packed uint16 pairs.  Recursion (stereo_matcher.cpp:258-330):
//   L(j) = C(j) + min(Lp(k), sat(min(Lp(k-1), Lp(k+1)) + P1), sat(mp + P2)) - mp
// with k = j + lo - lo_prev, mp the predecessor's minimum, Lp = 0xFFFF outside
// its labels and L = C at the first pixel.  The predecessor's path costs sit
// in a row with 0xFFFF sentinels around its labels, in LDS for pixels of at
// most kNarrow labels and in a global scratch row for wider ones; a lane reads
// its window of P + 2 values as aligned words (the window clamped into the
// sentinels) and shifts them into place.  Segments 0 and 1 use data prefetched
// kDist steps ahead; segments from 2 on (pixels of more than 96 labels) are
// loaded in the step, one segment ahead.  The pixel records, band-row offsets
// and intensities of a scanline are loaded 8 steps at a time (lane l of the
// group loads step 8 b + l) and unpacked into an 8-slot ring in LDS, read one
// step before the segment prefetch that needs them.
template <int DX, int DY, bool kAcc>
__global__ __launch_bounds__(64 * kWaves) void k_path(
    const JobDesc* __restrict__ jobs, const TileDesc* __restrict__ tiles, const std::uint64_t* __restrict__ gprefix,
    int ntiles, std::uint64_t groups, const std::uint64_t* __restrict__ bcorr,
    const std::uint64_t* __restrict__ pix, const std::uint8_t* __restrict__ C, std::uint16_t* __restrict__ S,
    std::uint16_t* __restrict__ scratch, int scratch_row, std::uint32_t* __restrict__ queue) {
  __shared__ __attribute__((aligned(16))) std::uint16_t lut[kWaves][256];
  __shared__ __attribute__((aligned(16))) std::uint32_t rows[kWaves][kScanPerWave][2][kRow / 2];
  __shared__ __attribute__((aligned(16))) u32x2 ring[kWaves][8][kScanPerWave];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63, grp = lane / kLanesPerScan, l8 = lane % kLanesPerScan;
  // Resident wave index: the wave keeps its LDS rows and its scratch rows for
  // every scanline group it takes from the queue.
  const std::uint64_t slot = std::uint64_t(blockIdx.x) * kWaves + wave;
  if (l8 < kPad / 2) {  // lead sentinels of both rows (never overwritten)
    rows[wave][grp][0][l8] = 0xFFFFFFFFU;
    rows[wave][grp][1][l8] = 0xFFFFFFFFU;
  }
  __attribute__((address_space(1))) std::uint32_t* srow =
      reinterpret_cast<__attribute__((address_space(1))) std::uint32_t*>(
          gptr_mut(scratch + (slot * kScanPerWave + grp) * 2 * std::uint64_t(scratch_row)));
  const int swords = scratch_row / 2;
  if (l8 < kPad / 2) {
    srow[l8] = 0xFFFFFFFFU;
    srow[swords + l8] = 0xFFFFFFFFU;
  }
  int lut_job = -1;
  for (;;) {
  // Next scanline group from the queue (one atomic per wave).
  std::uint32_t item = 0;
  if (lane == 0) item = atomicAdd(queue, 1U);
  item = static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(item), 0));
  if (item >= groups) break;
  const int t = find_tile(gprefix, ntiles, item);
  const TileDesc T = tiles[t];
  const JobDesc& J = jobs[T.job];
  if (T.job != lut_job) {
    for (int k = lane; k < 256; k += 64) lut[wave][k] = J.lut[k];
    lut_job = T.job;
  }
  __builtin_amdgcn_wave_barrier();
  const int ns = scanlines<DX, DY>(T.w, T.h);
  // Diagonal scanlines are longest in the middle of a tile: take those
  // groups first so that a launch ends with short ones.
  int rank = static_cast<int>(item - gprefix[t]);
  if constexpr (DX != 0 && DY != 0) {
    const int G = static_cast<int>(gprefix[t + 1] - gprefix[t]), mid = G / 2;
    rank = (rank & 1) ? mid - 1 - (rank >> 1) : mid + (rank >> 1);
  }
  const int s0 = rank * kScanPerWave;
  const int sl = min(s0 + grp, ns - 1);  // idle groups shadow the last scanline
  const bool grp_active = s0 + grp < ns;
  int x0, y0, len;
  scan_start<DX, DY>(sl, T.w, T.h, x0, y0, len);
  auto wave_max = [](int v) {
    v = max(v, __shfl_xor(v, 8));
    v = max(v, __shfl_xor(v, 16));
    v = max(v, __shfl_xor(v, 32));
    return __builtin_amdgcn_readfirstlane(v);
  };
  const int steps = wave_max(grp_active ? len : 0);

  const std::uint32_t P1x2 = static_cast<std::uint32_t>(J.p1) * 0x10001U;
  const std::uint64_t* tpix = pix + T.px_base;
  const std::uint64_t* tbc = bcorr + T.band_base;
  const std::uint16_t* img = J.img;
  const int img_stride = static_cast<int>(J.img_stride);
  // Base columns of a mirrored job read the unmirrored image backwards.
  const int img_x0 = J.mirrored != 0 ? J.width - 1 - T.x0 : T.x0, img_dx = J.mirrored != 0 ? -1 : 1;
  const std::uint8_t* Ct = C + T.cell_base;
  std::uint16_t* St = S + T.cell_base;

  // Batch of steps [base, base + 8): lane l8 loads step base + l8 (steps past
  // the scanline repeat its last pixel; they are never active).
  auto load_batch = [&](int base, StepBatch& b) {
    const int i = min(base + l8, len - 1);
    const int x = x0 + i * DX, y = y0 + i * DY;
    b.rec = ldb<std::uint64_t>(tpix, static_cast<std::uint32_t>(y * T.w + x) * 8U);
    b.rs = ldb<std::uint32_t>(tbc, static_cast<std::uint32_t>((x >> T.band_shift) * T.h + y) * 8U);
    b.ip = ldb<std::uint16_t>(img, static_cast<std::uint32_t>((T.y0 + y) * img_stride + img_x0 + img_dx * x) * 2U);
  };
  // Unpacks a batch into the ring (slot l8 of the group); last_ip carries the
  // intensity of the batch's last step to the next batch's first.
  std::uint32_t last_ip = 0;
  auto write_batch = [&](const StepBatch& b) {
    std::uint32_t prev = __shfl_up(b.ip, 1U, kLanesPerScan);
    prev = l8 == 0 ? last_ip : prev;
    const std::uint32_t d = static_cast<std::uint32_t>(abs(static_cast<int>(b.ip) - static_cast<int>(prev))) >> 8;
    ring[wave][l8][grp] = u32x2{b.rs + pix_off(b.rec), static_cast<std::uint32_t>(pix_lo(b.rec)) | d << 16 |
                                                           static_cast<std::uint32_t>(pix_num(b.rec) / kLabelAlign) << 24};
    last_ip = __shfl(b.ip, kLanesPerScan - 1, kLanesPerScan);
    __builtin_amdgcn_wave_barrier();
    // Keep the shuffles of b ahead of the loads that refill it: otherwise the
    // new intensity lands in a temporary whose copy into b waits for the load.
    __builtin_amdgcn_sched_barrier(0);
  };
  auto read_step = [&](int i, StepIn& u) {
    const u32x2 e = ring[wave][i & 7][grp];
    u.cell = e.x;
    u.meta = e.y;
  };
  // Costs and S of segment s (lanes past the pixel's labels read its last ones;
  // those values are never used).
  auto load_seg4 = [&](const StepIn& u, SegIn<4>& d) {
    const std::uint32_t cell = u.cell;
    const std::uint32_t j0 = static_cast<std::uint32_t>(min(4 * l8, step_num(u) - 4));
    d.c[0] = ldb<std::uint32_t>(Ct, cell + j0);
    if constexpr (kAcc) {
      const u32x2 v = ldb<u32x2>(St, (cell + j0) * 2U);
      d.s[0] = v.x;
      d.s[1] = v.y;
    }
  };
  auto load_seg8 = [&](const StepIn& u, int s, SegIn<8>& d) {
    const std::uint32_t cell = u.cell;
    const std::uint32_t j0 = static_cast<std::uint32_t>(min(seg_first<8>(s, l8), step_num(u) - 8));
    const u32x2 c = ldb<u32x2>(Ct, cell + j0);
    d.c[0] = c.x;
    d.c[1] = c.y;
    if constexpr (kAcc) {
      const u32x4 v = ldb<u32x4>(St, (cell + j0) * 2U);
      d.s[0] = v.x;
      d.s[1] = v.y;
      d.s[2] = v.z;
      d.s[3] = v.w;
    }
  };
  // Segment 1 only when a scanline of the wave needs it at that step (its
  // pixel has more than 32 labels); the step tests the same condition.
  // The label count / 16 is the top byte of meta: more than 32 labels is
  // meta > 0x02FFFFFF, never for idle groups.
  const std::uint32_t seg1_meta = grp_active ? 0x02FFFFFFU : 0xFFFFFFFFU;
  auto load_segs = [&](const StepIn& u, StepSegs& d) {
    load_seg4(u, d.a);
    if (__builtin_amdgcn_ballot_w64(u.meta > seg1_meta) != 0) load_seg8(u, 1, d.b);
  };

  int prev_lo = 0, prev_num = 0, flip = 0;
  std::uint32_t prev_min = 0;
  StepBatch B;
  // Step i: u, v are its pixel values and segments; it reads the pixel values
  // of step i + kDist + 1 into u_load (first unpacking the batch of that step
  // and loading the next one when it starts a batch), the P2 entry of step
  // i + kDist into u_seg and loads that step's segments into v_load.  The
  // batch of steps [8 b, 8 b + 8) is unpacked at step 8 b - kDist - 1, after
  // the last read of the slots it overwrites (steps 8 b - 8 .. 8 b - 1, read
  // by step 8 b - kDist - 2).
  auto step = [&](int i, const StepIn& u, const StepSegs& v, StepIn& u_load, StepIn& u_seg,
                  StepSegs& v_load) __attribute__((always_inline)) {
    u_seg.p2 = lut[wave][step_p2_index(u_seg)];
    if (((i + kDist + 1) & 7) == 0) {
      write_batch(B);
      load_batch(i + kDist + 1 + 8, B);
    }
    read_step(i + kDist + 1, u_load);
    load_segs(u_seg, v_load);
    const bool act = grp_active && i < len;
    const bool first = i == 0;
    const int lo = step_lo(u), num = step_num(u);
    const std::uint32_t cell = u.cell;
    const std::uint32_t pen = min(prev_min + u.p2, 0xFFFFU);
    const std::uint32_t PENx2 = pen * 0x10001U, MPx2 = prev_min * 0x10001U;
    const int delta = lo - prev_lo;
    const bool prev_wide = prev_num > kNarrow, cur_wide = num > kNarrow;
    const int n_lds = min(prev_num, kNarrow);
    // Label counts of active lanes (0 elsewhere): each ballot is one compare.
    const int num_act = act ? num : 0, prev_act = act ? prev_num : 0;
    const bool any_seg1 = __builtin_amdgcn_ballot_w64(num_act > 32) != 0;
    const bool any_seg2 = __builtin_amdgcn_ballot_w64(num_act > 96) != 0;
    const bool any_prev_wide = __builtin_amdgcn_ballot_w64(prev_act > kNarrow) != 0;
    const bool any_cur_wide = __builtin_amdgcn_ballot_w64(num_act > kNarrow) != 0;
    std::uint32_t* cur = rows[wave][grp][flip];
    const std::uint32_t* prv = rows[wave][grp][flip ^ 1];
    __attribute__((address_space(1))) std::uint32_t* scur = srow + flip * swords;
    const __attribute__((address_space(1))) std::uint32_t* sprv = srow + (flip ^ 1) * swords;
    std::uint32_t m = 0xFFFFFFFFU;

    // Byte position of predecessor label k - 1 for the window of a lane whose
    // first label is j0; k is clamped into [-(P + 1), n + 1] so that the
    // window stays inside the sentinels.
    auto wpos = [&](int P, int j0, int n) {
      return static_cast<std::uint32_t>(min(max(j0 + delta, -(P + 1)), n + 1) - 1 + kPad) * 2U;
    };
    // New path costs of P labels from their window words W (shift sh) and costs.
    auto recur = [&](auto Pc, const std::uint32_t* W, std::uint32_t sh, const std::uint32_t* c, std::uint32_t* L) {
      constexpr int P = decltype(Pc)::value;
      std::uint32_t Q[P / 2 + 1];
#pragma unroll
      for (int k = 0; k <= P / 2; ++k) Q[k] = __builtin_amdgcn_alignbyte(W[k + 1], W[k], sh);
#pragma unroll
      for (int k = 0; k < P / 2; ++k) {
        const std::uint32_t Cx2 = __builtin_amdgcn_perm(0U, c[k >> 1], (k & 1) ? 0x0c030c02U : 0x0c010c00U);
        if (first) {
          L[k] = Cx2;
        } else {
          const std::uint32_t p12 = __builtin_amdgcn_alignbyte(Q[k + 1], Q[k], 2U);
          const u16x2 nb =
              __builtin_elementwise_add_sat(__builtin_elementwise_min(as_pair(Q[k]), as_pair(Q[k + 1])), as_pair(P1x2));
          const u16x2 best = __builtin_elementwise_min(__builtin_elementwise_min(as_pair(p12), nb), as_pair(PENx2));
          L[k] = as_u32(best - as_pair(MPx2) + as_pair(Cx2));
        }
      }
    };
    // LDS window of a lane.
    auto lds_window = [&](auto Pc, int j0, std::uint32_t* W, std::uint32_t& sh) {
      constexpr int P = decltype(Pc)::value;
      sh = wpos(P, j0, n_lds);
      const std::uint32_t* wp = prv + (sh >> 2);
#pragma unroll
      for (int k = 0; k < P / 2 + 2; ++k) W[k] = wp[k];
    };
    // Scratch window of a lane (in range for every lane; used where prev_wide).
    auto scr_window = [&](auto Pc, int j0, std::uint32_t* W, std::uint32_t& sh) {
      constexpr int P = decltype(Pc)::value;
      sh = wpos(P, j0, prev_num);
      const auto wp = sprv + (sh >> 2);
#pragma unroll
      for (int k = 0; k < P / 2 + 2; ++k) W[k] = wp[k];
    };
    // Stores of a lane's new path costs (next step's row) and S; running min.
    // kLow: segments 0 and 1, whose labels (below 96) lie inside the LDS row
    // of every lane.  Their row store needs no condition: a lane past its
    // pixel's labels, or of a finished scanline, writes only positions the
    // next step does not read (the sentinels after the labels are written
    // later in the step, and reads stop inside them), and a wide pixel's row
    // is read from scratch.
    auto emit = [&](auto Pc, auto kLow, int j0, const std::uint32_t* L, const std::uint32_t* sv) {
      constexpr int P = decltype(Pc)::value;
      const bool on = act && j0 < num;
      const int w0 = (kPad + j0) >> 1;
      if (decltype(kLow)::value || (on && !cur_wide)) {
#pragma unroll
        for (int k = 0; k < P / 2; k += 2) *reinterpret_cast<u32x2*>(cur + w0 + k) = u32x2{L[k], L[k + 1]};
      }
      if (any_cur_wide) {
        if (on && cur_wide) {
#pragma unroll
          for (int k = 0; k < P / 2; ++k) scur[w0 + k] = L[k];
        }
      }
      std::uint32_t out[P / 2];
      std::uint32_t mm = L[0];
#pragma unroll
      for (int k = 0; k < P / 2; ++k) {
        out[k] = kAcc ? as_u32(as_pair(sv[k]) + as_pair(L[k])) : L[k];
        if (k > 0) mm = as_u32(__builtin_elementwise_min(as_pair(mm), as_pair(L[k])));
      }
      if (on) {
        const std::uint32_t off = (cell + static_cast<std::uint32_t>(j0)) * 2U;
        if constexpr (P == 4) {
          stb<u32x2>(St, off, u32x2{out[0], out[1]});
        } else {
          stb<u32x4>(St, off, u32x4{out[0], out[1], out[2], out[3]});
        }
      }
      m = on ? as_u32(__builtin_elementwise_min(as_pair(m), as_pair(mm))) : m;
    };
    using P4 = std::integral_constant<int, 4>;
    using P8 = std::integral_constant<int, 8>;
    // A lane's window: from LDS, or from its scratch row where prev_wide.  The
    // scratch loads are retired inside their block, so that the shared code
    // after it never waits on them.
    auto window = [&](auto Pc, int j0, std::uint32_t* W, std::uint32_t& sh) {
      constexpr int P = decltype(Pc)::value;
      lds_window(Pc, j0, W, sh);
      if (any_prev_wide) {
        std::uint32_t G[P / 2 + 2], gsh;
        scr_window(Pc, j0, G, gsh);
        __builtin_amdgcn_s_waitcnt(kVmcnt0);
#pragma unroll
        for (int k = 0; k < P / 2 + 2; ++k) W[k] = prev_wide ? G[k] : W[k];
        sh = prev_wide ? gsh : sh;
      }
    };
    auto seg8 = [&](int s, const SegIn<8>& d) {
      const int j0 = seg_first<8>(s, l8);
      std::uint32_t W[6], sh, L[4];
      window(P8{}, j0, W, sh);
      recur(P8{}, W, sh, d.c, L);
      emit(P8{}, std::true_type{}, j0, L, d.s);
    };
    auto seg4 = [&](const SegIn<4>& d) {
      const int j0 = seg_first<4>(0, l8);
      std::uint32_t W[4], sh, L[2];
      window(P4{}, j0, W, sh);
      recur(P4{}, W, sh, d.c, L);
      emit(P4{}, std::true_type{}, j0, L, d.s);
    };
    // Segments from 2 on (pixels of more than 96 labels): costs, S and the
    // scratch window of every lane (used where prev_wide) loaded one segment
    // ahead.  No load in the loop is conditional and the loop has one exit,
    // so the waits inside it only cover the segment being computed.
    auto wide = [&]() {
      const int segmax = wave_max(act ? segments(num) : 0);
      struct Ahead {
        SegIn<8> d;
        std::uint32_t G[6];
      } A, B;
      auto load = [&](int s, Ahead& x) {
        load_seg8(u, s, x.d);
        std::uint32_t gsh;
        scr_window(P8{}, seg_first<8>(s, l8), x.G, gsh);
      };
      auto run = [&](int s, const Ahead& x) {
        const int j0 = seg_first<8>(s, l8);
        std::uint32_t W[6], sh, L[4];
        lds_window(P8{}, j0, W, sh);
        const std::uint32_t gsh = wpos(8, j0, prev_num);
#pragma unroll
        for (int k = 0; k < 6; ++k) W[k] = prev_wide ? x.G[k] : W[k];
        sh = prev_wide ? gsh : sh;
        recur(P8{}, W, sh, x.d.c, L);
        emit(P8{}, std::false_type{}, j0, L, x.d.s);
      };
      load(2, A);
      for (int s = 2; s < segmax; s += 2) {
        load(s + 1, B);
        run(s, A);
        load(s + 2, A);
        if (s + 1 < segmax) run(s + 1, B);
      }
      // Retire the loads of the segments past the end.
      __builtin_amdgcn_s_waitcnt(kVmcnt0);
    };
    seg4(v.a);
    if (any_seg1) seg8(1, v.b);
    if (any_seg2) wide();
    // Sentinels after the labels (in the LDS row also for finished
    // scanlines, whose rows are not read again).
    if (l8 < kPad / 2) {
      const int w = ((kPad + num) >> 1) + l8;
      if (!cur_wide) cur[w] = 0xFFFFFFFFU;
      if (any_cur_wide && act && cur_wide) scur[w] = 0xFFFFFFFFU;
    }
    prev_min = min8(min(m & 0xFFFFU, m >> 16));
    prev_lo = lo;
    prev_num = num;
    flip ^= 1;
    __builtin_amdgcn_wave_barrier();  // this step's row writes before the next step's reads
  };

  // Register rotation: kDist + 2 sets of pixel values, kDist + 1 of segments,
  // the loop unrolled over a full period so every index is a constant.
  constexpr int NU = kDist + 2, NV = kDist + 1;
  StepIn U[NU];
  StepSegs V[NV];
  static_assert(kDist + 1 < 8, "the prologue reads the first batch only");
  load_batch(0, B);
  write_batch(B);
  load_batch(8, B);
  for (int j = 0; j <= kDist; ++j) read_step(j, U[j]);
  for (int j = 0; j < kDist; ++j) U[j].p2 = lut[wave][step_p2_index(U[j])];
  for (int j = 0; j < kDist; ++j) load_segs(U[j], V[j]);
  auto body = [&](auto Kc, int i) __attribute__((always_inline)) {
    constexpr int K = decltype(Kc)::value;
    step(i, U[K % NU], V[K % NV], U[(K + kDist + 1) % NU], U[(K + kDist) % NU], V[(K + kDist) % NV]);
  };
  for (int i = 0; i < steps; i += NU * NV) {
    [&]<int... K>(std::integer_sequence<int, K...>) __attribute__((always_inline)) {
      (body(std::integral_constant<int, K>{}, i + K), ...);
    }(std::make_integer_sequence<int, NU * NV>{});
  }
  // Retire this group's loads before the next group reuses their registers.
  __builtin_amdgcn_s_waitcnt(kVmcnt0);
  }
}

// ------------------------------------------------------------------ partial sums
// Sum of the path-cost slices of scans that ran side by side (uint16
// wrap-around, so the grouping of the 8 paths does not matter).
struct SumSrc {
  const std::uint16_t* p[4];
};
__global__ __launch_bounds__(256) void k_sum_u16(std::uint16_t* __restrict__ dst, SumSrc src, int nsrc,
                                                 std::uint64_t n8) {
  const std::uint64_t i = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n8) return;
  uint4 acc = reinterpret_cast<const uint4*>(src.p[0])[i];
  for (int k = 1; k < nsrc; ++k) {
    const uint4 v = reinterpret_cast<const uint4*>(src.p[k])[i];
    acc.x = as_u32(as_pair(acc.x) + as_pair(v.x));
    acc.y = as_u32(as_pair(acc.y) + as_pair(v.y));
    acc.z = as_u32(as_pair(acc.z) + as_pair(v.z));
    acc.w = as_u32(as_pair(acc.w) + as_pair(v.w));
  }
  reinterpret_cast<uint4*>(dst)[i] = acc;
}

// ------------------------------------------------------------------ selection
__device__ __forceinline__ float subpixel_offset(int method, float sm, float s0, float sp) {
  float delta = 0.0F;
  if (method == 1) {
    const float denom = sm - 2.0F * s0 + sp;
    delta = denom > 0.0F ? (sm - sp) / (2.0F * denom) : 0.0F;
  } else if (method == 2) {
    const float a = sm - s0, b = sp - s0;
    const float slope = a < b ? b : a;
    delta = slope > 0.0F ? (sm - sp) / (2.0F * slope) : 0.0F;
  }
  return delta < -0.5F ? -0.5F : (0.5F < delta ? 0.5F : delta);
}

// select_disparities (stereo_matcher.cpp:446-525) for one tile row per block,
// four lanes (a quad) per pixel.  Lane q of a quad holds labels
// [32 i + 8 q, 32 i + 8 q + 8) of chunk i, so a quad reads 64 contiguous
// bytes per chunk; the first two chunks stay in registers for the second
// pass of the winner-takes-all (sgm_kernels.cpp:405-426).
__device__ __forceinline__ std::uint32_t quad_min(std::uint32_t v) {
  const int i = static_cast<int>(v);
  v = min(v, static_cast<std::uint32_t>(__builtin_amdgcn_update_dpp(i, i, 0xB1, 0xF, 0xF, false)));
  const int j = static_cast<int>(v);
  return min(v, static_cast<std::uint32_t>(__builtin_amdgcn_update_dpp(j, j, 0x4E, 0xF, 0xF, false)));
}

// Minimum of (value << 16 | label) over the 8 labels of a chunk word group,
// labels k0..k0+7: the lowest value, the first label on ties.
__device__ __forceinline__ std::uint32_t key_min(const uint4& q, int k0, std::uint32_t m) {
  const std::uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const std::uint32_t v = (w[e >> 1] >> ((e & 1) * 16)) & 0xFFFFU;
    m = min(m, (v << 16) | static_cast<std::uint32_t>(k0 + e));
  }
  return m;
}

// Inserts the keys (value << 16 | label) of labels k0..k0+7 into the sorted
// four smallest keys t0 <= t1 <= t2 <= t3.
__device__ __forceinline__ void top4(const uint4& q, int k0, std::uint32_t& t0, std::uint32_t& t1,
                                     std::uint32_t& t2, std::uint32_t& t3) {
  const std::uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    std::uint32_t c = (((w[e >> 1] >> ((e & 1) * 16)) & 0xFFFFU) << 16) | static_cast<std::uint32_t>(k0 + e);
    std::uint32_t a = min(t0, c);
    c = max(t0, c);
    t0 = a;
    a = min(t1, c);
    c = max(t1, c);
    t1 = a;
    a = min(t2, c);
    c = max(t2, c);
    t2 = a;
    t3 = min(t3, c);
  }
}

// Lowest value of labels k0..k0+7 outside [kb - 1, kb + 1].
__device__ __forceinline__ std::uint32_t outside_min(const uint4& q, int k0, int kb, std::uint32_t m) {
  const std::uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const std::uint32_t v = (w[e >> 1] >> ((e & 1) * 16)) & 0xFFFFU;
    const int k = k0 + e;
    m = k + 1 < kb || k > kb + 1 ? min(m, v) : m;
  }
  return m;
}

// Label k of a lane's group of 8 starting at k0 (0xFFFFFFFF when the group
// does not hold k).
__device__ __forceinline__ std::uint32_t group_value(const uint4& q, int k0, int k) {
  const int e = k - k0;
  const std::uint64_t a = (std::uint64_t(q.y) << 32) | q.x, b = (std::uint64_t(q.w) << 32) | q.z;
  const std::uint32_t v = static_cast<std::uint32_t>((e < 4 ? a : b) >> ((e & 3) * 16)) & 0xFFFFU;
  return e >= 0 && e < 8 ? v : 0xFFFFFFFFU;
}

__global__ __launch_bounds__(256) void k_select(const JobDesc* __restrict__ jobs, const TileDesc* __restrict__ tiles,
                                                const std::uint64_t* __restrict__ row_prefix, int ntiles,
                                                const std::uint64_t* __restrict__ bcorr,
                                                const std::uint64_t* __restrict__ pix,
                                                SumSrc S, int nslices, SelectParams p,
                                                float* __restrict__ tdisp, float* __restrict__ tconf,
                                                std::int16_t* __restrict__ tlab) {
  constexpr int kStage = 1024;  // row pixels whose first cell and range are staged in LDS
  __shared__ std::uint64_t scell[kStage];
  __shared__ std::uint32_t srange[kStage];
  extern __shared__ std::uint32_t keys[];
  const std::uint64_t r = blockIdx.x;
  const int t = find_tile_wave(row_prefix, ntiles, r);
  const TileDesc T = tiles[t];
  const JobDesc& J = jobs[T.job];
  const int y = static_cast<int>(r - row_prefix[t]);
  const int Y = T.y0 + y;
  const auto rows = gptr(J.rows4) + 4 * Y;
  const int lf = rows[0], ll = rows[1], rf = rows[2], rl = rows[3];
  const int xb = min(max(lf - T.x0, 0), T.w);
  const int xe = min(max(ll + 1 - T.x0, xb), T.w);
  const std::uint64_t* bc = bcorr + T.band_base + y;
  // Aggregated costs: the sum of the slices of scans that ran side by side.
  auto ld8 = [&](std::uint64_t c) {
    uint4 a = *reinterpret_cast<const uint4*>(S.p[0] + c);
    for (int i = 1; i < nslices; ++i) {
      const uint4 v = *reinterpret_cast<const uint4*>(S.p[i] + c);
      a.x = as_u32(as_pair(a.x) + as_pair(v.x));
      a.y = as_u32(as_pair(a.y) + as_pair(v.y));
      a.z = as_u32(as_pair(a.z) + as_pair(v.z));
      a.w = as_u32(as_pair(a.w) + as_pair(v.w));
    }
    return a;
  };
  auto ld1 = [&](std::uint64_t c) {
    std::uint32_t a = S.p[0][c];
    for (int i = 1; i < nslices; ++i) a += S.p[i][c];
    return a & 0xFFFFU;
  };
  const std::uint64_t px = T.px_base + std::uint64_t(y) * T.w;
  const int D = J.D;
  const int q = threadIdx.x & 3;
  constexpr int kPerPass = 256 / 4;
  const bool staged = T.w <= kStage;
  if (staged) {
    // Every load of a thread before its first LDS store (256 threads).
    constexpr int kIt = kStage / 256;
    std::uint64_t rv[kIt], bv[kIt];
#pragma unroll
    for (int i = 0; i < kIt; ++i) {
      const int x = static_cast<int>(threadIdx.x) + 256 * i;
      rv[i] = x < T.w ? pix[px + x] : 0;
      bv[i] = x < T.w ? bc[std::size_t(x >> T.band_shift) * T.h] : 0;
    }
#pragma unroll
    for (int i = 0; i < kIt; ++i) {
      const int x = static_cast<int>(threadIdx.x) + 256 * i;
      if (x < T.w) {
        scell[x] = bv[i] + pix_off(rv[i]);
        srange[x] = static_cast<std::uint32_t>(rv[i] >> 32);
      }
    }
    __syncthreads();
  }
  // First cell, lo and label count of pixel x.
  auto pixel = [&](int x, int& lo, int& n) {
    std::uint64_t cell;
    std::uint32_t range;
    if (staged) {
      cell = scell[x];
      range = srange[x];
    } else {
      const std::uint64_t rec = pix[px + x];
      cell = bc[std::size_t(x >> T.band_shift) * T.h] + pix_off(rec);
      range = static_cast<std::uint32_t>(rec >> 32);
    }
    lo = static_cast<int>(range & 0xFFFFU);
    n = static_cast<int>(range >> 16);
    return T.cell_base + cell;
  };

  if (p.diagonal_lr) {
    // Right-image WTA along x - label (sgm_kernels.cpp:427-435): the first
    // (in x order) minimum of each diagonal, as a (cost, x) key minimum.
    for (int i = threadIdx.x; i < T.w + D; i += blockDim.x) keys[i] = 0xFFFFFFFFU;
    __syncthreads();
    for (int x = xb + (threadIdx.x >> 2); x < xe; x += kPerPass) {
      int lo, num;
      const std::uint64_t s = pixel(x, lo, num);
      for (int k0 = 8 * q; k0 < num; k0 += 32) {
#pragma unroll
        for (int e = 0; e < 8; ++e)
          atomicMin(&keys[x - (lo + k0 + e) + D - 1], (ld1(s + k0 + e) << 16) | std::uint32_t(x));
      }
    }
    __syncthreads();
  }

  for (int base = 0; base < T.w; base += kPerPass) {
    const int x = base + (threadIdx.x >> 2);
    if (x >= T.w) continue;  // whole quads leave together
    float d = kInf, conf = 0.0F;
    std::int16_t lab = INT16_MIN;
    if (x >= xb && x < xe) {
      const int X = T.x0 + x;
      int lo, n;
      const std::uint64_t s = pixel(x, lo, n);
      // Labels are a multiple of 16 per pixel, so each lane's group of 8 is
      // either wholly inside the pixel's labels or wholly past them.
      const int k0 = 8 * q, k1 = 32 + 8 * q;
      uint4 g0 = make_uint4(0xFFFFFFFFU, 0xFFFFFFFFU, 0xFFFFFFFFU, 0xFFFFFFFFU), g1 = g0;
      if (k0 < n) g0 = ld8(s + k0);
      if (k1 < n) g1 = ld8(s + k1);
      std::uint32_t m = 0xFFFFFFFFU;
      if (k0 < n) m = key_min(g0, k0, m);
      if (k1 < n) m = key_min(g1, k1, m);
      // Labels from 64 on are read once: the lane keeps its four smallest
      // keys among them (t0 <= t1 <= t2 <= t3).  At most three labels are
      // excluded from the second minimum (kb - 1, kb, kb + 1), so the first
      // of these keys outside that interval holds the lane's second minimum
      // over its labels from 64 on.
      std::uint32_t t0 = 0xFFFFFFFFU, t1 = t0, t2 = t0, t3 = t0;
      for (int k = 64 + 8 * q; k < n; k += 32) top4(ld8(s + k), k, t0, t1, t2, t3);
      m = quad_min(min(m, t0));
      const int kb = static_cast<int>(m & 0xFFFFU);
      std::uint32_t second = 0xFFFFU;
      if (k0 < n) second = outside_min(g0, k0, kb, second);
      if (k1 < n) second = outside_min(g1, k1, kb, second);
      auto outside = [&](std::uint32_t t) {
        const int k = static_cast<int>(t & 0xFFFFU);
        return k + 1 < kb || k > kb + 1;
      };
      const std::uint32_t tw = outside(t0) ? t0 : (outside(t1) ? t1 : (outside(t2) ? t2 : t3));
      second = min(second, tw >> 16);
      second = quad_min(second);
      // Costs at kb - 1 and kb + 1 from the registers of the quad when they
      // lie in the first 64 labels.
      const std::uint32_t sm = quad_min(min(group_value(g0, k0, kb - 1), group_value(g1, k1, kb - 1)));
      const std::uint32_t sp = quad_min(min(group_value(g0, k0, kb + 1), group_value(g1, k1, kb + 1)));
      if (q != 0) continue;
      const std::uint32_t bcost = m >> 16;
      const int best = lo + kb;
      const int j = kb;
      const int xr = X - J.dmin - best;
      bool ok = !(xr < rf || xr > rl);
      if (ok && ((j == 0 && lo > 0) || (j == n - 1 && lo + n < D))) ok = false;
      if (ok && p.diagonal_lr) {
        const std::uint32_t key = keys[x - best + D - 1];
        const int rbest = static_cast<int>(key & 0xFFFFU) - (x - best);
        if (abs(rbest - best) > p.lr_max_diff) ok = false;
      }
      const float best_cost = static_cast<float>(bcost);
      if (ok && p.uniqueness > 0.0F && second != 0xFFFFU &&
          best_cost > (1.0F - p.uniqueness) * static_cast<float>(second))
        ok = false;
      if (ok) {
        float delta = 0.0F;
        if (j > 0 && j < n - 1) {
          const std::uint32_t vm = j - 1 < 64 ? sm : ld1(s + j - 1);
          const std::uint32_t vp = j + 1 < 64 ? sp : ld1(s + j + 1);
          delta = subpixel_offset(p.subpixel, static_cast<float>(vm), best_cost, static_cast<float>(vp));
        }
        d = static_cast<float>(J.dmin + best) + delta;
        conf = second != 0xFFFFU && second > 0 ? 1.0F - best_cost / static_cast<float>(second) : 0.0F;
        lab = static_cast<std::int16_t>(J.dmin + best);
      }
    } else if (q != 0) {
      continue;
    }
    tdisp[px + x] = d;
    if (J.confidence) tconf[px + x] = conf;
    if (tlab != nullptr) tlab[px + x] = lab;
  }
}

// ------------------------------------------------------------------ blending
// axis_weight (stereo_matcher.cpp:530-543).
__device__ __forceinline__ float axis_weight(int v, int c0, int c1, int size, int blend_px) {
  const float blend = static_cast<float>(blend_px);
  float w = 1.0F;
  const float pos = static_cast<float>(v) + 0.5F;
  auto clamp01 = [](float a) { return a < 0.0F ? 0.0F : (1.0F < a ? 1.0F : a); };
  if (c0 > 0) {
    const float a = blend_px > 0 ? clamp01((pos - static_cast<float>(c0) + 0.5F * blend) / blend) : (v >= c0 ? 1.0F : 0.0F);
    w = a < w ? a : w;
  }
  if (c1 < size) {
    const float a = blend_px > 0 ? clamp01((static_cast<float>(c1) + 0.5F * blend - pos) / blend) : (v < c1 ? 1.0F : 0.0F);
    w = a < w ? a : w;
  }
  return w;
}

// First grid index whose extent end exceeds v (extents sorted by start and
// end); e(k) is extent k's end.
template <typename E>
__device__ __forceinline__ int first_covering(E e, int n, int v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (e(mid) > v) {
      hi = mid;
    } else {
      lo = mid + 1;
    }
  }
  return lo;
}

// Blend of one job's tiles (stereo_matcher.cpp:688-766): per pixel, the tiles
// that cover it in tile-index order, the same float operations as the
// reference (weighted mean if the valid weight is at least half of the total,
// the dominant tile's value if the mean is more than 1 away from it).  The
// block copies the grid's extents into LDS first (grids of at most kGrid
// tiles per axis), so a pixel's searches wait on one coalesced load.
__global__ __launch_bounds__(256) void k_blend(GridDesc g, const TileDesc* __restrict__ tiles,
                                               const float* __restrict__ tdisp, const float* __restrict__ tconf,
                                               const std::int16_t* __restrict__ tlab, float* __restrict__ disp,
                                               float* __restrict__ conf, std::int16_t* __restrict__ lab) {
  constexpr int kGrid = 256;
  __shared__ std::int32_t scols[4 * kGrid], srows[4 * kGrid];
  const bool staged = g.nx <= kGrid && g.ny <= kGrid;
  if (staged) {
    const auto gc = gptr(g.cols), gr = gptr(g.rows);
    for (int k = threadIdx.x; k < 4 * g.nx; k += blockDim.x) scols[k] = gc[k];
    for (int k = threadIdx.x; k < 4 * g.ny; k += blockDim.x) srows[k] = gr[k];
  }
  __syncthreads();
  const int X = blockIdx.x * blockDim.x + threadIdx.x, Y = blockIdx.y;
  if (X >= g.width) return;
  float wall = 0.0F, wsum = 0.0F, dsum = 0.0F, csum = 0.0F, wbest = 0.0F, dbest = kInf;
  std::int16_t lbest = INT16_MIN;
  // cols(i, k) / rows(j, k): entry k of extent i / j.
  auto blend = [&](auto cols, auto rows) {
    const int j0 = first_covering([&](int k) { return rows(k, 1); }, g.ny, Y);
    const int i0 = first_covering([&](int k) { return cols(k, 1); }, g.nx, X);
    for (int j = j0; j < g.ny && rows(j, 0) <= Y; ++j) {
      const float wy = axis_weight(Y, rows(j, 2), rows(j, 3), g.height, g.blend);
      if (wy <= 0.0F) continue;
      for (int i = i0; i < g.nx && cols(i, 0) <= X; ++i) {
        const float wgt = wy * axis_weight(X, cols(i, 2), cols(i, 3), g.width, g.blend);
        if (wgt <= 0.0F) continue;
        const TileDesc& T = tiles[g.tile_base + j * g.nx + i];
        const std::uint64_t ti = T.px_base + std::uint64_t(Y - T.y0) * T.w + (X - T.x0);
        wall += wgt;
        const float dv = tdisp[ti];
        if (__builtin_isfinite(dv)) {
          wsum += wgt;
          dsum += wgt * dv;
          if (conf != nullptr) csum += wgt * tconf[ti];
          if (wgt > wbest) {
            wbest = wgt;
            dbest = dv;
            if (lab != nullptr) lbest = tlab[ti];
          }
        }
      }
    }
  };
  if (staged) {
    blend([&](int i, int k) { return scols[4 * i + k]; }, [&](int j, int k) { return srows[4 * j + k]; });
  } else {
    const auto gc = gptr(g.cols), gr = gptr(g.rows);
    blend([&](int i, int k) { return gc[4 * i + k]; }, [&](int j, int k) { return gr[4 * j + k]; });
  }
  float dv = kInf;
  float cv = 0.0F;
  std::int16_t lv = INT16_MIN;
  if (wsum > 0.0F && wsum >= 0.5F * wall) {
    dv = dsum / wsum;
    if (fabsf(dv - dbest) > 1.0F) dv = dbest;
    cv = csum / wsum;
    lv = lbest;
  }
  const std::int64_t o = std::int64_t(Y) * g.width + X;
  disp[o] = dv;
  if (conf != nullptr) conf[o] = cv;
  if (lab != nullptr) lab[o] = lv;
}

__global__ void k_pack_rows(const std::int32_t* __restrict__ base, const std::int32_t* __restrict__ partner, int h,
                            std::int32_t* __restrict__ rows4) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= h) return;
  reinterpret_cast<int4*>(rows4)[y] = make_int4(base[2 * y], base[2 * y + 1], partner[2 * y], partner[2 * y + 1]);
}

// Blocks of a kernel that fit on the device at once.
template <typename K>
std::uint64_t resident_blocks(K kernel) {
  int dev = 0, cus = 0, per_cu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64 * kWaves, 0);
  return std::uint64_t(std::max(cus, 1)) * std::uint64_t(std::max(per_cu, 1));
}

template <int DX, int DY, bool kAcc>
std::uint64_t path_blocks() {
  static const std::uint64_t n = resident_blocks(k_path<DX, DY, kAcc>);
  return n;
}

template <int DX, int DY>
void path_launch(const JobDesc* jobs, const TileDesc* tiles, const std::uint64_t* gp, int ntiles,
                 std::uint64_t groups, const std::uint64_t* rs, const std::uint64_t* pix, const std::uint8_t* C,
                 std::uint16_t* S, std::uint16_t* scratch, int scratch_row, std::uint32_t* queue, std::uint64_t slots,
                 bool accumulate, hipStream_t s) {
  // Persistent waves: at most `slots` (scratch rows), at most one per group.
  const std::uint64_t waves = std::min(groups, slots);
  if (accumulate) {
    const std::uint64_t blocks = std::min((waves + kWaves - 1) / kWaves, path_blocks<DX, DY, true>());
    k_path<DX, DY, true><<<dim3(static_cast<unsigned>(blocks)), 64 * kWaves, 0, s>>>(
        jobs, tiles, gp, ntiles, groups, rs, pix, C, S, scratch, scratch_row, queue);
  } else {
    const std::uint64_t blocks = std::min((waves + kWaves - 1) / kWaves, path_blocks<DX, DY, false>());
    k_path<DX, DY, false><<<dim3(static_cast<unsigned>(blocks)), 64 * kWaves, 0, s>>>(
        jobs, tiles, gp, ntiles, groups, rs, pix, C, S, scratch, scratch_row, queue);
  }
}

}  // namespace

void launch_tile_row_scan(const JobDesc* jobs, const TileDesc* tiles, const std::uint64_t* row_prefix, int ntiles,
                          std::uint64_t total_rows, std::uint64_t* pix, std::uint64_t* row_cells,
                          std::uint32_t* row_max, std::uint64_t* bstart, hipStream_t s) {
  if (total_rows == 0) return;
  k_tile_row_scan<<<dim3(static_cast<unsigned>((total_rows + 3) / 4)), 256, 0, s>>>(jobs, tiles, row_prefix, ntiles,
                                                                                    total_rows, pix, row_cells, row_max,
                                                                                    bstart);
}

void launch_tile_cells(const TileDesc* tiles, int ntiles, const std::uint64_t* row_cells, const std::uint32_t* row_max,
                       const std::uint64_t* bstart, std::uint64_t* bcorr, std::uint64_t* tile_cells,
                       std::uint32_t* tile_max, hipStream_t s) {
  k_tile_cells<<<dim3(ntiles), kCellsThreads, 0, s>>>(tiles, row_cells, row_max, bstart, bcorr, tile_cells, tile_max);
}

void launch_cost(const JobDesc* jobs, const TileDesc* tiles, const std::uint64_t* row_prefix, int ntiles,
                 std::uint64_t total_rows, const std::uint64_t* bcorr, const std::uint64_t* pix, std::uint8_t* C,
                 hipStream_t s) {
  if (total_rows == 0) return;
  k_cost<<<dim3(static_cast<unsigned>(total_rows)), kCostThreads, 0, s>>>(jobs, tiles, row_prefix, ntiles, bcorr, pix, C);
}

int path_scratch_row(int max_num) { return kPad + (max_num + 15) / 16 * 16 + kPad; }

std::uint64_t path_slots() {
  std::uint64_t b = path_blocks<1, 0, true>();
  b = std::max(b, path_blocks<1, 0, false>());
  b = std::max(b, path_blocks<0, 1, true>());
  b = std::max(b, path_blocks<1, 1, true>());
  b = std::max(b, path_blocks<-1, 1, true>());
  return b * kWaves;
}

std::size_t path_scratch_elems(std::uint64_t slots, int max_num) {
  return std::size_t(slots) * kScanPerWave * 2 * std::size_t(path_scratch_row(max_num));
}

void launch_sgm_path(int path, const JobDesc* jobs, const TileDesc* tiles, const std::uint64_t* gp, int ntiles,
                     std::uint64_t groups, const std::uint64_t* rs, const std::uint64_t* pix, const std::uint8_t* C,
                     std::uint16_t* S, std::uint16_t* scratch, std::uint64_t slots, int max_num,
                     std::uint32_t* queue, bool accumulate, hipStream_t s) {
  if (groups == 0) return;
  const int srow = path_scratch_row(max_num);
  switch (path) {
    case 0: path_launch<1, 0>(jobs, tiles, gp, ntiles, groups, rs, pix, C, S, scratch, srow, queue, slots, accumulate, s); break;
    case 1: path_launch<-1, 0>(jobs, tiles, gp, ntiles, groups, rs, pix, C, S, scratch, srow, queue, slots, accumulate, s); break;
    case 2: path_launch<0, 1>(jobs, tiles, gp, ntiles, groups, rs, pix, C, S, scratch, srow, queue, slots, accumulate, s); break;
    case 3: path_launch<0, -1>(jobs, tiles, gp, ntiles, groups, rs, pix, C, S, scratch, srow, queue, slots, accumulate, s); break;
    case 4: path_launch<1, 1>(jobs, tiles, gp, ntiles, groups, rs, pix, C, S, scratch, srow, queue, slots, accumulate, s); break;
    case 5: path_launch<-1, -1>(jobs, tiles, gp, ntiles, groups, rs, pix, C, S, scratch, srow, queue, slots, accumulate, s); break;
    case 6: path_launch<-1, 1>(jobs, tiles, gp, ntiles, groups, rs, pix, C, S, scratch, srow, queue, slots, accumulate, s); break;
    default: path_launch<1, -1>(jobs, tiles, gp, ntiles, groups, rs, pix, C, S, scratch, srow, queue, slots, accumulate, s); break;
  }
}

void launch_select(const JobDesc* jobs, const TileDesc* tiles, const std::uint64_t* row_prefix, int ntiles,
                   std::uint64_t total_rows, const std::uint64_t* row_start, const std::uint64_t* col_off,
                   const std::uint16_t* const* S, int nslices, SelectParams p, int max_w, int max_D, float* tdisp,
                   float* tconf, std::int16_t* tlab, hipStream_t s) {
  if (total_rows == 0 || nslices < 1 || nslices > 4) return;
  SumSrc ss{};
  for (int k = 0; k < nslices; ++k) ss.p[k] = S[k];
  const std::size_t lds = p.diagonal_lr ? std::size_t(max_w + max_D) * sizeof(std::uint32_t) : 0;
  k_select<<<dim3(static_cast<unsigned>(total_rows)), 256, lds, s>>>(jobs, tiles, row_prefix, ntiles, row_start,
                                                                      col_off, ss, nslices, p, tdisp, tconf, tlab);
}

void launch_sum_u16(std::uint16_t* dst, const std::uint16_t* const* src, int nsrc, std::uint64_t n, hipStream_t s) {
  const std::uint64_t n8 = (n + 7) / 8;
  if (n8 == 0 || nsrc <= 0 || nsrc > 4) return;
  SumSrc ss{};
  for (int k = 0; k < nsrc; ++k) ss.p[k] = src[k];
  k_sum_u16<<<dim3(static_cast<unsigned>((n8 + 255) / 256)), 256, 0, s>>>(dst, ss, nsrc, n8);
}

void launch_pack_rows(const std::int32_t* base, const std::int32_t* partner, int h, std::int32_t* rows4,
                      hipStream_t s) {
  k_pack_rows<<<dim3((h + 255) / 256), 256, 0, s>>>(base, partner, h, rows4);
}

void launch_blend(GridDesc g, const TileDesc* tiles, const float* tdisp, const float* tconf, const std::int16_t* tlab,
                  float* disp, float* conf, std::int16_t* lab, hipStream_t s) {
  k_blend<<<dim3((g.width + 255) / 256, g.height), 256, 0, s>>>(g, tiles, tdisp, tconf, tlab, disp, conf, lab);
}

}  // namespace r3dg
