// gfx950 kernels of the per-pixel stages around the aggregation: Census,
// pyramid, masks, left-right check, speckle / median filters and the
// coarse-to-fine label ranges.  All are HBM-bound integer or exact float work
// with one thread per output pixel and row-coalesced accesses.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>
#include <type_traits>

#include "kernels.hpp"

namespace r3dg {
namespace {

constexpr float kInf = __builtin_huge_valf();

// std::min / std::max of the reference (first argument wins ties).
__device__ __forceinline__ float smin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ bool finite(float v) { return __builtin_isfinite(v); }

inline dim3 grid2(int w, int h, dim3 b) { return dim3((w + b.x - 1) / b.x, (h + b.y - 1) / b.y); }

// ------------------------------------------------------------------ Census
// 9x7 window (census.cpp:20-35, 104-113): raster-order neighbours, centre
// skipped, bit 61 - k set when neighbour k < centre; rows clamped, columns
// replicated (census.cpp:137-155).  A 64x4 block stages its (64+8)x(4+6) u16
// halo in LDS with coalesced row loads and writes one u64 per thread.
constexpr int kCw = 64, kCh = 4;
// A thread computes two horizontally adjacent pixels: a block covers
// (2 kCw) x kCh pixels, and each row of the two 9-wide windows (10 columns)
// is five aligned 32-bit LDS reads.
__global__ __launch_bounds__(256) void k_census(const std::uint16_t* __restrict__ img, std::int64_t is, int w, int h,
                                                std::uint64_t* __restrict__ out, std::int64_t os) {
  constexpr int kTw = 2 * kCw + 8;  // staged columns (even: rows stay 4-byte aligned)
  __shared__ __attribute__((aligned(4))) std::uint16_t t[kCh + 6][kTw];
  const int bx = blockIdx.x * 2 * kCw, by = blockIdx.y * kCh;
  const int tid = threadIdx.y * kCw + threadIdx.x;
  for (int i = tid; i < (kCh + 6) * kTw; i += 256) {
    const int ly = i / kTw, lx = i - ly * kTw;
    const int gy = min(max(by + ly - 3, 0), h - 1);
    const int gx = min(max(bx + lx - 4, 0), w - 1);
    t[ly][lx] = img[gy * is + gx];
  }
  __syncthreads();
  const int x = bx + 2 * threadIdx.x, y = by + threadIdx.y;
  if (x >= w || y >= h) return;
  const int tx = threadIdx.x, ty = threadIdx.y;
  // Columns 2 tx .. 2 tx + 9 of window row j: pixel x uses 0..8, x + 1 uses 1..9.
  auto row = [&](int j, std::uint32_t* v) {
    const std::uint32_t* r = reinterpret_cast<const std::uint32_t*>(&t[ty + j][0]) + tx;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const std::uint32_t u = r[k];
      v[2 * k] = u & 0xFFFFU;
      v[2 * k + 1] = u >> 16;
    }
  };
  std::uint32_t c[10];
  row(3, c);
  const std::uint32_t ca = c[4], cb = c[5];
  std::uint32_t ahi = 0, alo = 0, bhi = 0, blo = 0;  // bits 61..32 and 31..0
  int k = 0;
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    std::uint32_t v[10];
    row(j, v);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      if (j == 3 && i == 4) continue;
      const std::uint32_t a = v[i] < ca ? 1U : 0U, b = v[i + 1] < cb ? 1U : 0U;
      if (k < 30) {
        ahi = (ahi << 1) | a;
        bhi = (bhi << 1) | b;
      } else {
        alo = (alo << 1) | a;
        blo = (blo << 1) | b;
      }
      ++k;
    }
  }
  out[y * os + x] = (std::uint64_t(ahi) << 32) | alo;
  if (x + 1 < w) out[y * os + x + 1] = (std::uint64_t(bhi) << 32) | blo;
}

// ------------------------------------------------------------------ pyramid
// pyramid.cpp:34-85: [1 4 6 4 1] horizontally into uint32, then vertically,
// (sum + 128) >> 8, replicated borders.
template <typename T>
__global__ void k_pyr_h(const T* __restrict__ in, std::int64_t is, int w, int h, int dw,
                        std::uint32_t* __restrict__ tmp) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= dw) return;
  const T* r = in + y * is;
  const int sx = 2 * x;
  const std::uint32_t a = r[max(sx - 2, 0)], b = r[max(sx - 1, 0)], c = r[sx], d = r[min(sx + 1, w - 1)],
                      e = r[min(sx + 2, w - 1)];
  tmp[std::int64_t(y) * dw + x] = a + e + 4U * (b + d) + 6U * c;
}

template <typename T>
__global__ void k_pyr_v(const std::uint32_t* __restrict__ tmp, int h, int dw, int dh, T* __restrict__ out,
                        std::int64_t os, int threshold) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= dw) return;
  const int sy = 2 * y;
  auto at = [&](int yy) { return tmp[std::int64_t(yy) * dw + x]; };
  const std::uint32_t s = at(max(sy - 2, 0)) + at(min(sy + 2, h - 1)) + 4U * (at(max(sy - 1, 0)) + at(min(sy + 1, h - 1))) +
                          6U * at(sy);
  std::uint32_t v = (s + 128U) >> 8U;
  if (threshold) v = v >= 128U ? 255U : 0U;  // reduce_mask, stereo_matcher.cpp:892-900
  out[y * os + x] = static_cast<T>(v);
}

// pyramid.cpp:34-85 for float images: float sums, scaled by 1/256.
__global__ void k_pyr_hf(const float* __restrict__ in, std::int64_t is, int w, int dw, float* __restrict__ tmp) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= dw) return;
  const float* r = in + y * is;
  const int sx = 2 * x;
  const float a = r[max(sx - 2, 0)], b = r[max(sx - 1, 0)], c = r[sx], d = r[min(sx + 1, w - 1)],
              e = r[min(sx + 2, w - 1)];
  tmp[std::int64_t(y) * dw + x] = a + e + 4.0F * (b + d) + 6.0F * c;
}

__global__ void k_pyr_vf(const float* __restrict__ tmp, int h, int dw, int dh, float* __restrict__ out,
                         std::int64_t os) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= dw) return;
  const int sy = 2 * y;
  auto at = [&](int yy) { return tmp[std::int64_t(yy) * dw + x]; };
  const float s = at(max(sy - 2, 0)) + at(min(sy + 2, h - 1)) + 4.0F * (at(max(sy - 1, 0)) + at(min(sy + 1, h - 1))) +
                  6.0F * at(sy);
  out[y * os + x] = s * (1.0F / 256.0F);
}


// valid_intervals (stereo_matcher.cpp:43-67): first and last column with
// mask >= 128; {0, -1} for a row without one.
__global__ __launch_bounds__(256) void k_valid_intervals(const std::uint8_t* __restrict__ mask, std::int64_t ms, int w,
                                                         std::int32_t* __restrict__ rows) {
  const int y = blockIdx.x;
  int first = w, last = -1;
  if (mask == nullptr) {
    first = 0;
    last = w - 1;
  } else {
    const std::uint8_t* m = mask + y * ms;
    for (int x = threadIdx.x; x < w; x += 256) {
      if (m[x] >= 128) {
        first = min(first, x);
        last = max(last, x);
      }
    }
  }
  __shared__ int sf[256], sl[256];
  sf[threadIdx.x] = first;
  sl[threadIdx.x] = last;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      sf[threadIdx.x] = min(sf[threadIdx.x], sf[threadIdx.x + k]);
      sl[threadIdx.x] = max(sl[threadIdx.x], sl[threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const bool any = sl[0] >= 0;
    rows[2 * y] = any ? sf[0] : 0;
    rows[2 * y + 1] = any ? sl[0] : -1;
  }
}

// The row intervals of one level in one launch, one block per row: those of
// the left and right masks (as k_valid_intervals), with the mirrored pair's
// (the mirror of a row's valid columns [f, l] is [w - 1 - l, w - 1 - f]; a
// row without one stays {0, -1}) and the packed (base, partner) rows of both
// matching directions.  mbase == nullptr: no mirrored pair.
__global__ __launch_bounds__(256) void k_level_rows(const std::uint8_t* __restrict__ ml, std::int64_t msl, int wl,
                                                   const std::uint8_t* __restrict__ mr, std::int64_t msr, int wr,
                                                   std::int32_t* __restrict__ lrows, std::int32_t* __restrict__ rrows,
                                                   std::int32_t* __restrict__ mbase, std::int32_t* __restrict__ mpartner,
                                                   std::int32_t* __restrict__ rows4a, std::int32_t* __restrict__ rows4b) {
  const int y = blockIdx.x;
  auto scan = [&](const std::uint8_t* mask, std::int64_t ms, int w, int& first, int& last) {
    first = w;
    last = -1;
    if (mask == nullptr) {
      first = 0;
      last = w - 1;
      return;
    }
    const std::uint8_t* m = mask + y * ms;
    for (int x = threadIdx.x; x < w; x += 256) {
      if (m[x] >= 128) {
        first = min(first, x);
        last = max(last, x);
      }
    }
  };
  int lf, ll, rf, rl;
  scan(ml, msl, wl, lf, ll);
  scan(mr, msr, wr, rf, rl);
  __shared__ int s[4][256];
  s[0][threadIdx.x] = lf;
  s[1][threadIdx.x] = ll;
  s[2][threadIdx.x] = rf;
  s[3][threadIdx.x] = rl;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      s[0][threadIdx.x] = min(s[0][threadIdx.x], s[0][threadIdx.x + k]);
      s[1][threadIdx.x] = max(s[1][threadIdx.x], s[1][threadIdx.x + k]);
      s[2][threadIdx.x] = min(s[2][threadIdx.x], s[2][threadIdx.x + k]);
      s[3][threadIdx.x] = max(s[3][threadIdx.x], s[3][threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const bool lany = s[1][0] >= 0, rany = s[3][0] >= 0;
    lf = lany ? s[0][0] : 0;
    ll = lany ? s[1][0] : -1;
    rf = rany ? s[2][0] : 0;
    rl = rany ? s[3][0] : -1;
    lrows[2 * y] = lf;
    lrows[2 * y + 1] = ll;
    rrows[2 * y] = rf;
    rrows[2 * y + 1] = rl;
    reinterpret_cast<int4*>(rows4a)[y] = make_int4(lf, ll, rf, rl);
    if (mbase != nullptr) {
      const int bf = rany ? wr - 1 - rl : 0, bl = rany ? wr - 1 - rf : -1;
      const int pf = lany ? wl - 1 - ll : 0, pl = lany ? wl - 1 - lf : -1;
      mbase[2 * y] = bf;
      mbase[2 * y + 1] = bl;
      mpartner[2 * y] = pf;
      mpartner[2 * y + 1] = pl;
      reinterpret_cast<int4*>(rows4b)[y] = make_int4(bf, bl, pf, pl);
    }
  }
}

template <typename T>
__global__ void k_fill(T* p, std::size_t n, T v) {
  const std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ left-right check
__global__ void k_unmirror_right(const float* __restrict__ mirrored, int wr, int h, float off,
                                 float* __restrict__ right) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= wr) return;
  right[std::int64_t(y) * wr + x] = mirrored[std::int64_t(y) * wr + (wr - 1 - x)] - off;
}

// left_right_differ (stereo_matcher.cpp:777-782): other column
// lround(x + sign * d); invalid if outside, not finite or |d - d_other| > max_diff.
__global__ void k_lr_check(float* __restrict__ map, int w, int h, const float* __restrict__ other, int wo, float sign,
                           float max_diff, std::uint8_t* __restrict__ bad, int apply) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w) return;
  const std::int64_t i = std::int64_t(y) * w + x;
  const float d = map[i];
  bool differ = false;
  if (finite(d)) {
    const float pos = static_cast<float>(x) + sign * d;
    const int xo = static_cast<int>(roundf(pos));  // lround: halves away from zero
    const float dother = xo >= 0 && xo < wo ? other[std::int64_t(y) * wo + xo] : kInf;
    differ = !finite(dother) || fabsf(d - dother) > max_diff;
  }
  if (bad != nullptr) bad[i] = differ ? 1 : 0;
  if (apply && differ) map[i] = kInf;
}

__global__ void k_apply_mask(float* __restrict__ map, std::size_t n, const std::uint8_t* __restrict__ bad) {
  const std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n && bad[i] != 0) map[i] = kInf;
}

// ------------------------------------------------------------------ speckle filter
// remove_speckles (stereo_matcher.cpp:919-970) is a BFS whose segments are the
// connected components of the 4-neighbour graph with edges
// |d_j - d_i| <= max_diff (the relation is symmetric, so the segments do not
// depend on the visiting order).  Here: union-find in three steps, every
// parent pointer only ever moving to a smaller pixel index (no cycles):
//   1. each 32x32 block labels its horizontal runs with a wave ballot,
//      unions the runs vertically in LDS and points its pixels to the
//      block-local root;
//   2. pixels on block borders union across them in HBM (atomicMin links,
//      Playne & Hawick 2018, with path halving), except where the
//      neighbouring border edge already joins the same two runs;
//   3. every pixel finds its root and counts it, saturating at min_size (only
//      "fewer than min_size" matters, and big segments would serialise on one
//      counter), then small segments are invalidated.
constexpr int kCcW = 32, kCcH = 32;
static_assert(64 % kCcW == 0, "the local phase labels runs per wave: block rows must not straddle waves");

__device__ __forceinline__ int ld_relaxed(int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int cc_find(int* p, int i) {
  int par = ld_relaxed(&p[i]);
  while (par != i) {
    const int gp = ld_relaxed(&p[par]);
    if (gp != par) atomicMin(&p[i], gp);  // path halving: skip to the grandparent
    i = gp;
    par = ld_relaxed(&p[i]);
  }
  return i;
}

__device__ void cc_union(int* p, int a, int b) {
  for (;;) {
    a = cc_find(p, a);
    b = cc_find(p, b);
    if (a == b) return;
    if (a > b) {
      const int t = a;
      a = b;
      b = t;
    }
    const int old = atomicMin(&p[b], a);
    if (old == b) return;
    b = old;
  }
}

__device__ __forceinline__ int lds_find(int* lab, int i) {
  int par = __hip_atomic_load(&lab[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  while (par != i) {
    i = par;
    par = __hip_atomic_load(&lab[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  return i;
}

__device__ void lds_union(int* lab, int a, int b) {
  for (;;) {
    a = lds_find(lab, a);
    b = lds_find(lab, b);
    if (a == b) return;
    if (a > b) {
      const int t = a;
      a = b;
      b = t;
    }
    const int old = atomicMin(&lab[b], a);
    if (old == b) return;
    b = old;
  }
}

__global__ __launch_bounds__(1024) void k_cc_local(const float* __restrict__ d, int w, int h, float max_diff,
                                                  int* __restrict__ p, int* __restrict__ size) {
  __shared__ int lab[kCcW * kCcH];
  __shared__ float dv[kCcW * kCcH];
  const int tx = threadIdx.x, ty = threadIdx.y, t = ty * kCcW + tx;
  const int x = blockIdx.x * kCcW + tx, y = blockIdx.y * kCcH + ty;
  const bool in = x < w && y < h;
  dv[t] = in ? d[std::int64_t(y) * w + x] : __builtin_nanf("");
  __syncthreads();
  // Runs of horizontally joined pixels first: a wave holds two block rows,
  // and a pixel's parent is the first pixel of its run (the last lane at or
  // before it that does not join its left neighbour; lane 0 of each row
  // never does).
  const float v = dv[t];
  const bool left = in && tx > 0 && fabsf(v - dv[t - 1]) <= max_diff;
  const int lane = t & 63;
  const std::uint64_t starts = __builtin_amdgcn_ballot_w64(!left) & (~0ULL >> (63 - lane));
  lab[t] = t - lane + (63 - __builtin_clzll(starts));
  __syncthreads();
  // Vertical joins, skipped where the pixels to the left are joined to these
  // two and to each other (that union already holds both runs).
  if (in && ty + 1 < kCcH && y + 1 < h && fabsf(dv[t + kCcW] - v) <= max_diff) {
    const bool implied = left && fabsf(dv[t + kCcW] - dv[t + kCcW - 1]) <= max_diff &&
                         fabsf(dv[t + kCcW - 1] - dv[t - 1]) <= max_diff;
    if (!implied) lds_union(lab, t, t + kCcW);
  }
  __syncthreads();
  if (in) {
    const int r = lds_find(lab, t);
    const std::int64_t i = std::int64_t(y) * w + x;
    p[i] = (blockIdx.y * kCcH + r / kCcW) * w + blockIdx.x * kCcW + r % kCcW;
    size[i] = 0;
  }
}

__global__ __launch_bounds__(1024) void k_cc_border(const float* __restrict__ d, int w, int h, float max_diff, int* p) {
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int x = blockIdx.x * kCcW + tx, y = blockIdx.y * kCcH + ty;
  if (x >= w || y >= h) return;
  const int i = y * w + x;
  auto joined = [&](int a, int b) { return fabsf(d[b] - d[a]) <= max_diff; };  // b right of or below a
  // A border edge is skipped where the same edge one row up (right border)
  // or one column left (bottom border) joins the same two blocks and the
  // in-block edges between them are joined too: the partition already holds
  // it.  The in-block edges are exact after the local phase and each chain of
  // skips ends at the block's first row or column, so nothing is lost.
  if (tx == kCcW - 1 && x + 1 < w && joined(i, i + 1)) {
    const bool implied = ty > 0 && joined(i - w, i + 1 - w) && joined(i - w, i) && joined(i + 1 - w, i + 1);
    if (!implied) cc_union(p, i, i + 1);
  }
  if (ty == kCcH - 1 && y + 1 < h && joined(i, i + w)) {
    const bool implied = tx > 0 && joined(i - 1, i) && joined(i - 1 + w, i + w) && joined(i - 1, i - 1 + w);
    if (!implied) cc_union(p, i, i + w);
  }
}

// Component sizes, counted up to min_size: the lanes of a wave that share a
// root add their count with one atomic.  With a finite max_diff, non-finite
// pixels are components of their own (|inf - v| > max_diff) and stay at 0,
// below any min_size > 1, as their count of 1 would; with an infinite one
// they can join a segment and are counted.
__global__ __launch_bounds__(256) void k_cc_count(const float* __restrict__ d, int* p, int* __restrict__ size, int n,
                                                  int min_size, bool count_nonfinite) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int r = -1;
  bool live = false;
  if (i < n) {
    r = cc_find(p, i);
    __hip_atomic_store(&p[i], r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    live = count_nonfinite || finite(d[i]);
  }
  std::uint64_t todo = __builtin_amdgcn_ballot_w64(live);
  while (todo != 0) {
    const int leader = __builtin_ctzll(todo);
    const int lr = __shfl(r, leader);
    const std::uint64_t same = __builtin_amdgcn_ballot_w64(live && r == lr) & todo;
    if (lane == leader && ld_relaxed(&size[lr]) < min_size) atomicAdd(&size[lr], __builtin_popcountll(same));
    todo &= ~same;
  }
}

__global__ void k_cc_apply(float* __restrict__ d, const int* __restrict__ p, const int* __restrict__ size, int n,
                           int min_size) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && size[p[i]] < min_size) d[i] = kInf;
}

// ------------------------------------------------------------------ median
// median_filter_3x3 (stereo_matcher.cpp:972-1021): Paeth's median-of-9
// network for an all-valid interior window (median9, :903-915), else the median
// of the valid values (mean of the two middle ones for an even count).
__global__ void k_median3x3(const float* __restrict__ in, int w, int h, float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w) return;
  const float c = in[std::int64_t(y) * w + x];
  if (!finite(c)) {
    out[std::int64_t(y) * w + x] = c;
    return;
  }
  // The window in registers, outside or non-finite neighbours as +inf: after
  // a 9-input sorting network the n valid values are the n smallest, in
  // order, so their median is element n / 2 (and n / 2 - 1 for an even n).
  // For a full interior window this is the value Paeth's network returns.
  float v[9];
  int n = 0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int qx = x + k % 3 - 1, qy = y + k / 3 - 1;
    const bool inside = qx >= 0 && qy >= 0 && qx < w && qy < h;
    const float a = in[std::int64_t(min(max(qy, 0), h - 1)) * w + min(max(qx, 0), w - 1)];
    const bool ok = inside && finite(a);
    v[k] = ok ? a : kInf;
    n += ok ? 1 : 0;
  }
  auto s2 = [&](int i, int j) {
    const float a = v[i], b = v[j];
    v[i] = smin(a, b);
    v[j] = smax(a, b);
  };
  s2(0, 1); s2(3, 4); s2(6, 7); s2(1, 2); s2(4, 5); s2(7, 8); s2(0, 1); s2(3, 4); s2(6, 7);
  s2(0, 3); s2(3, 6); s2(0, 3); s2(1, 4); s2(4, 7); s2(1, 4); s2(2, 5); s2(5, 8); s2(2, 5);
  s2(1, 3); s2(5, 7); s2(2, 6); s2(4, 6); s2(2, 4); s2(2, 3); s2(5, 6);
  const int half = n / 2;
  float m = v[0], below = v[0];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    m = k == half ? v[k] : m;
    below = k == half - 1 ? v[k] : below;
  }
  if (n % 2 == 0) m = 0.5F * (m + below);
  out[std::int64_t(y) * w + x] = m;
}

// ------------------------------------------------------------------ rectification
// rectify_image (rectify/rectification.cpp:288-350), one thread per target
// pixel: the ray (x - cx) / fx, (y - cy) / fy, 1 of the target camera turned
// into the source camera, projected with its lens distortion
// (camera_model.cpp:68-103, OpenMVG formulas), and sampled bicubically
// (sampling.hpp:21-55).  Coordinates in double, weights and sums in float, in
// the reference's order of operations (the library is built without FMA
// contraction).
__device__ __forceinline__ void distort(const RectifyGeom& g, double ux, double uy, double& dx, double& dy) {
  const double r2 = ux * ux + uy * uy;
  const double* k = g.params;
  switch (g.model) {
    case 1: {
      const double f = 1.0 + k[0] * r2;
      dx = ux * f;
      dy = uy * f;
      return;
    }
    case 2: {
      const double f = 1.0 + r2 * (k[0] + r2 * (k[1] + r2 * k[2]));
      dx = ux * f;
      dy = uy * f;
      return;
    }
    case 3: {
      const double radial = r2 * (k[0] + r2 * (k[1] + r2 * k[2]));
      const double tx = k[4] * (r2 + 2.0 * ux * ux) + 2.0 * k[3] * ux * uy;
      const double ty = k[3] * (r2 + 2.0 * uy * uy) + 2.0 * k[4] * ux * uy;
      dx = ux + ux * radial + tx;
      dy = uy + uy * radial + ty;
      return;
    }
    case 4: {
      const double r = sqrt(r2);
      if (r <= 1e-8) break;
      const double theta = atan(r);
      const double t2 = theta * theta;
      const double theta_d = theta * (1.0 + t2 * (k[0] + t2 * (k[1] + t2 * (k[2] + t2 * k[3]))));
      const double f = theta_d / r;
      dx = ux * f;
      dy = uy * f;
      return;
    }
    default:
      break;
  }
  dx = ux;
  dy = uy;
}

// Catmull-Rom weights of the taps -1, 0, 1, 2 around floor(x) (sampling.hpp:21-28).
__device__ __forceinline__ void cubic_weights(float t, float* w) {
  const float t2 = t * t;
  const float t3 = t2 * t;
  w[0] = -0.5F * t3 + t2 - 0.5F * t;
  w[1] = 1.5F * t3 - 2.5F * t2 + 1.0F;
  w[2] = -1.5F * t3 + 2.0F * t2 + 0.5F * t;
  w[3] = 0.5F * t3 - 0.5F * t2;
}

template <typename T>
__device__ __forceinline__ T to_pixel(float v) {
  if constexpr (std::is_floating_point_v<T>) {
    return static_cast<T>(v);
  } else {
    const float r = roundf(v);
    const float hi = static_cast<float>(std::numeric_limits<T>::max());
    return static_cast<T>(r < 0.0F ? 0.0F : (hi < r ? hi : r));
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_rectify(const T* __restrict__ src, std::int64_t ss, RectifyGeom g,
                                                 T* __restrict__ out, std::int64_t os,
                                                 std::uint8_t* __restrict__ valid, std::int64_t vs) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= g.tw) return;
  const double ry = (y - g.tcy) / g.tfy;
  const double vx = (x - g.tcx) / g.tfx;
  const double r0 = g.m[0] * vx + g.m[1] * ry + g.m[2] * 1.0;
  const double r1 = g.m[3] * vx + g.m[4] * ry + g.m[5] * 1.0;
  const double r2 = g.m[6] * vx + g.m[7] * ry + g.m[8] * 1.0;
  bool inside = r2 > 1e-9 && g.sw > 0 && g.sh > 0;  // an empty source has no pixel inside
  double sx = 0.0, sy = 0.0;
  if (inside) {
    double dx, dy;
    distort(g, r0 / r2, r1 / r2, dx, dy);
    sx = g.fx * dx + g.cx;
    sy = g.fy * dy + g.cy;
    inside = sx >= -0.5 && sy >= -0.5 && sx <= g.sw - 0.5 && sy <= g.sh - 0.5;
  }
  T v = T{};
  if (inside) {
    const double fxl = floor(sx), fyl = floor(sy);
    float wx[4], wy[4];
    cubic_weights(static_cast<float>(sx - fxl), wx);
    cubic_weights(static_cast<float>(sy - fyl), wy);
    const int x0 = static_cast<int>(fxl) - 1, y0 = static_cast<int>(fyl) - 1;
    int xs[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) xs[i] = min(max(x0 + i, 0), g.sw - 1);
    float acc_out = 0.0F;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const T* row = src + std::int64_t(min(max(y0 + j, 0), g.sh - 1)) * ss;
      float acc = 0.0F;
#pragma unroll
      for (int i = 0; i < 4; ++i) acc += wx[i] * static_cast<float>(row[xs[i]]);
      acc_out += wy[j] * acc;
    }
    v = to_pixel<T>(acc_out);
  }
  out[std::int64_t(y) * os + x] = v;
  if (valid != nullptr) valid[std::int64_t(y) * vs + x] = inside ? 255 : 0;
}

// ------------------------------------------------------------------ combining pairs
// sample_disparity (multi_stereo.cpp:44-72): bilinear when the four
// neighbours are finite and within one pixel of each other, else the nearest
// pixel; +inf outside.
__device__ float sample_disparity(const float* __restrict__ d, std::int64_t s, int w, int h, double u, double v) {
  if (!(u > -0.5 && v > -0.5 && u < w - 0.5 && v < h - 0.5)) return kInf;
  const int x0 = min(max(static_cast<int>(floor(u)), 0), w - 1);
  const int y0 = min(max(static_cast<int>(floor(v)), 0), h - 1);
  const int x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);
  const float d00 = d[y0 * s + x0], d10 = d[y0 * s + x1], d01 = d[y1 * s + x0], d11 = d[y1 * s + x1];
  if (finite(d00) && finite(d10) && finite(d01) && finite(d11)) {
    const float lo = smin(smin(smin(d00, d10), d01), d11);
    const float hi = smax(smax(smax(d00, d10), d01), d11);
    if (hi - lo <= 1.0F) {
      const double ex = u - x0, ey = v - y0;
      const float fx = static_cast<float>(ex < 0.0 ? 0.0 : (1.0 < ex ? 1.0 : ex));
      const float fy = static_cast<float>(ey < 0.0 ? 0.0 : (1.0 < ey ? 1.0 : ey));
      const float top = d00 + fx * (d10 - d00);
      const float bottom = d01 + fx * (d11 - d01);
      return top + fy * (bottom - top);
    }
  }
  const int xn = min(max(static_cast<int>(lround(u)), 0), w - 1);
  const int yn = min(max(static_cast<int>(lround(v)), 0), h - 1);
  return d[yn * s + xn];
}

// Unit vector a / |a| (Eigen's normalized(): unchanged when |a| = 0).
__device__ __forceinline__ void unit3(double* a) {
  const double n2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
  if (n2 > 0.0) {
    const double n = sqrt(n2);
    a[0] = a[0] / n;
    a[1] = a[1] / n;
    a[2] = a[2] / n;
  }
}

// combine_pairs per grid pixel (multi_stereo.cpp:103-213), double precision
// in the reference's order of operations.
__global__ __launch_bounds__(256) void k_combine(const CombinePair* __restrict__ pairs, CombineGrid g,
                                                 float* __restrict__ depth, std::int64_t ds,
                                                 float* __restrict__ conf_out, std::int64_t cs,
                                                 std::uint16_t* __restrict__ mask_out, std::int64_t ms) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, yy = blockIdx.y;
  if (x >= g.w) return;
  double cw[kMaxCombinePairs], ck[kMaxCombinePairs];
  float cc[kMaxCombinePairs];
  int cp[kMaxCombinePairs], used[kMaxCombinePairs];
  int n = 0;
  for (int i = 0; i < g.npairs; ++i) {
    const CombinePair& P = pairs[i];
    const double rx = P.g[0] * x + P.g[1] * yy + P.g[2] * 1.0;
    const double ry = P.g[3] * x + P.g[4] * yy + P.g[5] * 1.0;
    const double c = P.g[6] * x + P.g[7] * yy + P.g[8] * 1.0;
    if (!(c > 1e-9)) continue;
    const double u = rx / c, v = ry / c;
    const float d = sample_disparity(P.disp, P.ds, P.dw, P.dh, u, v);
    if (!finite(d) || d <= P.offset) continue;
    ck[n] = P.fb / c;
    cw[n] = (d - P.offset) / ck[n];
    cp[n] = i;
    float cf = 0.0F;
    if (P.conf != nullptr) {
      const int xn = min(max(static_cast<int>(lround(u)), 0), P.cw - 1);
      const int yn = min(max(static_cast<int>(lround(v)), 0), P.ch - 1);
      cf = P.conf[yn * P.cs + xn];
    }
    cc[n] = cf;
    ++n;
  }
  float dep = 0.0F, cnf = 0.0F;
  std::uint16_t mask = 0;
  if (n >= g.min_pairs) {
    // The hypothesis with the most agreeing pairs (ties: higher summed confidence).
    int best = -1, best_support = 0;
    float best_conf = -1.0F;
    for (int i = 0; i < n; ++i) {
      int support = 0;
      float sc = 0.0F;
      for (int j = 0; j < n; ++j) {
        if (fabs(ck[j] * (cw[j] - cw[i])) <= g.inlier) {
          ++support;
          sc += cc[j];
        }
      }
      if (support > best_support || (support == best_support && sc > best_conf)) {
        best = i;
        best_support = support;
        best_conf = sc;
      }
    }
    int m = 0;
    for (int j = 0; j < n; ++j)
      if (fabs(ck[j] * (cw[j] - cw[best])) <= g.inlier) used[m++] = j;
    // Weighted least squares on w, dropping the worst pair while a residual is too large.
    double w_opt = 0.0;
    while (m > 0) {
      double sw = 0.0, swx = 0.0;
      for (int t = 0; t < m; ++t) {
        const int q = used[t];
        sw += ck[q] * ck[q];
        swx += ck[q] * ck[q] * cw[q];
      }
      w_opt = swx / sw;
      int worst = -1;
      double worst_res = g.max_res;
      for (int t = 0; t < m; ++t) {
        const int q = used[t];
        const double res = fabs(ck[q] * (cw[q] - w_opt));
        if (res > worst_res) {
          worst_res = res;
          worst = t;
        }
      }
      if (worst < 0) break;
      used[worst] = used[--m];
    }
    const bool majority = m > n - m || (m >= 2 && m >= n - m);
    if (m >= g.min_pairs && majority && w_opt > 0.0) {
      const double z = 1.0 / w_opt;
      const double px[3] = {static_cast<double>(x), static_cast<double>(yy), 1.0};
      double t[3], xw[3], tb[3];
      for (int i = 0; i < 3; ++i) t[i] = g.inv_k[3 * i] * px[0] + g.inv_k[3 * i + 1] * px[1] + g.inv_k[3 * i + 2] * px[2];
      for (int i = 0; i < 3; ++i) {
        const double ray = g.rot_t[3 * i] * t[0] + g.rot_t[3 * i + 1] * t[1] + g.rot_t[3 * i + 2] * t[2];
        xw[i] = g.center[i] + z * ray;
        tb[i] = g.center[i] - xw[i];
      }
      unit3(tb);
      bool wide = false;
      float sc = 0.0F;
      for (int k = 0; k < m; ++k) {
        const int q = used[k];
        mask = static_cast<std::uint16_t>(mask | (1U << cp[q]));
        const CombinePair& P = pairs[cp[q]];
        double tp[3] = {P.pc[0] - xw[0], P.pc[1] - xw[1], P.pc[2] - xw[2]};
        unit3(tp);
        wide = wide || tb[0] * tp[0] + tb[1] * tp[1] + tb[2] * tp[2] <= g.cos_min;
        sc += cc[q];
      }
      if (wide) {
        dep = static_cast<float>(z);
        cnf = sc / static_cast<float>(m);
      } else {
        mask = 0;
      }
    }
  }
  depth[yy * ds + x] = dep;
  conf_out[yy * cs + x] = cnf;
  mask_out[yy * ms + x] = mask;
}

// ------------------------------------------------------------------ label ranges
// ranges_from_coarse (stereo_matcher.cpp:1124-1267).
__global__ void k_rng_h(const float* __restrict__ coarse, int wc, int hc, int r, float* __restrict__ hmin,
                        float* __restrict__ hmax) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= wc) return;
  float mn = kInf, mx = -kInf;
  for (int k = max(0, x - r); k <= min(wc - 1, x + r); ++k) {
    const float v = coarse[std::int64_t(y) * wc + k];
    if (finite(v)) {
      mn = smin(mn, v);
      mx = smax(mx, v);
    }
  }
  hmin[std::int64_t(y) * wc + x] = mn;
  hmax[std::int64_t(y) * wc + x] = mx;
}

__global__ void k_rng_v(const float* __restrict__ hmin, const float* __restrict__ hmax, int wc, int hc, int r,
                        float* __restrict__ cmin, float* __restrict__ cmax) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= wc) return;
  float mn = kInf, mx = -kInf;
  for (int k = max(0, y - r); k <= min(hc - 1, y + r); ++k) {
    mn = smin(mn, hmin[std::int64_t(k) * wc + x]);
    mx = smax(mx, hmax[std::int64_t(k) * wc + x]);
  }
  cmin[std::int64_t(y) * wc + x] = mn;
  cmax[std::int64_t(y) * wc + x] = mx;
}

__global__ void k_rng_pyr(const float* __restrict__ amin, const float* __restrict__ amax, int aw, int ah, int nw,
                          int nh, float* __restrict__ bmin, float* __restrict__ bmax) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= nw) return;
  const int x0 = 2 * x, y0 = 2 * y, x1 = min(2 * x + 1, aw - 1), y1 = min(2 * y + 1, ah - 1);
  auto A = [&](const float* a, int xx, int yy) { return a[std::int64_t(yy) * aw + xx]; };
  bmin[std::int64_t(y) * nw + x] =
      smin(smin(A(amin, x0, y0), A(amin, x1, y0)), smin(A(amin, x0, y1), A(amin, x1, y1)));
  bmax[std::int64_t(y) * nw + x] =
      smax(smax(A(amax, x0, y0), A(amax, x1, y0)), smax(A(amax, x0, y1), A(amax, x1, y1)));
}

struct RangePyr {
  const float* mn[kMaxLevels];
  const float* mx[kMaxLevels];
  int w[kMaxLevels];
  int h[kMaxLevels];
  int levels;
};

// Levels 1 .. n (n <= 6) of the min / max pyramid in one launch.  The
// level-k pixel (x, y) is the min / max over the level-0 pixels [x 2^k,
// (x + 1) 2^k) x [y 2^k, (y + 1) 2^k) that lie in the grid (what the 2 x 2
// reduction of k_rng_pyr, repeating the last column and row, gives level by
// level), so for k <= 6 every pixel lies in one 64 x 64 region of level 0:
// a block reduces one region, level 1 from global memory and the others in
// LDS, where pixels outside the grid hold +inf / -inf and drop out of the
// minimum / maximum.
constexpr int kPyrRegion = 64;
constexpr int kPyrLevels = 6;
struct RangePyrOut {
  float* mn[kPyrLevels + 1];
  float* mx[kPyrLevels + 1];
  int w[kPyrLevels + 1];
  int h[kPyrLevels + 1];
  int n;  // levels 1 .. n are written
};
__global__ __launch_bounds__(256) void k_rng_pyr_all(const float* __restrict__ cmin, const float* __restrict__ cmax,
                                                    RangePyrOut P) {
  constexpr int kS = kPyrRegion / 2;  // level-1 pixels per side of a region
  __shared__ float smn[kS][kS + 1], smx[kS][kS + 1];
  const int tx = static_cast<int>(threadIdx.x) & 15, ty = static_cast<int>(threadIdx.x) >> 4;
  const int rx = static_cast<int>(blockIdx.x) * kPyrRegion, ry = static_cast<int>(blockIdx.y) * kPyrRegion;
  const int w0 = P.w[0], h0 = P.h[0];
  auto A = [&](const float* a, int xx, int yy) { return a[std::int64_t(yy) * w0 + xx]; };
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int lx = 2 * tx + (q & 1), ly = 2 * ty + (q >> 1);
    const int X = (rx >> 1) + lx, Y = (ry >> 1) + ly;
    float mn = __builtin_huge_valf(), mx = -__builtin_huge_valf();
    if (X < P.w[1] && Y < P.h[1]) {
      const int x0 = 2 * X, y0 = 2 * Y, x1 = min(2 * X + 1, w0 - 1), y1 = min(2 * Y + 1, h0 - 1);
      mn = smin(smin(A(cmin, x0, y0), A(cmin, x1, y0)), smin(A(cmin, x0, y1), A(cmin, x1, y1)));
      mx = smax(smax(A(cmax, x0, y0), A(cmax, x1, y0)), smax(A(cmax, x0, y1), A(cmax, x1, y1)));
      P.mn[1][std::int64_t(Y) * P.w[1] + X] = mn;
      P.mx[1][std::int64_t(Y) * P.w[1] + X] = mx;
    }
    smn[ly][lx] = mn;
    smx[ly][lx] = mx;
  }
  for (int k = 2; k <= P.n; ++k) {
    __syncthreads();
    const int side = kPyrRegion >> k;
    const int t = static_cast<int>(threadIdx.x);
    const int lx = t % side, ly = t / side;
    float mn = __builtin_huge_valf(), mx = -__builtin_huge_valf();
    const bool mine = t < side * side;
    if (mine) {
      mn = smin(smin(smn[2 * ly][2 * lx], smn[2 * ly][2 * lx + 1]), smin(smn[2 * ly + 1][2 * lx], smn[2 * ly + 1][2 * lx + 1]));
      mx = smax(smax(smx[2 * ly][2 * lx], smx[2 * ly][2 * lx + 1]), smax(smx[2 * ly + 1][2 * lx], smx[2 * ly + 1][2 * lx + 1]));
    }
    __syncthreads();
    if (mine) {
      const int X = (rx >> k) + lx, Y = (ry >> k) + ly;
      const bool in = X < P.w[k] && Y < P.h[k];
      smn[ly][lx] = in ? mn : __builtin_huge_valf();
      smx[ly][lx] = in ? mx : -__builtin_huge_valf();
      if (in) {
        P.mn[k][std::int64_t(Y) * P.w[k] + X] = mn;
        P.mx[k][std::int64_t(Y) * P.w[k] + X] = mx;
      }
    }
  }
}

// Pixels without a valid coarse disparity take the min / max of the first
// pyramid level with a valid value in the 3x3 block neighbourhood.
__global__ void k_rng_fill(float* __restrict__ cmin, float* __restrict__ cmax, int wc, int hc, RangePyr pyr) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= wc) return;
  const std::int64_t i = std::int64_t(y) * wc + x;
  if (cmin[i] != kInf) return;
  for (int k = 1; k < pyr.levels; ++k) {
    const int bx = x >> k, by = y >> k, aw = pyr.w[k], ah = pyr.h[k];
    float mn = kInf, mx = -kInf;
    for (int qy = max(0, by - 1); qy <= min(ah - 1, by + 1); ++qy) {
      for (int qx = max(0, bx - 1); qx <= min(aw - 1, bx + 1); ++qx) {
        mn = smin(mn, pyr.mn[k][std::int64_t(qy) * aw + qx]);
        mx = smax(mx, pyr.mx[k][std::int64_t(qy) * aw + qx]);
      }
    }
    if (mn != kInf) {
      cmin[i] = mn;
      cmax[i] = mx;
      return;
    }
  }
}

__global__ void k_rng_final(const float* __restrict__ cmin, const float* __restrict__ cmax, int wc, int hc, int width,
                            int height, const std::uint8_t* __restrict__ fine_valid, std::int64_t fvs,
                            const std::int32_t* __restrict__ prow, int dmin, int n_all, int margin, int partner_sign,
                            int mirrored, std::uint16_t* __restrict__ lo_out, std::uint16_t* __restrict__ num_out,
                            std::int64_t rs) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= width) return;
  const int pf = prow[2 * y], pl = prow[2 * y + 1];
  const int cy0 = y / 2, cy1 = min(cy0 + 1, hc - 1), cx0 = x / 2, cx1 = min(cx0 + 1, wc - 1);
  auto C = [&](const float* a, int xx, int yy) { return a[std::int64_t(yy) * wc + xx]; };
  const float mn = smin(smin(C(cmin, cx0, cy0), C(cmin, cx1, cy0)), smin(C(cmin, cx0, cy1), C(cmin, cx1, cy1)));
  const float mx = smax(smax(C(cmax, cx0, cy0), C(cmax, cx1, cy0)), smax(C(cmax, cx0, cy1), C(cmax, cx1, cy1)));
  int lo = 0, hi = n_all - 1;
  if (mn != kInf) {
    lo = static_cast<int>(floorf(2.0F * mn)) - dmin - margin;
    hi = static_cast<int>(ceilf(2.0F * mx)) - dmin + margin;
  }
  const int match_lo = (partner_sign < 0 ? x - pl : pf - x) - dmin;
  const int match_hi = (partner_sign < 0 ? x - pf : pl - x) - dmin;
  lo = max(lo, match_lo);
  hi = min(hi, match_hi);
  if (lo > hi || pl < pf || (fine_valid != nullptr && fine_valid[y * fvs + x] < 128)) hi = lo;
  lo = min(max(lo, 0), n_all - 1);
  hi = min(max(hi, lo), n_all - 1);
  const int n = min(n_all, (hi - lo + kLabelAlign) / kLabelAlign * kLabelAlign);
  lo = min(max(lo - (n - (hi - lo + 1)) / 2, 0), n_all - n);
  const int xo = mirrored ? width - 1 - x : x;
  lo_out[y * rs + xo] = static_cast<std::uint16_t>(lo);
  num_out[y * rs + xo] = static_cast<std::uint16_t>(n);
}

int range_levels(int wc, int hc, int fill, int* ws, int* hs) {
  int n = 1;
  ws[0] = wc;
  hs[0] = hc;
  while (n < kMaxLevels && (2 << n) <= 2 * fill && ws[n - 1] > 1) {
    ws[n] = (ws[n - 1] + 1) / 2;
    hs[n] = (hs[n - 1] + 1) / 2;
    ++n;
  }
  return n;
}

}  // namespace

void launch_census(const std::uint16_t* img, std::int64_t is, int w, int h, std::uint64_t* out, std::int64_t os,
                   hipStream_t s) {
  const dim3 b(kCw, kCh);
  k_census<<<dim3((w + 2 * kCw - 1) / (2 * kCw), (h + kCh - 1) / kCh), b, 0, s>>>(img, is, w, h, out, os);
}

void launch_pyramid_u16(const std::uint16_t* in, std::int64_t is, int w, int h, std::uint32_t* tmp,
                        std::uint16_t* out, std::int64_t os, hipStream_t s) {
  const int dw = (w + 1) / 2, dh = (h + 1) / 2;
  k_pyr_h<std::uint16_t><<<dim3((dw + 255) / 256, h), 256, 0, s>>>(in, is, w, h, dw, tmp);
  k_pyr_v<std::uint16_t><<<dim3((dw + 255) / 256, dh), 256, 0, s>>>(tmp, h, dw, dh, out, os, 0);
}

void launch_pyramid_u8(const std::uint8_t* in, std::int64_t is, int w, int h, std::uint32_t* tmp, std::uint8_t* out,
                       std::int64_t os, bool threshold, hipStream_t s) {
  const int dw = (w + 1) / 2, dh = (h + 1) / 2;
  k_pyr_h<std::uint8_t><<<dim3((dw + 255) / 256, h), 256, 0, s>>>(in, is, w, h, dw, tmp);
  k_pyr_v<std::uint8_t><<<dim3((dw + 255) / 256, dh), 256, 0, s>>>(tmp, h, dw, dh, out, os, threshold ? 1 : 0);
}

void launch_pyramid_f32(const float* in, std::int64_t is, int w, int h, float* tmp, float* out, std::int64_t os,
                        hipStream_t s) {
  const int dw = (w + 1) / 2, dh = (h + 1) / 2;
  k_pyr_hf<<<dim3((dw + 255) / 256, h), 256, 0, s>>>(in, is, w, dw, tmp);
  k_pyr_vf<<<dim3((dw + 255) / 256, dh), 256, 0, s>>>(tmp, h, dw, dh, out, os);
}

template <typename T>
void rectify_launch(const T* src, std::int64_t ss, const RectifyGeom& g, T* out, std::int64_t os,
                    std::uint8_t* valid, std::int64_t vs, hipStream_t s) {
  if (g.tw <= 0 || g.th <= 0) return;
  k_rectify<T><<<dim3((g.tw + 255) / 256, g.th), 256, 0, s>>>(src, ss, g, out, os, valid, vs);
}

void launch_rectify_u8(const std::uint8_t* src, std::int64_t ss, const RectifyGeom& g, std::uint8_t* out,
                       std::int64_t os, std::uint8_t* valid, std::int64_t vs, hipStream_t s) {
  rectify_launch(src, ss, g, out, os, valid, vs, s);
}

void launch_rectify_u16(const std::uint16_t* src, std::int64_t ss, const RectifyGeom& g, std::uint16_t* out,
                        std::int64_t os, std::uint8_t* valid, std::int64_t vs, hipStream_t s) {
  rectify_launch(src, ss, g, out, os, valid, vs, s);
}

void launch_rectify_f32(const float* src, std::int64_t ss, const RectifyGeom& g, float* out, std::int64_t os,
                        std::uint8_t* valid, std::int64_t vs, hipStream_t s) {
  rectify_launch(src, ss, g, out, os, valid, vs, s);
}

void launch_combine(const CombinePair* pairs, const CombineGrid& g, float* depth, std::int64_t ds, float* conf,
                    std::int64_t cs, std::uint16_t* mask, std::int64_t ms, hipStream_t s) {
  if (g.w <= 0 || g.h <= 0) return;
  k_combine<<<dim3((g.w + 255) / 256, g.h), 256, 0, s>>>(pairs, g, depth, ds, conf, cs, mask, ms);
}


// 8 pixels per thread: one 8-byte load where the row allows it, one 16-byte store.
__global__ __launch_bounds__(256) void k_widen_u8(const std::uint8_t* __restrict__ in, std::int64_t is, int w,
                                                  std::uint16_t* __restrict__ out, std::int64_t os) {
  const int x0 = 8 * (blockIdx.x * blockDim.x + threadIdx.x), y = blockIdx.y;
  if (x0 >= w) return;
  const std::uint8_t* src = in + std::int64_t(y) * is + x0;
  std::uint16_t* dst = out + std::int64_t(y) * os + x0;
  const bool vec = x0 + 8 <= w && (reinterpret_cast<std::uintptr_t>(src) & 7) == 0 &&
                   (reinterpret_cast<std::uintptr_t>(dst) & 15) == 0;
  if (vec) {
    const uint2 v = *reinterpret_cast<const uint2*>(src);
    auto wide = [](std::uint32_t b) {  // two bytes -> two u16 values * 257
      const std::uint32_t lo = b & 0xFFU, hi = (b >> 8) & 0xFFU;
      return (lo * 257U) | ((hi * 257U) << 16);
    };
    *reinterpret_cast<uint4*>(dst) = make_uint4(wide(v.x), wide(v.x >> 16), wide(v.y), wide(v.y >> 16));
    return;
  }
  for (int k = 0; k < 8 && x0 + k < w; ++k) dst[k] = static_cast<std::uint16_t>(src[k] * 257U);
}

void launch_widen_u8(const std::uint8_t* in, std::int64_t is, int w, int h, std::uint16_t* out, std::int64_t os,
                     hipStream_t s) {
  if (w <= 0 || h <= 0) return;
  k_widen_u8<<<dim3((w + 8 * 256 - 1) / (8 * 256), h), 256, 0, s>>>(in, is, w, out, os);
}

void launch_valid_intervals(const std::uint8_t* mask, std::int64_t ms, int w, int h, std::int32_t* rows,
                            hipStream_t s) {
  k_valid_intervals<<<h, 256, 0, s>>>(mask, ms, w, rows);
}

void launch_level_rows(const std::uint8_t* ml, std::int64_t msl, int wl, const std::uint8_t* mr, std::int64_t msr,
                       int wr, int h, std::int32_t* lrows, std::int32_t* rrows, std::int32_t* mbase,
                       std::int32_t* mpartner, std::int32_t* rows4a, std::int32_t* rows4b, hipStream_t s) {
  if (h <= 0) return;
  k_level_rows<<<h, 256, 0, s>>>(ml, msl, wl, mr, msr, wr, lrows, rrows, mbase, mpartner, rows4a, rows4b);
}

void launch_fill_f32(float* p, std::size_t n, float v, hipStream_t s) {
  if (n == 0) return;
  k_fill<float><<<(n + 255) / 256, 256, 0, s>>>(p, n, v);
}

void launch_fill_u8(std::uint8_t* p, std::size_t n, std::uint8_t v, hipStream_t s) {
  if (n == 0) return;
  k_fill<std::uint8_t><<<(n + 255) / 256, 256, 0, s>>>(p, n, v);
}

void launch_unmirror_right(const float* mirrored, int wr, int h, int wl, float* right, hipStream_t s) {
  k_unmirror_right<<<dim3((wr + 255) / 256, h), 256, 0, s>>>(mirrored, wr, h, static_cast<float>(wr - wl), right);
}

void launch_lr_check(float* map, int w, int h, const float* other, int wo, int sign, int max_diff,
                     std::uint8_t* bad, bool apply, hipStream_t s) {
  k_lr_check<<<dim3((w + 255) / 256, h), 256, 0, s>>>(map, w, h, other, wo, static_cast<float>(sign),
                                                     static_cast<float>(max_diff), bad, apply ? 1 : 0);
}

void launch_apply_mask(float* map, std::size_t n, const std::uint8_t* bad, hipStream_t s) {
  if (n == 0) return;
  k_apply_mask<<<(n + 255) / 256, 256, 0, s>>>(map, n, bad);
}

void launch_speckles(float* d, int w, int h, int min_size, float max_diff, std::int32_t* parent, std::int32_t* size,
                     hipStream_t s) {
  const int n = w * h;
  if (min_size <= 1 || n == 0) return;
  const dim3 b(kCcW, kCcH), g((w + kCcW - 1) / kCcW, (h + kCcH - 1) / kCcH);
  k_cc_local<<<g, b, 0, s>>>(d, w, h, max_diff, parent, size);
  k_cc_border<<<g, b, 0, s>>>(d, w, h, max_diff, parent);
  k_cc_count<<<(n + 255) / 256, 256, 0, s>>>(d, parent, size, n, min_size, !(max_diff <= 3.4e38F));
  k_cc_apply<<<(n + 255) / 256, 256, 0, s>>>(d, parent, size, n, min_size);
}

void launch_median3x3(const float* in, int w, int h, float* out, hipStream_t s) {
  k_median3x3<<<dim3((w + 255) / 256, h), 256, 0, s>>>(in, w, h, out);
}

std::size_t ranges_scratch_floats(int wc, int hc, int range_fill) {
  int ws[kMaxLevels], hs[kMaxLevels];
  const int n = range_levels(wc, hc, range_fill, ws, hs);
  std::size_t total = 4 * std::size_t(wc) * hc;  // hmin, hmax, cmin, cmax
  for (int k = 1; k < n; ++k) total += 2 * std::size_t(ws[k]) * hs[k];
  return total;
}

void launch_ranges_from_coarse(const float* coarse, int wc, int hc, int width, int height, const std::uint8_t* fine_valid,
                               std::int64_t fvs, const std::int32_t* partner_rows, int dmin, int n_all, int radius,
                               int margin, int fill, int partner_sign, bool mirrored, float* scratch,
                               std::uint16_t* lo, std::uint16_t* num, std::int64_t rs, hipStream_t s) {
  const std::size_t a = std::size_t(wc) * hc;
  float* hmin = scratch;
  float* hmax = scratch + a;
  float* cmin = scratch + 2 * a;
  float* cmax = scratch + 3 * a;
  const dim3 gc((wc + 255) / 256, hc);
  k_rng_h<<<gc, 256, 0, s>>>(coarse, wc, hc, radius, hmin, hmax);
  k_rng_v<<<gc, 256, 0, s>>>(hmin, hmax, wc, hc, radius, cmin, cmax);
  RangePyr pyr{};
  int ws[kMaxLevels], hs[kMaxLevels];
  pyr.levels = range_levels(wc, hc, fill, ws, hs);
  pyr.mn[0] = cmin;
  pyr.mx[0] = cmax;
  float* next = scratch + 4 * a;
  // Levels 1 .. 6 in one launch, any further ones (range_fill above 64) one
  // launch each.
  RangePyrOut po{};
  po.n = std::min(pyr.levels - 1, kPyrLevels);
  for (int k = 0; k < pyr.levels; ++k) {
    pyr.w[k] = ws[k];
    pyr.h[k] = hs[k];
    if (k == 0) continue;
    float* bmin = next;
    float* bmax = next + std::size_t(ws[k]) * hs[k];
    next = bmax + std::size_t(ws[k]) * hs[k];
    if (k > kPyrLevels)
      k_rng_pyr<<<dim3((ws[k] + 255) / 256, hs[k]), 256, 0, s>>>(pyr.mn[k - 1], pyr.mx[k - 1], ws[k - 1],
                                                                  hs[k - 1], ws[k], hs[k], bmin, bmax);
    pyr.mn[k] = bmin;
    pyr.mx[k] = bmax;
    if (k <= kPyrLevels) {
      po.mn[k] = bmin;
      po.mx[k] = bmax;
    }
    if (k == po.n) {  // the last level of the one-launch part
      for (int j = 0; j <= po.n; ++j) {
        po.w[j] = ws[j];
        po.h[j] = hs[j];
      }
      k_rng_pyr_all<<<dim3((wc + kPyrRegion - 1) / kPyrRegion, (hc + kPyrRegion - 1) / kPyrRegion), 256, 0, s>>>(
          cmin, cmax, po);
    }
  }
  if (pyr.levels > 1) k_rng_fill<<<gc, 256, 0, s>>>(cmin, cmax, wc, hc, pyr);
  k_rng_final<<<dim3((width + 255) / 256, height), 256, 0, s>>>(cmin, cmax, wc, hc, width, height, fine_valid, fvs,
                                                                 partner_rows, dmin, n_all, margin, partner_sign,
                                                                 mirrored ? 1 : 0, lo, num, rs);
}

}  // namespace r3dg
