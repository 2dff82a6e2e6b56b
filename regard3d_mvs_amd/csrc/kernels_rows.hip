// gfx950 kernels behind the reference's row-kernel interface
// (stereo/census.hpp:25-29, stereo/sgm_kernels.hpp:41-97): one image row of
// Census, matching costs, one 4-path SGM pass and winner-takes-all, on the
// caller's packed row layouts.  The matcher itself does not use them (it runs
// whole tiles per launch); they serve callers that drive the row loop
// themselves, with results identical to the reference's scalar kernels.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace r3dg {
namespace {

constexpr int kOff = 8;               // sgm::kLabelOffset
constexpr std::uint32_t kSent = 0xFFFFU;  // sgm::kSentinel

__device__ __forceinline__ std::uint32_t wave_min(std::uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = min(v, static_cast<std::uint32_t>(__shfl_xor(static_cast<int>(v), d)));
  return v;
}

// census_row (census.cpp:104-113): rows is 7 rows of n + 8 elements (padded by
// 4 replicated columns on both sides), row j = image row y - 3 + j.
__global__ __launch_bounds__(256) void k_census_row(const std::uint16_t* __restrict__ rows, std::uint64_t n,
                                                    std::uint64_t* __restrict__ out) {
  const std::uint64_t x = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const std::uint64_t stride = n + 8;
  const std::uint32_t center = rows[3 * stride + x + 4];
  std::uint64_t bits = 0;
  for (int j = 0; j < 7; ++j) {
    for (int c = 0; c < 9; ++c) {
      if (j == 3 && c == 4) continue;
      bits = (bits << 1) | (rows[j * stride + x + c] < center ? 1U : 0U);
    }
  }
  out[x] = bits;
}

// census_costs (sgm_kernels.cpp:377-385): a wave per pixel.  rev points at the
// caller's rev[rmin], so the reference's index w - 1 - x + lo - origin + d is
// shifted by rmin.
__global__ __launch_bounds__(64) void k_census_costs(const std::uint64_t* __restrict__ left,
                                                     const std::uint64_t* __restrict__ rev, std::int64_t rmin,
                                                     std::uint64_t width, const std::int32_t* __restrict__ lo,
                                                     const std::uint16_t* __restrict__ num,
                                                     const std::uint64_t* __restrict__ offset, std::int32_t origin,
                                                     std::uint16_t* __restrict__ cost) {
  const std::uint64_t x = blockIdx.x;
  const std::uint64_t c = left[x];
  const std::int64_t r0 = std::int64_t(width) - 1 - std::int64_t(x) + lo[x] - origin - rmin;
  const int n = num[x];
  for (int d = threadIdx.x; d < n; d += 64)
    cost[offset[x] + d] = static_cast<std::uint16_t>(__popcll(c ^ rev[r0 + d]));
}

// The predecessor of a pixel for one path, as align_predecessor
// (sgm_kernels.cpp:37-53) presents it: buf points at the predecessor's label
// q_lo in its padded path buffer.  When the predecessor's labels cover the
// pixel's range the buffer is read as it is (its padding supplies the d - 1 of
// the first and the d + 1 of the last label); otherwise labels outside the
// overlap read kSentinel.  `start`: the zero start buffer (read as it is).
struct Pred {
  const std::uint16_t* buf;
  int q_lo, q_num;
  bool start;
};

__device__ __forceinline__ std::uint32_t pred_at(const Pred& q, int lo, int num, int lab) {
  if (q.start) return q.buf[lab - lo];
  const int q_hi = q.q_lo + q.q_num, hi = lo + num;
  if (lo >= q.q_lo && hi <= q_hi) return q.buf[lab - q.q_lo];
  const int first = max(lo - 1, q.q_lo), last = min(hi + 1, q_hi);
  return lab >= first && lab < last ? q.buf[lab - q.q_lo] : kSent;
}

// L(d) = (min(Lq(d), sat(min(Lq(d-1), Lq(d+1)) + P1), penalty) - min_q + C) mod 2^16
// (sgm_kernels.cpp:391-397).
__device__ __forceinline__ std::uint32_t recur(const Pred& q, int lo, int num, int d, std::uint32_t p1,
                                               std::uint32_t penalty, std::uint32_t min_prev, std::uint32_t cost) {
  const int lab = lo + d;
  const std::uint32_t a = pred_at(q, lo, num, lab - 1), b = pred_at(q, lo, num, lab),
                      c = pred_at(q, lo, num, lab + 1);
  const std::uint32_t nb = min(min(a, c) + p1, 0xFFFFU);
  const std::uint32_t best = min(min(b, nb), penalty);
  return (best - min_prev + cost) & 0xFFFFU;
}

// Paths 1..3 of sgm_row (predecessors in the preceding row): a wave per
// pixel.  Writes the pixel's path costs and sentinels into cur, cur_min and
// the sum of the three paths into part.
__global__ __launch_bounds__(64) void k_sgm_row_vertical(SgmRowDev a, std::uint16_t* __restrict__ part) {
  const int w = static_cast<int>(a.width);
  const int i = blockIdx.x;
  const int x = a.direction > 0 ? i : w - 1 - i;
  const int lo = a.lo[x], num = a.num[x];
  const std::uint64_t off = a.offset[x];
  const std::uint64_t cur_size = a.offset[w] + std::uint64_t(w) * 2 * kOff;
  const std::uint64_t cur_off = off + std::uint64_t(x) * 2 * kOff + kOff;
  Pred q[3];
  std::uint32_t min_prev[3], pen[3];
  for (int k = 0; k < 3; ++k) {
    const int xp = k == 0 ? x : (k == 1 ? x - a.direction : x + a.direction);
    if (!a.has_prev || xp < 0 || xp >= w) {
      q[k] = Pred{a.start + kOff, lo, num, true};
      min_prev[k] = 0;
    } else {
      const std::uint64_t prev_size = a.prev_offset[w] + std::uint64_t(w) * 2 * kOff;
      const std::uint64_t pofs = a.prev_offset[xp] + std::uint64_t(xp) * 2 * kOff + kOff;
      q[k] = Pred{a.prev + k * prev_size + pofs, a.prev_lo[xp], a.prev_num[xp], false};
      min_prev[k] = a.prev_min[std::uint64_t(k) * w + xp];
    }
    pen[k] = min(min_prev[k] + a.p2[std::uint64_t(k + 1) * w + x], 0xFFFFU);
  }
  std::uint32_t m[3] = {0xFFFFU, 0xFFFFU, 0xFFFFU};
  for (int d = threadIdx.x; d < num; d += 64) {
    const std::uint32_t cst = a.cost[off + d];
    std::uint32_t total = 0;
    for (int k = 0; k < 3; ++k) {
      const std::uint32_t l = recur(q[k], lo, num, d, a.p1, pen[k], min_prev[k], cst);
      a.cur[k * cur_size + cur_off + d] = static_cast<std::uint16_t>(l);
      m[k] = min(m[k], l);
      total += l;
    }
    part[off + d] = static_cast<std::uint16_t>(total);
  }
  for (int k = 0; k < 3; ++k) {
    const std::uint32_t mk = wave_min(m[k]);
    if (threadIdx.x == 0) {
      a.cur_min[std::uint64_t(k) * w + x] = static_cast<std::uint16_t>(mk);
      a.cur[k * cur_size + cur_off - 1] = static_cast<std::uint16_t>(kSent);
      a.cur[k * cur_size + cur_off + num] = static_cast<std::uint16_t>(kSent);
    }
  }
}

// Path 0 of sgm_row (predecessor on the row itself): one block walks the row
// in pass order, labels across its threads; the previous pixel's path costs
// sit in one of two device buffers (hbuf, 2 x (max_num + 2) with a sentinel
// slot on both sides).  Adds the three other paths (part) into sum.
constexpr int kRowThreads = 256;
__global__ __launch_bounds__(kRowThreads) void k_sgm_row_horizontal(SgmRowDev a,
                                                                    const std::uint16_t* __restrict__ part,
                                                                    std::uint16_t* __restrict__ hbuf) {
  __shared__ std::uint32_t red[kRowThreads / 64];
  const int w = static_cast<int>(a.width);
  const std::uint64_t hs = a.max_num + 2;
  std::uint32_t horiz_min = 0;
  int q_lo = 0, q_num = 0;
  for (int i = 0; i < w; ++i) {
    const int x = a.direction > 0 ? i : w - 1 - i;
    const int lo = a.lo[x], num = a.num[x];
    const std::uint64_t off = a.offset[x];
    std::uint16_t* cur = hbuf + (i & 1) * hs + 1;
    const Pred q = i == 0 ? Pred{a.start + kOff, lo, num, true}
                          : Pred{hbuf + ((i + 1) & 1) * hs + 1, q_lo, q_num, false};
    const std::uint32_t min_prev = i == 0 ? 0U : horiz_min;
    const std::uint32_t pen = min(min_prev + a.p2[x], 0xFFFFU);
    std::uint32_t m = 0xFFFFU;
    for (int d = threadIdx.x; d < num; d += kRowThreads) {
      const std::uint32_t l = recur(q, lo, num, d, a.p1, pen, min_prev, a.cost[off + d]);
      cur[d] = static_cast<std::uint16_t>(l);
      m = min(m, l);
      const std::uint32_t total = l + part[off + d] + (a.accumulate ? a.sum[off + d] : 0U);
      a.sum[off + d] = static_cast<std::uint16_t>(total);
    }
    if (threadIdx.x == 0) {  // sentinels around this pixel's labels
      cur[-1] = static_cast<std::uint16_t>(kSent);
      cur[num] = static_cast<std::uint16_t>(kSent);
    }
    m = wave_min(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    horiz_min = min(min(red[0], red[1]), min(red[2], red[3]));
    q_lo = lo;
    q_num = num;
    __syncthreads();  // cur complete and red read before the next pixel
  }
}

// wta_row (sgm_kernels.cpp:405-426): a wave per pixel; first argmin as the
// minimum of (cost << 16 | label), second best outside best +- 1.
__global__ __launch_bounds__(64) void k_wta_best(const std::uint16_t* __restrict__ sum,
                                                 const std::int32_t* __restrict__ lo,
                                                 const std::uint16_t* __restrict__ num,
                                                 const std::uint64_t* __restrict__ offset,
                                                 std::uint16_t* __restrict__ best, std::uint16_t* __restrict__ cost,
                                                 std::uint16_t* __restrict__ second) {
  const std::uint64_t x = blockIdx.x;
  const std::uint16_t* s = sum + offset[x];
  const int n = num[x];
  std::uint32_t key = 0xFFFFFFFFU;
  for (int k = threadIdx.x; k < n; k += 64) key = min(key, (std::uint32_t(s[k]) << 16) | std::uint32_t(k));
  key = wave_min(key);
  const int kb = static_cast<int>(key & 0xFFFFU);
  std::uint32_t sec = 0xFFFFU;
  for (int k = threadIdx.x; k < n; k += 64)
    if (k + 1 < kb || k > kb + 1) sec = min(sec, std::uint32_t(s[k]));
  sec = wave_min(sec);
  if (threadIdx.x == 0) {
    best[x] = static_cast<std::uint16_t>(lo[x] + kb);
    cost[x] = static_cast<std::uint16_t>(key >> 16);
    second[x] = static_cast<std::uint16_t>(sec);
  }
}

// The right-image WTA of wta_row (sgm_kernels.cpp:427-435): a thread per
// right index i in [imin, imin + count), pixels in x order and labels in k
// order with a strict comparison, starting from the caller's values.
__global__ __launch_bounds__(256) void k_wta_right(const std::uint16_t* __restrict__ sum, std::uint64_t width,
                                                   const std::int32_t* __restrict__ lo,
                                                   const std::uint16_t* __restrict__ num,
                                                   const std::uint64_t* __restrict__ offset, std::int64_t shift,
                                                   std::int64_t imin, std::uint64_t count,
                                                   std::uint16_t* __restrict__ rc, std::uint16_t* __restrict__ rb) {
  const std::uint64_t t = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const std::int64_t i = imin + std::int64_t(t);
  std::uint32_t c = rc[t], b = rb[t];
  for (std::uint64_t x = 0; x < width; ++x) {
    const std::int64_t k = std::int64_t(x) - lo[x] + shift - i;
    if (k >= 0 && k < num[x]) {
      const std::uint32_t v = sum[offset[x] + k];
      if (v < c) {
        c = v;
        b = static_cast<std::uint32_t>(lo[x] + k) & 0xFFFFU;
      }
    }
  }
  rc[t] = static_cast<std::uint16_t>(c);
  rb[t] = static_cast<std::uint16_t>(b);
}

}  // namespace

void launch_census_row(const std::uint16_t* rows, std::uint64_t n, std::uint64_t* out, hipStream_t s) {
  if (n == 0) return;
  k_census_row<<<dim3(static_cast<unsigned>((n + 255) / 256)), 256, 0, s>>>(rows, n, out);
}

void launch_census_costs(const std::uint64_t* left, const std::uint64_t* rev, std::int64_t rmin, std::uint64_t width,
                         const std::int32_t* lo, const std::uint16_t* num, const std::uint64_t* offset,
                         std::int32_t origin, std::uint16_t* cost, hipStream_t s) {
  if (width == 0) return;
  k_census_costs<<<dim3(static_cast<unsigned>(width)), 64, 0, s>>>(left, rev, rmin, width, lo, num, offset, origin,
                                                                   cost);
}

void launch_sgm_row(const SgmRowDev& a, std::uint16_t* part, std::uint16_t* hbuf, hipStream_t s) {
  if (a.width == 0) return;
  k_sgm_row_vertical<<<dim3(static_cast<unsigned>(a.width)), 64, 0, s>>>(a, part);
  k_sgm_row_horizontal<<<dim3(1), kRowThreads, 0, s>>>(a, part, hbuf);
}

void launch_wta_row(const std::uint16_t* sum, std::uint64_t width, const std::int32_t* lo, const std::uint16_t* num,
                    const std::uint64_t* offset, std::uint16_t* best, std::uint16_t* cost, std::uint16_t* second,
                    std::int64_t shift, std::int64_t imin, std::uint64_t count, std::uint16_t* rc, std::uint16_t* rb,
                    hipStream_t s) {
  if (width == 0) return;
  k_wta_best<<<dim3(static_cast<unsigned>(width)), 64, 0, s>>>(sum, lo, num, offset, best, cost, second);
  if (rc != nullptr && count > 0)
    k_wta_right<<<dim3(static_cast<unsigned>((count + 255) / 256)), 256, 0, s>>>(sum, width, lo, num, offset, shift,
                                                                                imin, count, rc, rb);
}

}  // namespace r3dg
