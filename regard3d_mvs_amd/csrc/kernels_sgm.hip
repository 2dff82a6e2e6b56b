// gfx950 kernels of the SGM core: tile label layout, the 8 path scans,
// winner-takes-all selection and tile blending.
//
// Aggregation dataflow (HBM-bound, no MFMA: nothing here is a contraction):
// one launch per path direction over every tile of every job.  A wave64 scans
// four adjacent scanlines of one tile at once, 16 lanes per scanline, labels
// across the 16 lanes (label j on lane j mod 16).  The predecessor's path costs
// live in LDS (two uint16 buffers per scanline, re-aligned to the current
// pixel's range by index arithmetic, labels outside the predecessor's range
// read as 0xFFFF), the matching cost is recomputed on the fly as
// popcount(CL ^ CR) from the Census images, and the min over labels is a
// 16-lane butterfly.  The first path stores S, the others add to it (uint16
// wrap-around, so the order of the paths does not matter).
//
// Recurrence (sgm_kernels.cpp:176-228, 387-403; stereo_matcher.cpp:1280-1362):
//   L(p,d) = C(p,d) + min(L(q,d), sat16(min(L(q,d-1), L(q,d+1)) + P1),
//                         sat16(min_q + P2(p,q))) - min_q            (mod 2^16)
// with L = C at the first pixel of a path inside the tile.
#include <hip/hip_runtime.h>

#include <cstdint>

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
// Exclusive scan of the label counts of one tile row (block-wide, in chunks
// of 256 pixels): col_off = cell offset of each pixel inside its row.
__global__ __launch_bounds__(256) void k_tile_row_scan(const JobDesc* __restrict__ jobs,
                                                       const TileDesc* __restrict__ tiles,
                                                       const std::uint64_t* __restrict__ row_prefix, int ntiles,
                                                       std::uint32_t* __restrict__ col_off,
                                                       std::uint64_t* __restrict__ row_cells,
                                                       std::uint32_t* __restrict__ row_max) {
  const std::uint64_t r = blockIdx.x;
  const int t = find_tile(row_prefix, ntiles, r);
  const TileDesc T = tiles[t];
  const JobDesc& J = jobs[T.job];
  const int y = static_cast<int>(r - row_prefix[t]);
  const int Y = T.y0 + y;
  __shared__ std::uint32_t sc[256];
  __shared__ std::uint32_t smax;
  if (threadIdx.x == 0) smax = 0;
  __syncthreads();
  std::uint32_t carry = 0;
  for (int base = 0; base < T.w; base += 256) {
    const int x = base + threadIdx.x;
    std::uint32_t v = 0;
    if (x < T.w) {
      int lo, num;
      label_range(J, T.x0 + x, Y, lo, num);
      v = static_cast<std::uint32_t>(num);
      atomicMax(&smax, v);
    }
    sc[threadIdx.x] = v;
    __syncthreads();
    for (int k = 1; k < 256; k <<= 1) {
      const std::uint32_t add = threadIdx.x >= static_cast<unsigned>(k) ? sc[threadIdx.x - k] : 0U;
      __syncthreads();
      sc[threadIdx.x] += add;
      __syncthreads();
    }
    if (x < T.w) col_off[T.px_base + std::uint64_t(y) * T.w + x] = carry + sc[threadIdx.x] - v;
    carry += sc[255];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    row_cells[T.row_base + y] = carry;
    row_max[T.row_base + y] = smax;
  }
}

// Exclusive scan of the row cell counts of one tile: row_start, tile_cells.
__global__ __launch_bounds__(256) void k_tile_cells(const TileDesc* __restrict__ tiles,
                                                    const std::uint64_t* __restrict__ row_cells,
                                                    const std::uint32_t* __restrict__ row_max,
                                                    std::uint64_t* __restrict__ row_start,
                                                    std::uint64_t* __restrict__ tile_cells,
                                                    std::uint32_t* __restrict__ tile_max) {
  const TileDesc T = tiles[blockIdx.x];
  __shared__ std::uint64_t sc[256];
  __shared__ std::uint32_t smax;
  if (threadIdx.x == 0) smax = 0;
  __syncthreads();
  std::uint64_t carry = 0;
  for (int base = 0; base < T.h; base += 256) {
    const int y = base + threadIdx.x;
    const std::uint64_t v = y < T.h ? row_cells[T.row_base + y] : 0;
    if (y < T.h) atomicMax(&smax, row_max[T.row_base + y]);
    sc[threadIdx.x] = v;
    __syncthreads();
    for (int k = 1; k < 256; k <<= 1) {
      const std::uint64_t add = threadIdx.x >= static_cast<unsigned>(k) ? sc[threadIdx.x - k] : 0;
      __syncthreads();
      sc[threadIdx.x] += add;
      __syncthreads();
    }
    if (y < T.h) row_start[T.row_base + y] = carry + sc[threadIdx.x] - v;
    carry += sc[255];
    __syncthreads();
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

__device__ __forceinline__ std::uint32_t min16(std::uint32_t v) {
  v = min(v, static_cast<std::uint32_t>(__shfl_xor(static_cast<int>(v), 8, 16)));
  v = min(v, static_cast<std::uint32_t>(__shfl_xor(static_cast<int>(v), 4, 16)));
  v = min(v, static_cast<std::uint32_t>(__shfl_xor(static_cast<int>(v), 2, 16)));
  v = min(v, static_cast<std::uint32_t>(__shfl_xor(static_cast<int>(v), 1, 16)));
  return v;
}

template <int DX, int DY>
__global__ __launch_bounds__(256) void k_path(const JobDesc* __restrict__ jobs, const TileDesc* __restrict__ tiles,
                                              const std::uint64_t* __restrict__ gprefix, int ntiles,
                                              std::uint64_t groups, const std::uint64_t* __restrict__ row_start,
                                              const std::uint32_t* __restrict__ col_off, std::uint16_t* __restrict__ S,
                                              int accumulate, int max_num) {
  extern __shared__ std::uint16_t lds[];
  const int waves = blockDim.x >> 6;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63, grp = lane >> 4, l16 = lane & 15;
  const std::uint64_t gid = std::uint64_t(blockIdx.x) * waves + wave;
  if (gid >= groups) return;
  const int t = find_tile(gprefix, ntiles, gid);
  const TileDesc T = tiles[t];
  const JobDesc& J = jobs[T.job];
  const int s = static_cast<int>(gid - gprefix[t]) * 4 + grp;
  if (s >= scanlines<DX, DY>(T.w, T.h)) return;

  std::uint16_t* buf = lds + std::size_t(wave * 4 + grp) * 2 * max_num;
  int x, y, len;
  scan_start<DX, DY>(s, T.w, T.h, x, y, len);
  const std::uint32_t P1 = static_cast<std::uint32_t>(J.p1);
  const std::uint32_t invalid = static_cast<std::uint32_t>(J.invalid_cost);
  int prev_lo = 0, prev_num = 0;
  std::uint32_t prev_min = 0;
  for (int i = 0; i < len; ++i, x += DX, y += DY) {
    const int X = T.x0 + x, Y = T.y0 + y;
    int lo, num;
    label_range(J, X, Y, lo, num);
    const std::uint64_t cell = T.cell_base + row_start[T.row_base + y] + col_off[T.px_base + std::uint64_t(y) * T.w + x];
    const std::uint64_t cl = J.cl[Y * J.cl_stride + X];
    const int lf = J.lrows[2 * Y], ll = J.lrows[2 * Y + 1];
    const int rf = J.rrows[2 * Y], rl = J.rrows[2 * Y + 1];
    const bool lvalid = X >= lf && X <= ll;
    std::uint32_t pen = 0;
    if (i > 0) {
      const int ip = J.img[Y * J.img_stride + X];
      const int iq = J.img[(Y - DY) * J.img_stride + (X - DX)];
      pen = min(prev_min + J.lut[abs(ip - iq) >> 8], 0xFFFFU);
    }
    std::uint16_t* cur = buf + (i & 1) * max_num;
    const std::uint16_t* prv = buf + ((i + 1) & 1) * max_num;
    const std::uint64_t* crow = J.cr + Y * J.cr_stride;
    const int xr0 = X - J.dmin - lo;
    std::uint16_t* srow = S + cell;
    std::uint32_t m = 0xFFFFU;
    for (int j = l16; j < num; j += 16) {
      std::uint32_t c = 0;
      if (lvalid) {
        const int xr = xr0 - j;
        c = xr >= rf && xr <= rl ? static_cast<std::uint32_t>(__popcll(cl ^ crow[xr])) : invalid;
      }
      std::uint32_t L = c;
      if (i > 0) {
        const int k = lo + j - prev_lo;
        const std::uint32_t same = static_cast<unsigned>(k) < static_cast<unsigned>(prev_num) ? prv[k] : 0xFFFFU;
        const std::uint32_t a = static_cast<unsigned>(k - 1) < static_cast<unsigned>(prev_num) ? prv[k - 1] : 0xFFFFU;
        const std::uint32_t b = static_cast<unsigned>(k + 1) < static_cast<unsigned>(prev_num) ? prv[k + 1] : 0xFFFFU;
        const std::uint32_t nb = min(min(a, b) + P1, 0xFFFFU);
        const std::uint32_t best = min(min(same, nb), pen);
        L = (best - prev_min + c) & 0xFFFFU;
      }
      cur[j] = static_cast<std::uint16_t>(L);
      m = min(m, L);
      if (accumulate) {
        srow[j] = static_cast<std::uint16_t>(srow[j] + L);
      } else {
        srow[j] = static_cast<std::uint16_t>(L);
      }
    }
    prev_min = min16(m);
    prev_lo = lo;
    prev_num = num;
    // The next step reads this step's LDS row (other lanes' values).
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
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

// First argmin k_best, its cost, and the lowest cost outside k_best +- 1
// (sgm_kernels.cpp:405-426).
__device__ __forceinline__ void wta(const std::uint16_t* __restrict__ s, int n, int& kb, std::uint32_t& bc,
                                    std::uint32_t& sec) {
  kb = 0;
  bc = 0xFFFFFFFFU;
  for (int k0 = 0; k0 < n; k0 += 8) {
    const uint4 q = *reinterpret_cast<const uint4*>(s + k0);
    const std::uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const std::uint32_t v = (w[e >> 1] >> ((e & 1) * 16)) & 0xFFFFU;
      if (v < bc) {
        bc = v;
        kb = k0 + e;
      }
    }
  }
  sec = 0xFFFFU;
  for (int k0 = 0; k0 < n; k0 += 8) {
    const uint4 q = *reinterpret_cast<const uint4*>(s + k0);
    const std::uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = k0 + e;
      const std::uint32_t v = (w[e >> 1] >> ((e & 1) * 16)) & 0xFFFFU;
      if (k + 1 < kb || k > kb + 1) sec = min(sec, v);
    }
  }
}

// select_disparities (stereo_matcher.cpp:446-525) for one tile row per block.
__global__ __launch_bounds__(256) void k_select(const JobDesc* __restrict__ jobs, const TileDesc* __restrict__ tiles,
                                                const std::uint64_t* __restrict__ row_prefix, int ntiles,
                                                const std::uint64_t* __restrict__ row_start,
                                                const std::uint32_t* __restrict__ col_off,
                                                const std::uint16_t* __restrict__ S, SelectParams p,
                                                float* __restrict__ tdisp, float* __restrict__ tconf,
                                                std::int16_t* __restrict__ tlab) {
  extern __shared__ std::uint32_t keys[];
  const std::uint64_t r = blockIdx.x;
  const int t = find_tile(row_prefix, ntiles, r);
  const TileDesc T = tiles[t];
  const JobDesc& J = jobs[T.job];
  const int y = static_cast<int>(r - row_prefix[t]);
  const int Y = T.y0 + y;
  const int lf = J.lrows[2 * Y], ll = J.lrows[2 * Y + 1];
  const int rf = J.rrows[2 * Y], rl = J.rrows[2 * Y + 1];
  const int xb = min(max(lf - T.x0, 0), T.w);
  const int xe = min(max(ll + 1 - T.x0, xb), T.w);
  const std::uint64_t rowcell = T.cell_base + row_start[T.row_base + y];
  const std::uint64_t px = T.px_base + std::uint64_t(y) * T.w;
  const int D = J.D;

  if (p.diagonal_lr) {
    // Right-image WTA along x - label (sgm_kernels.cpp:427-435): the first
    // (in x order) minimum of each diagonal, as a (cost, x) key minimum.
    for (int i = threadIdx.x; i < T.w + D; i += blockDim.x) keys[i] = 0xFFFFFFFFU;
    __syncthreads();
    for (int x = xb + threadIdx.x; x < xe; x += blockDim.x) {
      int lo, num;
      label_range(J, T.x0 + x, Y, lo, num);
      const std::uint16_t* s = S + rowcell + col_off[px + x];
      for (int k = 0; k < num; ++k) {
        atomicMin(&keys[x - (lo + k) + D - 1], (std::uint32_t(s[k]) << 16) | std::uint32_t(x));
      }
    }
    __syncthreads();
  }

  for (int x = threadIdx.x; x < T.w; x += blockDim.x) {
    float d = kInf, conf = 0.0F;
    std::int16_t lab = INT16_MIN;
    if (x >= xb && x < xe) {
      const int X = T.x0 + x;
      int lo, n;
      label_range(J, X, Y, lo, n);
      const std::uint16_t* s = S + rowcell + col_off[px + x];
      int kb;
      std::uint32_t bc, second;
      wta(s, n, kb, bc, second);
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
      const float best_cost = static_cast<float>(bc);
      if (ok && p.uniqueness > 0.0F && second != 0xFFFFU &&
          best_cost > (1.0F - p.uniqueness) * static_cast<float>(second))
        ok = false;
      if (ok) {
        float delta = 0.0F;
        if (j > 0 && j < n - 1) delta = subpixel_offset(p.subpixel, static_cast<float>(s[j - 1]), best_cost,
                                                         static_cast<float>(s[j + 1]));
        d = static_cast<float>(J.dmin + best) + delta;
        conf = second != 0xFFFFU && second > 0 ? 1.0F - best_cost / static_cast<float>(second) : 0.0F;
        lab = static_cast<std::int16_t>(J.dmin + best);
      }
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

// First grid index whose extent end exceeds v (extents sorted by start and end).
__device__ __forceinline__ int first_covering(const std::int32_t* e, int n, int v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (e[4 * mid + 1] > v) {
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
// the dominant tile's value if the mean is more than 1 away from it).
__global__ void k_blend(GridDesc g, const TileDesc* __restrict__ tiles, const float* __restrict__ tdisp,
                        const float* __restrict__ tconf, const std::int16_t* __restrict__ tlab,
                        float* __restrict__ disp, float* __restrict__ conf, std::int16_t* __restrict__ lab) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x, Y = blockIdx.y;
  if (X >= g.width) return;
  float wall = 0.0F, wsum = 0.0F, dsum = 0.0F, csum = 0.0F, wbest = 0.0F, dbest = kInf;
  std::int16_t lbest = INT16_MIN;
  const int j0 = first_covering(g.rows, g.ny, Y);
  const int i0 = first_covering(g.cols, g.nx, X);
  for (int j = j0; j < g.ny && g.rows[4 * j] <= Y; ++j) {
    const float wy = axis_weight(Y, g.rows[4 * j + 2], g.rows[4 * j + 3], g.height, g.blend);
    if (wy <= 0.0F) continue;
    for (int i = i0; i < g.nx && g.cols[4 * i] <= X; ++i) {
      const float wgt = wy * axis_weight(X, g.cols[4 * i + 2], g.cols[4 * i + 3], g.width, g.blend);
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

template <int DX, int DY>
void path_launch(const JobDesc* jobs, const TileDesc* tiles, const std::uint64_t* gp, int ntiles,
                 std::uint64_t groups, const std::uint64_t* rs, const std::uint32_t* co, std::uint16_t* S,
                 bool accumulate, int max_num, hipStream_t s) {
  // Four scanlines per wave; 2 LDS rows of max_num labels per scanline.
  const int waves = max_num <= 1024 ? 4 : 1;
  const std::size_t lds = std::size_t(waves) * 4 * 2 * max_num * sizeof(std::uint16_t);
  const std::uint64_t blocks = (groups + waves - 1) / waves;
  k_path<DX, DY><<<dim3(static_cast<unsigned>(blocks)), 64 * waves, lds, s>>>(jobs, tiles, gp, ntiles, groups, rs, co,
                                                                              S, accumulate ? 1 : 0, max_num);
}

}  // namespace

void launch_tile_row_scan(const JobDesc* jobs, const TileDesc* tiles, const std::uint64_t* row_prefix, int ntiles,
                          std::uint64_t total_rows, std::uint32_t* col_off, std::uint64_t* row_cells,
                          std::uint32_t* row_max, hipStream_t s) {
  k_tile_row_scan<<<dim3(static_cast<unsigned>(total_rows)), 256, 0, s>>>(jobs, tiles, row_prefix, ntiles, col_off,
                                                                          row_cells, row_max);
}

void launch_tile_cells(const TileDesc* tiles, int ntiles, const std::uint64_t* row_cells, const std::uint32_t* row_max,
                       std::uint64_t* row_start, std::uint64_t* tile_cells, std::uint32_t* tile_max, hipStream_t s) {
  k_tile_cells<<<dim3(ntiles), 256, 0, s>>>(tiles, row_cells, row_max, row_start, tile_cells, tile_max);
}

void launch_sgm_path(int path, const JobDesc* jobs, const TileDesc* tiles, const std::uint64_t* gp, int ntiles,
                     std::uint64_t groups, const std::uint64_t* rs, const std::uint32_t* co, std::uint16_t* S,
                     bool accumulate, int max_num, hipStream_t s) {
  if (groups == 0) return;
  switch (path) {
    case 0: path_launch<1, 0>(jobs, tiles, gp, ntiles, groups, rs, co, S, accumulate, max_num, s); break;
    case 1: path_launch<-1, 0>(jobs, tiles, gp, ntiles, groups, rs, co, S, accumulate, max_num, s); break;
    case 2: path_launch<0, 1>(jobs, tiles, gp, ntiles, groups, rs, co, S, accumulate, max_num, s); break;
    case 3: path_launch<0, -1>(jobs, tiles, gp, ntiles, groups, rs, co, S, accumulate, max_num, s); break;
    case 4: path_launch<1, 1>(jobs, tiles, gp, ntiles, groups, rs, co, S, accumulate, max_num, s); break;
    case 5: path_launch<-1, -1>(jobs, tiles, gp, ntiles, groups, rs, co, S, accumulate, max_num, s); break;
    case 6: path_launch<-1, 1>(jobs, tiles, gp, ntiles, groups, rs, co, S, accumulate, max_num, s); break;
    default: path_launch<1, -1>(jobs, tiles, gp, ntiles, groups, rs, co, S, accumulate, max_num, s); break;
  }
}

void launch_select(const JobDesc* jobs, const TileDesc* tiles, const std::uint64_t* row_prefix, int ntiles,
                   std::uint64_t total_rows, const std::uint64_t* row_start, const std::uint32_t* col_off,
                   const std::uint16_t* S, SelectParams p, int max_w, int max_D, float* tdisp, float* tconf,
                   std::int16_t* tlab, hipStream_t s) {
  const std::size_t lds = p.diagonal_lr ? std::size_t(max_w + max_D) * sizeof(std::uint32_t) : 0;
  k_select<<<dim3(static_cast<unsigned>(total_rows)), 256, lds, s>>>(jobs, tiles, row_prefix, ntiles, row_start,
                                                                      col_off, S, p, tdisp, tconf, tlab);
}

void launch_blend(GridDesc g, const TileDesc* tiles, const float* tdisp, const float* tconf, const std::int16_t* tlab,
                  float* disp, float* conf, std::int16_t* lab, hipStream_t s) {
  k_blend<<<dim3((g.width + 255) / 256, g.height), 256, 0, s>>>(g, tiles, tdisp, tconf, tlab, disp, conf, lab);
}

}  // namespace r3dg
