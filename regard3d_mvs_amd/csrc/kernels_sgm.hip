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
// Exclusive scan of the label counts of one tile row (block-wide, in chunks
// of 256 pixels) into the pixel records: cell offset inside the row, lo, num.
__global__ __launch_bounds__(256) void k_tile_row_scan(const JobDesc* __restrict__ jobs,
                                                       const TileDesc* __restrict__ tiles,
                                                       const std::uint64_t* __restrict__ row_prefix, int ntiles,
                                                       std::uint64_t* __restrict__ pix,
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
    int lo = 0, num = 0;
    if (x < T.w) {
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
    if (x < T.w) pix[T.px_base + std::uint64_t(y) * T.w + x] = pix_record(carry + sc[threadIdx.x] - v, lo, num);
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

// Min over the 16 lanes of a DPP row, broadcast to all of them (row_ror).
__device__ __forceinline__ std::uint32_t row_min16(std::uint32_t v) {
  v = min(v, static_cast<std::uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x128, 0xF, 0xF, false)));
  v = min(v, static_cast<std::uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x124, 0xF, 0xF, false)));
  v = min(v, static_cast<std::uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x122, 0xF, 0xF, false)));
  v = min(v, static_cast<std::uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x121, 0xF, 0xF, false)));
  return v;
}

// Per-step values of one scanline pixel that do not depend on the recursion.
struct StepIn {
  std::uint64_t rec;  // pixel record
  std::uint64_t rs;   // row start of the pixel's tile row
  std::uint64_t cl;   // base Census
  int4 rows;          // valid columns of the row (base first/last, partner first/last)
  int ip;             // base intensity
};

// Labels j = lane + 16 m of the first kPre chunks, loaded one step ahead.
template <int kPre>
struct StepLanes {
  std::uint32_t s[kPre];   // current S (accumulating paths)
  std::uint64_t cr[kPre];  // partner Census at column X - dmin - lo - j (clamped into the image)
};

// One path direction over every tile.  The loop over the scanline's pixels is
// software-pipelined: the pixel records, Census and row data of step i + 2
// and the S / partner-Census values of step i + 1 are in flight while step i
// computes, with no branch around them, so the compiler's wait counters stay
// exact.  Steps past the end of a scanline (scanlines of a wave differ in
// length) and label chunks past a pixel's range run with clamped loads and
// send their stores to a per-lane dummy slot past the end of S.
template <int DX, int DY, int kPre, bool kAcc>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_path(const JobDesc* __restrict__ jobs, const TileDesc* __restrict__ tiles,
                                              const std::uint64_t* __restrict__ gprefix, int ntiles,
                                              std::uint64_t groups, const std::uint64_t* __restrict__ row_start,
                                              const std::uint64_t* __restrict__ pix, std::uint16_t* __restrict__ S,
                                              std::uint64_t total_cells, int max_num) {
  extern __shared__ std::uint16_t lds[];
  const int waves = blockDim.x >> 6;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63, grp = lane >> 4, l16 = lane & 15;
  const std::uint64_t gid = std::uint64_t(blockIdx.x) * waves + wave;
  const bool wave_active = gid < groups;
  const int t = find_tile(gprefix, ntiles, wave_active ? gid : gprefix[ntiles] - 1);
  const TileDesc T = tiles[t];
  const JobDesc& J = jobs[T.job];
  std::uint16_t* my_lut = lds + wave * 256;  // P2 table of the wave's job
  for (int k = lane; k < 256; k += 64) my_lut[k] = J.lut[k];
  __syncthreads();
  if (!wave_active) return;
  const int ns = scanlines<DX, DY>(T.w, T.h);
  const int s0 = static_cast<int>(gid - gprefix[t]) * 4;
  const int s = min(s0 + grp, ns - 1);     // idle groups shadow the last scanline
  const bool grp_active = s0 + grp < ns;
  int x0, y0, len;
  scan_start<DX, DY>(s, T.w, T.h, x0, y0, len);
  // Steps of the wave: the longest of its scanlines.
  int steps = grp_active ? len : 0;
  steps = max(steps, __shfl_xor(steps, 16));
  steps = max(steps, __shfl_xor(steps, 32));
  steps = __builtin_amdgcn_readfirstlane(steps);

  // LDS row of a scanline: the labels plus room for the masked-off chunks of a
  // 4-chunk batch (their writes land past the row's labels and are never read).
  const int rowlen = (max_num > 16 * kPre ? max_num : 16 * kPre) + 48;
  std::uint16_t* buf = lds + waves * 256 + std::size_t(wave * 4 + grp) * 2 * rowlen;
  const std::uint32_t P1 = static_cast<std::uint32_t>(J.p1);
  const std::uint32_t invalid = static_cast<std::uint32_t>(J.invalid_cost);
  const int pw1 = J.pwidth - 1;
  const auto tpix = gptr(pix + T.px_base);
  const auto trs = gptr(row_start + T.row_base);
  const auto gcl = gptr(J.cl);
  const auto gcr = gptr(J.cr);
  const auto gimg = gptr(J.img);
  const auto grows = gptr(J.rows4);
  const auto gS = gptr_mut(S);
  const auto dummy = gS + total_cells + lane;  // inside the padding of S

  auto load_in = [&](int i, StepIn& u) {
    i = min(i, len - 1);
    const int x = x0 + i * DX, y = y0 + i * DY;
    const int X = T.x0 + x, Y = T.y0 + y;
    u.rec = tpix[std::int64_t(y) * T.w + x];
    u.rs = trs[y];
    u.cl = gcl[Y * J.cl_stride + X];
    u.rows = make_int4(grows[4 * Y], grows[4 * Y + 1], grows[4 * Y + 2], grows[4 * Y + 3]);
    u.ip = gimg[Y * J.img_stride + X];
  };
  auto load_lanes = [&](int i, const StepIn& u, StepLanes<kPre>& v) {
    i = min(i, len - 1);
    const int X = T.x0 + x0 + i * DX, Y = T.y0 + y0 + i * DY;
    const std::uint64_t cell = T.cell_base + u.rs + pix_off(u.rec);
    const int xr0 = X - J.dmin - pix_lo(u.rec) - l16;
    const auto crow = gcr + Y * J.cr_stride;
#pragma unroll
    for (int m = 0; m < kPre; ++m) {
      // Chunks past the pixel's labels read neighbouring (padded) cells and a
      // clamped column; their values are never used.
      if constexpr (kAcc) v.s[m] = gS[cell + l16 + 16 * m];
      v.cr[m] = crow[min(max(xr0 - 16 * m, 0), pw1)];
    }
  };

  int prev_lo = 0, prev_num = 0, prev_ip = 0;
  std::uint32_t prev_min = 0;
  int flip = 0;
  // Step i with its inputs u (this step), un (next) and un2 (receives step
  // i + 2), v (this step) and vn (receives step i + 1).
  auto step = [&](int i, StepIn& u, StepIn& un, StepIn& un2, StepLanes<kPre>& v, StepLanes<kPre>& vn) {
    load_in(i + 2, un2);
    load_lanes(i + 1, un, vn);
    const bool act = grp_active && i < len;
    const int ic = min(i, len - 1);
    const int X = T.x0 + x0 + ic * DX;
    const int lo = pix_lo(u.rec), num = pix_num(u.rec);
    const std::uint64_t cell = T.cell_base + u.rs + pix_off(u.rec);
    const bool lvalid = X >= u.rows.x && X <= u.rows.y;
    const int rf = u.rows.z, rl = u.rows.w;
    const bool first = i == 0;
    const std::uint32_t pen = min(prev_min + my_lut[abs(u.ip - prev_ip) >> 8], 0xFFFFU);
    std::uint16_t* cur = buf + flip * rowlen;
    const std::uint16_t* prv = buf + (flip ^ 1) * rowlen;
    const int xr0 = X - J.dmin - lo;
    const int delta = lo - prev_lo;
    std::uint32_t m = 0xFFFFU;
    auto label = [&](int j, bool on, std::uint32_t sv, std::uint64_t crv) {
      const int xr = xr0 - j;
      std::uint32_t c = xr >= rf && xr <= rl ? static_cast<std::uint32_t>(__popcll(u.cl ^ crv)) : invalid;
      c = lvalid ? c : 0U;
      const int k = j + delta;
      const std::uint32_t same = static_cast<unsigned>(k) < static_cast<unsigned>(prev_num) ? prv[k] : 0xFFFFU;
      const std::uint32_t a = static_cast<unsigned>(k - 1) < static_cast<unsigned>(prev_num) ? prv[k - 1] : 0xFFFFU;
      const std::uint32_t b = static_cast<unsigned>(k + 1) < static_cast<unsigned>(prev_num) ? prv[k + 1] : 0xFFFFU;
      const std::uint32_t nb = min(min(a, b) + P1, 0xFFFFU);
      const std::uint32_t best = min(min(same, nb), pen);
      const std::uint32_t L = first ? c : ((best - prev_min + c) & 0xFFFFU);
      cur[j] = static_cast<std::uint16_t>(L);  // j < rowlen; unused past num
      m = on ? min(m, L) : m;
      const auto dst = on ? gS + cell + j : dummy;
      *dst = static_cast<std::uint16_t>(kAcc ? sv + L : L);
    };
    const int chunks = num >> 4;
#pragma unroll
    for (int q = 0; q < kPre; ++q) label(l16 + 16 * q, act && q < chunks, v.s[q], v.cr[q]);
    if (chunks > kPre) {  // wide ranges: batches of 4 chunks, loads issued together
      const auto crow = gcr + (T.y0 + y0 + ic * DY) * J.cr_stride;
      for (int q = kPre; q < chunks; q += 4) {
        std::uint32_t sv[4];
        std::uint64_t cv[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int j = min(l16 + 16 * (q + b), num - 1);
          sv[b] = kAcc ? static_cast<std::uint32_t>(gS[cell + j]) : 0U;
          cv[b] = crow[min(max(xr0 - j, 0), pw1)];
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) label(l16 + 16 * (q + b), act && q + b < chunks, sv[b], cv[b]);
      }
    }
    prev_min = row_min16(m);
    prev_lo = lo;
    prev_num = num;
    prev_ip = u.ip;
    flip ^= 1;
    __builtin_amdgcn_wave_barrier();  // keeps this step's LDS writes before the next step's reads
  };

  StepIn ua, ub, uc;
  StepLanes<kPre> va, vb;
  load_in(0, ua);
  load_in(1, ub);
  load_lanes(0, ua, va);
  for (int i = 0; i < steps; i += 6) {
    step(i + 0, ua, ub, uc, va, vb);
    step(i + 1, ub, uc, ua, vb, va);
    step(i + 2, uc, ua, ub, va, vb);
    step(i + 3, ua, ub, uc, vb, va);
    step(i + 4, ub, uc, ua, va, vb);
    step(i + 5, uc, ua, ub, vb, va);
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
                                                const std::uint64_t* __restrict__ pix,
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
  const int4 rows = reinterpret_cast<const int4*>(J.rows4)[Y];
  const int lf = rows.x, ll = rows.y, rf = rows.z, rl = rows.w;
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
      const std::uint64_t rec = pix[px + x];
      const int lo = pix_lo(rec), num = pix_num(rec);
      const std::uint16_t* s = S + rowcell + pix_off(rec);
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
      const std::uint64_t rec = pix[px + x];
      const int lo = pix_lo(rec), n = pix_num(rec);
      const std::uint16_t* s = S + rowcell + pix_off(rec);
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

__global__ void k_pack_rows(const std::int32_t* __restrict__ base, const std::int32_t* __restrict__ partner, int h,
                            std::int32_t* __restrict__ rows4) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= h) return;
  reinterpret_cast<int4*>(rows4)[y] = make_int4(base[2 * y], base[2 * y + 1], partner[2 * y], partner[2 * y + 1]);
}

template <int DX, int DY>
void path_launch(const JobDesc* jobs, const TileDesc* tiles, const std::uint64_t* gp, int ntiles,
                 std::uint64_t groups, const std::uint64_t* rs, const std::uint64_t* pix, std::uint16_t* S,
                 std::uint64_t total_cells, bool accumulate, int max_num, hipStream_t s) {
  // Four scanlines per wave; 2 LDS rows of max_num labels per scanline plus
  // one 256-entry P2 table per wave.
  constexpr int kPre = 2;
  const int rowlen = std::max(max_num, 16 * kPre) + 48;  // as in k_path
  const auto lds_of = [&](int w) { return std::size_t(w) * (256 + 4 * 2 * std::size_t(rowlen)) * sizeof(std::uint16_t); };
  const int waves = lds_of(4) <= 65536 ? 4 : 1;
  const std::size_t lds = lds_of(waves);
  const std::uint64_t blocks = (groups + waves - 1) / waves;
  if (accumulate) {
    k_path<DX, DY, kPre, true><<<dim3(static_cast<unsigned>(blocks)), 64 * waves, lds, s>>>(
        jobs, tiles, gp, ntiles, groups, rs, pix, S, total_cells, max_num);
  } else {
    k_path<DX, DY, kPre, false><<<dim3(static_cast<unsigned>(blocks)), 64 * waves, lds, s>>>(
        jobs, tiles, gp, ntiles, groups, rs, pix, S, total_cells, max_num);
  }
}

}  // namespace

void launch_tile_row_scan(const JobDesc* jobs, const TileDesc* tiles, const std::uint64_t* row_prefix, int ntiles,
                          std::uint64_t total_rows, std::uint64_t* pix, std::uint64_t* row_cells,
                          std::uint32_t* row_max, hipStream_t s) {
  k_tile_row_scan<<<dim3(static_cast<unsigned>(total_rows)), 256, 0, s>>>(jobs, tiles, row_prefix, ntiles, pix,
                                                                          row_cells, row_max);
}

void launch_tile_cells(const TileDesc* tiles, int ntiles, const std::uint64_t* row_cells, const std::uint32_t* row_max,
                       std::uint64_t* row_start, std::uint64_t* tile_cells, std::uint32_t* tile_max, hipStream_t s) {
  k_tile_cells<<<dim3(ntiles), 256, 0, s>>>(tiles, row_cells, row_max, row_start, tile_cells, tile_max);
}

void launch_sgm_path(int path, const JobDesc* jobs, const TileDesc* tiles, const std::uint64_t* gp, int ntiles,
                     std::uint64_t groups, const std::uint64_t* rs, const std::uint64_t* co, std::uint16_t* S,
                     std::uint64_t total_cells, bool accumulate, int max_num, hipStream_t s) {
  if (groups == 0) return;
  switch (path) {
    case 0: path_launch<1, 0>(jobs, tiles, gp, ntiles, groups, rs, co, S, total_cells, accumulate, max_num, s); break;
    case 1: path_launch<-1, 0>(jobs, tiles, gp, ntiles, groups, rs, co, S, total_cells, accumulate, max_num, s); break;
    case 2: path_launch<0, 1>(jobs, tiles, gp, ntiles, groups, rs, co, S, total_cells, accumulate, max_num, s); break;
    case 3: path_launch<0, -1>(jobs, tiles, gp, ntiles, groups, rs, co, S, total_cells, accumulate, max_num, s); break;
    case 4: path_launch<1, 1>(jobs, tiles, gp, ntiles, groups, rs, co, S, total_cells, accumulate, max_num, s); break;
    case 5: path_launch<-1, -1>(jobs, tiles, gp, ntiles, groups, rs, co, S, total_cells, accumulate, max_num, s); break;
    case 6: path_launch<-1, 1>(jobs, tiles, gp, ntiles, groups, rs, co, S, total_cells, accumulate, max_num, s); break;
    default: path_launch<1, -1>(jobs, tiles, gp, ntiles, groups, rs, co, S, total_cells, accumulate, max_num, s); break;
  }
}

void launch_select(const JobDesc* jobs, const TileDesc* tiles, const std::uint64_t* row_prefix, int ntiles,
                   std::uint64_t total_rows, const std::uint64_t* row_start, const std::uint64_t* col_off,
                   const std::uint16_t* S, SelectParams p, int max_w, int max_D, float* tdisp, float* tconf,
                   std::int16_t* tlab, hipStream_t s) {
  const std::size_t lds = p.diagonal_lr ? std::size_t(max_w + max_D) * sizeof(std::uint32_t) : 0;
  k_select<<<dim3(static_cast<unsigned>(total_rows)), 256, lds, s>>>(jobs, tiles, row_prefix, ntiles, row_start,
                                                                      col_off, S, p, tdisp, tconf, tlab);
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
