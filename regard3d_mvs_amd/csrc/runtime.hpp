// Host runtime of the GPU matcher: context, device memory arena, errors.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cstdint>
#include <thread>
#include <chrono>
#include <cstring>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "r3dg.h"

namespace r3dg {

// Exception carrying an r3dg_status; converted to a return code at the C ABI.
struct Failure : std::runtime_error {
  int32_t status;
  Failure(int32_t s, const std::string& m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] inline void fail(int32_t status, const std::string& m) { throw Failure(status, m); }

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
      (void)hipGetLastError();
      fail(R3DG_ERROR_OUT_OF_MEMORY, std::string(what) + ": " + hipGetErrorString(e));
    }
    fail(R3DG_ERROR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  }
}

// Caching device allocator bounded by the context's HBM budget.  Freed blocks
// are kept for reuse (every matcher call allocates the same shapes again);
// blocks are rounded to 64 KiB (small) or 2 MiB (large) so shapes of
// different pairs share them.
class Arena {
 public:
  explicit Arena(std::uint64_t budget) : budget_(budget) {}
  Arena(const Arena&) = delete;
  Arena& operator=(const Arena&) = delete;
  ~Arena() { trim(); }

  void* alloc(std::size_t bytes) {
    if (fail_after_ >= 0 && fail_after_-- == 0)
      fail(R3DG_ERROR_OUT_OF_MEMORY, "hipMalloc: injected allocation failure (r3dg_inject_alloc_failure)");
    const std::size_t size = round(bytes);
    auto it = free_.lower_bound(size);
    if (it != free_.end() && it->first <= 2 * size) {
      void* p = it->second;
      live_[p] = it->first;
      live_bytes_ += it->first;
      free_.erase(it);
      return p;
    }
    if (held_ + size > budget_) release_cached(held_ + size - budget_);
    if (held_ + size > budget_) {
      fail(R3DG_ERROR_BUDGET, "device memory budget exceeded: need " + std::to_string(size >> 20) + " MiB more, " +
                                  std::to_string(held_ >> 20) + " of " + std::to_string(budget_ >> 20) +
                                  " MiB held");
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, size);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      trim();
      e = hipMalloc(&p, size);
    }
    hip_check(e, "hipMalloc");
    held_ += size;
    live_[p] = size;
    live_bytes_ += size;
    return p;
  }

  void release(void* p) {
    if (p == nullptr) return;
    auto it = live_.find(p);
    if (it == live_.end()) return;
    free_.emplace(it->second, p);
    live_bytes_ -= it->second;
    live_.erase(it);
  }

  void trim() {
    for (auto& [size, p] : free_) {
      (void)hipFree(p);
      held_ -= size;
    }
    free_.clear();
  }

  // Frees cached blocks, largest first, until `bytes` are returned (or none
  // is left); smaller blocks stay cached for the shapes that reuse them.
  void release_cached(std::uint64_t bytes) {
    std::uint64_t done = 0;
    while (done < bytes && !free_.empty()) {
      auto it = std::prev(free_.end());
      (void)hipFree(it->second);
      held_ -= it->first;
      done += it->first;
      free_.erase(it);
    }
  }

  [[nodiscard]] std::uint64_t held() const { return held_; }
  // Bytes of blocks in use (held minus the cached free blocks).
  [[nodiscard]] std::uint64_t live() const { return live_bytes_; }
  [[nodiscard]] std::uint64_t budget() const { return budget_; }
  // Test hook: the allocation after `n` more succeed fails once as if the
  // device were out of memory (n < 0: none).
  void inject_failure(std::int64_t n) { fail_after_ = n; }

 private:
  static std::size_t round(std::size_t b) {
    const std::size_t g = b >= (std::size_t{4} << 20) ? (std::size_t{2} << 20) : (std::size_t{64} << 10);
    return (std::max<std::size_t>(b, 1) + g - 1) / g * g;
  }
  std::multimap<std::size_t, void*> free_;
  std::unordered_map<void*, std::size_t> live_;
  std::uint64_t held_ = 0;
  std::uint64_t live_bytes_ = 0;
  std::uint64_t budget_;
  std::int64_t fail_after_ = -1;
};

// Cached pinned host blocks (hipHostMalloc), for staging transfers of host
// images: pinned copies run at DMA speed, pageable strided ones do not.
class PinnedPool {
 public:
  PinnedPool() = default;
  PinnedPool(const PinnedPool&) = delete;
  PinnedPool& operator=(const PinnedPool&) = delete;
  ~PinnedPool() {
    for (auto& [size, p] : free_) (void)hipHostFree(p);
  }
  void* get(std::size_t bytes) {
    const std::size_t g = bytes >= (std::size_t{1} << 20) ? (std::size_t{1} << 20) : (std::size_t{64} << 10);
    const std::size_t size = (std::max<std::size_t>(bytes, 1) + g - 1) / g * g;
    auto it = free_.lower_bound(size);
    if (it != free_.end() && it->first <= 2 * size) {
      void* p = it->second;
      sizes_[p] = it->first;
      free_.erase(it);
      return p;
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, size, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      for (auto& [sz, q] : free_) (void)hipHostFree(q);
      free_.clear();
      if (hipHostMalloc(&p, size, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
      }
    }
    sizes_[p] = size;
    return p;
  }
  void put(void* p) {
    auto it = sizes_.find(p);
    if (it == sizes_.end()) return;
    free_.emplace(it->second, p);
    sizes_.erase(it);
  }

 private:
  std::multimap<std::size_t, void*> free_;
  std::unordered_map<void*, std::size_t> sizes_;
};

// Copies `rows` rows of `row_bytes` between strided buffers, split over a few
// threads when large.
inline void copy_rows(void* dst, std::size_t dst_stride, const void* src, std::size_t src_stride,
                      std::size_t row_bytes, std::size_t rows) {
  auto part = [&](std::size_t r0, std::size_t r1) {
    for (std::size_t r = r0; r < r1; ++r)
      std::memcpy(static_cast<char*>(dst) + r * dst_stride, static_cast<const char*>(src) + r * src_stride, row_bytes);
  };
  const std::size_t total = row_bytes * rows;
  const std::size_t nt = std::min<std::size_t>(8, std::max<std::size_t>(1, total >> 21));  // 2 MiB or more a thread
  if (nt <= 1 || rows < nt) {
    part(0, rows);
    return;
  }
  std::vector<std::thread> th;
  const std::size_t per = (rows + nt - 1) / nt;
  for (std::size_t k = 1; k < nt; ++k) th.emplace_back(part, std::min(rows, k * per), std::min(rows, (k + 1) * per));
  part(0, std::min(rows, per));
  for (auto& t : th) t.join();
}

// RAII device buffer of n elements of T from an arena.
template <typename T>
class DBuf {
 public:
  DBuf() = default;
  DBuf(Arena& a, std::size_t n) : arena_(&a), n_(n) { p_ = static_cast<T*>(a.alloc(n * sizeof(T))); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : arena_(o.arena_), p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      reset();
      arena_ = o.arena_;
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  ~DBuf() { reset(); }
  void reset() {
    if (p_ != nullptr && arena_ != nullptr) arena_->release(p_);
    p_ = nullptr;
    n_ = 0;
  }
  [[nodiscard]] T* get() const { return p_; }
  [[nodiscard]] std::size_t size() const { return n_; }

 private:
  Arena* arena_ = nullptr;
  T* p_ = nullptr;
  std::size_t n_ = 0;
};

// Device image: a strided view plus optional ownership.
template <typename T>
struct DImg {
  T* p = nullptr;
  int w = 0;
  int h = 0;
  std::int64_t s = 0;
  DBuf<T> own;
  [[nodiscard]] bool empty() const { return p == nullptr || w == 0 || h == 0; }
  [[nodiscard]] std::size_t pixels() const { return std::size_t(w) * std::size_t(h); }
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  Arena arena;
  std::mutex mutex;
  std::atomic<int> cancel{0};
  // Caller's cancellation test (r3dg_set_cancel_poll), polled with `cancel`.
  int32_t (*cancel_poll)(void*) = nullptr;
  void* cancel_user = nullptr;
  std::string error;
  int32_t error_status = R3DG_OK;
  // Profiling: HIP event pairs around stages, read back at the end of a call.
  int profiling = 0;
  struct Mark {
    int stage;
    hipEvent_t start, stop;
  };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> event_pool;
  std::uint64_t agg_launches = 0;
  std::uint64_t agg_bytes = 0;
  // Host transfers staged through pinned blocks: outputs are copied to the
  // caller's buffers, and the blocks returned, when the stream is next synced.
  PinnedPool pinned;
  struct StagedOut {
    void* pinned;
    void* dst;
    std::size_t row_bytes, rows, dst_stride;
  };
  std::vector<StagedOut> staged_out;
  std::vector<void*> staged_blocks;
  double wait_seconds = 0;  // time spent in sync(), for the request loop's counters
  // Waits for the stream, then completes the staged outputs.
  void sync() {
    const auto w0 = std::chrono::steady_clock::now();
    const hipError_t e = hipStreamSynchronize(stream);
    wait_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
    if (e == hipSuccess) {
      for (const StagedOut& o : staged_out) copy_rows(o.dst, o.dst_stride, o.pinned, o.row_bytes, o.row_bytes, o.rows);
    }
    staged_out.clear();
    for (void* p : staged_blocks) pinned.put(p);
    staged_blocks.clear();
    hip_check(e, "hipStreamSynchronize");
  }
  // After a failure: drop staged outputs without copying them.
  void drop_staged() {
    staged_out.clear();
    for (void* p : staged_blocks) pinned.put(p);
    staged_blocks.clear();
  }
  // Cached views and the request loop (types private to the matcher).  The
  // loop pointer has its own lock: submitting must not wait for a step.
  std::shared_ptr<void> views;
  std::mutex server_mutex;
  std::shared_ptr<void> server;

  // Extra streams for path scans that run side by side (levels whose scans do
  // not fill the device); created on first use, destroyed with the context.
  std::array<hipStream_t, 3> side{};
  // Waits for the context's stream and every side stream.
  void sync_all() {
    (void)hipStreamSynchronize(stream);
    for (hipStream_t s : side) {
      if (s != nullptr) (void)hipStreamSynchronize(s);
    }
  }
  hipStream_t side_stream(int i) {
    if (side[std::size_t(i)] == nullptr)
      hip_check(hipStreamCreateWithFlags(&side[std::size_t(i)], hipStreamNonBlocking), "hipStreamCreate");
    return side[std::size_t(i)];
  }

  explicit Context(std::uint64_t budget) : arena(budget) {}
  ~Context() {
    for (hipStream_t s : side) {
      if (s != nullptr) (void)hipStreamDestroy(s);
    }
    for (auto& m : marks) {
      event_pool.push_back(m.start);
      event_pool.push_back(m.stop);
    }
    for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
  }
  hipEvent_t event() {
    if (!event_pool.empty()) {
      hipEvent_t e = event_pool.back();
      event_pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
  }
  // Adds the recorded stage durations to `seconds` (the stream must be idle).
  void collect(double* seconds, int stages) {
    for (auto& m : marks) {
      float ms = 0.0F;
      if (hipEventElapsedTime(&ms, m.start, m.stop) == hipSuccess && m.stage >= 0 && m.stage < stages)
        seconds[m.stage] += 1e-3 * ms;
      event_pool.push_back(m.start);
      event_pool.push_back(m.stop);
    }
    marks.clear();
  }
};

// Records a stage's start and stop on the context's stream when profiling.
class StageTimer {
 public:
  StageTimer(Context& c, int stage) : c_(c), stage_(stage) {
    if (c_.profiling != 0) {
      start_ = c_.event();
      hip_check(hipEventRecord(start_, c_.stream), "hipEventRecord");
    }
  }
  StageTimer(const StageTimer&) = delete;
  StageTimer& operator=(const StageTimer&) = delete;
  ~StageTimer() {
    if (start_ == nullptr) return;
    hipEvent_t stop = nullptr;
    try {
      stop = c_.event();
    } catch (...) {
      c_.event_pool.push_back(start_);
      return;
    }
    if (hipEventRecord(stop, c_.stream) == hipSuccess) {
      c_.marks.push_back({stage_, start_, stop});
    } else {
      c_.event_pool.push_back(start_);
      c_.event_pool.push_back(stop);
    }
  }

 private:
  Context& c_;
  int stage_;
  hipEvent_t start_ = nullptr;
};

}  // namespace r3dg
