// Cancellation of the r3d API (reference: src/r3d/core/cancel.hpp:8-58).
// The calling thread's token is handed to the GPU context for the length of
// each call (r3dg_set_cancel_poll), which polls it between pyramid levels and
// tile chunks and stops the call with r3d::Cancelled.
#pragma once

#include <atomic>

#include "r3d/core/error.hpp"

namespace r3d {

class CancellationToken {
 public:
  CancellationToken() = default;
  CancellationToken(const CancellationToken&) = delete;
  CancellationToken& operator=(const CancellationToken&) = delete;

  void request() noexcept { requested_.store(true, std::memory_order_relaxed); }
  void reset() noexcept { requested_.store(false, std::memory_order_relaxed); }
  [[nodiscard]] bool requested() const noexcept { return requested_.load(std::memory_order_relaxed); }
  void throw_if_requested() const {
    if (requested()) throw Cancelled();
  }

 private:
  std::atomic<bool> requested_{false};
};

// The calling thread's token (null if none).
[[nodiscard]] const CancellationToken* current_cancellation() noexcept;

// Throws Cancelled if the calling thread's token is requested.
void throw_if_cancelled();

// Makes `token` the calling thread's token for the scope (null: none).
class CancellationScope {
 public:
  explicit CancellationScope(const CancellationToken* token) noexcept;
  CancellationScope(const CancellationScope&) = delete;
  CancellationScope& operator=(const CancellationScope&) = delete;
  ~CancellationScope();

 private:
  const CancellationToken* previous_;
};

}  // namespace r3d
