// Thread-count controls of the r3d API (reference: src/r3d/core/thread_pool.hpp:28-79)
// that the matcher's callers use.  Matching runs on the GPU, whose results do
// not depend on a host thread count (as the reference's do not, D5); the count
// is recorded and reported, and sizes nothing.
#pragma once

namespace r3d {

class ThreadPool {
 public:
  explicit ThreadPool(unsigned num_threads = 0);
  ThreadPool(const ThreadPool&) = delete;
  ThreadPool& operator=(const ThreadPool&) = delete;
  [[nodiscard]] unsigned num_threads() const noexcept { return num_threads_; }

 private:
  unsigned num_threads_;
};

ThreadPool& default_thread_pool();
void set_default_thread_count(unsigned num_threads);

}  // namespace r3d
