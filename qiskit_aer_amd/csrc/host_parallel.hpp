// Host-side parallel loop for per-circuit work (planning, queueing).
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace aerhip {

// Runs body(i) for i in [0, count) on up to 16 host threads.  The first
// exception thrown by any call is rethrown after every thread has joined;
// the remaining indices are skipped once one call has failed.
template <typename F>
void parallel_for(uint64_t count, F &&body) {
  const uint64_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const uint64_t nt = std::min<uint64_t>(count, hw);
  if (nt <= 1) {
    for (uint64_t i = 0; i < count; ++i) body(i);
    return;
  }
  std::atomic<uint64_t> next{0};
  std::exception_ptr err;
  std::mutex err_mu;
  std::vector<std::thread> pool;
  pool.reserve(nt);
  for (uint64_t t = 0; t < nt; ++t)
    pool.emplace_back([&] {
      for (;;) {
        const uint64_t i = next.fetch_add(1);
        if (i >= count) return;
        try {
          body(i);
        } catch (...) {
          std::lock_guard<std::mutex> g(err_mu);
          if (!err) err = std::current_exception();
          next.store(count);
        }
      }
    });
  for (std::thread &th : pool) th.join();
  if (err) std::rethrow_exception(err);
}

}  // namespace aerhip
