// Optional per-stage wall-clock trace (MHGP_TIMING=1): synchronises the
// stream at each mark and prints the elapsed time to stderr.
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "common.hpp"

namespace mhgp {

struct StageTimer {
  bool on;
  hipStream_t stream;
  std::chrono::steady_clock::time_point t;
  explicit StageTimer(hipStream_t s) : on(std::getenv("MHGP_TIMING") != nullptr), stream(s) {
    if (on) {
      (void)hipStreamSynchronize(stream);
      t = std::chrono::steady_clock::now();
    }
  }
  void mark(const char *name) {
    if (!on) return;
    (void)hipStreamSynchronize(stream);
    auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[mhgp] %-28s %9.3f ms\n", name,
                 std::chrono::duration<double, std::milli>(now - t).count());
    t = now;
  }
};

}  // namespace mhgp
