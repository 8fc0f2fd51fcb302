// Interval rule that splits the level-sorted catalogue into exact level
// classes (host and device).
//
// Each record carries an FP64 level and a bound e on its relative error, so
// its exact level lies in [level (1 - 2e), level (1 + 2e)] (the factor 2
// covers the rounding of the bound itself).  Records are sorted by the FP64
// level; record i provably lies above every earlier record, and so starts a
// new class, exactly when its lower bound exceeds the largest upper bound of
// the records before it.  The widths differ from record to record, so the
// neighbour's upper bound alone is not enough.
#pragma once

#include <hip/hip_runtime.h>

namespace mhgp {

__host__ __device__ inline double level_upper(double level, float err) { return level * (1.0 + 2.0 * (double)err); }
__host__ __device__ inline double level_lower(double level, float err) { return level * (1.0 - 2.0 * (double)err); }
__host__ __device__ inline bool level_link(double lower_i, double prefmax_upper_before) {
  return lower_i > prefmax_upper_before;
}

// Host form over a sorted sequence (tests and small host runs): link[0] = 1.
inline void level_links_host(const double *level, const float *err, long n, unsigned *link) {
  double top = 0.0;
  for (long i = 0; i < n; ++i) {
    link[i] = i == 0 ? 1u : (level_link(level_lower(level[i], err[i]), top) ? 1u : 0u);
    const double up = level_upper(level[i], err[i]);
    top = i == 0 || up > top ? up : top;
  }
}

}  // namespace mhgp
