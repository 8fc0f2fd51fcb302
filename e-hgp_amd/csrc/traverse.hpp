// LBVH traversal helpers (device).
#pragma once

#include "common.hpp"
#include "exact_device.hpp"

namespace mhgp {

// Squared distance lower bound from q to an FP64 box.  Rounding makes the
// value approximate; callers compare it against cut-offs inflated by kSlack.
__device__ inline double box_d2(const double q[3], const double lo[3], const double hi[3]) {
  double s = 0.0;
  for (int a = 0; a < 3; ++a) {
    double d = fmax(fmax(lo[a] - q[a], q[a] - hi[a]), 0.0);
    s += d * d;
  }
  return s;
}

// Relative slack dominating every FP64 rounding error of the distance and
// bound computations used for pruning (errors are below 1e-15 relative),
// and an absolute one for the underflow of squares in the subnormal range.
constexpr double kSlack = 1e-12;
constexpr double kSlackAbs = 1e-300;

// FP64 squared distance with an exactness flag: exact when every difference,
// square and partial sum is error-free.
__device__ inline double d2_flag(const double p[3], double x, double y, double z, bool &exact) {
  double v[3] = {x, y, z};
  double s = 0.0;
  exact = true;
  for (int a = 0; a < 3; ++a) {
    double d, de;
    two_diff(v[a], p[a], d, de);
    double sq = d * d;
    double se = __builtin_fma(d, d, -sq);
    double t, te;
    two_sum(s, sq, t, te);
    exact = exact && de == 0.0 && se == 0.0 && te == 0.0;
    s = t;
  }
  // below the normal range the residuals themselves underflow to zero
  exact = exact && (s == 0.0 || s >= 0x1p-960);
  return s;
}

}  // namespace mhgp
