// Error-free FP64 transforms shared by device kernels and host code.
//
// Knuth's TwoSum and Dekker's TwoDiff give the rounding error of a sum
// exactly; the FMA form of TwoProduct gives the rounding error of a product
// exactly (barring underflow).  Shewchuk's grow-expansion turns a short list
// of doubles into a non-overlapping expansion whose largest non-zero
// component carries the sign of the exact sum.  (J. R. Shewchuk, "Adaptive
// Precision Floating-Point Arithmetic and Fast Robust Geometric Predicates",
// Discrete Comput. Geom. 18, 1997.)
#pragma once

#include <hip/hip_runtime.h>

namespace mhgp {

__host__ __device__ inline void two_sum(double a, double b, double &x, double &y) {
#pragma clang fp contract(off)
  x = a + b;
  double bv = x - a;
  double av = x - bv;
  y = (a - av) + (b - bv);
}

__host__ __device__ inline void two_diff(double a, double b, double &x, double &y) {
#pragma clang fp contract(off)
  x = a - b;
  double bv = a - x;
  double av = x + bv;
  y = (a - av) + (bv - b);
}

__host__ __device__ inline void two_prod(double a, double b, double &x, double &y) {
  x = a * b;
  y = __builtin_fma(a, b, -x);
}

// Sign of the exact sum of k <= 16 doubles.
__host__ __device__ inline int exact_sum_sign(const double *t, int k) {
  double h[16];
  int m = 0;
  for (int i = 0; i < k; ++i) {
    double q = t[i];
    for (int j = 0; j < m; ++j) {
      double s, e;
      two_sum(q, h[j], s, e);
      h[j] = e;
      q = s;
    }
    h[m++] = q;
  }
  for (int j = m - 1; j >= 0; --j)
    if (h[j] != 0.0) return h[j] > 0 ? 1 : -1;
  return 0;
}

}  // namespace mhgp
