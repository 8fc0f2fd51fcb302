// Host build of the double-double level evaluation, for the CPU tests.
#include <cstdint>

#include "level_dd.hpp"

extern "C" {

// out = {hi, lo, err} of the squared circumradius of the support ids[0..m)
void dd_level_check(const double *X, const double *Y, const double *Z, const uint32_t *ids, int m,
                    double *out) {
  mhgp::DDE l = mhgp::dd_level(X, Y, Z, ids, m);
  out[0] = l.hi;
  out[1] = l.lo;
  out[2] = l.err;
}

int dd_order_check(const double *x, const double *y) {
  return mhgp::dd_order({x[0], x[1], x[2]}, {y[0], y[1], y[2]});
}
}
