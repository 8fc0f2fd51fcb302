// Host build of the level-class link rule (level_links.hpp), for the CPU tests.
#include "level_links.hpp"

extern "C" void links_check(const double *level, const float *err, long n, unsigned *link) {
  mhgp::level_links_host(level, err, n, link);
}
