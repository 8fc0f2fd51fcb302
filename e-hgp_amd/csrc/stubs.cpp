#include "pipeline.hpp"
namespace mhgp {
void tower(Context &, int, mhgp_tower_info *) { throw StateError("tower not built yet"); }
void tower_fetch(Context &, int32_t *, int32_t *, int64_t *, double *, int64_t *, int64_t *) { throw StateError("tower not built yet"); }
}
