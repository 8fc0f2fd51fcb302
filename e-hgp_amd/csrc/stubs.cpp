#include "pipeline.hpp"
namespace mhgp {
void supports(Context &, int, mhgp_catalog_info *) { throw StateError("supports not built yet"); }
void supports_fetch(Context &, uint32_t *, uint32_t *, uint32_t *, double *, int64_t *) { throw StateError("supports not built yet"); }
void tower(Context &, int, mhgp_tower_info *) { throw StateError("tower not built yet"); }
void tower_fetch(Context &, int32_t *, int32_t *, int64_t *, double *, int64_t *, int64_t *) { throw StateError("tower not built yet"); }
}
