// C ABI of libmhgp.so (declarations and contract in include/mhgp.h).
#include <algorithm>
#include <cmath>
#include <exception>
#include <numeric>
#include <string>

#include "../../include/mhgp.h"
#include "../host/certify.hpp"
#include "../host/exact.hpp"
#include "context.hpp"
#include "pipeline.hpp"

namespace {

thread_local std::string g_last_error;

int fail(mhgp::Context *c, int code, const std::string &msg, bool poison) {
  g_last_error = msg;
  if (c) {
    c->last_error = msg;
    if (poison) c->poisoned = true;
  }
  return code;
}

// Runs f, maps exceptions to error codes.  Argument errors leave the context
// usable; device and internal failures poison it.
template <class F>
int guarded(mhgp::Context *c, F &&f) {
  if (!c) return fail(nullptr, MHGP_E_ARGUMENT, "null context", false);
  if (c->poisoned) return fail(c, MHGP_E_POISONED, "context poisoned by an earlier failure: " + c->last_error, false);
  try {
    MHGP_HIP(hipSetDevice(c->device));
    int rc = f();
    if (rc == MHGP_OK && c->capacity_bytes > 0 && mhgp::device_bytes(*c) > c->capacity_bytes) {
      const int64_t held = mhgp::device_bytes(*c);
      mhgp::release_device_state(*c);
      return fail(c, MHGP_E_CAPACITY, "the call needed " + std::to_string(held) + " device bytes, more than the " +
                  "context's capacity of " + std::to_string(c->capacity_bytes) + "; its device buffers were released",
                  false);
    }
    if (rc == MHGP_OK) g_last_error.clear();
    return rc;
  } catch (const mhgp::DuplicateError &e) {
    return fail(c, MHGP_E_DUPLICATE, e.what(), false);
  } catch (const mhgp::NonFiniteError &e) {
    return fail(c, MHGP_E_NONFINITE, e.what(), false);
  } catch (const std::invalid_argument &e) {
    return fail(c, MHGP_E_ARGUMENT, e.what(), false);
  } catch (const mhgp::StateError &e) {
    return fail(c, MHGP_E_STATE, e.what(), false);
  } catch (const mhgp::UnsupportedError &e) {
    return fail(c, MHGP_E_UNSUPPORTED, e.what(), false);
  } catch (const mhgp::CapacityError &e) {
    return fail(c, MHGP_E_CAPACITY, e.what(), false);
  } catch (const mhgp::HipError &e) {
    return fail(c, MHGP_E_DEVICE, e.what(), true);
  } catch (const std::exception &e) {
    return fail(c, MHGP_E_INTERNAL, e.what(), true);
  }
}

}  // namespace

struct mhgp_ctx {
  mhgp::Context c;
};

extern "C" {

const char *mhgp_last_error(void) { return g_last_error.c_str(); }

int mhgp_ctx_create(int device, int64_t capacity_bytes, mhgp_ctx **out) {
  if (!out) return fail(nullptr, MHGP_E_ARGUMENT, "null output pointer", false);
  if (capacity_bytes < 0) return fail(nullptr, MHGP_E_ARGUMENT, "negative capacity", false);
  *out = nullptr;
  try {
    int count = 0;
    MHGP_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(nullptr, MHGP_E_ARGUMENT, "no such HIP device", false);
    MHGP_HIP(hipSetDevice(device));
    auto *h = new mhgp_ctx();
    h->c.device = device;
    h->c.capacity_bytes = capacity_bytes;
    MHGP_HIP(hipStreamCreateWithFlags(&h->c.stream, hipStreamNonBlocking));
    *out = h;
    g_last_error.clear();
    return MHGP_OK;
  } catch (const std::exception &e) {
    return fail(nullptr, MHGP_E_DEVICE, e.what(), false);
  }
}

void mhgp_ctx_destroy(mhgp_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->c.device);
  if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
  delete ctx;
}

int64_t mhgp_ctx_device_bytes(const mhgp_ctx *ctx) { return ctx ? mhgp::device_bytes(ctx->c) : 0; }

int mhgp_cloud_upload(mhgp_ctx *ctx, const double *xyz, int64_t n) {
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    if (!xyz || n <= 0 || n >= (int64_t)0x7fffffff) throw std::invalid_argument("invalid point array");
    mhgp::upload_host(ctx->c, xyz, n);
    return MHGP_OK;
  });
}

int mhgp_cloud_upload_device(mhgp_ctx *ctx, const double *d_xyz, int64_t n) {
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    if (!d_xyz || n <= 0 || n >= (int64_t)0x7fffffff) throw std::invalid_argument("invalid point array");
    mhgp::upload_device(ctx->c, d_xyz, n);
    return MHGP_OK;
  });
}

int64_t mhgp_point_count(const mhgp_ctx *ctx) { return ctx ? ctx->c.n : 0; }

int mhgp_canonical_points(mhgp_ctx *ctx, double *xyz_out, uint32_t *source_index_out) {
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    mhgp::require_cloud(ctx->c);
    mhgp::canonical_points(ctx->c, xyz_out, source_index_out);
    return MHGP_OK;
  });
}

int mhgp_lbvh_snapshot(mhgp_ctx *ctx, uint64_t *leaves, uint64_t *nodes, uint64_t *counters) {
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    mhgp::require_cloud(ctx->c);
    if (!leaves || !nodes || !counters) throw std::invalid_argument("null output buffer");
    mhgp::export_snapshot(ctx->c, leaves, nodes, counters);
    return MHGP_OK;
  });
}

int mhgp_knn(mhgp_ctx *ctx, int k, uint32_t *ids_out) {
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    mhgp::require_cloud(ctx->c);
    if (!ids_out) throw std::invalid_argument("null output buffer");
    mhgp::knn_host(ctx->c, k, ids_out);
    return MHGP_OK;
  });
}

int mhgp_lbvh_certify(const double *xyz, int64_t n, const uint64_t *leaves, const uint64_t *nodes,
                      const uint64_t *counters) {
  if (!xyz || !leaves || !nodes || !counters || n <= 0)
    return fail(nullptr, MHGP_E_ARGUMENT, "invalid snapshot arrays", false);
  try {
    const std::string why = mhgp::certify_snapshot(xyz, n, leaves, nodes, counters);
    if (!why.empty()) return fail(nullptr, MHGP_E_ARGUMENT, "snapshot rejected: " + why, false);
    g_last_error.clear();
    return MHGP_OK;
  } catch (const std::exception &e) {
    return fail(nullptr, MHGP_E_INTERNAL, e.what(), false);
  }
}

int mhgp_knn_partition(mhgp_ctx *ctx, const int64_t *query_num, const int64_t *query_den, int64_t queries, int k,
                       const uint32_t *excluded, const int64_t *excluded_offsets, uint32_t shell_cap,
                       uint32_t *choice_ids, uint32_t *strict_ids, uint32_t *strict_count, uint32_t *shell_ids,
                       uint32_t *shell_count, uint32_t *cutoff_id, double *cutoff_sq) {
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    mhgp::knn_partition(ctx->c, query_num, query_den, queries, k, excluded, excluded_offsets, shell_cap, choice_ids,
                        strict_ids, strict_count, shell_ids, shell_count, cutoff_id, cutoff_sq);
    return MHGP_OK;
  });
}

int mhgp_emst(mhgp_ctx *ctx, uint32_t *edges_out) {
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    mhgp::require_cloud(ctx->c);
    if (!edges_out && ctx->c.n > 1) throw std::invalid_argument("null output buffer");
    mhgp::emst_host(ctx->c, edges_out);
    return MHGP_OK;
  });
}

int mhgp_supports(mhgp_ctx *ctx, int k_max, mhgp_catalog_info *info) {
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    mhgp::require_cloud(ctx->c);
    if (!info) throw std::invalid_argument("null info");
    mhgp::supports(ctx->c, k_max, info);
    return MHGP_OK;
  });
}

int mhgp_closed_ball(mhgp_ctx *ctx, const uint32_t *support_ids, const uint8_t *support_sizes, int64_t queries,
                     uint32_t cap, uint32_t *interior_ids, uint32_t *interior_count, uint32_t *shell_count) {
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    mhgp::closed_ball(ctx->c, support_ids, support_sizes, queries, cap, interior_ids, interior_count, shell_count);
    return MHGP_OK;
  });
}

int mhgp_closed_ball_noalloc(mhgp_ctx *ctx, const uint32_t *support_ids, const uint8_t *support_sizes,
                             int64_t queries, uint32_t cap, uint32_t *interior_ids, uint32_t *interior_count,
                             uint32_t *shell_count) {
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    mhgp::closed_ball_noalloc(ctx->c, support_ids, support_sizes, queries, cap, interior_ids, interior_count,
                              shell_count);
    return MHGP_OK;
  });
}

int64_t mhgp_closed_ball_noalloc_scratch_bytes(void) { return (int64_t)mhgp::closed_ball_scratch_bytes(); }

int mhgp_supports_fetch(mhgp_ctx *ctx, uint32_t *ids, uint32_t *meta, uint32_t *interior,
                        double *level, int64_t *level_class) {
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    mhgp::supports_fetch(ctx->c, ids, meta, interior, level, level_class);
    return MHGP_OK;
  });
}

int mhgp_supports_radii(mhgp_ctx *ctx, double *radius_out) {
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    mhgp::supports_radii(ctx->c, radius_out);
    return MHGP_OK;
  });
}

int mhgp_supports_diagnostics(mhgp_ctx *ctx, uint32_t *shell_count, int64_t *witness, int64_t *exterior_count) {
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    mhgp::supports_diagnostics(ctx->c, shell_count, witness, exterior_count);
    return MHGP_OK;
  });
}

int mhgp_tower(mhgp_ctx *ctx, int k_max, mhgp_tower_info *info) {
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    if (!info) throw std::invalid_argument("null info");
    mhgp::tower(ctx->c, k_max, info);
    return MHGP_OK;
  });
}

int mhgp_tower_fetch(mhgp_ctx *ctx, int32_t *order, int32_t *kind, int64_t *level_class,
                     double *level, int64_t *child_offset, int64_t *children) {
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    mhgp::tower_fetch(ctx->c, order, kind, level_class, level, child_offset, children);
    return MHGP_OK;
  });
}

int mhgp_tower_cut(mhgp_ctx *ctx, int order, int64_t level_class, int closed, int64_t *roots, int64_t *root_count,
                   int64_t *root_of_point) {
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    mhgp::tower_cut(ctx->c, order, level_class, closed, roots, root_count, root_of_point);
    return MHGP_OK;
  });
}

int mhgp_tower_vertical(mhgp_ctx *ctx, int64_t *target) {
  return guarded(ctx ? &ctx->c : nullptr, [&] {
    mhgp::tower_vertical(ctx->c, target);
    return MHGP_OK;
  });
}

}  // extern "C"

extern "C" int mhgp_ctx_stats(mhgp_ctx *ctx, mhgp_stats *stats, int reset) {
  if (!ctx || !stats) return fail(ctx ? &ctx->c : nullptr, MHGP_E_ARGUMENT, "null argument", false);
  *stats = ctx->c.stats;
  if (reset) ctx->c.stats = mhgp_stats{};
  return MHGP_OK;
}

extern "C" int mhgp_ctx_set_region_limit(mhgp_ctx *ctx, int64_t limit) {
  if (!ctx) return fail(nullptr, MHGP_E_ARGUMENT, "null context", false);
  if (limit < 640 || limit > 2048) return fail(&ctx->c, MHGP_E_ARGUMENT, "region limit must lie in 640..2048", false);
  ctx->c.region_limit = (uint32_t)limit;
  return MHGP_OK;
}

extern "C" int mhgp_ctx_count_flops(mhgp_ctx *ctx, int on) {
  if (!ctx) return fail(nullptr, MHGP_E_ARGUMENT, "null context", false);
  ctx->c.count_flops = on != 0;
  return MHGP_OK;
}
