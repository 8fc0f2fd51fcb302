// Host-side orchestration of the device stages (C++ layer under the C ABI).
#pragma once

#include <stdexcept>

#include "../../include/mhgp.h"
#include "context.hpp"

namespace mhgp {

struct DuplicateError : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};
struct NonFiniteError : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};
struct StateError : std::logic_error {
  using std::logic_error::logic_error;
};
// Work beyond a documented device capacity.
struct CapacityError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
// A well-defined input for which the requested object does not exist
// (e.g. the direct tower of a cloud with extra-shell diagnostics).
struct UnsupportedError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

int64_t device_bytes(const Context &c);
// frees every device buffer of the context (its cloud, catalogue and tower
// are gone; the stream stays)
void release_device_state(Context &c);
void upload_host(Context &c, const double *xyz, int64_t n);
void upload_device(Context &c, const double *d_xyz, int64_t n);
void require_cloud(const Context &c);
void canonical_points(Context &c, double *xyz_out, uint32_t *source_index_out);

// ball.hip
void closed_ball(Context &c, const uint32_t *ids, const uint8_t *sizes, int64_t Q, uint32_t cap,
                 uint32_t *interior, uint32_t *nin, uint32_t *nshell);
// the same with every device buffer in a fixed file-scope scratch array
void closed_ball_noalloc(Context &c, const uint32_t *ids, const uint8_t *sizes, int64_t Q, uint32_t cap,
                         uint32_t *interior, uint32_t *nin, uint32_t *nshell);
size_t closed_ball_scratch_bytes();

// knn.hip
// on `stream` (null: the context stream)
void knn_all(Context &c, int k, uint32_t *d_out, uint8_t *d_flags, hipStream_t stream = nullptr);
// k nearest neighbours to host memory, near ties settled exactly on the host
void knn_host(Context &c, int k, uint32_t *ids_out);
// exact top-k partitions of rational query points
void knn_partition(Context &c, const int64_t *num, const int64_t *den, int64_t Q, int k, const uint32_t *excl,
                   const int64_t *eoff, uint32_t shell_cap, uint32_t *choice, uint32_t *strict,
                   uint32_t *strict_count, uint32_t *shell, uint32_t *shell_count, uint32_t *cutoff_id,
                   double *cutoff_sq);

// emst.hip
void emst_host(Context &c, uint32_t *edges_out);

// supports.hip
void supports(Context &c, int k_max, mhgp_catalog_info *info);
void supports_fetch(Context &c, uint32_t *ids, uint32_t *meta, uint32_t *interior, double *level,
                    int64_t *level_class);
// per-PointId enumeration radius of the last catalogue (input unit)
void supports_radii(Context &c, double *radius_out);
// extra-shell diagnostics of the last catalogue's records (canonical order)
void supports_diagnostics(Context &c, uint32_t *shell_count, int64_t *witness, int64_t *exterior_count);

// tower.cpp
void tower(Context &c, int k_max, mhgp_tower_info *info);
void tower_fetch(Context &c, int32_t *order, int32_t *kind, int64_t *level_class, double *level,
                 int64_t *child_offset, int64_t *children);
void tower_vertical(Context &c, int64_t *target);
void tower_cut(Context &c, int order, int64_t cls, int closed, int64_t *roots, int64_t *count,
               int64_t *root_of_point);

// exact host mirror of the canonical cloud, loaded lazily
const class ExactCloud &exact_cloud(Context &c);

}  // namespace mhgp
