/* C ABI of the MI355X-native MorseHGP3D backend (libmhgp.so).
 *
 * The backend replaces the reference's GPU launcher seam, the
 * `detail::*_on_gpu` functions declared in morsehgp3d/src/cuda/<stage>_internal.hpp,
 * which the reference host adapters (morsehgp3d/src/gpu/<adapter>.cpp) call.  Those are
 * C++ functions over std::span of POD records; here the same work is exposed
 * as plain C over opaque contexts, pointers and sizes.
 *
 * Conventions
 *  - Every function returns MHGP_OK (0) or a negative MHGP_E_* code; the
 *    message of the last failure on the calling thread is mhgp_last_error().
 *  - A context owns one HIP stream and all device state of one cloud.  Calls
 *    on one context are synchronous and must be serialised by the caller
 *    (reference: morsehgp3d/include/morsehgp3d/gpu/morton_lbvh_build.hpp:181-184).
 *  - Any failure inside a device call poisons the context: every later call
 *    except mhgp_ctx_destroy returns MHGP_E_POISONED (reference fail-closed
 *    rule, phase14_morton_lbvh_build_internal.hpp:96-110).
 *  - Point identifiers are canonical PointIds: ranks in the lexicographic
 *    binary64 total order of the coordinates (point_cloud.cpp:40-95).
 */
#ifndef MHGP_H
#define MHGP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MHGP_OK 0
#define MHGP_E_ARGUMENT -1
#define MHGP_E_DUPLICATE -2
#define MHGP_E_NONFINITE -3
#define MHGP_E_DEVICE -4
#define MHGP_E_POISONED -5
#define MHGP_E_STATE -6
#define MHGP_E_CAPACITY -7
#define MHGP_E_UNSUPPORTED -8
#define MHGP_E_INTERNAL -9

typedef struct mhgp_ctx mhgp_ctx;

/* Message of the last failure on this thread ("" when none). */
const char *mhgp_last_error(void);

/* Creates a context on HIP device `device` with its own stream.
 * capacity_bytes > 0 bounds the device memory the context may hold (the
 * capacity of the reference's launcher contexts, e.g.
 * MortonLbvhBuildContext's device capacity, morton_lbvh_build.hpp:41-118):
 * the catalogue refuses record queues that would exceed it, and any call
 * that leaves the context holding more returns MHGP_E_CAPACITY after
 * releasing the context's device buffers (the context stays usable; upload
 * again).  0: no bound beyond the device's own memory. */
int mhgp_ctx_create(int device, int64_t capacity_bytes, mhgp_ctx **out);
/* Releases every device buffer of the context. */
void mhgp_ctx_destroy(mhgp_ctx *ctx);
/* Device bytes currently held by the context. */
int64_t mhgp_ctx_device_bytes(const mhgp_ctx *ctx);

/* Uploads n points (xyz interleaved, binary64, source order, host memory),
 * assigns canonical PointIds and builds the exact Morton LBVH.
 * Replaces propose_phase14_morton_bins_on_gpu and
 * build_phase14_morton_lbvh_snapshot_on_gpu
 * (morsehgp3d/src/cuda/phase14_morton_lbvh_build_internal.hpp:137,149).
 * Returns MHGP_E_DUPLICATE for repeated points, MHGP_E_NONFINITE for NaN/inf. */
int mhgp_cloud_upload(mhgp_ctx *ctx, const double *xyz, int64_t n);
/* Same with the input already in device memory. */
int mhgp_cloud_upload_device(mhgp_ctx *ctx, const double *d_xyz, int64_t n);

int64_t mhgp_point_count(const mhgp_ctx *ctx);

/* Canonical coordinates (n*3, PointId order) and PointId -> source index. */
int mhgp_canonical_points(mhgp_ctx *ctx, double *xyz_out, uint32_t *source_index_out);

/* Reference snapshot layout (morsehgp3d/include/morsehgp3d/spatial/lbvh.hpp:
 * 43-103): leaves 2n u64 (Morton code, PointId) in Morton order; nodes 10 u64
 * each (lower witness ids xyz, upper witness ids xyz, left, right,
 * leaf_begin, leaf_end) for 2n-1 nodes in strict postorder, leaves included;
 * counters: point_count, node_count, maximum_depth,
 * morton_collision_group_count, maximum_morton_collision_size. */
int mhgp_lbvh_snapshot(mhgp_ctx *ctx, uint64_t *leaves, uint64_t *nodes, uint64_t *counters);

/* Exact re-certification of an LBVH snapshot against a canonical cloud
 * (xyz: n x 3 binary64 in PointId order, as mhgp_canonical_points returns
 * it), on the host, with no context or device: the role of the reference's
 * MortonLbvhIndex::import_certified_snapshot (morsehgp3d/src/cpu/spatial/
 * lbvh.cpp:562-829).  Every code is re-quantised exactly, the leaf order,
 * every split, postorder index, leaf range and witness, and the five
 * counters are checked.  MHGP_OK when the snapshot is exactly the certified
 * LBVH; MHGP_E_ARGUMENT naming the first violation otherwise. */
int mhgp_lbvh_certify(const double *xyz, int64_t n, const uint64_t *leaves, const uint64_t *nodes,
                      const uint64_t *counters);

/* k nearest other points of every point, 1 <= k <= 32, n*k PointIds ordered
 * by exact squared distance then PointId.  Replaces the top-k proposals of
 * propose_phase14_facet_top_k_candidates_on_gpu
 * (phase14_facet_top_k_proposal_internal.hpp:183) for data-point queries;
 * ties are settled exactly, no proposal is left to the caller. */
int mhgp_knn(mhgp_ctx *ctx, int k, uint32_t *ids_out);

/* Exact top-k partitions of rational query points: reference lbvh_top_k
 * (morsehgp3d/include/morsehgp3d/spatial/lbvh.hpp:424, lbvh.cpp:1098-1198)
 * and its TopKPartition (spatial/query.hpp:109-169; brute_force.cpp:320-370).
 * Query q is (query_num[3q..3q+2]) / query_den[q] (den > 0, one common
 * denominator per query); its eligible points are every PointId except
 * excluded[excluded_offsets[q] .. excluded_offsets[q+1]) (excluded_offsets
 * has queries + 1 entries starting at 0; both may be NULL for none).
 * requested rank k in 1..32, at most the eligible count.  Per query:
 *  cutoff_sq[q]: the exact k-th smallest squared distance (FP64 value; the
 *    exact rational is |p - q|^2 for p = cutoff_id[q]);
 *  cutoff_id[q]: the least PointId on the cutoff shell;
 *  strict_ids[k*q ..]: the points strictly closer than the cutoff, by exact
 *    (squared distance, PointId), strict_count[q] < k of them, UINT32_MAX padded;
 *  shell_ids[shell_cap*q ..]: the complete cutoff shell ascending (complete
 *    when shell_count[q] <= shell_cap), shell_count[q] its size;
 *  choice_ids[k*q ..]: the canonical choice, the strict points plus the
 *    smallest shell ids up to k, ascending.
 * The device proposes a superset of the shell; the host decides exactly. */
int mhgp_knn_partition(mhgp_ctx *ctx, const int64_t *query_num, const int64_t *query_den, int64_t queries, int k,
                       const uint32_t *excluded, const int64_t *excluded_offsets, uint32_t shell_cap,
                       uint32_t *choice_ids, uint32_t *strict_ids, uint32_t *strict_count, uint32_t *shell_ids,
                       uint32_t *shell_count, uint32_t *cutoff_id, double *cutoff_sq);

/* Exact Euclidean minimum spanning tree under kappa = (d^2, u, v): n-1 edges
 * (u < v) in kappa order.  Replaces propose_k1_boruvka_candidates_on_gpu
 * (phase5_k1_boruvka_internal.hpp:201) together with the host's exact
 * Borůvka contraction (morsehgp3d/src/cpu/hierarchy/boruvka.cpp:893). */
int mhgp_emst(mhgp_ctx *ctx, uint32_t *edges_out);

/* Closed-ball partition of the cloud against the circumspheres of `queries`
 * supports (reference lbvh_closed_ball, morsehgp3d/src/cpu/spatial/
 * lbvh.cpp:1210, and the sparse classification of the support streams,
 * pair_support_stream.cpp:1781-1952).  support_ids: 4 canonical PointIds per
 * query, the first support_sizes[q] (1..4, affinely independent) used; the
 * sphere is their circumsphere with centre in their affine hull.
 * interior_count[q]: points strictly inside; shell_count[q]: points on the
 * sphere, the support included; exterior = n - interior - shell.
 * interior_ids: cap per query, the cap smallest interior PointIds in
 * ascending order, UINT32_MAX padded; complete when interior_count[q] <= cap.
 * Exact: undecided points are settled on the host. */
int mhgp_closed_ball(mhgp_ctx *ctx, const uint32_t *support_ids, const uint8_t *support_sizes,
                     int64_t queries, uint32_t cap, uint32_t *interior_ids, uint32_t *interior_count,
                     uint32_t *shell_count);

/* mhgp_closed_ball without device allocation.  No counterpart in the
 * reference: it exists for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and errors as mhgp_closed_ball; the
 * query staging, per-query counts and flags and the undecided-point list live
 * in one fixed device scratch array of mhgp_closed_ball_noalloc_scratch_bytes()
 * bytes per GPU, held until the process ends.  A call needs about
 * 93 + 4 * max(cap, 1) bytes per query (at least 8 KiB for the undecided
 * list); beyond the scratch it returns MHGP_E_CAPACITY.  Every call on one GPU
 * shares that GPU's scratch: calls that overlap on one GPU (several contexts
 * on several host threads) race and must be serialised by the caller.  The
 * rare exact settlement of undecided points runs on the host, in host memory. */
int mhgp_closed_ball_noalloc(mhgp_ctx *ctx, const uint32_t *support_ids, const uint8_t *support_sizes,
                             int64_t queries, uint32_t cap, uint32_t *interior_ids, uint32_t *interior_count,
                             uint32_t *shell_count);
int64_t mhgp_closed_ball_noalloc_scratch_bytes(void);

/* Critical-sphere catalogue for orders 1..k_max (s_max = min(min(k_max,n)+1, n)):
 * every support of 2..4 points whose circumcentre lies strictly inside its
 * convex hull, whose sphere holds no further point and whose closed rank is
 * <= s_max, plus relevant extra-shell diagnostics.  Singletons (rank 1,
 * level 0) are implicit.  Replaces propose_pair_support_phi_on_gpu
 * (phase9_pair_support_phi_internal.hpp:110) and the direct pair/higher
 * support streams (pair_support_stream.cpp:2372, higher_support_stream.cpp:2397)
 * behind build_exact_direct_support_terminal_facade
 * (direct_support_terminal.cpp:361). */
typedef struct {
  int64_t event_count;       /* regular events (sizes 2..4) */
  int64_t diagnostic_count;  /* extra-shell diagnostics */
  int64_t candidate_count;   /* supports examined exactly */
  int64_t fallback_count;    /* predicates settled by the host exact path */
  int64_t anchors_widened;   /* anchors whose gather radius had to grow */
  int32_t s_max;
  int32_t status;            /* 0 exact; 1 conditional: anchors_open anchors were not enumerated */
  int64_t anchors_open;      /* anchors whose candidate region exceeds the enumeration bound (their
                                supports are missing; every other anchor's are complete and exact) */
  int64_t largest_open_region;  /* largest candidate region among them (points) */
} mhgp_catalog_info;

int mhgp_supports(mhgp_ctx *ctx, int k_max, mhgp_catalog_info *info);

/* Event records of the last mhgp_supports call, in canonical order
 * (level, rank, support).  ids: 4 per event (unused = UINT32_MAX);
 * meta: size | rank << 8 | interior_count << 16 | degenerate << 24;
 * interior: s_max-1 ids per event (unused = UINT32_MAX);
 * level: FP64 approximation of the exact squared radius (relative error
 *   below 1e-12, a deterministic function of the support; exact order and
 *   equality are carried by level_class);
 * level_class: dense rank of the exact squared radius (equal classes share it). */
int mhgp_supports_fetch(mhgp_ctx *ctx, uint32_t *ids, uint32_t *meta, uint32_t *interior,
                        double *level, int64_t *level_class);

/* Shell diagnostics of the last catalogue's records, in the order of
 * mhgp_supports_fetch (reference: the extra-shell fields of the pair and
 * higher records, pair_support_stream.hpp:272-284):
 *  shell_count: points on the sphere, the support included
 *    (observed closed rank = interior count + shell_count; minimum possible
 *    closed rank = interior count + support size = the record's rank);
 *  witness: the least PointId on the sphere outside the support
 *    (canonical extra-shell witness), -1 for a regular event;
 *  exterior_count: points strictly outside the sphere.
 * Any argument may be NULL. */
int mhgp_supports_diagnostics(mhgp_ctx *ctx, uint32_t *shell_count, int64_t *witness, int64_t *exterior_count);

/* Enumeration radius of every PointId u in the last catalogue (n doubles,
 * input unit): every point of every critical sphere whose least PointId is u
 * lies within radius[u] of u.  The catalogue proves this per anchor with its
 * direction-cell certificate (an upper bound on the radius of any critical
 * sphere through u, doubled) before it enumerates; an exhaustive search among
 * the points within radius[u] of u therefore finds every record of u.  No
 * reference counterpart: the reference's streams bound their frontier by
 * node-box prunes instead (pair_support_stream.cpp:446). */
int mhgp_supports_radii(mhgp_ctx *ctx, double *radius_out);

/* hgp_reduced merge-forest tower for orders 1..k_max from the last catalogue
 * (reference build_exact_direct_morse_forest_journal,
 * morsehgp3d/src/cpu/hierarchy/direct_morse_forest_journal.cpp:1042). */
typedef struct {
  int64_t node_count;
  int64_t child_count;
  int64_t descent_steps;
  int32_t orders;
  int32_t status;  /* 0 certified for the catalogue, 1 conditional */
} mhgp_tower_info;

int mhgp_tower(mhgp_ctx *ctx, int k_max, mhgp_tower_info *info);

/* Nodes of the last tower: order, kind (0 order-one birth, 1 reduced birth,
 * 2 multifusion), level class (as in mhgp_supports_fetch), FP64 level, and a
 * CSR child list (child_offset has node_count + 1 entries). */
int mhgp_tower_fetch(mhgp_ctx *ctx, int32_t *order, int32_t *kind, int64_t *level_class,
                     double *level, int64_t *child_offset, int64_t *children);

/* Components of order `order` of the last tower at a cut of level class
 * `level_class` (as in mhgp_supports_fetch; -1 is level zero): closed != 0
 * keeps every node at or below the level, closed == 0 only those strictly
 * below (the order-one leaves, born at zero, are then absent at level zero).
 * roots (capacity node_count, may be NULL): the components' root nodes
 * ascending, root_count their number.  root_of_point (order one only, n
 * entries, may be NULL): each PointId's root, -1 when not yet born.
 * Reference: K1CompactForest::cut_strict / cut_closed
 * (morsehgp3d/include/morsehgp3d/hierarchy/k1_forest.hpp:85-91; k1_forest.cpp:665),
 * here for every order. */
int mhgp_tower_cut(mhgp_ctx *ctx, int order, int64_t level_class, int closed, int64_t *roots, int64_t *root_count,
                   int64_t *root_of_point);

/* Vertical maps of the last tower: for every node (numbered as in
 * mhgp_tower_fetch) of order k + 1 >= 2, the node of order k onto which its
 * component maps at the node's own level (closed cut): the root, at that
 * level, of the order-k component holding the k-subsets of the node's
 * facets; -1 for order-one nodes.  Reference: the profile vertical map
 * family (reference/morsehgp3d_oracle/hierarchy.py:656-780, serialised as
 * VerticalMap assignments, schemas/morsehgp3d-contract-v2.schema.json:904)
 * and build_exact_direct_morse_vertical_journal
 * (direct_morse_vertical_journal.cpp:822). */
int mhgp_tower_vertical(mhgp_ctx *ctx, int64_t *target);

/* Device-time accounting of the context's hot kernels since the last reset,
 * measured with HIP events on the context stream.  No reference counterpart:
 * it mirrors the audit counter blocks the reference attaches to its results
 * (e.g. MortonLbvhDeviceBuildAudit, morton_lbvh_build.hpp:41-118). */
typedef struct {
  double anchor_ms;          /* catalogue anchor kernel(s) */
  int64_t anchor_launches;
  int64_t anchor_bytes;      /* algorithmic bytes: gathered points read + records written */
  double arm_ms;             /* tower arm-resolution kernel(s) */
  int64_t arm_launches;
  int64_t arm_bytes;         /* algorithmic bytes: saddle records read + carriers written */
  int64_t gathered_points;   /* neighbour entries gathered for the anchors */
  int64_t events;            /* catalogue records written by the device */
  int64_t anchor_flops;      /* FP64 operations of the certified predicates the anchor kernel evaluated
                                (per evaluation: the value, its magnitude bound, the certified sign);
                                counted only while mhgp_ctx_count_flops is on */
  int64_t anchor_flops32;    /* FP32 operations of its FP32 tests, counted the same way (certificate and
                                membership tests, capsule tests, the FP32 predicates and key intervals);
                                over the same launches, while mhgp_ctx_count_flops is on */
} mhgp_stats;

int mhgp_ctx_stats(mhgp_ctx *ctx, mhgp_stats *stats, int reset);
/* Turns the FP64 and FP32 operation counts of the anchor kernel on (1) or off
 * (0, the default: the kernel then skips the per-call LDS atomics of the
 * counts). */
int mhgp_ctx_count_flops(mhgp_ctx *ctx, int on);

/* Enumeration budget of the catalogue: the largest candidate region (points)
 * the global-memory anchor path enumerates, 640 <= limit <= 2048 (default
 * 2048).  Anchors whose region exceeds it stay open and mhgp_supports reports
 * status 1 (conditional) with their count and largest region, never an exact
 * status.  Role of the reference's work budgets that end a stream with
 * budget_exhausted (pair_support_stream.hpp:48-62). */
int mhgp_ctx_set_region_limit(mhgp_ctx *ctx, int64_t limit);

/* Continuous-batching server: a pool of worker slots (each one context and
 * one stream) that runs upload -> EMST -> catalogue -> tower for submitted
 * clouds.  A request is admitted when a slot is free and its predicted
 * device memory fits under memory_cap_bytes (<= 0: 80% of the device's
 * HBM); a request larger than the cap runs alone.  The reference has no
 * server; this is the serving loop that sits above its launchers. */
typedef struct mhgp_server mhgp_server;

typedef struct {
  int32_t code;         /* MHGP_OK or the error of the failed stage */
  int32_t status;       /* 0 exact, 1 conditional */
  int64_t event_count;  /* catalogue records */
  int64_t node_count;   /* tower nodes (sizes for mhgp_server_fetch) */
  int64_t child_count;
  double queue_ms;      /* submit -> admitted */
  double service_ms;    /* admitted -> tower on the host */
  double latency_ms;    /* submit -> tower on the host */
} mhgp_request_result;

int mhgp_server_create(int device, int slots, int64_t memory_cap_bytes, mhgp_server **out);
/* Enumeration budget (as mhgp_ctx_set_region_limit) of the requests
 * submitted after the call. */
int mhgp_server_set_region_limit(mhgp_server *srv, int64_t limit);
/* Copies xyz (n x 3 binary64); k_max in 1..10. */
int mhgp_server_submit(mhgp_server *srv, const double *xyz, int64_t n, int k_max,
                       int64_t *ticket);
/* Blocks until the request has finished. */
int mhgp_server_wait(mhgp_server *srv, int64_t ticket, mhgp_request_result *result);
/* Tower of a finished request, laid out as mhgp_tower_fetch. */
int mhgp_server_fetch(mhgp_server *srv, int64_t ticket, int32_t *order, int32_t *kind,
                      int64_t *level_class, double *level, int64_t *child_offset,
                      int64_t *children);
int mhgp_server_release(mhgp_server *srv, int64_t ticket);
void mhgp_server_destroy(mhgp_server *srv);

/* Batched processing of many clouds in one call.  Clouds run in groups, side
 * by side (one host thread and one context per cloud of a group); the first
 * anchor pass of the catalogue of every cloud of a group is ONE persistent
 * launch over all of their anchors (workgroups take anchors from a shared
 * counter across clouds), then each cloud continues as the plain path does:
 * the remaining anchor passes, the tower unless the catalogue is conditional.
 * A group holds at most `group` clouds (<= 0: 32) whose predicted device
 * bytes fit memory_cap_bytes (<= 0: 80% of the device's memory).  Results per
 * cloud as for the server; a failing cloud fails alone.  No reference
 * counterpart: the batching layer above the launchers.
 * Currently switched off: mhgp_batch_run returns MHGP_E_UNSUPPORTED (a
 * process running it ended with a host segmentation fault whose cause has
 * not been found); batch through the server meanwhile. */
typedef struct mhgp_batch mhgp_batch;

typedef struct {
  int64_t clouds;
  int64_t groups;     /* persistent launches */
  int64_t anchors;    /* anchors those launches processed */
  double anchor_ms;   /* their device time, summed */
  double wall_ms;     /* the whole call */
} mhgp_batch_info;

int mhgp_batch_run(int device, int64_t count, const double *const *xyz, const int64_t *n, const int32_t *k_max,
                   int32_t group, int64_t memory_cap_bytes, mhgp_batch **out);
int mhgp_batch_info_get(const mhgp_batch *b, mhgp_batch_info *info);
/* Result of cloud i (code, status, sizes; service_ms is the cloud's own
 * time, latency_ms the time from the call's start to its completion). */
int mhgp_batch_result(const mhgp_batch *b, int64_t cloud, mhgp_request_result *result);
/* Tower of cloud i, laid out as mhgp_tower_fetch (MHGP_E_STATE: none). */
int mhgp_batch_fetch(const mhgp_batch *b, int64_t cloud, int32_t *order, int32_t *kind, int64_t *level_class,
                     double *level, int64_t *child_offset, int64_t *children);
void mhgp_batch_destroy(mhgp_batch *b);

#ifdef __cplusplus
}
#endif

#endif
