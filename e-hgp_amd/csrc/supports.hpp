// Device buffers and host records of the critical-sphere catalogue.
#pragma once

#include <vector>

#include "stages.hpp"

namespace mhgp {

constexpr int kCells = 384;      // 6 cube faces x 8 x 8 direction cells
constexpr int kGatherCap = 600;  // gathered neighbours held in LDS
constexpr int kRegionCap = 384;  // region members held in LDS
constexpr uint32_t kRegionLimitMax = 2048;  // largest candidate region any path enumerates
constexpr int kGatherCapWide = 1024;  // the same for the wide LDS path
constexpr int kRegionCapWide = 640;
constexpr int kMaxInterior = 10; // s_max - 2 for s_max <= 11 (K_max <= 10)

// One catalogue record: support ids (unused = kInvalid), meta
// (size | rank << 8 | interior count << 16 | degenerate << 24), interior ids,
// FP64 squared radius and a bound on its relative error.
struct EventRec {
  uint32_t ids[4];
  uint32_t meta;
  uint32_t interior[kMaxInterior];
  float level_err;
  double level;
};
static_assert(sizeof(EventRec) == 72, "event layout");

// Work the device could not certify: the host settles it exactly.
// kind 2: pair (u, v); 3: face (u, v, w) incl. its triangle and apexes;
// 4: tetrahedron (u, v, w, z).  region_off/len locate the anchor's region.
struct FallbackRec {
  uint32_t kind;
  uint32_t ids[4];
  uint32_t region_off;
  uint32_t region_len;
  uint32_t pad;
};

struct CatalogBuffers {
  bool valid = false;
  int k_max = 0;
  int s_max = 0;
  DevBuf<double> radius;          // gather radius per anchor
  DevBuf<uint32_t> gcount;        // gathered count per anchor
  DevBuf<uint64_t> goff;          // CSR offsets
  DevBuf<uint32_t> glist;         // gathered ids
  DevBuf<double> need;            // radius required by the certificate (0 = satisfied)
  DevBuf<uint32_t> status;        // per anchor: 0 pending, 1 done, 2 widen, 3 big
  DevBuf<EventRec> events;
  DevBuf<FallbackRec> fallbacks;
  DevBuf<uint32_t> fb_region;     // region member ids of anchors with fallbacks
  DevBuf<uint64_t> counters;      // events, fallbacks, region words, candidates
  DevBuf<double> cells;           // per cell: dir xyz, cos d, sin d, chord
  DevBuf<float> cellf;            // per cell: the FP32 cell of the certificate and the masks
  DevBuf<double> scratch;         // global slabs for anchors too big for LDS
  DevBuf<uint32_t> anchors;       // anchor list of the current launch
  // canonical order on the device: records (device emitted + host settled),
  // their order by (exact level, rank, support) and dense exact level classes
  int64_t count = 0;
  DevBuf<uint64_t> sort_keys;
  DevBuf<uint32_t> order;
  DevBuf<uint32_t> cls;
  DevBuf<uint32_t> ambiguous;     // open-run starts, and the runs left to the host (start, length)
  // small per-call device scratch, kept so that no call frees device memory
  // (a free waits for the whole device, side streams included)
  DevBuf<uint32_t> overflow;      // gather traversal overflow flag
  DevBuf<uint64_t> total;         // gathered entries of a pass
  DevBuf<uint64_t> run_count;     // open runs, runs left to the host
  DevBuf<uint32_t> run_pos, run_gidx, run_new, run_link;
  DevBuf<EventRec> run_rec;       // records of the runs ordered on the host
  int64_t open_runs = 0;          // runs the FP64 intervals left open
  int64_t host_runs = 0;          // of those, runs ordered exactly on the host
  mhgp_catalog_info info{};
};

int64_t catalog_device_bytes(const CatalogBuffers &b);
// order.hip: canonical order and level classes of B.events[0, B.count)
void order_catalogue(Context &c, CatalogBuffers &B);

}  // namespace mhgp
