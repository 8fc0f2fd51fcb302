// Shared device/host definitions for the MI355X MorseHGP3D backend.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace mhgp {

struct HipError : std::runtime_error {
  explicit HipError(const std::string &what) : std::runtime_error(what) {}
};

#define MHGP_HIP(call)                                                              \
  do {                                                                              \
    hipError_t mhgp_e_ = (call);                                                    \
    if (mhgp_e_ != hipSuccess)                                                      \
      throw ::mhgp::HipError(std::string(#call) + ": " + hipGetErrorString(mhgp_e_)); \
  } while (0)

#define MHGP_LAUNCH_CHECK() MHGP_HIP(hipGetLastError())

constexpr int kWave = 64;
constexpr uint32_t kInvalid = 0xffffffffu;

// Grow-only device buffer owned by a context.
template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  T *ensure(size_t n) {
    if (n == 0) n = 1;
    if (n > cap) {
      release();
      MHGP_HIP(hipMalloc(&p, n * sizeof(T)));
      cap = n;
    }
    return p;
  }
  size_t bytes() const { return cap * sizeof(T); }
};

// Page-locked host buffer (hipHostMalloc), grown on demand and reused, for
// large device-to-host results.
struct PinnedBuf {
  void *p = nullptr;
  size_t cap = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf &) = delete;
  PinnedBuf &operator=(const PinnedBuf &) = delete;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  void *ensure(size_t bytes) {
    if (bytes > cap) {
      if (p) (void)hipHostFree(p);
      p = nullptr;
      cap = 0;
      MHGP_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault));
      cap = bytes;
    }
    return p;
  }
};

// Internal LBVH node: FP64 box (exact, coordinates are binary64), children,
// leaf range.  64 bytes, one cache line per node.  A child index with the high
// bit set denotes a leaf position in Morton order.
struct alignas(64) Node {
  double lo[3];
  double hi[3];
  uint32_t left, right;
  uint32_t begin, end;  // leaf positions [begin, end)
};
static_assert(sizeof(Node) == 64, "node layout");

constexpr uint32_t kLeafBit = 0x80000000u;

// LBVH internal node for FP32 ball queries, two levels collapsed into one
// 128-byte line: the node's grandchildren (or a child that is a leaf), their
// boxes rounded outward, children as in Node::left/right (kInvalid: none).
struct alignas(128) WideNode {
  float lo[4][3], hi[4][3];
  uint32_t child[4];
};
static_assert(sizeof(WideNode) == 128, "wide node layout");

__host__ __device__ inline uint64_t order_key(uint64_t b) {
  if (b == 0x8000000000000000ull) b = 0;  // -0 -> +0
  return (b >> 63) ? ~b : (b ^ 0x8000000000000000ull);
}

__host__ __device__ inline uint64_t dbits(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  return b;
}

inline int blocks_for(int64_t n, int t) { return (int)((n + t - 1) / t); }

// Copy on the context stream and wait for it.  Context streams are created
// non-blocking, so a plain hipMemcpy (null stream) would not be ordered after
// the kernels queued on them.
inline void copy_wait(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s) {
  MHGP_HIP(hipMemcpyAsync(dst, src, bytes, kind, s));
  MHGP_HIP(hipStreamSynchronize(s));
}

}  // namespace mhgp
