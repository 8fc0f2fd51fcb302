// Three-phase inclusive scan: per-tile scans with tile totals, a scan of the
// totals (recursively), then a uniform add of the tile offsets.  A tile is
// 4 096 values: 256 lanes hold 16 consecutive values each (four 16-byte
// loads), scan them in registers, and combine the lane totals with a
// shuffle scan per wave and one LDS exchange of the four wave totals.
#include <stdexcept>

#include "scan.hpp"

namespace mhgp {

namespace {

constexpr int kScanBlock = 256;
constexpr int kPerThread = 16;
constexpr int kTileItems = kScanBlock * kPerThread;

template <bool VEC>
__device__ inline void load16(const uint32_t *data, int64_t base, int64_t n, uint32_t (&v)[kPerThread]) {
  if (VEC && base + kPerThread <= n) {
    const uint4 *p = reinterpret_cast<const uint4 *>(data + base);
#pragma unroll
    for (int j = 0; j < kPerThread / 4; ++j) {
      const uint4 q = p[j];
      v[4 * j] = q.x;
      v[4 * j + 1] = q.y;
      v[4 * j + 2] = q.z;
      v[4 * j + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) v[i] = base + i < n ? data[base + i] : 0u;
  }
}

template <bool VEC>
__device__ inline void store16(uint32_t *data, int64_t base, int64_t n, const uint32_t (&v)[kPerThread]) {
  if (VEC && base + kPerThread <= n) {
    uint4 *p = reinterpret_cast<uint4 *>(data + base);
#pragma unroll
    for (int j = 0; j < kPerThread / 4; ++j) p[j] = make_uint4(v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]);
  } else {
#pragma unroll
    for (int i = 0; i < kPerThread; ++i)
      if (base + i < n) data[base + i] = v[i];
  }
}

template <bool VEC>
__global__ void __launch_bounds__(kScanBlock) k_block_scan(uint32_t *data, int64_t n, uint32_t *sums) {
  __shared__ uint32_t wsum[kScanBlock / 64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t base = (int64_t)blockIdx.x * kTileItems + (int64_t)t * kPerThread;
  uint32_t v[kPerThread];
  load16<VEC>(data, base, n, v);
#pragma unroll
  for (int i = 1; i < kPerThread; ++i) v[i] += v[i - 1];
  // inclusive scan of the lane totals across the wave
  uint32_t x = v[kPerThread - 1];
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, (unsigned)d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  uint32_t before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kScanBlock / 64; ++w) {
    before += w < wave ? wsum[w] : 0u;
    total += wsum[w];
  }
  const uint32_t prefix = before + x - v[kPerThread - 1];
#pragma unroll
  for (int i = 0; i < kPerThread; ++i) v[i] += prefix;
  store16<VEC>(data, base, n, v);
  if (t == 0) sums[blockIdx.x] = total;
}

template <bool VEC>
__global__ void __launch_bounds__(kScanBlock) k_add_offsets(uint32_t *data, int64_t n, const uint32_t *sums) {
  if (blockIdx.x == 0) return;
  const uint32_t add = sums[blockIdx.x - 1];
  const int64_t base = (int64_t)blockIdx.x * kTileItems + (int64_t)threadIdx.x * kPerThread;
  uint32_t v[kPerThread];
  load16<VEC>(data, base, n, v);
#pragma unroll
  for (int i = 0; i < kPerThread; ++i) v[i] += add;
  store16<VEC>(data, base, n, v);
}

// Prefix maximum over uint64: 256 lanes x 8 consecutive values per tile,
// the same three phases as the sum.
constexpr int kMaxPerThread = 8;
constexpr int kMaxTileItems = kScanBlock * kMaxPerThread;

__device__ inline uint64_t umax64(uint64_t a, uint64_t b) { return a > b ? a : b; }

__global__ void __launch_bounds__(kScanBlock) k_block_max(uint64_t *data, int64_t n, uint64_t *tops) {
  __shared__ uint64_t wtop[kScanBlock / 64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t base = (int64_t)blockIdx.x * kMaxTileItems + (int64_t)t * kMaxPerThread;
  uint64_t v[kMaxPerThread];
#pragma unroll
  for (int i = 0; i < kMaxPerThread; ++i) v[i] = base + i < n ? data[base + i] : 0ull;
#pragma unroll
  for (int i = 1; i < kMaxPerThread; ++i) v[i] = umax64(v[i], v[i - 1]);
  uint64_t x = v[kMaxPerThread - 1];
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t y = __shfl_up(x, (unsigned)d, 64);
    if (lane >= d) x = umax64(x, y);
  }
  if (lane == 63) wtop[wave] = x;
  __syncthreads();
  uint64_t before = 0, top = 0;
#pragma unroll
  for (int w = 0; w < kScanBlock / 64; ++w) {
    before = w < wave ? umax64(before, wtop[w]) : before;
    top = umax64(top, wtop[w]);
  }
  // lanes before this one in the wave
  uint64_t prev = __shfl_up(x, 1u, 64);
  if (lane == 0) prev = 0;
  const uint64_t prefix = umax64(before, prev);
#pragma unroll
  for (int i = 0; i < kMaxPerThread; ++i)
    if (base + i < n) data[base + i] = umax64(v[i], prefix);
  if (t == 0) tops[blockIdx.x] = top;
}

__global__ void __launch_bounds__(kScanBlock) k_apply_max(uint64_t *data, int64_t n, const uint64_t *tops) {
  if (blockIdx.x == 0) return;
  const uint64_t m = tops[blockIdx.x - 1];
  const int64_t base = (int64_t)blockIdx.x * kMaxTileItems + (int64_t)threadIdx.x * kMaxPerThread;
#pragma unroll
  for (int i = 0; i < kMaxPerThread; ++i)
    if (base + i < n) data[base + i] = umax64(data[base + i], m);
}

}  // namespace

void Scanner::inclusive_max(uint64_t *data, int64_t n, hipStream_t s, int depth) {
  if (n <= 0) return;
  if (depth >= 4) throw std::runtime_error("scan too deep");
  const int64_t blocks = (n + kMaxTileItems - 1) / kMaxTileItems;
  uint64_t *tops = level64[depth].ensure(blocks);
  hipLaunchKernelGGL(k_block_max, dim3(blocks), dim3(kScanBlock), 0, s, data, n, tops);
  if (blocks > 1) {
    inclusive_max(tops, blocks, s, depth + 1);
    hipLaunchKernelGGL(k_apply_max, dim3(blocks), dim3(kScanBlock), 0, s, data, n, tops);
  }
  MHGP_LAUNCH_CHECK();
}

void Scanner::inclusive(uint32_t *data, int64_t n, hipStream_t s, int depth) {
  if (n <= 0) return;
  if (depth >= 4) throw std::runtime_error("scan too deep");
  const int64_t blocks = (n + kTileItems - 1) / kTileItems;
  uint32_t *bs = level[depth].ensure(blocks);
  const bool vec = ((uintptr_t)data & 15u) == 0;
  if (vec)
    hipLaunchKernelGGL(k_block_scan<true>, dim3(blocks), dim3(kScanBlock), 0, s, data, n, bs);
  else
    hipLaunchKernelGGL(k_block_scan<false>, dim3(blocks), dim3(kScanBlock), 0, s, data, n, bs);
  if (blocks > 1) {
    inclusive(bs, blocks, s, depth + 1);
    if (vec)
      hipLaunchKernelGGL(k_add_offsets<true>, dim3(blocks), dim3(kScanBlock), 0, s, data, n, bs);
    else
      hipLaunchKernelGGL(k_add_offsets<false>, dim3(blocks), dim3(kScanBlock), 0, s, data, n, bs);
  }
  MHGP_LAUNCH_CHECK();
}

}  // namespace mhgp
