// Three-phase inclusive scan: per-block scans with block totals, a scan of
// the totals (recursively), then a uniform add of the block offsets.
#include <stdexcept>

#include "scan.hpp"

namespace mhgp {

namespace {

constexpr int kScanBlock = 1024;
constexpr int kPerThread = 4;
constexpr int kTileItems = kScanBlock * kPerThread;

__global__ void __launch_bounds__(kScanBlock) k_block_scan(uint32_t *data, int64_t n, uint32_t *sums) {
  __shared__ uint32_t s[kScanBlock];
  const int t = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * kTileItems + (int64_t)t * kPerThread;
  uint32_t v[kPerThread];
  uint32_t run = 0;
  for (int i = 0; i < kPerThread; ++i) {
    v[i] = (base + i < n) ? data[base + i] : 0;
    run += v[i];
    v[i] = run;
  }
  s[t] = run;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    uint32_t x = t >= off ? s[t - off] : 0;
    __syncthreads();
    s[t] += x;
    __syncthreads();
  }
  const uint32_t prefix = s[t] - run;
  for (int i = 0; i < kPerThread; ++i)
    if (base + i < n) data[base + i] = v[i] + prefix;
  if (t == kScanBlock - 1) sums[blockIdx.x] = s[t];
}

__global__ void k_add_offsets(uint32_t *data, int64_t n, const uint32_t *sums) {
  if (blockIdx.x == 0) return;
  const uint32_t add = sums[blockIdx.x - 1];
  const int64_t base = (int64_t)blockIdx.x * kTileItems;
  for (int i = threadIdx.x; i < kTileItems; i += blockDim.x)
    if (base + i < n) data[base + i] += add;
}

}  // namespace

void Scanner::inclusive(uint32_t *data, int64_t n, hipStream_t s, int depth) {
  if (n <= 0) return;
  if (depth >= 4) throw std::runtime_error("scan too deep");
  const int64_t blocks = (n + kTileItems - 1) / kTileItems;
  uint32_t *bs = level[depth].ensure(blocks);
  hipLaunchKernelGGL(k_block_scan, dim3(blocks), dim3(kScanBlock), 0, s, data, n, bs);
  if (blocks > 1) {
    inclusive(bs, blocks, s, depth + 1);
    hipLaunchKernelGGL(k_add_offsets, dim3(blocks), dim3(256), 0, s, data, n, bs);
  }
  MHGP_LAUNCH_CHECK();
}

}  // namespace mhgp
