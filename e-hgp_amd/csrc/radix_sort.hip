// Stable LSD radix sort of (uint64 key, uint32 value) pairs, 8-bit digits.
//
// Pass structure (one pass per digit):
//   1. tile histograms: each 1024-key tile counts its digits in LDS;
//   2. multi-block inclusive scan over [digit][tile]; the entry before
//      (digit, tile) is that tile's base offset for the digit (digit-major,
//      so equal digits keep tile order);
//   3. scatter: one 64-lane wavefront per tile walks its keys in order, ranks
//      equal digits inside each 64-key chunk with 8 ballots, and keeps a
//      per-digit running count in LDS, which makes the pass stable.
// Passes whose digit is the same for every key (detected on the device by
// the scan) copy instead of scattering, so ping-pong buffers stay consistent.
#include "radix_sort.hpp"

namespace mhgp {

namespace {

constexpr int kTile = 1024;
constexpr int kDigits = 256;

__global__ void rs_histogram(const uint64_t *__restrict__ keys, int64_t n, int shift,
                             uint32_t *__restrict__ counts, int tiles) {
  __shared__ uint32_t h[kDigits];
  for (int i = threadIdx.x; i < kDigits; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
  for (int i = threadIdx.x; i < kTile; i += blockDim.x) {
    int64_t g = base + i;
    if (g < n) atomicAdd(&h[(keys[g] >> shift) & 0xff], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kDigits; d += blockDim.x) counts[(size_t)d * tiles + blockIdx.x] = h[d];
}

// flags[0] = 1 when one digit holds every key (the pass is the identity
// permutation); incl is the inclusive scan of the digit-major counts.
__global__ void rs_identity(const uint32_t *__restrict__ incl, int64_t n, int tiles, uint32_t *flags) {
  __shared__ uint32_t full;
  const int d = threadIdx.x;
  if (d == 0) full = 0;
  __syncthreads();
  const uint32_t lo = d == 0 ? 0u : incl[(size_t)d * tiles - 1];
  const uint32_t hi = incl[(size_t)(d + 1) * tiles - 1];
  if (hi - lo == (uint32_t)n) full = 1;
  __syncthreads();
  if (d == 0) flags[0] = full;
}

__global__ void __launch_bounds__(64) rs_scatter(const uint64_t *__restrict__ kin,
                                                 const uint32_t *__restrict__ vin,
                                                 uint64_t *__restrict__ kout,
                                                 uint32_t *__restrict__ vout, int64_t n, int shift,
                                                 const uint32_t *__restrict__ offsets, int tiles,
                                                 const uint32_t *__restrict__ flags) {
  __shared__ uint32_t run[kDigits];
  const int lane = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * kTile;
  if (flags[0]) {  // identity pass
    for (int i = lane; i < kTile; i += 64) {
      int64_t g = base + i;
      if (g < n) { kout[g] = kin[g]; vout[g] = vin[g]; }
    }
    return;
  }
  // exclusive offset of (digit, tile) = inclusive scan at the previous entry
  for (int d = lane; d < kDigits; d += 64) {
    const size_t e = (size_t)d * tiles + blockIdx.x;
    run[d] = e ? offsets[e - 1] : 0u;
  }
  __syncthreads();
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int c = 0; c < kTile; c += 64) {
    const int64_t g = base + c + lane;
    const bool valid = g < n;
    uint64_t key = valid ? kin[g] : 0;
    uint32_t val = valid ? vin[g] : 0;
    uint32_t dig = valid ? (uint32_t)((key >> shift) & 0xff) : 0;
    uint64_t match = __ballot(valid);
    for (int bit = 0; bit < 8; ++bit) {
      uint64_t bal = __ballot(valid && ((dig >> bit) & 1u));
      match &= ((dig >> bit) & 1u) ? bal : ~bal;
    }
    const uint32_t rank = (uint32_t)__popcll(match & below);
    const uint32_t cnt = (uint32_t)__popcll(match);
    uint32_t pos = 0;
    if (valid) pos = run[dig] + rank;
    __syncthreads();
    if (valid) {
      kout[pos] = key;
      vout[pos] = val;
      // the last lane of each digit group advances the running count
      if (rank + 1 == cnt) run[dig] += cnt;
    }
    __syncthreads();
  }
}

}  // namespace

void RadixSorter::sort(uint64_t *keys, uint32_t *vals, int64_t n, int begin_bit, int end_bit,
                       hipStream_t stream) {
  if (n <= 1) return;
  const int tiles = blocks_for(n, kTile);
  tmp_keys.ensure(n);
  tmp_vals.ensure(n);
  counts.ensure((size_t)kDigits * tiles);
  flags.ensure(1);
  uint64_t *ka = keys, *kb = tmp_keys.p;
  uint32_t *va = vals, *vb = tmp_vals.p;
  int passes = 0;
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    hipLaunchKernelGGL(rs_histogram, dim3(tiles), dim3(256), 0, stream, ka, n, shift, counts.p,
                       tiles);
    scanner.inclusive(counts.p, (int64_t)kDigits * tiles, stream);
    hipLaunchKernelGGL(rs_identity, dim3(1), dim3(kDigits), 0, stream, counts.p, n, tiles, flags.p);
    hipLaunchKernelGGL(rs_scatter, dim3(tiles), dim3(64), 0, stream, ka, va, kb, vb, n, shift,
                       counts.p, tiles, flags.p);
    MHGP_LAUNCH_CHECK();
    std::swap(ka, kb);
    std::swap(va, vb);
    passes++;
  }
  if (passes & 1) {
    MHGP_HIP(hipMemcpyAsync(keys, ka, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, stream));
    MHGP_HIP(hipMemcpyAsync(vals, va, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
  }
}

}  // namespace mhgp
