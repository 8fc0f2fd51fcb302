// Stable LSD radix sort of (uint64 key, uint32 value) pairs, 8-bit digits.
//
// Pass structure (one pass per digit), on tiles of 4 096 keys:
//   1. tile histograms: 64-key groups of equal digits are found with eight
//      ballots and counted with one LDS atomic per group;
//   2. multi-block inclusive scan over [digit][tile]; the entry before
//      (digit, tile) is that tile's base offset for the digit (digit-major,
//      so equal digits keep tile order);
//   3. scatter: each of the four waves of a tile ranks its 1 024 keys in
//      order (64-key chunks, eight ballots, a per-digit running count in
//      LDS), the tile is reordered by digit in LDS, and the reordered tile is
//      written out in index order, so the stores of a digit run are
//      contiguous.  Key order inside a digit is (wave, chunk, lane), the
//      input order: the pass is stable.
// Passes whose digit is the same for every key (detected on the device by
// the scan) copy instead of scattering, so ping-pong buffers stay consistent.
#include "radix_sort.hpp"

namespace mhgp {

namespace {

constexpr int kTile = 4096;
constexpr int kDigits = 256;
constexpr int kSortBlock = 256;
constexpr int kWaves = kSortBlock / 64;
constexpr int kChunks = kTile / kSortBlock;  // 64-key chunks per wave

__device__ inline uint64_t lanes_below() {
  const int lane = (int)(threadIdx.x & 63u);
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Lanes of the wave holding a valid key with the same digit as this lane.
__device__ inline uint64_t digit_match(bool valid, uint32_t dig) {
  uint64_t match = __ballot(valid);
#pragma unroll
  for (int bit = 0; bit < 8; ++bit) {
    const uint64_t bal = __ballot(valid && ((dig >> bit) & 1u));
    match &= ((dig >> bit) & 1u) ? bal : ~bal;
  }
  return match;
}

__global__ void __launch_bounds__(kSortBlock) rs_histogram(const uint64_t *__restrict__ keys, int64_t n, int shift,
                                                           uint32_t *__restrict__ counts, int tiles) {
  __shared__ uint32_t h[kDigits];
  const int t = threadIdx.x;
  h[t] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
  const uint64_t below = lanes_below();
  for (int i = t; i < kTile; i += kSortBlock) {
    const int64_t g = base + i;
    const bool valid = g < n;
    const uint32_t dig = valid ? (uint32_t)((keys[g] >> shift) & 0xff) : 0u;
    const uint64_t match = digit_match(valid, dig);
    // the lowest lane of each group adds the group
    if (valid && (match & below) == 0) atomicAdd(&h[dig], (uint32_t)__popcll(match));
  }
  __syncthreads();
  counts[(size_t)t * tiles + blockIdx.x] = h[t];
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

__global__ void __launch_bounds__(kSortBlock) rs_scatter(const uint64_t *__restrict__ kin,
                                                         const uint32_t *__restrict__ vin,
                                                         uint64_t *__restrict__ kout,
                                                         uint32_t *__restrict__ vout, int64_t n, int shift,
                                                         const uint32_t *__restrict__ offsets, int tiles,
                                                         const uint32_t *__restrict__ flags) {
  __shared__ uint64_t skey[kTile];
  __shared__ uint32_t sval[kTile];
  __shared__ uint32_t wrun[kWaves][kDigits];  // per wave: running count, then the wave's base
  __shared__ uint32_t lbase[kDigits];         // tile-local start of each digit
  __shared__ uint32_t gbase[kDigits];         // global start of the tile's keys of each digit
  __shared__ uint32_t wtot[kWaves];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t base = (int64_t)blockIdx.x * kTile;
  if (flags[0]) {  // identity pass
    for (int i = t; i < kTile; i += kSortBlock) {
      const int64_t g = base + i;
      if (g < n) {
        kout[g] = kin[g];
        vout[g] = vin[g];
      }
    }
    return;
  }
#pragma unroll
  for (int w = 0; w < kWaves; ++w) wrun[w][t] = 0;
  __syncthreads();
  uint64_t key[kChunks];
  uint32_t val[kChunks], pos[kChunks];
  const int64_t wbase = base + (int64_t)wave * (kChunks * 64) + lane;
#pragma unroll
  for (int c = 0; c < kChunks; ++c) {
    const int64_t g = wbase + c * 64;
    key[c] = g < n ? kin[g] : 0ull;
    val[c] = g < n ? vin[g] : 0u;
  }
  const uint64_t below = lanes_below();
  // rank inside the wave's keys of the same digit, in input order; the
  // running counts are per wave, and one wave's LDS operations take effect
  // in program order, so each chunk reads the counts the previous one left
#pragma unroll
  for (int c = 0; c < kChunks; ++c) {
    const bool valid = wbase + c * 64 < n;
    const uint32_t dig = (uint32_t)((key[c] >> shift) & 0xff);
    const uint64_t match = digit_match(valid, dig);
    const uint32_t rank = (uint32_t)__popcll(match & below);
    const uint32_t cnt = (uint32_t)__popcll(match);
    const uint32_t r0 = valid ? wrun[wave][dig] : 0u;
    pos[c] = r0 + rank;
    if (valid && rank + 1 == cnt) wrun[wave][dig] = r0 + cnt;
  }
  __syncthreads();
  // digit t: wave bases, tile-local start (exclusive scan over the digits)
  // and global start
  uint32_t tot = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const uint32_t c = wrun[w][t];
    wrun[w][t] = tot;
    tot += c;
  }
  uint32_t x = tot;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, (unsigned)d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wtot[wave] = x;
  const size_t e = (size_t)t * tiles + blockIdx.x;
  gbase[t] = e ? offsets[e - 1] : 0u;
  __syncthreads();
  uint32_t before = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) before += w < wave ? wtot[w] : 0u;
  lbase[t] = before + x - tot;
  __syncthreads();
#pragma unroll
  for (int c = 0; c < kChunks; ++c) {
    if (wbase + c * 64 >= n) continue;
    const uint32_t dig = (uint32_t)((key[c] >> shift) & 0xff);
    const uint32_t lp = lbase[dig] + wrun[wave][dig] + pos[c];
    skey[lp] = key[c];
    sval[lp] = val[c];
  }
  __syncthreads();
  const int valid_n = (int)min((int64_t)kTile, n - base);
  for (int i = t; i < valid_n; i += kSortBlock) {
    const uint64_t k = skey[i];
    const uint32_t dig = (uint32_t)((k >> shift) & 0xff);
    const uint32_t gp = gbase[dig] + (uint32_t)i - lbase[dig];
    kout[gp] = k;
    vout[gp] = sval[i];
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
    hipLaunchKernelGGL(rs_histogram, dim3(tiles), dim3(kSortBlock), 0, stream, ka, n, shift, counts.p, tiles);
    scanner.inclusive(counts.p, (int64_t)kDigits * tiles, stream);
    hipLaunchKernelGGL(rs_identity, dim3(1), dim3(kDigits), 0, stream, counts.p, n, tiles, flags.p);
    hipLaunchKernelGGL(rs_scatter, dim3(tiles), dim3(kSortBlock), 0, stream, ka, va, kb, vb, n, shift, counts.p,
                       tiles, flags.p);
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
