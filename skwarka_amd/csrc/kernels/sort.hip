// Stable key-value radix sort of 32-bit keys and values on bits [begin_bit, end_bit) (gfx950).
// Orders the final gather's irradiance lookups by their Morton cells and a shooting round's
// photons by path index; keys equal on the sorted bits keep their input order.
//
// LSD passes of up to 8 bits. Each pass:
//  1. k_sort_hist: one 256-thread block per tile of 4096 keys counts the tile's digits (per-wave
//     LDS counters, summed) into counts[digit * tiles + tile];
//  2. an exclusive scan of that digit-major array (k_scan_reduce / k_scan_top / k_scan_apply)
//     gives every (digit, tile) its first output position;
//  3. k_sort_scatter: the block reads its tile again in 16 rounds of 256 keys (round r holds keys
//     r * 256 .. r * 256 + 255 of the tile, so rounds are in input order). Within a round each
//     wave ranks its lanes by digit with ballots (lanes of equal digit, lower lanes first), the
//     waves' digit counts give each key its rank within the round, and a running count per digit
//     carries the rounds before it: a stable rank within the tile. The tile is placed in LDS in
//     digit order, then written out with consecutive threads on consecutive positions of each
//     digit's run, so the global writes coalesce.
// Passes alternate between the output and a temporary buffer so that the last pass writes the
// output; the input is never written.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"

namespace skw {
namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kWaves = kThreads / 64;
constexpr uint32_t kItems = 16;
constexpr uint32_t kTile = kThreads * kItems;  // keys per block
constexpr uint32_t kBins = 256;
constexpr uint32_t kMaxBits = 8;
constexpr uint32_t kScanTile = 4096;           // scan elements per block
constexpr uint32_t kScanTopThreads = 1024;

__device__ __forceinline__ uint32_t digit_of(uint32_t key, uint32_t shift, uint32_t mask) {
  return (key >> shift) & mask;
}

__global__ __launch_bounds__(kThreads) void k_sort_hist(const uint32_t *__restrict__ keys, uint32_t n,
                                                        uint32_t shift, uint32_t mask, uint32_t tiles,
                                                        uint32_t *__restrict__ counts) {
  __shared__ uint32_t h[kWaves][kBins];
  const uint32_t t = threadIdx.x, wave = t >> 6;
  for (uint32_t w = 0; w < kWaves; ++w) h[w][t] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  for (uint32_t r = 0; r < kItems; ++r) {
    const uint64_t i = base + r * kThreads + t;
    if (i < n) atomicAdd(&h[wave][digit_of(keys[i], shift, mask)], 1u);
  }
  __syncthreads();
  uint32_t c = 0;
  for (uint32_t w = 0; w < kWaves; ++w) c += h[w][t];
  counts[(uint64_t)t * tiles + blockIdx.x] = c;
}

// Block-wide exclusive scan of one value per thread (kThreads or kScanTopThreads threads);
// returns the thread's prefix and sets `total`. `tmp` holds one word per wave.
template <uint32_t NT>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *tmp, uint32_t &total) {
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  uint32_t x = v;
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) tmp[wave] = x;
  __syncthreads();
  uint32_t wsum = 0, all = 0;
  for (uint32_t w = 0; w < NT / 64; ++w) {
    const uint32_t s = tmp[w];
    if (w < wave) wsum += s;
    all += s;
  }
  __syncthreads();
  total = all;
  return wsum + x - v;
}

// Sum of each kScanTile-element chunk.
__global__ __launch_bounds__(kThreads) void k_scan_reduce(const uint32_t *__restrict__ in, uint64_t m,
                                                          uint32_t *__restrict__ partial) {
  __shared__ uint32_t tmp[kWaves];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  uint32_t s = 0;
  for (uint32_t r = 0; r < kScanTile / kThreads; ++r) {
    const uint64_t i = base + r * kThreads + threadIdx.x;
    if (i < m) s += in[i];
  }
  uint32_t total;
  (void)block_exclusive_scan<kThreads>(s, tmp, total);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// Exclusive scan of the chunk sums in place (one block, any count).
__global__ __launch_bounds__(kScanTopThreads) void k_scan_top(uint32_t *__restrict__ partial, uint32_t p) {
  __shared__ uint32_t tmp[kScanTopThreads / 64];
  uint32_t carry = 0;
  for (uint32_t b = 0; b < p; b += kScanTopThreads) {
    const uint32_t i = b + threadIdx.x;
    const uint32_t v = i < p ? partial[i] : 0u;
    uint32_t total;
    const uint32_t x = block_exclusive_scan<kScanTopThreads>(v, tmp, total);
    if (i < p) partial[i] = carry + x;
    carry += total;
  }
}

// Exclusive scan of each chunk, offset by its chunk's prefix. Thread t scans elements
// [t * 16, t * 16 + 16) of the chunk sequentially.
__global__ __launch_bounds__(kThreads) void k_scan_apply(const uint32_t *__restrict__ in, uint64_t m,
                                                         const uint32_t *__restrict__ partial,
                                                         uint32_t *__restrict__ out) {
  __shared__ uint32_t tmp[kWaves];
  constexpr uint32_t kPer = kScanTile / kThreads;
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kPer;
  uint32_t v[kPer];
  uint32_t s = 0;
  for (uint32_t j = 0; j < kPer; ++j) {
    v[j] = base + j < m ? in[base + j] : 0u;
    s += v[j];
  }
  uint32_t total;
  uint32_t run = partial[blockIdx.x] + block_exclusive_scan<kThreads>(s, tmp, total);
  for (uint32_t j = 0; j < kPer; ++j) {
    if (base + j < m) out[base + j] = run;
    run += v[j];
  }
}

__global__ __launch_bounds__(kThreads) void k_sort_scatter(const uint32_t *__restrict__ keys_in,
                                                           const uint32_t *__restrict__ vals_in, uint32_t n,
                                                           uint32_t shift, uint32_t bits, uint32_t tiles,
                                                           const uint32_t *__restrict__ counts,
                                                           const uint32_t *__restrict__ offsets,
                                                           uint32_t *__restrict__ keys_out,
                                                           uint32_t *__restrict__ vals_out) {
  __shared__ uint32_t skeys[kTile];
  __shared__ uint32_t svals[kTile];
  __shared__ uint32_t wcnt[kWaves][kBins];
  __shared__ uint32_t running[kBins];
  __shared__ uint32_t tile_start[kBins];
  __shared__ uint32_t goff[kBins];
  __shared__ uint32_t tmp[kWaves];
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const uint32_t mask = (1u << bits) - 1u;
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  const uint32_t tile_n = (uint32_t)((uint64_t)n - base < kTile ? (uint64_t)n - base : kTile);
  // The tile's digit runs: where each digit starts in the tile (LDS) and in the output.
  const uint32_t cnt = counts[(uint64_t)t * tiles + blockIdx.x];
  uint32_t total;
  tile_start[t] = block_exclusive_scan<kThreads>(cnt, tmp, total);
  goff[t] = offsets[(uint64_t)t * tiles + blockIdx.x];
  running[t] = 0;
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint32_t r = 0; r < kItems; ++r) {
    const uint32_t j = r * kThreads + t;  // position in the tile
    const bool valid = j < tile_n;
    uint32_t key = 0, val = 0, d = 0;
    if (valid) {
      key = keys_in[base + j];
      val = vals_in[base + j];
      d = digit_of(key, shift, mask);
    }
    // lanes of this wave holding a key of the same digit
    uint64_t same = __ballot(valid);
    for (uint32_t b = 0; b < bits; ++b) {
      const uint64_t ones = __ballot(valid && ((d >> b) & 1u));
      same &= ((d >> b) & 1u) ? ones : ~ones;
    }
    for (uint32_t w = 0; w < kWaves; ++w) wcnt[w][t] = 0;
    __syncthreads();
    if (valid && (same & below) == 0) wcnt[wave][d] = (uint32_t)__popcll(same);
    __syncthreads();
    if (valid) {
      uint32_t pre = 0;
      for (uint32_t w = 0; w < wave; ++w) pre += wcnt[w][d];
      const uint32_t pos = tile_start[d] + running[d] + pre + (uint32_t)__popcll(same & below);
      skeys[pos] = key;
      svals[pos] = val;
    }
    __syncthreads();
    uint32_t add = 0;
    for (uint32_t w = 0; w < kWaves; ++w) add += wcnt[w][t];
    running[t] += add;
    __syncthreads();
  }
  for (uint32_t j = t; j < tile_n; j += kThreads) {
    const uint32_t key = skeys[j];
    const uint32_t d = digit_of(key, shift, mask);
    const uint32_t pos = goff[d] + (j - tile_start[d]);
    keys_out[pos] = key;
    vals_out[pos] = svals[j];
  }
}

__global__ __launch_bounds__(256) void k_copy_pairs(const uint32_t *__restrict__ ki, const uint32_t *__restrict__ vi,
                                                    uint32_t *__restrict__ ko, uint32_t *__restrict__ vo, uint32_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ko[i] = ki[i];
  vo[i] = vi[i];
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

hipError_t sort_pairs_u32(void *temp, size_t &temp_bytes, const uint32_t *keys_in, uint32_t *keys_out,
                          const uint32_t *vals_in, uint32_t *vals_out, uint64_t n, int begin_bit, int end_bit,
                          hipStream_t stream) {
  if (n > 0xFFFFFFFFull || begin_bit < 0 || end_bit > 32 || begin_bit > end_bit) return hipErrorInvalidValue;
  const uint32_t nn = (uint32_t)n;
  const uint32_t tiles = (uint32_t)((n + kTile - 1) / kTile);
  const uint64_t m = (uint64_t)kBins * tiles;  // scanned (digit, tile) counts
  const uint32_t chunks = (uint32_t)((m + kScanTile - 1) / kScanTile);
  const size_t b_keys = align256(n * sizeof(uint32_t)), b_counts = align256(m * sizeof(uint32_t));
  const size_t need = 2 * b_keys + 2 * b_counts + align256((size_t)chunks * sizeof(uint32_t)) + 256;
  if (temp == nullptr) {
    temp_bytes = need;
    return hipSuccess;
  }
  if (temp_bytes < need) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  uint8_t *p = reinterpret_cast<uint8_t *>((reinterpret_cast<uintptr_t>(temp) + 255) & ~uintptr_t(255));
  uint32_t *tkeys = reinterpret_cast<uint32_t *>(p);
  uint32_t *tvals = reinterpret_cast<uint32_t *>(p + b_keys);
  uint32_t *counts = reinterpret_cast<uint32_t *>(p + 2 * b_keys);
  uint32_t *offsets = reinterpret_cast<uint32_t *>(p + 2 * b_keys + b_counts);
  uint32_t *partial = reinterpret_cast<uint32_t *>(p + 2 * b_keys + 2 * b_counts);
  const uint32_t span = (uint32_t)(end_bit - begin_bit);
  const uint32_t passes = (span + kMaxBits - 1) / kMaxBits;
  if (passes == 0) {
    hipLaunchKernelGGL(k_copy_pairs, dim3((nn + 255) / 256), dim3(256), 0, stream, keys_in, vals_in, keys_out,
                       vals_out, nn);
    return hipGetLastError();
  }
  const uint32_t *ki = keys_in, *vi = vals_in;
  for (uint32_t ps = 0; ps < passes; ++ps) {
    const uint32_t shift = (uint32_t)begin_bit + ps * kMaxBits;
    const uint32_t bits = std::min<uint32_t>(kMaxBits, (uint32_t)end_bit - shift);
    // the last pass writes the output
    const bool to_out = ((passes - 1 - ps) & 1u) == 0;
    uint32_t *ko = to_out ? keys_out : tkeys, *vo = to_out ? vals_out : tvals;
    hipLaunchKernelGGL(k_sort_hist, dim3(tiles), dim3(kThreads), 0, stream, ki, nn, shift, (1u << bits) - 1u, tiles,
                       counts);
    hipLaunchKernelGGL(k_scan_reduce, dim3(chunks), dim3(kThreads), 0, stream, counts, m, partial);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanTopThreads), 0, stream, partial, chunks);
    hipLaunchKernelGGL(k_scan_apply, dim3(chunks), dim3(kThreads), 0, stream, counts, m, partial, offsets);
    hipLaunchKernelGGL(k_sort_scatter, dim3(tiles), dim3(kThreads), 0, stream, ki, vi, nn, shift, bits, tiles, counts,
                       offsets, ko, vo);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    ki = ko;
    vi = vo;
  }
  return hipSuccess;
}

}  // namespace skw
