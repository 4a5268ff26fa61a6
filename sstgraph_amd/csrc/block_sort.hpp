// One-workgroup sort of up to kBlockSortMax 64-bit keys held in LDS.
//
// The keys carry no payload, so equal keys need no stable order and the
// sort compares whole keys.  Two steps:
//   1. runs of 64: each wave sorts its runs across its lanes with a bitonic
//      network (21 compare-exchange steps over lane shuffles), all of its
//      runs side by side;
//   2. the runs are merged pairwise by rank, doubling the run length per
//      round: a key's place in the merged pair is its index in its own run
//      plus the number of keys of the other run before it (strictly smaller
//      keys for a key of the left run, smaller or equal for the right run,
//      so equal keys keep left before right and every place is taken once).
//      A thread owns up to four consecutive keys of one run: the first and
//      the last are ranked by a fixed-step search over the other run (a
//      power of two long), the ones between by a search inside the range
//      those two ranks bound (nearly always a few keys).
// The length is padded to a power of two (at least 64) with all-ones keys.
// They sort to the end, and since they equal any all-ones key of the input,
// the first n keys afterwards are the input sorted whatever it holds.
// Replaces an 8-bit LSD radix sort of six passes (four barriers and a
// sixteen-step column scan each) for these sizes.
#pragma once

#include "common.hpp"

namespace sstg {

constexpr uint32_t kBlockSortMax = 4096;

// number of a[0..len) before x: a[i] < x, or a[i] <= x when !strict;
// a nondecreasing, len a power of two
__device__ __forceinline__ uint32_t rank_in_run(const uint64_t *a, uint32_t len,
                                                uint64_t x, bool strict) {
  uint32_t b = 0;
  for (uint32_t h = len >> 1; h > 0; h >>= 1) {
    const uint64_t v = a[b + h - 1];
    if (v < x || (!strict && v == x)) b += h;
  }
  const uint64_t v = a[b];
  return b + (v < x || (!strict && v == x));
}

// the same count, known to lie in [lo, hi]
__device__ __forceinline__ uint32_t rank_between(const uint64_t *a,
                                                 uint32_t lo, uint32_t hi,
                                                 uint64_t x, bool strict) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    const uint64_t v = a[mid];
    if (v < x || (!strict && v == x))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// The bitonic network over the 64 lanes of a wave for the first R of a
// wave's runs (one key per lane each), all R side by side.  With kPay the
// payload px follows its key: a lane takes its partner's key and payload
// only when the partner's key is strictly smaller (larger) than its own,
// so equal keys stay put and the pairs remain a permutation.
template <int R, bool kPay, int N>
__device__ __forceinline__ void wave_bitonic(uint64_t (&x)[N], uint32_t (&px)[N],
                                             uint32_t lane) {
#pragma unroll
  for (uint32_t k = 2; k <= (uint32_t)kWave; k <<= 1) {
#pragma unroll
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      const bool take_min = ((lane & j) == 0) == ((lane & k) == 0);
#pragma unroll
      for (int q = 0; q < R; q++) {
        const uint64_t y = __shfl_xor(x[q], (int)j, kWave);
        const bool take = take_min ? y < x[q] : y > x[q];
        if (kPay) {
          const uint32_t py = __shfl_xor(px[q], (int)j, kWave);
          px[q] = take ? py : px[q];
        }
        x[q] = take ? y : x[q];
      }
    }
  }
}

template <int T, bool kPay>
__device__ void block_sort_impl(uint64_t *s, uint16_t *v, uint32_t n) {
  static_assert(T % kWave == 0 && kBlockSortMax % T == 0, "block shape");
  constexpr int kWaves = T / kWave;
  constexpr int kItems = kBlockSortMax / T;          // keys per thread, at most
  constexpr int kRuns = kBlockSortMax / kWave / kWaves;  // 64-key runs per wave
  uint32_t np = kWave;
  while (np < n) np <<= 1;
  for (uint32_t i = n + threadIdx.x; i < np; i += T) {
    s[i] = ~0ull;
    if (kPay) v[i] = 0;
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  // 1. a wave's runs (w, w + kWaves, ...) go through the network together,
  // so their shuffles overlap.  The shuffles (two lane permutes per step
  // for a 64-bit key) bound this step, so runs a wave does not own are
  // skipped (a branch uniform across the wave)
  {
    const uint32_t runs = np / kWave;
    const uint32_t mine = __builtin_amdgcn_readfirstlane(
        runs > w ? (runs - w + kWaves - 1) / kWaves : 0);
    uint64_t x[kRuns];
    uint32_t px[kRuns];
#pragma unroll
    for (int q = 0; q < kRuns; q++) {
      const uint32_t r = w + (uint32_t)q * kWaves;
      x[q] = r < runs ? s[r * kWave + lane] : ~0ull;
      px[q] = kPay && r < runs ? v[r * kWave + lane] : 0;
    }
    static_assert(kRuns >= 1 && kRuns <= 4, "runs per wave");
    if (mine >= 4)
      wave_bitonic<kRuns, kPay>(x, px, lane);
    else if (mine == 3)
      wave_bitonic<kRuns < 3 ? kRuns : 3, kPay>(x, px, lane);
    else if (mine == 2)
      wave_bitonic<kRuns < 2 ? kRuns : 2, kPay>(x, px, lane);
    else if (mine == 1)
      wave_bitonic<1, kPay>(x, px, lane);
#pragma unroll
    for (int q = 0; q < kRuns; q++) {
      const uint32_t r = w + (uint32_t)q * kWaves;
      if (r < runs) {
        s[r * kWave + lane] = x[q];
        if (kPay) v[r * kWave + lane] = (uint16_t)px[q];
      }
    }
  }
  __syncthreads();
  // 2. thread t owns the `per` consecutive keys [t * per, +per), all in one
  // run: the first and last are ranked by a full search of the other run,
  // the ones between inside the range those two bound
  const uint32_t per = np > (uint32_t)T ? np / T : 1;
  const uint32_t i0 = threadIdx.x * per;
  for (uint32_t len = kWave; len < np; len <<= 1) {
    uint64_t x[kItems];
    uint16_t px[kItems];
    uint32_t at[kItems];
    if (i0 < np) {
      const uint32_t pair = i0 & ~(2 * len - 1);
      const bool left = (i0 & len) == 0;
      const uint64_t *other = s + (left ? pair + len : pair);
      // a key's place: its index less len for the right run, plus its rank
      const uint32_t own = i0 - (left ? 0 : len);
#pragma unroll
      for (int j = 0; j < kItems; j++)
        if ((uint32_t)j < per) {
          x[j] = s[i0 + j];
          if (kPay) px[j] = v[i0 + j];
        }
      uint64_t xl = x[0];  // the last key (no dynamic register index)
#pragma unroll
      for (int j = 1; j < kItems; j++)
        if ((uint32_t)j == per - 1) xl = x[j];
      const uint32_t c0 = rank_in_run(other, len, x[0], left);
      const uint32_t c1 = per > 1 ? rank_in_run(other, len, xl, left) : c0;
#pragma unroll
      for (int j = 0; j < kItems; j++) {
        if ((uint32_t)j >= per) break;
        const uint32_t c = j == 0 ? c0
                           : (uint32_t)j == per - 1
                               ? c1
                               : rank_between(other, c0, c1, x[j], left);
        at[j] = own + (uint32_t)j + c;
      }
    }
    __syncthreads();
    if (i0 < np) {
#pragma unroll
      for (int j = 0; j < kItems; j++)
        if ((uint32_t)j < per) {
          s[at[j]] = x[j];
          if (kPay) v[at[j]] = px[j];
        }
    }
    __syncthreads();
  }
}

// Sorts s[0, n) ascending, n <= kBlockSortMax, with all T threads of the
// block (every thread must call it).  s must have room for the padded
// length: the next power of two of max(n, 64).  Ends with a barrier.
template <int T>
__device__ void block_sort_u64(uint64_t *s, uint32_t n) {
  block_sort_impl<T, false>(s, nullptr, n);
}

// The same, moving the 16-bit payload v[i] with key s[i] (v has the same
// room; padding entries get payload 0).
template <int T>
__device__ void block_sort_u64_u16(uint64_t *s, uint16_t *v, uint32_t n) {
  block_sort_impl<T, true>(s, v, n);
}

}  // namespace sstg
