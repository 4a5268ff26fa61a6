// One-workgroup sort of up to kBlockSortMax 64-bit keys held in LDS.
//
// The keys carry no payload, so equal keys need no stable order and the
// sort compares whole keys.  Two steps:
//   1. runs: each wave sorts one run of 64 * R keys with a bitonic network,
//      R = 1, 2 or 4 consecutive keys per lane (R = n / 1024 rounded up to
//      a power of two; below 1024 keys, runs of 64 on the first waves);
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

// The value of v held by lane (this lane) ^ J, J a power of two below 64,
// exchanged without LDS: DPP within quads (J = 1, 2) and rows (J = 4, 8: a
// shift up or down by J, chosen by the lane's bit J), and the gfx950 row
// and half-wave swaps for J = 16 and 32 (a swap of v with itself leaves
// each lane's partner value in one of the two results).
template <int J>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v) {
  static_assert(J == 1 || J == 2 || J == 4 || J == 8 || J == 16 || J == 32,
                "lane distance");
  const uint32_t lane = __lane_id();
  if constexpr (J == 1) {
    return __builtin_amdgcn_update_dpp(0u, v, 0xB1, 0xF, 0xF, false);  // [1,0,3,2]
  } else if constexpr (J == 2) {
    return __builtin_amdgcn_update_dpp(0u, v, 0x4E, 0xF, 0xF, false);  // [2,3,0,1]
  } else if constexpr (J == 4 || J == 8) {
    const uint32_t up = __builtin_amdgcn_update_dpp(0u, v, 0x100 + J, 0xF, 0xF, false);
    const uint32_t dn = __builtin_amdgcn_update_dpp(0u, v, 0x110 + J, 0xF, 0xF, false);
    return (lane & J) ? dn : up;  // row_shl:J reads lane + J, row_shr:J lane - J
  } else if constexpr (J == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane & 16) ? r[0] : r[1];
  } else {
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane & 32) ? r[0] : r[1];
  }
}

// lane_xor for a distance known once the caller's loops are unrolled
__device__ __forceinline__ uint32_t lane_xor_at(uint32_t v, uint32_t d) {
  switch (d) {
    case 1: return lane_xor<1>(v);
    case 2: return lane_xor<2>(v);
    case 4: return lane_xor<4>(v);
    case 8: return lane_xor<8>(v);
    case 16: return lane_xor<16>(v);
    default: return lane_xor<32>(v);
  }
}

__device__ __forceinline__ uint64_t lane_xor_at(uint64_t v, uint32_t d) {
  const uint32_t lo = lane_xor_at((uint32_t)v, d);
  const uint32_t hi = lane_xor_at((uint32_t)(v >> 32), d);
  return ((uint64_t)hi << 32) | lo;
}

// A wave's run of 64 * R keys, R consecutive keys per lane (lane l holds
// elements R*l .. R*l + R - 1), through a bitonic network.  Steps whose
// partners are less than R apart stay within a lane's registers; the rest
// pair lane l with lane l ^ (j / R), one exchange per register (lane_xor,
// no LDS).  So a run of 256 keys costs the exchanges of four runs of 64 and
// saves two merge rounds.
// With kPay the payload px follows its key: a pair is exchanged only when
// its keys are out of order, so equal keys stay put and the pairs remain a
// permutation.
template <int R, bool kPay>
__device__ __forceinline__ void wave_bitonic(uint64_t (&x)[R], uint32_t (&px)[R],
                                             uint32_t lane) {
#pragma unroll
  for (uint32_t k = 2; k <= (uint32_t)(kWave * R); k <<= 1) {
#pragma unroll
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      if (j < (uint32_t)R) {
#pragma unroll
        for (int q = 0; q < R; q++) {
          if ((q & j) != 0) continue;
          const int q2 = q | (int)j;
          const uint32_t e = lane * R + (uint32_t)q;
          const bool up = (e & k) == 0;
          const bool swap = up ? x[q2] < x[q] : x[q] < x[q2];
          const uint64_t a = x[q], b = x[q2];
          x[q] = swap ? b : a;
          x[q2] = swap ? a : b;
          if (kPay) {
            const uint32_t pa = px[q], pb = px[q2];
            px[q] = swap ? pb : pa;
            px[q2] = swap ? pa : pb;
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < R; q++) {
          const uint32_t e = lane * R + (uint32_t)q;
          const bool take_min = ((e & j) == 0) == ((e & k) == 0);
          const uint64_t y = lane_xor_at(x[q], j / R);
          const bool take = take_min ? y < x[q] : y > x[q];
          if (kPay) {
            const uint32_t py = lane_xor_at(px[q], j / R);
            px[q] = take ? py : px[q];
          }
          x[q] = take ? y : x[q];
        }
      }
    }
  }
}

// Wave w sorts the run s[w * 64R, +64R) (and its payload) in place.
template <int R, bool kPay>
__device__ __forceinline__ void wave_sort_run(uint64_t *s, uint16_t *v,
                                              uint32_t w, uint32_t lane) {
  uint64_t x[R];
  uint32_t px[R];
  const uint32_t base = w * (uint32_t)(kWave * R) + lane * R;
#pragma unroll
  for (int q = 0; q < R; q++) {
    x[q] = s[base + q];
    px[q] = kPay ? v[base + q] : 0;
  }
  wave_bitonic<R, kPay>(x, px, lane);
#pragma unroll
  for (int q = 0; q < R; q++) {
    s[base + q] = x[q];
    if (kPay) v[base + q] = (uint16_t)px[q];
  }
}

template <int T, bool kPay>
__device__ void block_sort_impl(uint64_t *s, uint16_t *v, uint32_t n) {
  static_assert(T % kWave == 0 && kBlockSortMax % T == 0, "block shape");
  constexpr int kItems = kBlockSortMax / T;  // keys per thread, at most
  static_assert(kItems == 1 || kItems == 2 || kItems == 4, "keys per lane");
  uint32_t np = kWave;
  while (np < n) np <<= 1;
  for (uint32_t i = n + threadIdx.x; i < np; i += T) {
    s[i] = ~0ull;
    if (kPay) v[i] = 0;
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  // 1. runs: one per wave of 64 * np / T keys once np > T (np / T keys per
  // lane), else np / 64 runs of 64 on the first waves
  const uint32_t per = np > (uint32_t)T ? np / T : 1;
  if (per == 1) {
    if (w < np / kWave) wave_sort_run<1, kPay>(s, v, w, lane);
  } else if (per == 2) {
    wave_sort_run<kItems >= 2 ? 2 : 1, kPay>(s, v, w, lane);
  } else {
    wave_sort_run<kItems, kPay>(s, v, w, lane);
  }
  __syncthreads();
  // 2. thread t owns the `per` consecutive keys [t * per, +per), all in one
  // run: the first and last are ranked by a full search of the other run,
  // the ones between inside the range those two bound
  const uint32_t i0 = threadIdx.x * per;
  for (uint32_t len = kWave * per; len < np; len <<= 1) {
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
