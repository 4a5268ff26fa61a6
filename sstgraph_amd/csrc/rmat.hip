// Device RMAT edge generator, bit-exact with rMat<uint32_t>
// (include/SSTGraph/internal/rmat_util.h:82-131).  A normalised variant
// draws r = hash32(x) / 2^32 instead, so the quadrant probabilities are the
// a, b, c given rather than half of them.
//
// The reference draws r = hash32(x) / INT32_MAX in double and compares it
// with a, a+b, a+b+c.  The quotient is monotone in hash32(x), so each
// comparison is equivalent to hash32(x) < T for an integer threshold T that
// the host finds with the same double expression; the device then compares
// integers only.
#include "graph.hpp"

namespace sstg {

namespace {

__device__ __forceinline__ uint32_t hash32_dev(uint32_t a) {
  a = (a + 0x7ed55d16u) + (a << 12u);
  a = (a ^ 0xc761c23cu) ^ (a >> 19u);
  a = (a + 0x165667b1u) + (a << 5u);
  a = (a + 0xd3a2646cu) ^ (a << 9u);
  a = (a + 0xfd7046c5u) + (a << 3u);
  a = (a ^ 0xb55a4f09u) ^ (a >> 16u);
  return a;
}

uint32_t hash32_host(uint32_t a) {
  a = (a + 0x7ed55d16u) + (a << 12u);
  a = (a ^ 0xc761c23cu) ^ (a >> 19u);
  a = (a + 0x165667b1u) + (a << 5u);
  a = (a + 0xd3a2646cu) ^ (a << 9u);
  a = (a + 0xfd7046c5u) + (a << 3u);
  a = (a ^ 0xb55a4f09u) ^ (a >> 16u);
  return a;
}

// smallest h in [0, 2^32] with (double)h / D >= p (2^32: none), where D is
// INT32_MAX (the reference's draw) or 2^32 (the normalised one); mid < 2^32
// throughout the search
uint64_t threshold(double p, bool normalized) {
  const double D = normalized ? 4294967296.0 : (double)2147483647;
  uint64_t lo = 0, hi = 1ull << 32;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if ((double)(uint32_t)mid / D >= p)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

__global__ void k_rmat(uint32_t n, uint32_t lg, uint32_t h, uint64_t ta,
                       uint64_t tab, uint64_t tabc, uint64_t first,
                       uint64_t count, uint32_t *src, uint32_t *dst) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  uint32_t i = (uint32_t)(first + t);
  uint32_t start = hash32_dev((uint32_t)(2u * i) * h);
  uint32_t stride = hash32_dev((uint32_t)(2u * i + 1u) * h);
  uint32_t r = 0, c = 0, half = n;
  for (uint32_t k = 0; k < lg; k++) {
    half >>= 1;
    uint64_t hv = hash32_dev(start + k * stride);
    if (hv < ta) {
    } else if (hv < tab) {
      c += half;
    } else if (hv < tabc) {
      r += half;
    } else {
      r += half;
      c += half;
    }
  }
  src[t] = r;
  dst[t] = c;
}

__device__ __forceinline__ uint64_t hash64_dev(uint64_t u) {
  uint64_t v = u * 3935559000370003845ull + 2691343689449507681ull;
  v ^= v >> 21u;
  v ^= v << 37u;
  v ^= v >> 4u;
  v *= 4768777513237032717ull;
  v ^= v << 20u;
  v ^= v >> 41u;
  v ^= v << 5u;
  return v;
}

// Seeded Erdos-Renyi G(n, m) edge stream: edge i = (hash64(seed<<32 | 2i)
// mod n, hash64(seed<<32 | 2i+1) mod n).
__global__ void k_er(uint32_t n, uint64_t seed, uint64_t first, uint64_t count,
                     uint32_t *src, uint32_t *dst) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  uint64_t i = first + t;
  uint64_t base = seed << 32;
  src[t] = (uint32_t)(hash64_dev(base + 2 * i) % n);
  dst[t] = (uint32_t)(hash64_dev(base + 2 * i + 1) % n);
}

}  // namespace

int er_generate(uint32_t n, uint32_t seed, uint64_t first, uint64_t count,
                uint32_t *src, uint32_t *dst, hipStream_t s) {
  if (n == 0) return SSTG_E_INVAL;
  if (count == 0) return SSTG_OK;
  SSTG_LAUNCH(k_er, ceil_div(count, 256), 256, 0, s, n, (uint64_t)seed, first,
              count, src, dst);
  return SSTG_OK;
}

int rmat_generate(uint32_t n, uint32_t seed, double a, double b, double c,
                  uint64_t first, uint64_t count, uint32_t *src, uint32_t *dst,
                  hipStream_t s, bool normalized) {
  double ab = a + b, abc = a + b + c;
  if (abc > 1) return SSTG_E_INVAL;                      // rmat_util.h:92-95
  if (n == 0 || (n & (n - 1)) != 0) return SSTG_E_INVAL;  // rmat_util.h:96-99
  if (count == 0) return SSTG_OK;
  uint32_t lg = 0;
  while ((1u << lg) < n) lg++;
  uint32_t h = hash32_host(seed);
  SSTG_LAUNCH(k_rmat, ceil_div(count, 256), 256, 0, s, n, lg, h,
              threshold(a, normalized), threshold(ab, normalized),
              threshold(abc, normalized), first, count, src, dst);
  return SSTG_OK;
}

}  // namespace sstg
