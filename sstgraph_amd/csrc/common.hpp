// Shared definitions of the sstgraph_amd HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "sstg.h"

namespace sstg {

constexpr int kWave = 64;

// ---------------------------------------------------------------- errors
struct Status {
  int code = SSTG_OK;
};

#define SSTG_HIP(expr)                                                        \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::sstg::note_hip_error(e_, #expr, __FILE__, __LINE__);                  \
      return e_ == hipErrorOutOfMemory ? SSTG_E_NOMEM : SSTG_E_DEVICE;        \
    }                                                                         \
  } while (0)

#define SSTG_TRY(expr)                                                        \
  do {                                                                        \
    int rc_ = (expr);                                                         \
    if (rc_ != SSTG_OK) return rc_;                                           \
  } while (0)

void note_hip_error(hipError_t e, const char *what, const char *file, int line);

// ------------------------------------------------------------- profiler
// Brackets kernel launches with HIP events on the launch stream when enabled.
struct Profiler {
  struct Rec {
    const char *name;
    hipEvent_t a, b;
  };
  struct Acc {
    std::string name;
    double ms = 0;
    uint64_t n = 0;
  };
  bool on = false;
  std::mutex mu;
  std::vector<Rec> pending;
  std::vector<hipEvent_t> free_events;
  std::vector<Acc> acc;
  hipEvent_t get_event();
  void begin(const char *name, hipStream_t s, hipEvent_t *a);
  void end(const char *name, hipStream_t s, hipEvent_t a);
  void flush();
  void reset();
};
Profiler &profiler();

// Launch helper: LAUNCH(name, kernel, grid, block, shmem, stream, args...)
#define SSTG_LAUNCH(kernel, grid, block, shmem, stream, ...)                  \
  do {                                                                        \
    ::sstg::Profiler &p_ = ::sstg::profiler();                                \
    hipEvent_t ev_ = nullptr;                                                 \
    if (p_.on) p_.begin(#kernel, stream, &ev_);                               \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), shmem, stream,        \
                       __VA_ARGS__);                                          \
    if (p_.on) p_.end(#kernel, stream, ev_);                                  \
  } while (0)

// ---------------------------------------------------------- workspace
// Growable device scratch slots, reused across calls on one graph.  All use
// is ordered on the graph's stream; a slot is only regrown after the stream
// has drained.
struct Workspace {
  struct Slot {
    void *p = nullptr;
    size_t bytes = 0;
  };
  std::vector<Slot> slots;
  hipStream_t stream = nullptr;
  size_t total = 0;
  // Arena mode (begin_arena .. end_arena): get() carves each slot from one
  // fixed device range instead of allocating, and fails with SSTG_E_NOMEM
  // when the range is spent.  A slot asked for again with no more bytes
  // keeps its memory; a larger request takes a fresh piece (its contents
  // are lost, as when an allocated slot grows).
  uint8_t *arena = nullptr;
  size_t arena_bytes = 0, arena_used = 0;
  // bumped whenever a slot gets new memory (or all are released): a
  // recorded launch sequence that holds slot addresses is stale after it
  uint64_t generation = 0;
  std::vector<Slot> arena_slots;
  void begin_arena(void *base, size_t bytes) {
    arena = static_cast<uint8_t *>(base);
    arena_bytes = bytes;
    arena_used = 0;
    arena_slots.clear();
  }
  void end_arena() {
    arena = nullptr;
    arena_bytes = arena_used = 0;
    arena_slots.clear();
  }
  int get(int slot, size_t bytes, void **out);
  template <class T> int get(int slot, size_t count, T **out) {
    void *p = nullptr;
    int rc = get(slot, count * sizeof(T) + 256, &p);
    *out = static_cast<T *>(p);
    return rc;
  }
  void release_all();
};

// ------------------------------------------------------------- helpers
inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
inline uint32_t bit_width_u32(uint32_t x) {
  return x ? 32u - (uint32_t)__builtin_clz(x) : 0u;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// popcount(mask & lanes-below-me)
__device__ __forceinline__ uint32_t mask_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

template <class T>
__device__ __forceinline__ uint64_t lower_bound_dev(const T *a, uint64_t len,
                                                    T x) {
  uint64_t lo = 0, hi = len;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Whether x is in the sorted row a[0..len), by a binary search that keeps
// the value at its upper bound, so membership needs no load after the
// search (the row's degree and offset can be loaded together beforehand).
// *at, if given, gets the first index with a[i] >= x.
__device__ __forceinline__ bool row_has(const uint32_t *a, uint32_t len,
                                        uint32_t x, uint32_t *at = nullptr) {
  uint32_t lo = 0, hi = len, vhi = 0;  // answer in [lo, hi]; a[hi] == vhi
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    const uint32_t v = a[mid];
    if (v < x) {
      lo = mid + 1;
    } else {
      hi = mid;
      vhi = v;
    }
  }
  if (at) *at = lo;
  return lo < len && vhi == x;
}

// index of the segment containing flat position x, given exclusive offsets
// off[0..k) (off[0] == 0, non-decreasing): largest s with off[s] <= x.
__device__ __forceinline__ uint64_t find_segment(const uint64_t *off,
                                                 uint64_t k, uint64_t x) {
  uint64_t lo = 0, hi = k;
  while (hi - lo > 1) {
    uint64_t mid = (lo + hi) >> 1;
    if (off[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// MI355X dispatches workgroups round-robin over its 8 XCDs, each with a
// private L2.  xcd_block() renumbers block b of nb so that each XCD runs a
// contiguous range of logical blocks (XCD x gets the x-th slice), which keeps
// neighbouring vertex ranges -- and the gathers they share -- in one L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t q = nb / kXcds, r = nb % kXcds;
  const uint32_t x = b % kXcds, i = b / kXcds;
  return x * q + (x < r ? x : r) + i;
}

// wave-wide inclusive sum (wave64)
template <class T> __device__ __forceinline__ T wave_inclusive_sum(T v) {
  const uint32_t l = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    T t = __shfl_up(v, o, kWave);
    if (l >= (uint32_t)o) v += t;
  }
  return v;
}

template <class T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// ------------------------------------------------------- device graph
// The tinyset store of one graph in HBM.  Row v holds deg[v] sorted unique
// column ids at pool[off[v] .. off[v] + deg[v]); the slot has room for cap[v]
// ids (a power of two, at least 4).  A row that outgrows its slot moves to a
// fresh slot at the pool tail; the old slot becomes garbage until the next
// rebuild compacts the pool.
struct DevGraph {
  uint32_t n;       // rows
  uint64_t *off;    // [n]
  uint32_t *deg;    // [n]
  uint32_t *cap;    // [n]
  uint32_t *pool;   // [pool_cap]
};

}  // namespace sstg
