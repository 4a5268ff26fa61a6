// EdgeMap/VertexMap algorithms over the device-resident tinyset store.
//
// The reference calls EdgeMapVertexMap::{BFS, PR_S, BC, CC, TC, BF, TouchAll}
// (include/SSTGraph/internal/test.hpp:476-618), which scan neighbours through
// SparseMatrixV::map_neighbors / map_range (SparseMatrix.hpp:174-315).  Here a
// row's neighbours are pool[off[v] .. off[v] + deg[v]) in ascending order.
#include <cstdlib>

#include <chrono>

#include "algo.hpp"
#include "cc.hpp"
#include "edge_map.hpp"
#include "primitives.hpp"

namespace sstg {

namespace {

constexpr int kT = 256;
constexpr uint32_t kInf = 0xFFFFFFFFu;

inline uint64_t blocks(uint64_t n) { return ceil_div(n ? n : 1, kT); }

// ----------------------------------------------------------------- BFS
// Top-down (sparse push) step of a BFS level as an EdgeMap functor: the
// first writer of an unvisited neighbour claims it.
struct BfsPush {
  uint32_t *level;
  int32_t *parent;
  uint32_t next_level;
  __device__ __forceinline__ uint32_t peek(uint32_t d) const {
    return level[d];
  }
  __device__ static __forceinline__ bool open(uint32_t x) { return x == kInf; }
  __device__ __forceinline__ uint32_t try_claim(uint32_t d) const {
    return atomicCAS(&level[d], kInf, next_level);
  }
  __device__ __forceinline__ void claimed(uint32_t s, uint32_t d) const {
    if (parent) parent[d] = (int32_t)s;
  }
};

// Bottom-up (dense pull) step: an unvisited vertex takes the first frontier
// neighbour it finds.
struct BfsPull {
  uint32_t *level;
  int32_t *parent;
  uint32_t next_level;
  __device__ __forceinline__ bool cond(uint32_t d) const {
    return level[d] == kInf;
  }
  __device__ __forceinline__ void update(uint32_t s, uint32_t d) const {
    level[d] = next_level;
    if (parent) parent[d] = (int32_t)s;
  }
  __device__ __forceinline__ bool in_output(uint32_t v) const {
    return level[v] == next_level;
  }
};

// frontier and visited predicates of the BFS bitmaps
struct LevelIs {
  const uint32_t *level;
  uint32_t L;
  __device__ __forceinline__ bool operator()(uint32_t v) const {
    return level[v] == L;
  }
};
struct Reached {
  const uint32_t *level;
  __device__ __forceinline__ bool operator()(uint32_t v) const {
    return level[v] != kInf;
  }
};

// Start of a BFS from src: levels and parents reset, the queue holding src,
// the visited filter holding only src (threads cover max(n, words)); info
// gets src's degree and the degree prefix at n / 2 and at n.
__global__ void k_bfs_start(uint32_t n, uint32_t *level, int32_t *parent,
                            uint32_t src, uint32_t *queue, uint32_t *vis,
                            uint64_t words, const uint32_t *deg,
                            const uint64_t *prefix,
                            unsigned long long *info) {
  const uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v < n) {
    level[v] = (v == src) ? 0u : kInf;
    if (parent) parent[v] = (v == src) ? (int32_t)src : -1;
  }
  if (v < words) vis[v] = (v == (src >> 5)) ? 1u << (src & 31) : 0u;
  if (v == 0) {
    queue[0] = src;
    info[0] = deg[src];
    info[1] = prefix[n / 2];
    info[2] = prefix[n];
  }
}

// The tail kernel's starting state, written by a kernel so that a captured
// launch sequence holds the values themselves.
struct TailState {
  uint64_t w[7];
};
__global__ void k_tail_state(uint64_t *st, TailState t) {
  if (threadIdx.x < 7) st[threadIdx.x] = t.w[threadIdx.x];
}

__global__ void k_level_dist(const uint32_t *level, uint32_t n, int32_t *out) {
  uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v < n) out[v] = level[v] == kInf ? (int32_t)(0x7FFFFFFF / 2)
                                       : (int32_t)level[v];
}

// ------------------------------------------------------------ PageRank
template <class Real>
__global__ void k_pr_init(uint32_t n, const uint32_t *deg, Real x0, Real *rank,
                          Real *contrib) {
  uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v >= n) return;
  rank[v] = x0;
  uint32_t d = deg[v];
  contrib[v] = d ? x0 / (Real)d : (Real)0;
}

// ------------------------------------------------------------------ CC
__global__ void k_cc_final(uint32_t n, uint32_t *p, uint32_t *out) {
  uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v >= n) return;
  out[v] = cc::find(p, (uint32_t)v);
}

// ------------------------------------------------------------------ TC
// lower[v] = number of neighbours of v below v.
__global__ void k_lower_degree(DevGraph g, uint32_t *lower) {
  uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v >= g.n) return;
  uint32_t d = g.deg[v];
  lower[v] = d ? (uint32_t)lower_bound_dev(g.pool + g.off[v], (uint64_t)d,
                                           (uint32_t)v)
               : 0;
}

// |a[0..na) n b[0..nb)| for sorted unique arrays: binary-search the shorter
// into the longer with a moving lower bound, or a linear merge when the two
// are of similar length.
__device__ __forceinline__ uint64_t intersect_count(const uint32_t *a,
                                                    uint32_t na,
                                                    const uint32_t *b,
                                                    uint32_t nb) {
  if (na > nb) {
    const uint32_t *t = a;
    a = b;
    b = t;
    uint32_t tn = na;
    na = nb;
    nb = tn;
  }
  uint64_t c = 0;
  if (na == 0) return 0;
  if ((uint64_t)na * 16 < nb) {
    uint32_t lo = 0;
    for (uint32_t i = 0; i < na && lo < nb; i++) {
      uint32_t x = a[i];
      uint32_t hi = nb;
      while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (b[mid] < x)
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo < nb && b[lo] == x) c++;
    }
  } else {
    uint32_t i = 0, j = 0;
    while (i < na && j < nb) {
      uint32_t x = a[i], y = b[j];
      c += (x == y);
      i += (x <= y);
      j += (y <= x);
    }
  }
  return c;
}

// One thread per stored edge (u, v): for v < u at position k of N(u), the
// neighbours of u below v are N(u)[0..k); count those also in
// N(v)[0..lower[v]) (common_neighbors(u, v, early_end=true), TinySet_small.
// hpp:357-407).
__global__ void k_tc_edges(DevGraph g, const uint64_t *prefix, uint64_t m,
                           const uint32_t *lower, unsigned long long *out) {
  uint64_t e0 = ((uint64_t)blockIdx.x * kT + threadIdx.x) * 4;
  uint64_t acc = 0;
  if (e0 < m) {
    uint64_t end = e0 + 4 < m ? e0 + 4 : m;
    uint64_t u = find_segment(prefix, g.n, e0);
    for (uint64_t e = e0; e < end; e++) {
      while (e >= prefix[u + 1]) u++;
      uint32_t k = (uint32_t)(e - prefix[u]);
      const uint32_t *nu = g.pool + g.off[u];
      uint32_t v = nu[k];
      if (v >= (uint32_t)u || k == 0) continue;
      uint32_t lv = lower[v];
      if (lv == 0) continue;
      acc += intersect_count(nu, k, g.pool + g.off[v], lv);
    }
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & (kWave - 1)) == 0 && acc)
    atomicAdd(out, (unsigned long long)acc);
}

__global__ void k_common(DevGraph g, uint32_t u, uint32_t v, int early_end,
                         unsigned long long *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint64_t ea = early_end ? u : 0xFFFFFFFFFFFFFFFFull;
  uint64_t eb = early_end ? v : 0xFFFFFFFFFFFFFFFFull;
  const uint32_t *a = g.pool + g.off[u], *b = g.pool + g.off[v];
  uint32_t na = g.deg[u], nb = g.deg[v], i = 0, j = 0;
  unsigned long long c = 0;
  while (i < na && j < nb) {
    uint64_t x = a[i], y = b[j];
    if (x >= ea || y >= eb) break;
    c += (x == y);
    i += (x <= y);
    j += (y <= x);
  }
  *out = c;
}

// ------------------------------------------------------------------ BC
// sigma(w) = sum of sigma over neighbours one level up; vertex list = the
// BFS order slice of one level.
struct SigmaPull {
  const uint32_t *level;
  double *sigma;
  uint32_t L;  // level of the vertices being computed
  using Acc = double;
  __device__ __forceinline__ Acc zero() const { return 0.0; }
  __device__ __forceinline__ void add(Acc &a, uint32_t, uint32_t u) const {
    if (level[u] + 1 == L) a += sigma[u];
  }
  __device__ __forceinline__ Acc merge(Acc a, Acc b) const { return a + b; }
  __device__ static __forceinline__ Acc shfl_xor(Acc a, int o) {
    return __shfl_xor(a, o, kWave);
  }
  __device__ __forceinline__ void store(uint32_t v, Acc s) const {
    sigma[v] = s;
  }
};

// delta(v) = sum over children w of sigma(v)/sigma(w) * (1 + delta(w)).
struct DeltaPull {
  const uint32_t *level;
  const double *sigma;
  double *delta;
  uint32_t L;
  using Acc = double;
  __device__ __forceinline__ Acc zero() const { return 0.0; }
  __device__ __forceinline__ void add(Acc &a, uint32_t v, uint32_t w) const {
    if (level[w] == L + 1) a += sigma[v] / sigma[w] * (1.0 + delta[w]);
  }
  __device__ __forceinline__ Acc merge(Acc a, Acc b) const { return a + b; }
  __device__ static __forceinline__ Acc shfl_xor(Acc a, int o) {
    return __shfl_xor(a, o, kWave);
  }
  __device__ __forceinline__ void store(uint32_t v, Acc s) const {
    delta[v] = s;
  }
};

// Brandes pulls over one BFS level's vertices (a graph of fewer than 64
// levels) against a bitmap per level: the per-edge test reads a 2 MB bitmap
// at rMat scale 24, which stays in an XCD's L2, where a byte per vertex
// (16 MB) or the u32 levels (64 MB) miss it; sigma and the quotients are
// gathered only for the parents (forward) and children (backward) it
// selects.  (Keeping the values of each level densely by rank within the
// level, one more prefix load per selected edge, measured slower: 0.64
// against 0.57 ms for the 16-lane sigma pull of a level at scale 24.)
struct SigmaPullBm {
  const uint32_t *prev;  // bitmap of level L - 1
  double *sigma;
  using Acc = double;
  __device__ __forceinline__ Acc zero() const { return 0.0; }
  // branch-free: a non-parent reads v's own (cached) slot, so the gathers
  // of the unrolled neighbours can all be in flight at once
  __device__ __forceinline__ void add(Acc &a, uint32_t v, uint32_t u) const {
    const bool p = (prev[u >> 5] >> (u & 31)) & 1u;
    const double x = sigma[p ? u : v];
    a += p ? x : 0.0;
  }
  __device__ __forceinline__ Acc merge(Acc a, Acc b) const { return a + b; }
  __device__ static __forceinline__ Acc shfl_xor(Acc a, int o) {
    return __shfl_xor(a, o, kWave);
  }
  __device__ __forceinline__ void store(uint32_t v, Acc s) const {
    sigma[v] = s;
  }
};

// delta(v) = sigma(v) * sum over children w of (1 + delta(w)) / sigma(w).
// Once a vertex's delta is final its sigma slot takes the quotient
// (1 + delta) / sigma, which its parents' pull (one level up, the next
// launch) gathers: one 8-byte value and no divide per child.  Within a
// launch only level-L slots are written and only level-(L+1) slots read.
struct DeltaPullBm {
  const uint32_t *next;  // bitmap of level L + 1
  double *q;  // sigma below level L+1, (1 + delta) / sigma from L+1 down
  double *delta;
  using Acc = double;
  __device__ __forceinline__ Acc zero() const { return 0.0; }
  __device__ __forceinline__ void add(Acc &a, uint32_t v, uint32_t w) const {
    const bool c = (next[w >> 5] >> (w & 31)) & 1u;
    const double x = q[c ? w : v];
    a += c ? x : 0.0;
  }
  __device__ __forceinline__ Acc merge(Acc a, Acc b) const { return a + b; }
  __device__ static __forceinline__ Acc shfl_xor(Acc a, int o) {
    return __shfl_xor(a, o, kWave);
  }
  __device__ __forceinline__ void store(uint32_t v, Acc s) const {
    const double sv = q[v], d = sv * s;
    delta[v] = d;
    q[v] = (1.0 + d) / sv;
  }
};

// the deepest level's quotients (1 + 0) / sigma
__global__ void k_bc_last_quotients(const uint32_t *list, uint64_t count,
                                    double *q) {
  const uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i < count) q[list[i]] = 1.0 / q[list[i]];
}

// bm + l * words: the bitmap of level l, for l < levels; a wave covers
// kLevelRounds x 64 vertices, one lane per vertex and round, a ballot per
// level and round (no atomics).  All of a lane's level loads are issued
// before the first ballot (one load per thread waited for each in turn):
// past n a lane reads vertex n - 1 and counts it on no level.
constexpr int kLevelRounds = 8;
__global__ void k_level_bitmaps(const uint32_t *level, uint32_t n,
                                uint32_t levels, uint32_t *bm,
                                uint64_t words) {
  if (n == 0) return;
  const uint32_t lane = lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * kT + threadIdx.x) / kWave;
  const uint64_t v0 = wave * kLevelRounds * kWave + lane;
  uint32_t lv[kLevelRounds];
#pragma unroll
  for (int r = 0; r < kLevelRounds; r++) {
    const uint64_t v = v0 + (uint64_t)r * kWave;
    const uint32_t x = level[v < n ? v : n - 1];
    lv[r] = v < n ? x : kInf;
  }
  for (uint32_t l = 0; l < levels; l++) {
#pragma unroll
    for (int r = 0; r < kLevelRounds; r++) {
      const uint64_t m = __ballot(lv[r] == l);
      const uint64_t w = (v0 + (uint64_t)r * kWave) >> 5;
      if ((lane & 31) == 0 && w < words)
        bm[l * words + w] = (lane == 0) ? (uint32_t)m : (uint32_t)(m >> 32);
    }
  }
}

// Sort key of the i-th reached vertex: (level * 4 + lane class) << 32 | v,
// the classes of the reductions' degree bins: 0 degree <= t0 (1 lane),
// 1 <= t1 (8 lanes), 2 <= t2 (16 lanes), 3 larger (a block)
__global__ void k_bc_keys(const uint32_t *order, uint64_t count,
                          const uint32_t *level, const uint32_t *deg,
                          uint32_t t0, uint32_t t1, uint32_t t2,
                          uint64_t *key) {
  const uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= count) return;
  const uint32_t v = order[i], d = deg[v];
  const uint32_t c = d <= t0 ? 0u : d <= t1 ? 1u : d <= t2 ? 2u : 3u;
  key[i] = ((uint64_t)(level[v] * 4u + c) << 32) | v;
}

// first index of each (level, class) bucket in the sorted keys (lower bound
// of bucket << 32), buckets 0 .. nb
__global__ void k_bc_bucket_starts(const uint64_t *sorted, uint64_t count,
                                   uint32_t nb, uint64_t *start) {
  const uint32_t b = blockIdx.x * kT + threadIdx.x;
  if (b > nb) return;
  const uint64_t want = (uint64_t)b << 32;
  uint64_t lo = 0, hi = count;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (sorted[mid] < want)
      lo = mid + 1;
    else
      hi = mid;
  }
  start[b] = lo;
}

// sigma push for directed graphs: every vertex v of level L-1 adds sigma(v)
// to its out-neighbours on level L (integer-valued sums, exact below 2^53).
__global__ void k_sigma_push(DevGraph g, const uint32_t *list, uint64_t count,
                             const uint32_t *level, uint32_t L,
                             double *sigma) {
  uint64_t gid = (uint64_t)blockIdx.x * kT + threadIdx.x;
  uint64_t i = gid / kWave;
  if (i >= count) return;
  uint32_t v = list[i];
  uint32_t d = g.deg[v];
  const uint32_t *row = g.pool + g.off[v];
  double sv = sigma[v];
  for (uint32_t k = (uint32_t)(gid % kWave); k < d; k += kWave) {
    uint32_t w = row[k];
    if (level[w] == L) atomicAdd(&sigma[w], sv);
  }
}

// Clears *flag if some stored edge (u, v) has no reverse (v, u).
__global__ void k_symmetry(DevGraph g, const uint64_t *prefix, uint64_t m,
                           uint32_t *flag) {
  uint64_t e = ((uint64_t)blockIdx.x * kT + threadIdx.x) * 8;
  if (e >= m) return;
  uint64_t end = e + 8 < m ? e + 8 : m;
  uint64_t u = find_segment(prefix, g.n, e);
  for (; e < end; e++) {
    while (e >= prefix[u + 1]) u++;
    uint32_t v = g.pool[g.off[u] + (e - prefix[u])];
    bool ok = false;
    if (v < g.n) ok = row_has(g.pool + g.off[v], g.deg[v], (uint32_t)u);
    if (!ok) *flag = 0;
  }
}

__global__ void k_bc_init(uint32_t n, double *sigma, double *delta,
                          uint32_t src) {
  uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v >= n) return;
  sigma[v] = v == src ? 1.0 : 0.0;
  delta[v] = 0.0;
}

}  // namespace

// ====================================================================== BFS
// Direction-optimising BFS (Beamer): top-down push over a compacted queue
// while the frontier is small, bottom-up pull over a frontier bitmap once the
// frontier grows and its edges exceed 1/alpha of the unexplored edges, back
// to top-down when the frontier shrinks below n/beta (alpha and beta are
// per-graph settings, defaults 15 and 18).  Records the BFS order per level
// in `order` (level starts in level_start) when requested (BC uses it).
int graph_is_symmetric(sstg_graph *g, bool *out) {
  if (g->sym_version == g->version) {
    *out = g->symmetric;
    return SSTG_OK;
  }
  hipStream_t s = g->stream;
  uint32_t *flag = (uint32_t *)(g->dev_ctr + 11);
  uint32_t one = 1;
  SSTG_HIP(hipMemcpyAsync(flag, &one, sizeof(uint32_t), hipMemcpyHostToDevice,
                          s));
  if (g->m) {
    const uint64_t *prefix = nullptr;
    SSTG_TRY(graph_degree_prefix(g, &prefix));
    SSTG_LAUNCH(k_symmetry, ceil_div(ceil_div(g->m, 8), kT), kT, 0, s,
                g->view(), prefix, g->m, flag);
  }
  SSTG_HIP(hipMemcpyAsync(g->host_rb + 20, flag, sizeof(uint32_t),
                          hipMemcpyDeviceToHost, s));
  SSTG_HIP(hipStreamSynchronize(s));
  g->symmetric = (uint32_t)g->host_rb[20] != 0;
  g->sym_version = g->version;
  *out = g->symmetric;
  return SSTG_OK;
}

struct BfsRun {
  uint32_t levels = 0;
  uint64_t reached = 0;
  std::vector<uint64_t> level_start;  // into order, levels + 1 entries
};

// The tail of a BFS in one block.  Once the frontier and its edges are small
// the per-level launches and host read-backs cost more than the work, so the
// remaining top-down levels run in a single 1024-thread block: per level it
// scans the frontier's degrees in LDS, walks the frontier's edges flat (one
// edge per thread per step), claims unvisited neighbours as BfsPush does and
// appends them to the queue through an LDS counter.  It stops when the
// frontier empties, grows past the bounds (the host takes over again), or
// after kTailLevels levels.  state: [0] start of the current frontier in the
// queue, [1] its size, [2] its level, [3] levels run, [4] degree sum of every
// frontier found, [5] size of the frontier before the current one, [6]
// degree sum of the current frontier; ends[k] = end of the frontier found by
// the k-th level run.
constexpr uint32_t kTailThreads = 1024;
constexpr uint32_t kTailFrontier = 16384;
constexpr uint32_t kTailEdges = 1u << 17;
constexpr uint32_t kTailLevels = 1024;
constexpr uint32_t kTailPer = kTailFrontier / kTailThreads;
constexpr uint32_t kTailBatch = 4;
// a source of larger degree starts with a grid-wide push instead
constexpr uint32_t kTailStartEdges = 2048;

__global__ void __launch_bounds__(kTailThreads)
    k_bfs_tail(DevGraph g, uint32_t *level, int32_t *parent, uint32_t *queue,
               uint64_t *state, uint64_t *ends) {
  __shared__ uint32_t pre[kTailFrontier + 1];
  __shared__ uint32_t wsum[kTailThreads / kWave];
  __shared__ uint32_t s_next;
  __shared__ unsigned long long s_edges;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  uint64_t head = state[0];
  uint32_t cnt = (uint32_t)state[1];
  uint32_t L = (uint32_t)state[2];
  uint64_t prev = state[5], found_edges = 0, cur_edges = state[6];
  uint32_t k = 0;
  while (cnt > 0 && k < kTailLevels) {  // cnt and k are block-uniform
    // exclusive prefix of the frontier's degrees (kTailPer per thread)
    uint32_t dl[kTailPer];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < kTailPer; j++) {
      const uint32_t idx = tid * kTailPer + j;
      dl[j] = idx < cnt ? g.deg[queue[head + idx]] : 0u;
      sum += dl[j];
    }
    const uint32_t incl = wave_inclusive_sum(sum);
    if (lane == kWave - 1) wsum[wave] = incl;
    if (tid == 0) {
      s_next = 0;
      s_edges = 0;
    }
    __syncthreads();
    uint32_t run = incl - sum, total = 0;
    for (uint32_t w = 0; w < kTailThreads / kWave; w++) {
      const uint32_t x = wsum[w];
      if (w < wave) run += x;
      total += x;
    }
#pragma unroll
    for (uint32_t j = 0; j < kTailPer; j++) {
      const uint32_t idx = tid * kTailPer + j;
      if (idx < cnt) pre[idx] = run;
      run += dl[j];
    }
    if (tid == 0) pre[cnt] = total;
    __syncthreads();
    // the frontier's edges, flat: kTailBatch consecutive edges per thread
    // per step, each load of the batch issued before any is used
    const uint64_t tail = head + cnt;
    for (uint32_t e0 = tid * kTailBatch; e0 < total;
         e0 += kTailThreads * kTailBatch) {
      uint32_t u[kTailBatch], idx[kTailBatch], w[kTailBatch], lv[kTailBatch];
      uint64_t rb[kTailBatch];
      uint32_t lo = 0, hi = cnt;  // largest j with pre[j] <= e0
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= e0)
          lo = mid;
        else
          hi = mid;
      }
#pragma unroll
      for (uint32_t j = 0; j < kTailBatch; j++) {
        const uint32_t e = e0 + j;
        if (e < total) {
          while (pre[lo + 1] <= e) lo++;  // pre[cnt] = total > e
          u[j] = queue[head + lo];
          idx[j] = e - pre[lo];
        }
      }
#pragma unroll
      for (uint32_t j = 0; j < kTailBatch; j++)
        rb[j] = e0 + j < total ? g.off[u[j]] : 0;
#pragma unroll
      for (uint32_t j = 0; j < kTailBatch; j++)
        w[j] = e0 + j < total ? g.pool[rb[j] + idx[j]] : 0u;
#pragma unroll
      for (uint32_t j = 0; j < kTailBatch; j++)
        lv[j] = e0 + j < total ? level[w[j]] : 0u;
#pragma unroll
      for (uint32_t j = 0; j < kTailBatch; j++) {
        if (e0 + j < total && lv[j] == kInf &&
            atomicCAS(&level[w[j]], kInf, L + 1) == kInf) {
          if (parent) parent[w[j]] = (int32_t)u[j];
          const uint32_t pos = atomicAdd(&s_next, 1u);
          queue[tail + pos] = w[j];
          atomicAdd(&s_edges, (unsigned long long)g.deg[w[j]]);
        }
      }
    }
    __syncthreads();
    const uint32_t nxt = s_next;
    const uint64_t ned = s_edges;
    if (tid == 0) ends[k] = tail + nxt;
    prev = cnt;
    head = tail;
    cnt = nxt;
    cur_edges = ned;
    found_edges += ned;
    L++;
    k++;
    __syncthreads();  // s_next and s_edges are reset by the next level
    if (cnt > kTailFrontier || ned > kTailEdges) break;
  }
  if (tid == 0) {
    state[0] = head;
    state[1] = cnt;
    state[2] = L;
    state[3] = k;
    state[4] = found_edges;
    state[5] = prev;
    state[6] = cur_edges;
  }
}

// Every level's vertices are appended to the queue (order, or a workspace
// buffer) whatever the direction, so level L's vertices are
// queue[level_start[L] .. level_start[L + 1]).  Per level the host reads one
// small counter block and decides the next step; counters and bitmaps are
// cleared by the kernels themselves.
static int bfs_core(sstg_graph *g, uint32_t src, uint32_t *level,
                    int32_t *parent, uint32_t *order, bool top_down_only,
                    BfsRun *run, ReadLog *log = nullptr) {
  hipStream_t s = g->stream;
  Workspace &ws = g->ws;
  const uint32_t n = g->n_rows;
  const uint64_t *prefix = nullptr;
  SSTG_TRY(graph_degree_prefix(g, &prefix));
  EdgeMapCtx ctx;
  SSTG_TRY(edge_map_ctx(g, &ctx, kWsAlgo));
  ctx.log = log;
  uint32_t *queue = order;
  if (!queue) SSTG_TRY(ws.get(kWsAlgo + 10, (uint64_t)n + 1, &queue));
  unsigned long long *info = ctx.ctrs + 2 * em::kStepCtrs;
  SSTG_LAUNCH(k_bfs_start, blocks(std::max<uint64_t>(n, ctx.words)), kT, 0, s,
              n, level, parent, src, queue, ctx.vis, ctx.words,
              (const uint32_t *)g->deg, prefix, info);
  // the source's degree, and the degree mass of the upper half of the ids
  // against the whole (rows are sorted by id, so when the high ids carry
  // more edges -- as in the reference's rMat, whose hubs have mostly 1
  // bits -- a frontier vertex is met sooner from a row's end)
  SSTG_TRY(read_back(g, log, g->host_rb + 8, info, 3 * sizeof(uint64_t)));
  const uint32_t src_deg = (uint32_t)g->host_rb[8];
  ctx.rev = 2 * g->host_rb[9] < g->host_rb[10];
  uint64_t qs = 0, nf = 1, mf = src_deg;
  uint64_t explored = src_deg;  // edges of visited vertices
  bool bottom_up = false;
  bool vis_stale = false;  // bottom-up steps do not maintain the filter
  const double alpha = g->bfs_alpha, beta = g->bfs_beta;
  uint64_t prev_nf = 0;
  if (run) {  // level 0 is queue[0, 1); each step appends the end of the next
    run->level_start.clear();
    run->level_start.push_back(0);
    run->level_start.push_back(1);
  }
  // a bottom-up step pulls over out-neighbours, which equals pulling over
  // in-neighbours only when every edge has its reverse
  bool symmetric = false;
  if (!top_down_only) SSTG_TRY(graph_is_symmetric(g, &symmetric));
  top_down_only = top_down_only || !symmetric;
  uint32_t L = 0;
  const bool verbose = std::getenv("SSTG_VERBOSE") != nullptr;
  auto t_level = std::chrono::steady_clock::now();
  while (nf > 0) {
    uint64_t unexplored = g->m > explored ? g->m - explored : 0;
    // the unexplored-edge estimate runs out in the tail, so a shrinking
    // frontier never switches to bottom-up
    bool want_bu = !top_down_only &&
                   (bottom_up ? (double)nf >= (double)n / beta
                              : nf > prev_nf &&
                                    (double)mf > (double)unexplored / alpha);
    if (!want_bu && mf == ~0ull) {
      // a bottom-up step left the frontier's degree sum open: its prefix,
      // which the push below reuses
      SSTG_TRY(frontier_prefix(g, ctx, queue + qs, nf, &mf));
      explored += mf;
    }
    if (!want_bu && nf <= kTailFrontier && mf <= kTailEdges &&
        (L > 0 || mf <= kTailStartEdges)) {
      // the remaining top-down levels in one block (k_bfs_tail)
      uint64_t *st = nullptr, *ends = nullptr;
      SSTG_TRY(ws.get(kWsAlgo + 8, 8, &st));
      SSTG_TRY(ws.get(kWsAlgo + 9, kTailLevels, &ends));
      uint64_t *h = g->host_rb + 48;  // pinned staging
      const TailState t0{{qs, nf, L, 0, 0, prev_nf, mf}};
      SSTG_LAUNCH(k_tail_state, 1, kWave, 0, s, st, t0);
      SSTG_LAUNCH(k_bfs_tail, 1, kTailThreads, 0, s, g->view(), level, parent,
                  queue, st, ends);
      SSTG_TRY(read_back(g, log, h, st, 7 * sizeof(uint64_t)));
      const uint64_t levels_run = h[3];
      if (run && levels_run) {
        std::vector<uint64_t> e(levels_run);
        SSTG_TRY(read_back(g, log, e.data(), ends,
                           levels_run * sizeof(uint64_t)));
        for (uint64_t x : e) run->level_start.push_back(x);
      }
      if (verbose) {
        const auto t1 = std::chrono::steady_clock::now();
        std::fprintf(stderr,
                     "bfs levels %u..%u in one block, frontier %llu  %.3f ms\n", L,
                     (uint32_t)(L + levels_run - 1), (unsigned long long)h[1],
                     std::chrono::duration<double, std::milli>(t1 - t_level)
                         .count());
        t_level = t1;
      }
      qs = h[0];
      nf = h[1];
      L = (uint32_t)h[2];
      explored += h[4];
      prev_nf = h[5];
      mf = h[6];
      bottom_up = false;
      continue;
    }
    EdgeMapResult res;
    uint32_t *next = queue + qs + nf;
    if (want_bu) {
      SSTG_TRY(edge_map_dense(g, ctx, BfsPull{level, parent, L + 1},
                              LevelIs{level, L}, next, &res));
      bottom_up = true;
      vis_stale = true;
    } else {
      if (vis_stale) {  // the filter again from the levels
        SSTG_TRY(edge_map_bits(g, ctx, Reached{level}, ctx.vis, nullptr));
        vis_stale = false;
      }
      SSTG_TRY(edge_map_sparse(g, ctx, queue + qs, nf, mf,
                               BfsPush{level, parent, L + 1}, next, &res));
      bottom_up = false;
    }
    if (verbose) {
      auto t1 = std::chrono::steady_clock::now();
      std::fprintf(stderr,
                   "bfs level %u %s nf=%llu mf=%lld -> next %llu  %.3f ms\n", L,
                   want_bu ? "pull" : "push", (unsigned long long)nf,
                   (long long)mf, (unsigned long long)res.count,
                   std::chrono::duration<double, std::milli>(t1 - t_level)
                       .count());
      t_level = t1;
    }
    qs += nf;
    prev_nf = nf;
    nf = res.count;
    mf = res.degree_sum;  // ~0 after a bottom-up step: not computed
    if (mf != ~0ull) explored += mf;
    L++;
    if (run) run->level_start.push_back(qs + nf);
  }
  if (run) {
    run->levels = L;
    run->reached = qs;
  }
  return SSTG_OK;
}

// A recorded launch sequence of a graph: hipGraph, its executable and the
// key (a RecKey) it was recorded for.
struct RecSlot {
  void **graph, **exec, **key;
};

static void rec_release(RecSlot r) {
  if (*r.exec) hipGraphExecDestroy((hipGraphExec_t)*r.exec);
  if (*r.graph) hipGraphDestroy((hipGraph_t)*r.graph);
  *r.exec = *r.graph = nullptr;
  delete static_cast<RecKey *>(*r.key);
  *r.key = nullptr;
}

void bfs_release(sstg_graph *g) {
  rec_release(RecSlot{&g->bfs_graph, &g->bfs_exec, &g->bfs_key});
  rec_release(RecSlot{&g->bc_graph, &g->bc_exec, &g->bc_key});
}

static RecKey rec_key_of(const sstg_graph *g, uint32_t src, const void *out0,
                         const void *out1) {
  RecKey k{};
  k.version = g->version;
  k.tune_epoch = g->tune_epoch;
  k.ws_generation = g->ws.generation;
  k.src = src;
  k.n = g->n_rows;
  const void *p[5] = {out0, out1, g->off, g->deg, g->pool};
  for (int i = 0; i < 5; i++) k.ptr[i] = p[i];
  k.alpha = g->bfs_alpha;
  k.beta = g->bfs_beta;
  return k;
}

// A traversal from a given source on a given graph version takes the same
// steps every time: every level's frontier is the same set, so the counts
// the host reads back, and with them every launch's shape, repeat (only the
// order of the vertices inside a top-down level's queue differs).  The
// first call runs plainly (run(log), no final wait), logging each read-back;
// the same run is then captured with the logged values standing in for the
// reads (no waits), and later calls with the same key replay the recording
// in one launch.
template <class Run>
static int run_recorded(sstg_graph *g, RecSlot slot, uint32_t src,
                        const void *out0, const void *out1, Run run) {
  hipStream_t s = g->stream;
  RecKey key = rec_key_of(g, src, out0, out1);
  const RecKey *have = static_cast<const RecKey *>(*slot.key);
  if (*slot.exec && have && std::memcmp(have, &key, sizeof(key)) == 0) {
    SSTG_HIP(hipGraphLaunch((hipGraphExec_t)*slot.exec, s));
    SSTG_HIP(hipStreamSynchronize(s));
    return SSTG_OK;
  }
  rec_release(slot);
  ReadLog log;
  log.mode = 1;
  SSTG_TRY(run(&log));
  SSTG_HIP(hipStreamSynchronize(s));
  // after the run: it may have allocated workspace or built cached data
  key = rec_key_of(g, src, out0, out1);
  log.mode = 2;
  log.pos = 0;
  Profiler &prof = profiler();
  const bool prof_on = prof.on;
  prof.on = false;  // no event records inside the capture
  SSTG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int rc = run(&log);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &graph);
  prof.on = prof_on;
  // keep the recording only if it consumed exactly the logged reads and
  // allocated nothing (the results are those of the plain run either way)
  if (rc != SSTG_OK || e != hipSuccess || log.pos != log.words.size() ||
      g->ws.generation != key.ws_generation) {
    if (graph) hipGraphDestroy(graph);
    if (rc != SSTG_OK) return rc;
    SSTG_HIP(e);
    return SSTG_OK;
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    hipGraphDestroy(graph);
    SSTG_HIP(ei);
  }
  *slot.graph = graph;
  *slot.exec = exec;
  *slot.key = new RecKey(key);
  return SSTG_OK;
}

int algo_bfs(sstg_graph *g, uint32_t src, int32_t *parent, uint32_t *level) {
  uint32_t *lv = level;
  if (!lv) SSTG_TRY(g->ws.get(kWsAlgo + 13, g->n_rows, &lv));
  if (iteration_capture_enabled(g))
    return run_recorded(
        g, RecSlot{&g->bfs_graph, &g->bfs_exec, &g->bfs_key}, src, lv, parent,
        [&](ReadLog *log) {
          return bfs_core(g, src, lv, parent, nullptr, false, nullptr, log);
        });
  SSTG_TRY(bfs_core(g, src, lv, parent, nullptr, false, nullptr));
  SSTG_HIP(hipStreamSynchronize(g->stream));
  return SSTG_OK;
}

int algo_bf(sstg_graph *g, uint32_t src, int32_t *out) {
  uint32_t *lv = nullptr;
  SSTG_TRY(g->ws.get(kWsAlgo + 13, g->n_rows, &lv));
  SSTG_TRY(bfs_core(g, src, lv, nullptr, nullptr, false, nullptr));
  SSTG_LAUNCH(k_level_dist, blocks(g->n_rows), kT, 0, g->stream,
              (const uint32_t *)lv, g->n_rows, out);
  SSTG_HIP(hipStreamSynchronize(g->stream));
  return SSTG_OK;
}

// ================================================================ PageRank
template <class Real, bool kKahan>
static int pagerank_impl(sstg_graph *g, uint32_t iters, double damping,
                         Real *out) {
  hipStream_t s = g->stream;
  const uint32_t n = g->n_rows;
  if (n == 0) return SSTG_OK;
  DegreeBins bins;
  SSTG_TRY(degree_bins(g, &bins));
  Real *contrib = nullptr, *contrib2 = nullptr;
  SSTG_TRY(g->ws.get(kWsAlgo + 20, n, &contrib));
  SSTG_TRY(g->ws.get(kWsAlgo + 21, n, &contrib2));
  const Real x0 = (Real)(1.0 / (double)n);
  SSTG_LAUNCH(k_pr_init<Real>, blocks(n), kT, 0, s, n, (const uint32_t *)g->deg,
              x0, out, contrib);
  const Real base = (Real)((1.0 - damping) / (double)n);
  const Real damp = (Real)damping;
  // pairs of iterations replayed from one recording (contributions back in
  // their first buffer after two); an odd last iteration launched plainly
  uint32_t plain = iters;
  if (iters >= 2 && iteration_capture_enabled(g)) {
    IterPairKey key{};
    key.path = kKahan ? 2 : 1;
    key.version = g->version;
    key.setting = g->tune_epoch;
    key.n = n;
    const void *ptrs[10] = {contrib, contrib2, out, g->deg, g->off, g->pool,
                            bins.list[0], bins.list[3], nullptr, nullptr};
    for (int i = 0; i < 10; i++) key.ptr[i] = ptrs[i];
    key.damping = damping;
    SSTG_TRY(replay_iteration_pairs(g, key, iters / 2, [&]() {
      SSTG_TRY(vertex_pull_all(
          g, bins, PrPull<Real, kKahan>{contrib, contrib2, out, g->deg, damp, base}));
      return vertex_pull_all(
          g, bins, PrPull<Real, kKahan>{contrib2, contrib, out, g->deg, damp, base});
    }));
    plain = iters & 1u;
  }
  for (uint32_t it = 0; it < plain; it++) {
    PrPull<Real, kKahan> f{contrib, contrib2, out, g->deg, damp, base};
    SSTG_TRY(vertex_pull_all(g, bins, f));
    Real *t = contrib;
    contrib = contrib2;
    contrib2 = t;
  }
  SSTG_HIP(hipStreamSynchronize(s));
  return SSTG_OK;
}

int algo_pagerank(sstg_graph *g, uint32_t iters, double damping, double *out,
                  bool allow_segments) {
  if (allow_segments && pagerank_segmented_applies(g))
    return pagerank_segmented(g, iters, damping, out);
  return pagerank_impl<double, false>(g, iters, damping, out);
}

int algo_pagerank_f32k(sstg_graph *g, uint32_t iters, double damping,
                       float *out) {
  return pagerank_impl<float, true>(g, iters, damping, out);
}

// ====================================================================== CC
int algo_cc(sstg_graph *g, uint32_t *out) {
  hipStream_t s = g->stream;
  const uint32_t n = g->n_rows;
  if (n == 0) return SSTG_OK;
  const uint64_t *prefix = nullptr;
  SSTG_TRY(graph_degree_prefix(g, &prefix));
  uint32_t *p = nullptr, *giant = nullptr;
  SSTG_TRY(g->ws.get(kWsAlgo + 22, n, &p));
  SSTG_TRY(g->ws.get(kWsAlgo + 30, 1, &giant));
  bool symmetric = false;
  SSTG_TRY(graph_is_symmetric(g, &symmetric));
  const DevGraph v = g->view();
  if (g->m && symmetric) {
    SSTG_LAUNCH(cc::k_identity, blocks(n), kT, 0, s, n, p);
    SSTG_LAUNCH(cc::k_link_sample, blocks(n), kT, 0, s, v, p);
    SSTG_LAUNCH(cc::k_compress, blocks(n), kT, 0, s, n, p);
    SSTG_LAUNCH(cc::k_mode, 1, kT, 0, s, (const uint32_t *)nullptr, 1u, n, p,
                giant);
    SSTG_LAUNCH(cc::k_link_rest, blocks(n), kT, 0, s, v, p,
                (const uint32_t *)giant, (const uint32_t *)nullptr);
  } else {
    SSTG_LAUNCH(cc::k_init, blocks(n), kT, 0, s, v, p);
    if (g->m)
      SSTG_LAUNCH(cc::k_hook_all, ceil_div(ceil_div(g->m, 8), kT), kT, 0, s, v,
                prefix, g->m, p);
  }
  SSTG_LAUNCH(k_cc_final, blocks(n), kT, 0, s, n, p, out);
  SSTG_HIP(hipStreamSynchronize(s));
  return SSTG_OK;
}

// ====================================================================== TC
int algo_tc(sstg_graph *g, uint64_t *out) {
  hipStream_t s = g->stream;
  const uint32_t n = g->n_rows;
  *out = 0;
  if (n == 0 || g->m == 0) return SSTG_OK;
  bool symmetric = false;
  SSTG_TRY(graph_is_symmetric(g, &symmetric));
  if (symmetric) {
    int rc = tc_oriented(g, out);
    if (rc != SSTG_E_UNSUPPORTED) return rc;
  }
  // rows as stored: L(x) = ids of row x below x, a prefix of the row
  uint32_t *lower = nullptr;
  SSTG_TRY(g->ws.get(kWsAlgo + 23, n, &lower));
  SSTG_LAUNCH(k_lower_degree, blocks(n), kT, 0, s, g->view(), lower);
  if (!symmetric) {
    int rc = count_list_triangles(g, g->off, g->pool, lower, out);
    if (rc != SSTG_E_UNSUPPORTED) return rc;
  }
  // lists of 2^23 ids or more: one thread per stored edge
  const uint64_t *prefix = nullptr;
  SSTG_TRY(graph_degree_prefix(g, &prefix));
  unsigned long long *acc = (unsigned long long *)(g->dev_ctr + 8);
  SSTG_HIP(hipMemsetAsync(acc, 0, sizeof(uint64_t), s));
  SSTG_LAUNCH(k_tc_edges, ceil_div(ceil_div(g->m, 4), kT), kT, 0, s, g->view(),
              prefix, g->m, (const uint32_t *)lower, acc);
  SSTG_HIP(hipMemcpyAsync(g->host_rb, acc, sizeof(uint64_t),
                          hipMemcpyDeviceToHost, s));
  SSTG_HIP(hipStreamSynchronize(s));
  *out = g->host_rb[0];
  return SSTG_OK;
}

int algo_common_neighbors(sstg_graph *g, uint32_t u, uint32_t v,
                          int early_end, uint64_t *out) {
  hipStream_t s = g->stream;
  unsigned long long *acc = (unsigned long long *)(g->dev_ctr + 8);
  SSTG_LAUNCH(k_common, 1, 64, 0, s, g->view(), u, v, early_end, acc);
  SSTG_HIP(hipMemcpyAsync(g->host_rb, acc, sizeof(uint64_t),
                          hipMemcpyDeviceToHost, s));
  SSTG_HIP(hipStreamSynchronize(s));
  *out = g->host_rb[0];
  return SSTG_OK;
}

// ====================================================================== BC
// Brandes from src on a graph of fewer than 64 BFS levels: the reached
// vertices grouped by (level, lane class) with one radix pass, so each
// level's pulls run over exactly its vertices (1, 8, 16 lanes or a block by
// the reductions' degree bins) instead of whole degree bins filtered by
// level.
static int bc_leveled(sstg_graph *g, uint32_t src, const uint32_t *level,
                      const uint32_t *order, const BfsRun &run, double *sigma,
                      double *out, ReadLog *log) {
  hipStream_t s = g->stream;
  const uint32_t n = g->n_rows;
  const uint64_t reached = run.reached;
  const uint32_t nb = run.levels * 4;
  uint64_t *key = nullptr, *tmp = nullptr, *start = nullptr;
  uint32_t *bm = nullptr;
  const uint64_t words = ((uint64_t)n + 31) / 32 + 1;
  SSTG_TRY(g->ws.get(kWsAlgo + 16, reached + 1, &key));
  SSTG_TRY(g->ws.get(kWsAlgo + 17, reached + 1, &tmp));
  SSTG_TRY(g->ws.get(kWsAlgo + 18, (uint64_t)nb + 1, &start));
  SSTG_TRY(g->ws.get(kWsAlgo + 19, run.levels * words, &bm));
  const uint32_t t0 = g->bin_hi[kBinsPull][0], t1 = g->bin_hi[kBinsPull][1],
                 t2 = g->bin_hi[kBinsPull][2];
  SSTG_LAUNCH(k_level_bitmaps,
              ceil_div(words * 32, (uint64_t)kT * kLevelRounds), kT, 0, s,
              level, n, run.levels, bm, words);
  SSTG_LAUNCH(k_bc_keys, blocks(reached), kT, 0, s, order, reached, level,
              (const uint32_t *)g->deg, t0, t1, t2, key);
  uint64_t *sorted = key;
  const int shifts[1] = {32};
  SSTG_TRY(radix_sort_u64(key, tmp, reached, shifts, 1, g->ws, kWsPrim, s,
                          &sorted));
  SSTG_LAUNCH(k_bc_bucket_starts, ceil_div(nb + 1, kT), kT, 0, s,
              (const uint64_t *)sorted, reached, nb, start);
  // the vertex ids of the sorted keys, as u32 over the other buffer
  uint32_t *list = reinterpret_cast<uint32_t *>(sorted == key ? tmp : key);
  SSTG_TRY(narrow_u64_to_u32(sorted, reached, list, s));
  std::vector<uint64_t> st(nb + 1);
  SSTG_TRY(read_back(g, log, st.data(), start, (nb + 1) * sizeof(uint64_t)));
  const DevGraph v = g->view();
  auto pull = [&](uint32_t L, auto f) -> int {
    using Fn = decltype(f);
    const uint64_t *b = st.data() + L * 4;  // b[0..4]: the level's classes
    const uint32_t *lst = list;
    if (b[1] > b[0])
      SSTG_LAUNCH((em::k_vertex_pull<Fn, 1>), ceil_div(b[1] - b[0], kT), kT, 0,
                  s, v, lst + b[0], b[1] - b[0], f);
    if (b[2] > b[1])
      SSTG_LAUNCH((em::k_vertex_pull<Fn, 8>), ceil_div((b[2] - b[1]) * 8, kT),
                  kT, 0, s, v, lst + b[1], b[2] - b[1], f);
    if (b[3] > b[2])
      SSTG_LAUNCH((em::k_vertex_pull<Fn, 16>),
                  ceil_div((b[3] - b[2]) * 16, kT), kT, 0, s, v, lst + b[2],
                  b[3] - b[2], f);
    if (b[4] > b[3])
      SSTG_LAUNCH((em::k_vertex_pull<Fn, 256>), b[4] - b[3], kT, 0, s, v,
                  lst + b[3], b[4] - b[3], f);
    return SSTG_OK;
  };
  SSTG_LAUNCH(k_bc_init, blocks(n), kT, 0, s, n, sigma, out, src);
  // forward: sigma from the parents, level by level
  for (uint32_t L = 1; L < run.levels; L++)
    SSTG_TRY(pull(L, SigmaPullBm{bm + (L - 1) * words, sigma}));
  // backward: dependencies from the children, the deepest level (no
  // children, delta 0) only turned into quotients
  const uint32_t last = run.levels - 1;
  const uint64_t lb = st[last * 4], le = st[last * 4 + 4];
  if (le > lb)
    SSTG_LAUNCH(k_bc_last_quotients, ceil_div(le - lb, kT), kT, 0, s,
                (const uint32_t *)(list + lb), le - lb, sigma);
  for (int64_t L = (int64_t)run.levels - 2; L >= 0; L--)
    SSTG_TRY(pull((uint32_t)L, DeltaPullBm{bm + (L + 1) * words, sigma, out}));
  return SSTG_OK;
}

int algo_bc(sstg_graph *g, uint32_t src, double *out) {
  hipStream_t s = g->stream;
  const uint32_t n = g->n_rows;
  uint32_t *level = nullptr, *order = nullptr;
  double *sigma = nullptr;
  SSTG_TRY(g->ws.get(kWsAlgo + 13, n, &level));
  SSTG_TRY(g->ws.get(kWsAlgo + 14, (uint64_t)n + 1, &order));
  SSTG_TRY(g->ws.get(kWsAlgo + 15, n, &sigma));
  bool symmetric = false;
  SSTG_TRY(graph_is_symmetric(g, &symmetric));
  // symmetric graphs of fewer than 64 levels (every benchmark graph): the
  // leveled Brandes, recorded and replayed like BFS when asked for
  auto leveled = [&](ReadLog *log, bool *done) -> int {
    BfsRun run;
    SSTG_TRY(bfs_core(g, src, level, nullptr, order, false, &run, log));
    *done = run.levels < 64;
    if (!*done) return SSTG_OK;
    return bc_leveled(g, src, level, order, run, sigma, out, log);
  };
  if (symmetric && iteration_capture_enabled(g)) {
    // the level count is fixed for a key, so a recording is always leveled;
    // a graph of 64 levels or more records nothing and runs plainly below
    bool deep = false;
    const int rc = run_recorded(
        g, RecSlot{&g->bc_graph, &g->bc_exec, &g->bc_key}, src, out, nullptr,
        [&](ReadLog *log) {
          bool done = false;
          SSTG_TRY(leveled(log, &done));
          if (done) return (int)SSTG_OK;
          deep = true;
          return (int)SSTG_E_UNSUPPORTED;
        });
    if (!deep) return rc;
    SSTG_HIP(hipStreamSynchronize(s));
  }
  BfsRun run;
  SSTG_TRY(bfs_core(g, src, level, nullptr, order, false, &run));
  if (symmetric && run.levels < 64) {
    SSTG_TRY(bc_leveled(g, src, level, order, run, sigma, out, nullptr));
    SSTG_HIP(hipStreamSynchronize(s));
    return SSTG_OK;
  }
  SSTG_LAUNCH(k_bc_init, blocks(n), kT, 0, s, n, sigma, out, src);
  // forward: sigma level by level (level 0 is the source, sigma = 1); pulled
  // from the parents on symmetric graphs, pushed along out-edges otherwise
  for (uint32_t L = 1; L < run.levels; L++) {
    if (symmetric) {
      uint64_t b = run.level_start[L], e = run.level_start[L + 1];
      if (e <= b) continue;
      SigmaPull f{level, sigma, L};
      SSTG_TRY(vertex_pull_level(g, order + b, e - b, level, L, f));
    } else {
      uint64_t b = run.level_start[L - 1], e = run.level_start[L];
      if (e <= b) continue;
      SSTG_LAUNCH(k_sigma_push, ceil_div((e - b) * kWave, kT), kT, 0, s,
                  g->view(), (const uint32_t *)(order + b), e - b,
                  (const uint32_t *)level, L, sigma);
    }
  }
  // backward: dependencies from the deepest level up to the source
  for (int64_t L = (int64_t)run.levels - 2; L >= 0; L--) {
    uint64_t b = run.level_start[L], e = run.level_start[L + 1];
    if (e <= b) continue;
    DeltaPull f{level, sigma, out, (uint32_t)L};
    SSTG_TRY(vertex_pull_level(g, order + b, e - b, level, (uint32_t)L, f));
  }
  SSTG_HIP(hipStreamSynchronize(s));
  return SSTG_OK;
}

}  // namespace sstg
