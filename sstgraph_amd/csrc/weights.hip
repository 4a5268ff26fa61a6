// Edge weights and weighted Bellman-Ford.
//
// The reference's valued tinysets (SparseMatrixV<true, uint32_t>, values in
// the PMA next to each key, PMA.hpp:104-116) update a present key's value in
// place: the last write wins (PMA.hpp:1095-1146).  Here the binary store and
// its update kernels stay as they are; the weights of a graph live in a
// device hash table beside it, keyed by (row << 32 | col), created by the
// first weighted insert:
//  * a weighted batch inserts its edges into the store (set semantics,
//    self-loops skipped, as insert_batch) and upserts their weights; each
//    slot's value is (tag << 32 | weight) updated with atomicMax, and tags
//    rise with the position in the batch and from batch to batch, so the
//    last write of a key wins whatever the thread order;
//  * removing edges erases their weights (tombstones, dropped on rehash);
//  * an edge that has no weight entry (inserted by an unweighted batch) has
//    weight 1, as the reference reads an unweighted file into a weighted
//    graph (helpers.hpp:557-566).
// Weighted algorithms read a weight array aligned with the store's flat edge
// order, filled from the table once per graph version.
//
// Bellman-Ford (called at test.hpp:618 and test.hpp:811, weights read as
// int32): dist = INT32_MAX / 2 where unreached, relaxations in 64-bit, rounds
// over the vertices whose distance fell in the previous round, at most n
// rounds.  With no negative cycle the result is the exact shortest distance.
#include <algorithm>

#include "algo.hpp"
#include "edge_map.hpp"
#include "graph.hpp"
#include "primitives.hpp"

namespace sstg {

namespace {

constexpr int kT = 256;
constexpr uint64_t kWtEmpty = ~0ull;
constexpr uint64_t kWtTomb = ~0ull - 1;  // keys are below: row < 2^32 - 1
constexpr int32_t kDistInf = 0x7FFFFFFF / 2;

__device__ __forceinline__ uint64_t wt_hash(uint64_t u) {
  uint64_t v = u * 3935559000370003845ull + 2691343689449507681ull;
  v ^= v >> 21;
  v ^= v << 37;
  v ^= v >> 4;
  v *= 4768777513237032717ull;
  v ^= v << 20;
  v ^= v >> 41;
  v ^= v << 5;
  return v;
}

struct WtTable {
  unsigned long long *keys;
  unsigned long long *vals;
  uint64_t mask;
};

__device__ __forceinline__ int64_t wt_find(const WtTable &t, uint64_t key) {
  uint64_t h = wt_hash(key) & t.mask;
  for (;;) {
    const uint64_t k = t.keys[h];
    if (k == key) return (int64_t)h;
    if (k == kWtEmpty) return -1;
    h = (h + 1) & t.mask;
  }
}

__global__ void k_wt_fill(unsigned long long *p, uint64_t n,
                          unsigned long long v) {
  const uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) p[i] = v;
}

// per-block sums of a 0/1 per thread into part[block]
__device__ __forceinline__ void block_partial_count(uint32_t c,
                                                    uint64_t *part) {
  __shared__ uint64_t ws[kT / kWave];
  const uint64_t s = wave_sum((uint64_t)c);
  if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < kT / kWave; w++) t += ws[w];
    part[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kT)
    k_wt_upsert(WtTable t, const uint32_t *src, const uint32_t *dst,
                const uint32_t *w, uint64_t n, uint32_t tag0, int keep_self,
                uint64_t *part) {
  const uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  uint32_t added = 0;
  if (i < n && (keep_self || src[i] != dst[i])) {
    const uint64_t key = ((uint64_t)src[i] << 32) | dst[i];
    uint64_t h = wt_hash(key) & t.mask;
    for (;;) {
      const uint64_t k = t.keys[h];
      if (k == key) break;
      if (k == kWtEmpty) {
        const uint64_t old = atomicCAS(&t.keys[h], kWtEmpty, key);
        if (old == kWtEmpty) {
          added = 1;
          break;
        }
        if (old == key) break;
      }
      h = (h + 1) & t.mask;
    }
    atomicMax(&t.vals[h],
              ((unsigned long long)(tag0 + (uint32_t)i) << 32) | w[i]);
  }
  block_partial_count(added, part);
}

__global__ void __launch_bounds__(kT)
    k_wt_erase(WtTable t, const uint32_t *src, const uint32_t *dst, uint64_t n,
               uint64_t *part) {
  const uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  uint32_t removed = 0;
  if (i < n) {
    const uint64_t key = ((uint64_t)src[i] << 32) | dst[i];
    const int64_t h = wt_find(t, key);
    if (h >= 0 && atomicCAS(&t.keys[h], key, kWtTomb) == key) removed = 1;
  }
  block_partial_count(removed, part);
}

// live entries of the old table into the new one (all keys distinct)
__global__ void k_wt_rehash(WtTable o, uint64_t ocap, WtTable t) {
  const uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= ocap) return;
  const uint64_t key = o.keys[i];
  if (key == kWtEmpty || key == kWtTomb) return;
  uint64_t h = wt_hash(key) & t.mask;
  while (atomicCAS(&t.keys[h], kWtEmpty, key) != kWtEmpty)
    h = (h + 1) & t.mask;
  t.vals[h] = o.vals[i];
}

// drop the tags (keep the weights) before the tag counter wraps
__global__ void k_wt_untag(unsigned long long *vals, uint64_t cap) {
  const uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i < cap) vals[i] &= 0xFFFFFFFFull;
}

__global__ void k_sum_parts(const uint64_t *part, uint64_t nb, uint64_t *out) {
  __shared__ uint64_t ws[kT / kWave];
  uint64_t t = 0;
  for (uint64_t i = threadIdx.x; i < nb; i += kT) t += part[i];
  t = wave_sum(t);
  if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t a = 0;
    for (int w = 0; w < kT / kWave; w++) a += ws[w];
    *out = a;
  }
}

// weight of every stored edge in flat order (8 per thread)
__global__ void k_wt_flat(DevGraph g, const uint64_t *prefix, uint64_t m,
                          WtTable t, int has_table, int32_t *out) {
  uint64_t e = ((uint64_t)blockIdx.x * kT + threadIdx.x) * 8;
  if (e >= m) return;
  const uint64_t end = e + 8 < m ? e + 8 : m;
  uint64_t u = find_segment(prefix, g.n, e);
  for (; e < end; e++) {
    while (e >= prefix[u + 1]) u++;
    const uint32_t c = g.pool[g.off[u] + (e - prefix[u])];
    int32_t w = 1;
    if (has_table) {
      const int64_t h = wt_find(t, (u << 32) | c);
      if (h >= 0) w = (int32_t)(uint32_t)t.vals[h];
    }
    out[e] = w;
  }
}

// value(row, col): 0 for an absent edge (PMA.hpp:1397-1420 returns {}), the
// stored weight, or 1 for an edge that has none
__global__ void k_wt_get(DevGraph g, WtTable t, int has_table,
                         const uint32_t *rows, const uint32_t *cols,
                         uint64_t n, uint32_t *out) {
  const uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = rows[i], c = cols[i];
  uint32_t w = 0;
  if (r < g.n) {
    if (row_has(g.pool + g.off[r], g.deg[r], c)) {
      w = 1;
      if (has_table) {
        const int64_t h = wt_find(t, ((uint64_t)r << 32) | c);
        if (h >= 0) w = (uint32_t)t.vals[h];
      }
    }
  }
  out[i] = w;
}

// ---- Bellman-Ford
__global__ void k_bf_init(uint32_t n, uint32_t src, int32_t *dist,
                          uint8_t *flag) {
  const uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v >= n) return;
  dist[v] = v == src ? 0 : kDistInf;
  flag[v] = 0;
}

// relax every edge of the frontier (flat over its edges, 8 per thread); a
// vertex whose distance fell is flagged for the next round
__global__ void k_bf_relax(DevGraph g, const uint64_t *deg_prefix,
                           const int32_t *wflat, const uint32_t *front,
                           const uint64_t *fpre, uint64_t nf, uint64_t mf,
                           int32_t *dist, uint8_t *flag) {
  uint64_t e = ((uint64_t)blockIdx.x * kT + threadIdx.x) * 8;
  if (e >= mf) return;
  const uint64_t end = e + 8 < mf ? e + 8 : mf;
  uint64_t f = find_segment(fpre, nf, e);
  for (; e < end; e++) {
    while (e >= fpre[f + 1]) f++;
    const uint32_t u = front[f];
    const uint64_t k = e - fpre[f];
    const uint32_t v = g.pool[g.off[u] + k];
    const int64_t nd = (int64_t)dist[u] + wflat[deg_prefix[u] + k];
    if (nd < (int64_t)dist[v]) {
      const int32_t old = atomicMin(&dist[v], (int32_t)nd);
      if (nd < (int64_t)old) flag[v] = 1;
    }
  }
}

__global__ void k_clear_flags(uint8_t *flag, const uint32_t *q, uint64_t nq) {
  const uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i < nq) flag[q[i]] = 0;
}

inline uint64_t nblocks(uint64_t n) { return n ? (n + kT - 1) / kT : 1; }

WtTable table_of(const sstg_graph *g) {
  return WtTable{(unsigned long long *)g->wt_keys,
                 (unsigned long long *)g->wt_vals, g->wt_cap - 1};
}

// counts of the last upsert/erase (sum of per-block partials)
int read_count(sstg_graph *g, const uint64_t *part, uint64_t nb,
               uint64_t *out) {
  hipStream_t s = g->stream;
  SSTG_LAUNCH(k_sum_parts, 1, kT, 0, s, part, nb, g->dev_ctr + 7);
  SSTG_HIP(hipMemcpyAsync(g->host_rb + 18, g->dev_ctr + 7, sizeof(uint64_t),
                          hipMemcpyDeviceToHost, s));
  SSTG_HIP(hipStreamSynchronize(s));
  *out = g->host_rb[18];
  return SSTG_OK;
}

// room for `extra` more keys at load <= 1/2 (rehash drops tombstones)
int wt_reserve(sstg_graph *g, uint64_t extra) {
  hipStream_t s = g->stream;
  if (g->wt_keys && 2 * (g->wt_used + extra) <= g->wt_cap) return SSTG_OK;
  uint64_t cap = 1024;
  while (cap < 4 * (g->wt_live + extra)) cap <<= 1;
  unsigned long long *keys = nullptr, *vals = nullptr;
  SSTG_HIP(hipMalloc(&keys, cap * 8));
  if (hipMalloc(&vals, cap * 8) != hipSuccess) {
    (void)hipGetLastError();
    hipFree(keys);
    return SSTG_E_NOMEM;
  }
  SSTG_LAUNCH(k_wt_fill, nblocks(cap), kT, 0, s, keys, cap, kWtEmpty);
  SSTG_LAUNCH(k_wt_fill, nblocks(cap), kT, 0, s, vals, cap, 0ull);
  const WtTable nt{keys, vals, cap - 1};
  if (g->wt_keys) {
    SSTG_LAUNCH(k_wt_rehash, nblocks(g->wt_cap), kT, 0, s, table_of(g),
                g->wt_cap, nt);
    SSTG_HIP(hipStreamSynchronize(s));
    hipFree(g->wt_keys);
    hipFree(g->wt_vals);
  }
  g->wt_keys = (uint64_t *)keys;
  g->wt_vals = (uint64_t *)vals;
  g->wt_cap = cap;
  g->wt_used = g->wt_live;
  return SSTG_OK;
}

}  // namespace

int weights_upsert(sstg_graph *g, const uint32_t *src, const uint32_t *dst,
                   const uint32_t *w, uint64_t n, bool keep_self) {
  if (n == 0) return SSTG_OK;
  if (n >= (1ull << 31)) return SSTG_E_INVAL;  // callers pass chunks
  hipStream_t s = g->stream;
  SSTG_TRY(wt_reserve(g, n));
  if ((uint64_t)g->wt_tag + n >= 0xFFFFFFFFull) {
    SSTG_LAUNCH(k_wt_untag, nblocks(g->wt_cap), kT, 0, s,
                (unsigned long long *)g->wt_vals, g->wt_cap);
    g->wt_tag = 1;
  }
  uint64_t *part = nullptr;
  const uint64_t nb = nblocks(n);
  SSTG_TRY(g->ws.get(kWsWeights, nb, &part));
  SSTG_LAUNCH(k_wt_upsert, nb, kT, 0, s, table_of(g), src, dst, w, n,
              g->wt_tag, keep_self ? 1 : 0, part);
  g->wt_tag += (uint32_t)n;
  uint64_t added = 0;
  SSTG_TRY(read_count(g, part, nb, &added));
  g->wt_live += added;
  g->wt_used += added;
  g->wtflat_version = ~0ull;
  return SSTG_OK;
}

int weights_erase(sstg_graph *g, const uint32_t *src, const uint32_t *dst,
                  uint64_t n) {
  if (!g->wt_keys || n == 0) return SSTG_OK;
  uint64_t *part = nullptr;
  const uint64_t nb = nblocks(n);
  SSTG_TRY(g->ws.get(kWsWeights, nb, &part));
  SSTG_LAUNCH(k_wt_erase, nb, kT, 0, g->stream, table_of(g), src, dst, n,
              part);
  uint64_t removed = 0;
  SSTG_TRY(read_count(g, part, nb, &removed));
  g->wt_live -= removed;
  g->wtflat_version = ~0ull;
  return SSTG_OK;
}

int weights_get(sstg_graph *g, const uint32_t *rows, const uint32_t *cols,
                uint64_t n, uint32_t *out) {
  if (n == 0) return SSTG_OK;
  SSTG_LAUNCH(k_wt_get, nblocks(n), kT, 0, g->stream, g->view(), table_of(g),
              g->wt_keys ? 1 : 0, rows, cols, n, out);
  return SSTG_OK;
}

// weights of the stored edges in flat (degree prefix) order, cached per
// graph version
static int weights_flat(sstg_graph *g, const int32_t **out) {
  int32_t *flat = nullptr;
  SSTG_TRY(g->ws.get(kWsWeightFlat, g->m ? g->m : 1, &flat));
  *out = flat;
  if (g->wtflat_version == g->version) return SSTG_OK;
  const uint64_t *prefix = nullptr;
  SSTG_TRY(graph_degree_prefix(g, &prefix));
  if (g->m)
    SSTG_LAUNCH(k_wt_flat, nblocks((g->m + 7) / 8), kT, 0, g->stream,
                g->view(), prefix, g->m, table_of(g), g->wt_keys ? 1 : 0,
                flat);
  g->wtflat_version = g->version;
  return SSTG_OK;
}

int algo_bf_weighted(sstg_graph *g, uint32_t src, int32_t *dist) {
  hipStream_t s = g->stream;
  const uint32_t n = g->n_rows;
  if (n == 0) return SSTG_OK;
  const int32_t *wflat = nullptr;
  SSTG_TRY(weights_flat(g, &wflat));
  const uint64_t *prefix = nullptr;
  SSTG_TRY(graph_degree_prefix(g, &prefix));
  uint8_t *flag = nullptr;
  uint32_t *q = nullptr;
  uint64_t *fpre = nullptr;
  SSTG_TRY(g->ws.get(kWsAlgo + 11, n, &flag));
  SSTG_TRY(g->ws.get(kWsAlgo + 12, (uint64_t)n + 1, &q));
  SSTG_TRY(g->ws.get(kWsAlgo + 10, (uint64_t)n + 1, &fpre));
  SSTG_LAUNCH(k_bf_init, nblocks(n), kT, 0, s, n, src, dist, flag);
  if (src >= n) {
    SSTG_HIP(hipStreamSynchronize(s));
    return SSTG_OK;
  }
  uint32_t src_h = src;
  SSTG_HIP(hipMemcpyAsync(q, &src_h, sizeof(uint32_t), hipMemcpyHostToDevice,
                          s));
  uint64_t nf = 1;
  for (uint32_t round = 0; round < n && nf; round++) {
    SSTG_LAUNCH(em::k_frontier_degrees, nblocks(nf + 1), kT, 0, s, g->view(),
                (const uint32_t *)q, nf, fpre);
    SSTG_TRY(scan_exclusive_u64(fpre, fpre, nf + 1, g->dev_ctr + 8, g->ws,
                                kWsPrim, s));
    SSTG_HIP(hipMemcpyAsync(g->host_rb + 19, g->dev_ctr + 8, sizeof(uint64_t),
                            hipMemcpyDeviceToHost, s));
    SSTG_HIP(hipStreamSynchronize(s));
    const uint64_t mf = g->host_rb[19];
    if (mf)
      SSTG_LAUNCH(k_bf_relax, nblocks((mf + 7) / 8), kT, 0, s, g->view(),
                  prefix, wflat, (const uint32_t *)q, (const uint64_t *)fpre,
                  nf, mf, dist, flag);
    // next frontier: the flagged vertices, flags cleared for the round after
    SSTG_TRY(select_flagged_u32(flag, nullptr, n, q, g->dev_ctr + 8, g->ws,
                                kWsPrim, s));
    SSTG_HIP(hipMemcpyAsync(g->host_rb + 19, g->dev_ctr + 8, sizeof(uint64_t),
                            hipMemcpyDeviceToHost, s));
    SSTG_HIP(hipStreamSynchronize(s));
    nf = g->host_rb[19];
    if (nf)
      SSTG_LAUNCH(k_clear_flags, nblocks(nf), kT, 0, s, flag,
                  (const uint32_t *)q, nf);
  }
  SSTG_HIP(hipStreamSynchronize(s));
  return SSTG_OK;
}

void weights_free(sstg_graph *g) {
  hipFree(g->wt_keys);
  hipFree(g->wt_vals);
  g->wt_keys = g->wt_vals = nullptr;
  g->wt_cap = g->wt_used = g->wt_live = 0;
}

}  // namespace sstg
