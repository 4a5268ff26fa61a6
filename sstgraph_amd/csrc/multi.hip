// Many independent graphs as one device-resident store, processed together.
//
// k graphs are packed as the disjoint union of their vertex sets: graph g
// owns global ids [vbase[g], vbase[g+1]).  Because no edge crosses graphs,
//   * one BFS seeded with every graph's source gives each graph's levels,
//   * one union-find pass gives each graph's components (min-id roots, so
//     local label = global root - vbase[g]),
//   * PageRank with per-vertex 1/n_g normalisation gives each graph's ranks,
// all from single launches over the union instead of k launches per step.
// The whole BFS + CC + PR sequence is captured once into a hipGraph (fixed
// kernel sequence: BFS levels up to a bound, each level exits early once the
// frontier is empty) and replayed.  No memset nodes are recorded: buffers
// are cleared by fill kernels.
#include <algorithm>
#include <vector>

#include "algo.hpp"
#include "cc.hpp"
#include "edge_map.hpp"
#include "primitives.hpp"

struct sstg_multi {
  sstg_graph *u = nullptr;  // the union graph
  uint32_t k = 0;
  std::vector<uint32_t> vbase;  // k + 1 entries (host)
  uint32_t *d_vbase = nullptr;  // device copy
  uint32_t *d_graph_of = nullptr;  // graph id of each vertex
  // captured execution
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t exec_version = ~0ull, exec_epoch = ~0ull;
  uint32_t exec_iters = 0, exec_levels = 0, levels_run = 0;
  double exec_damping = 0;
  int exec_f32 = -1;
  // outputs of the captured run (device)
  uint32_t *level = nullptr, *label = nullptr, *sources = nullptr,
           *uf = nullptr, *counts = nullptr, *giant = nullptr;
  double *rank64 = nullptr, *c64a = nullptr, *c64b = nullptr;
  float *rank32 = nullptr, *c32a = nullptr, *c32b = nullptr;
};

namespace sstg {
namespace {

constexpr int kT = 256;
constexpr uint32_t kInf = 0xFFFFFFFFu;
inline uint64_t blocks(uint64_t n) { return ceil_div(n ? n : 1, kT); }

}  // namespace
constexpr uint32_t kMultiCounts = 1u << 16;  // BFS level counters
namespace {

__global__ void k_graph_of(const uint32_t *vbase, uint32_t k, uint32_t n,
                           uint32_t *graph_of) {
  uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v >= n) return;
  uint32_t lo = 0, hi = k;  // largest g with vbase[g] <= v
  while (hi - lo > 1) {
    uint32_t m = (lo + hi) >> 1;
    if (vbase[m] <= v)
      lo = m;
    else
      hi = m;
  }
  graph_of[v] = lo;
}

// Local ids -> global ids; a local id outside the graph becomes the invalid
// global id `bad`, which the batch update rejects with SSTG_E_RANGE.
__global__ void k_localize(const uint32_t *src, const uint32_t *dst,
                           uint64_t n, uint32_t base, uint32_t ng, uint32_t bad,
                           uint32_t *gsrc, uint32_t *gdst) {
  uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  uint32_t s = src[i], d = dst[i];
  gsrc[i] = s < ng ? s + base : bad;
  gdst[i] = d < ng ? d + base : bad;
}

// Batches of several graphs, concatenated: element i of batch j (off[j] <= i
// < off[j+1]) is shifted by its graph's vertex base bbase[j]; ids outside that
// graph (>= bsize[j]) become `bad`, so that the update rejects the whole call.
__global__ void k_localize_batches(const uint32_t *src, const uint32_t *dst,
                                   uint64_t n, const uint64_t *off,
                                   uint32_t nb, const uint32_t *bbase,
                                   const uint32_t *bsize, uint32_t bad,
                                   uint32_t *gsrc, uint32_t *gdst) {
  uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  // the batch holding i: largest j < nb with off[j] <= i (empty batches
  // share their offset with the next one, which is larger)
  uint32_t lo = 0, hi = nb;  // off[lo] <= i < off[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  const uint32_t base = bbase[lo], ng = bsize[lo];
  const uint32_t s = src[i], d = dst[i];
  gsrc[i] = s < ng ? s + base : bad;
  gdst[i] = d < ng ? d + base : bad;
}

// ---- multi-source BFS, one kernel per level (top-down over all vertices)
__global__ void k_mbfs_init(uint32_t n, const uint32_t *vbase,
                            const uint32_t *graph_of, const uint32_t *sources,
                            uint32_t *level, uint32_t *counts,
                            uint32_t max_levels) {
  uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v < n) {
    uint32_t g = graph_of[v];
    level[v] = (v - vbase[g] == sources[g]) ? 0u : kInf;
  }
  if (v < max_levels) counts[v] = (v == 0) ? 1u : 0u;  // counts[0] != 0
}

// Level L: every vertex of level L claims its unvisited neighbours, with G
// lanes per vertex for the vertices of one degree bin.  counts[L + 1] is set
// non-zero once a vertex of level L + 1 is found (read first, so only the
// first few waves store to it); every launch exits at once when level L is
// empty.
template <int G>
__global__ void __launch_bounds__(kT)
    k_mbfs_level(DevGraph g, const uint32_t *list, uint64_t count, uint32_t L,
                 uint32_t *level, uint32_t *counts) {
  if (counts[L] == 0) return;
  const uint64_t gid = (uint64_t)blockIdx.x * kT + threadIdx.x;
  const uint64_t i = gid / G;
  const uint32_t lig = (uint32_t)(gid % G);
  bool found = false;
  if (i < count) {
    const uint32_t v = list[i];
    if (level[v] == L) {
      const uint32_t d = g.deg[v];
      const uint32_t *row = g.pool + g.off[v];
      for (uint32_t k = lig; k < d; k += G) {
        const uint32_t w = row[k];
        if (level[w] == kInf && atomicCAS(&level[w], kInf, L + 1) == kInf)
          found = true;
      }
    }
  }
  if (__ballot(found) && lane_id() == 0 && counts[L + 1] == 0) counts[L + 1] = 1;
}

// ---- CC over the union: the single-graph union-find kernels (cc.hpp)
// with one large component per graph
__global__ void k_mcc_final(uint32_t n, const uint32_t *vbase,
                            const uint32_t *graph_of, uint32_t *p,
                            uint32_t *label) {
  uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v >= n) return;
  label[v] = cc::find(p, (uint32_t)v) - vbase[graph_of[v]];
}

// ---- PageRank over the union with per-graph normalisation
template <class Real>
__global__ void k_mpr_init(uint32_t n, const uint32_t *deg,
                           const uint32_t *vbase, const uint32_t *graph_of,
                           Real *rank, Real *contrib) {
  uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v >= n) return;
  uint32_t g = graph_of[v];
  Real x0 = (Real)(1.0 / (double)(vbase[g + 1] - vbase[g]));
  rank[v] = x0;
  uint32_t d = deg[v];
  contrib[v] = d ? x0 / (Real)d : (Real)0;
}

template <class Real, bool kKahan> struct MultiPr;

template <class Real> struct MultiPr<Real, false> {
  const Real *contrib;
  Real *contrib_next, *rank;
  const uint32_t *deg, *vbase, *graph_of;
  Real damp, one_minus;
  using Acc = Real;
  __device__ __forceinline__ Acc zero() const { return (Real)0; }
  __device__ __forceinline__ void add(Acc &a, uint32_t, uint32_t u) const {
    a += contrib[u];
  }
  __device__ __forceinline__ Acc merge(Acc a, Acc b) const { return a + b; }
  __device__ static __forceinline__ Acc shfl_xor(Acc a, int o) {
    return __shfl_xor(a, o, kWave);
  }
  __device__ __forceinline__ void store(uint32_t v, Acc a) const {
    uint32_t g = graph_of[v];
    Real base = one_minus / (Real)(vbase[g + 1] - vbase[g]);
    Real r = damp * a + base;
    rank[v] = r;
    uint32_t d = deg[v];
    contrib_next[v] = d ? r / (Real)d : (Real)0;
  }
};

template <> struct MultiPr<float, true> {
  const float *contrib;
  float *contrib_next, *rank;
  const uint32_t *deg, *vbase, *graph_of;
  float damp, one_minus;
  using Acc = KahanF;
  __device__ __forceinline__ Acc zero() const { return KahanF{0.f, 0.f}; }
  __device__ __forceinline__ void add(Acc &a, uint32_t, uint32_t u) const {
    float y = contrib[u] - a.c;
    float t = a.s + y;
    a.c = (t - a.s) - y;
    a.s = t;
  }
  __device__ __forceinline__ Acc merge(Acc a, Acc b) const {
    float t = a.s + b.s;
    float e = (fabsf(a.s) >= fabsf(b.s)) ? (a.s - t) + b.s : (b.s - t) + a.s;
    return KahanF{t, a.c + b.c - e};
  }
  __device__ static __forceinline__ Acc shfl_xor(Acc a, int o) {
    return KahanF{__shfl_xor(a.s, o, kWave), __shfl_xor(a.c, o, kWave)};
  }
  __device__ __forceinline__ void store(uint32_t v, Acc a) const {
    uint32_t g = graph_of[v];
    float base = one_minus / (float)(vbase[g + 1] - vbase[g]);
    float r = damp * (a.s - a.c) + base;
    rank[v] = r;
    uint32_t d = deg[v];
    contrib_next[v] = d ? r / (float)d : 0.f;
  }
};

template <class Real, bool kKahan>
int record_pagerank(sstg_multi *M, const DegreeBins &bins, uint32_t iters,
                    double damping, Real *rank, Real *ca, Real *cb) {
  sstg_graph *g = M->u;
  hipStream_t s = g->stream;
  const uint32_t n = g->n_rows;
  SSTG_LAUNCH(k_mpr_init<Real>, blocks(n), kT, 0, s, n,
              (const uint32_t *)g->deg, (const uint32_t *)M->d_vbase,
              (const uint32_t *)M->d_graph_of, rank, ca);
  for (uint32_t it = 0; it < iters; it++) {
    MultiPr<Real, kKahan> f{ca,          cb,           rank,
                            g->deg,      M->d_vbase,   M->d_graph_of,
                            (Real)damping, (Real)(1.0 - damping)};
    SSTG_TRY(vertex_pull_all(g, bins, f));
    std::swap(ca, cb);
  }
  return SSTG_OK;
}

}  // namespace

// One BFS level over all graphs, by degree bin: 1, 4 and 16 lanes per vertex
// and a block (16 rather than 64 lanes in the third bin took the batched
// pass from 66.7 to 62.6 ms, 4 rather than 8 in the second from 62.2 to
// 60.3 ms).
static void launch_level(sstg_multi *M, const DegreeBins &bins, uint32_t L,
                         hipStream_t s) {
  const DevGraph v = M->u->view();
  if (bins.count[0])
    hipLaunchKernelGGL(k_mbfs_level<1>, dim3(ceil_div(bins.count[0], kT)),
                       dim3(kT), 0, s, v, (const uint32_t *)bins.list[0],
                       bins.count[0], L, M->level, M->counts);
  if (bins.count[1])
    hipLaunchKernelGGL(k_mbfs_level<4>, dim3(ceil_div(bins.count[1] * 4, kT)),
                       dim3(kT), 0, s, v, (const uint32_t *)bins.list[1],
                       bins.count[1], L, M->level, M->counts);
  if (bins.count[2])
    hipLaunchKernelGGL(k_mbfs_level<16>,
                       dim3(ceil_div(bins.count[2] * 16, kT)), dim3(kT), 0, s,
                       v, (const uint32_t *)bins.list[2], bins.count[2], L,
                       M->level, M->counts);
  if (bins.count[3])
    hipLaunchKernelGGL(k_mbfs_level<256>,
                       dim3(ceil_div(bins.count[3] * 256, kT)), dim3(kT), 0, s,
                       v, (const uint32_t *)bins.list[3], bins.count[3], L,
                       M->level, M->counts);
}

// Records BFS (levels 0..max_levels), CC and PR into one hipGraph.
static int capture(sstg_multi *M, uint32_t max_levels, uint32_t iters,
                   double damping, bool f32) {
  sstg_graph *g = M->u;
  hipStream_t s = g->stream;
  const uint32_t n = g->n_rows;
  // everything that allocates or reads back happens before capture
  const uint64_t *prefix = nullptr;
  SSTG_TRY(graph_degree_prefix(g, &prefix));
  // the level kernels and PageRank share the reductions' threshold set,
  // which MultiGraph tuning sets on PageRank
  DegreeBins bins;
  SSTG_TRY(degree_bins(g, &bins, kBinsPull));
  bool symmetric = false;
  SSTG_TRY(graph_is_symmetric(g, &symmetric));
  SSTG_HIP(hipStreamSynchronize(s));
  // the event profiler must not record into a capturing stream
  Profiler &prof = profiler();
  const bool prof_on = prof.on;
  prof.on = false;
  if (M->graph) {
    hipGraphExecDestroy(M->exec);
    hipGraphDestroy(M->graph);
    M->graph = nullptr;
    M->exec = nullptr;
  }
  SSTG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = SSTG_OK;
  do {
    // BFS from every graph's source at once
    hipLaunchKernelGGL(k_mbfs_init,
                       dim3(blocks(std::max<uint64_t>(n, kMultiCounts))),
                       dim3(kT), 0, s, n, (const uint32_t *)M->d_vbase,
                       (const uint32_t *)M->d_graph_of,
                       (const uint32_t *)M->sources, M->level, M->counts,
                       kMultiCounts);
    for (uint32_t L = 0; L < max_levels; L++) launch_level(M, bins, L, s);
    // CC
    const DevGraph v = g->view();
    if (g->m && symmetric) {
      hipLaunchKernelGGL(cc::k_identity, dim3(blocks(n)), dim3(kT), 0, s, n,
                         M->uf);
      hipLaunchKernelGGL(cc::k_link_sample, dim3(blocks(n)), dim3(kT), 0, s, v,
                         M->uf);
      hipLaunchKernelGGL(cc::k_compress, dim3(blocks(n)), dim3(kT), 0, s, n,
                         M->uf);
      hipLaunchKernelGGL(cc::k_mode,
                         dim3(ceil_div((uint64_t)M->k * kWave, kT)), dim3(kT),
                         0, s, (const uint32_t *)M->d_vbase, M->k, n, M->uf,
                         M->giant);
      hipLaunchKernelGGL(cc::k_link_rest, dim3(blocks(n)), dim3(kT), 0, s, v,
                         M->uf, (const uint32_t *)M->giant,
                         (const uint32_t *)M->d_graph_of);
    } else {
      hipLaunchKernelGGL(cc::k_init, dim3(blocks(n)), dim3(kT), 0, s, v,
                         M->uf);
      if (g->m)
        hipLaunchKernelGGL(cc::k_hook_all,
                           dim3(ceil_div(ceil_div(g->m, 8), kT)), dim3(kT), 0,
                           s, v, prefix, g->m, M->uf);
    }
    hipLaunchKernelGGL(k_mcc_final, dim3(blocks(n)), dim3(kT), 0, s, n,
                       (const uint32_t *)M->d_vbase,
                       (const uint32_t *)M->d_graph_of, M->uf, M->label);
    // PageRank
    if (f32)
      rc = record_pagerank<float, true>(M, bins, iters, damping, M->rank32,
                                        M->c32a, M->c32b);
    else
      rc = record_pagerank<double, false>(M, bins, iters, damping, M->rank64,
                                          M->c64a, M->c64b);
  } while (0);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(s, &graph);
  prof.on = prof_on;
  if (rc != SSTG_OK) {
    if (graph) hipGraphDestroy(graph);
    return rc;
  }
  SSTG_HIP(e);
  SSTG_HIP(hipGraphInstantiate(&M->exec, graph, nullptr, nullptr, 0));
  M->graph = graph;
  M->exec_version = g->version;
  M->exec_epoch = g->tune_epoch;
  M->exec_iters = iters;
  M->exec_levels = max_levels;
  M->exec_damping = damping;
  M->exec_f32 = f32 ? 1 : 0;
  return SSTG_OK;
}

// One update pass over the union store for nb batches of uneven size: the
// batches are localised into union ids by one kernel and merged by one
// insert (or remove) of their concatenation.  Invalid graph ids are rejected
// before anything runs; out-of-range vertex ids fail the whole call with
// SSTG_E_RANGE and leave every graph unchanged.
static int multi_update_batches(sstg_multi *M, uint32_t nb, const uint32_t *graph,
                                const uint64_t *count, const uint32_t *src,
                                const uint32_t *dst, uint32_t flags, bool insert) {
  if (!M || (nb && (!graph || !count))) return SSTG_E_INVAL;
  std::vector<uint64_t> off(nb + 1);
  std::vector<uint32_t> bb(2 * (size_t)nb + 1);  // bases, then sizes
  uint64_t n = 0;
  for (uint32_t j = 0; j < nb; j++) {
    if (graph[j] >= M->k) return SSTG_E_INVAL;
    off[j] = n;
    n += count[j];
    bb[j] = M->vbase[graph[j]];
    bb[nb + j] = M->vbase[graph[j] + 1] - M->vbase[graph[j]];
  }
  off[nb] = n;
  if (n == 0) return SSTG_OK;
  if (!src || !dst) return SSTG_E_INVAL;
  sstg_graph *g = M->u;
  hipSetDevice(g->device);
  hipStream_t s = g->stream;
  uint32_t *ds = nullptr, *dd = nullptr, *gs = nullptr, *gd = nullptr,
           *d_bb = nullptr;
  uint64_t *d_off = nullptr;
  SSTG_TRY(g->ws.get(kWsMulti, n, &gs));
  SSTG_TRY(g->ws.get(kWsMulti + 1, n, &gd));
  SSTG_TRY(g->ws.get(kWsMultiOff, nb + 1, &d_off));
  SSTG_TRY(g->ws.get(kWsMultiBase, 2 * (uint64_t)nb + 1, &d_bb));
  SSTG_HIP(hipMemcpyAsync(d_off, off.data(), (nb + 1) * sizeof(uint64_t),
                          hipMemcpyHostToDevice, s));
  SSTG_HIP(hipMemcpyAsync(d_bb, bb.data(), bb.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice, s));
  if (flags & SSTG_DEVICE_PTRS) {
    ds = (uint32_t *)src;
    dd = (uint32_t *)dst;
  } else {
    SSTG_TRY(g->ws.get(kWsMulti + 2, n, &ds));
    SSTG_TRY(g->ws.get(kWsMulti + 3, n, &dd));
    SSTG_HIP(hipMemcpyAsync(ds, src, n * 4, hipMemcpyHostToDevice, s));
    SSTG_HIP(hipMemcpyAsync(dd, dst, n * 4, hipMemcpyHostToDevice, s));
  }
  SSTG_LAUNCH(k_localize_batches, blocks(n), kT, 0, s, (const uint32_t *)ds,
              (const uint32_t *)dd, n, (const uint64_t *)d_off, nb,
              (const uint32_t *)d_bb, (const uint32_t *)(d_bb + nb),
              g->n_rows, gs, gd);
  const bool skip_self = insert && !(flags & SSTG_KEEP_SELF_LOOPS);
  return graph_update(g, gs, gd, n, insert, skip_self, true);
}

}  // namespace sstg

using namespace sstg;

extern "C" {

int sstg_multi_create(uint32_t k, const uint32_t *n_per_graph,
                      sstg_multi **out) {
  if (!out || !n_per_graph || k == 0) return SSTG_E_INVAL;
  *out = nullptr;
  uint64_t total = 0;
  std::vector<uint32_t> vbase(k + 1);
  for (uint32_t g = 0; g < k; g++) {
    if (n_per_graph[g] == 0) return SSTG_E_INVAL;
    vbase[g] = (uint32_t)total;
    total += n_per_graph[g];
  }
  if (total >= 0xFFFFFFFFull) return SSTG_E_RANGE;
  vbase[k] = (uint32_t)total;
  sstg_multi *M = new (std::nothrow) sstg_multi();
  if (!M) return SSTG_E_NOMEM;
  int rc = sstg_create((uint32_t)total, (uint32_t)total, 0, &M->u);
  if (rc) {
    delete M;
    return rc;
  }
  M->k = k;
  M->vbase = vbase;
  sstg_graph *g = M->u;
  const uint64_t n = total;
  auto fail = [&](int code) {
    sstg_multi_destroy(M);
    return code;
  };
  if (hipMalloc(&M->d_vbase, (k + 1) * sizeof(uint32_t)) ||
      hipMalloc(&M->d_graph_of, n * sizeof(uint32_t)) ||
      hipMalloc(&M->level, n * sizeof(uint32_t)) ||
      hipMalloc(&M->label, n * sizeof(uint32_t)) ||
      hipMalloc(&M->uf, n * sizeof(uint32_t)) ||
      hipMalloc(&M->sources, k * sizeof(uint32_t)) ||
      hipMalloc(&M->counts, kMultiCounts * sizeof(uint32_t)) ||
      hipMalloc(&M->giant, k * sizeof(uint32_t)) ||
      hipMalloc(&M->rank64, n * sizeof(double)) ||
      hipMalloc(&M->c64a, n * sizeof(double)) ||
      hipMalloc(&M->c64b, n * sizeof(double)) ||
      hipMalloc(&M->rank32, n * sizeof(float)) ||
      hipMalloc(&M->c32a, n * sizeof(float)) ||
      hipMalloc(&M->c32b, n * sizeof(float)))
    return fail(SSTG_E_NOMEM);
  // on the stream the kernel below runs on (a copy on the null stream is
  // not ordered before work on it)
  if (hipMemcpyAsync(M->d_vbase, vbase.data(), (k + 1) * sizeof(uint32_t),
                     hipMemcpyHostToDevice, g->stream))
    return fail(SSTG_E_DEVICE);
  hipLaunchKernelGGL(k_graph_of, dim3(blocks(n)), dim3(kT), 0, g->stream,
                     (const uint32_t *)M->d_vbase, k, (uint32_t)n,
                     M->d_graph_of);
  if (hipStreamSynchronize(g->stream)) return fail(SSTG_E_DEVICE);
  *out = M;
  return SSTG_OK;
}

int sstg_multi_destroy(sstg_multi *M) {
  if (!M) return SSTG_E_INVAL;
  if (M->exec) hipGraphExecDestroy(M->exec);
  if (M->graph) hipGraphDestroy(M->graph);
  for (void *p : {(void *)M->d_vbase, (void *)M->d_graph_of, (void *)M->level,
                  (void *)M->label, (void *)M->uf, (void *)M->sources,
                  (void *)M->counts, (void *)M->giant, (void *)M->rank64,
                  (void *)M->c64a,
                  (void *)M->c64b, (void *)M->rank32, (void *)M->c32a,
                  (void *)M->c32b})
    if (p) hipFree(p);
  if (M->u) sstg_destroy(M->u);
  delete M;
  return SSTG_OK;
}

int sstg_multi_graph(sstg_multi *M, sstg_graph **out) {
  if (!M || !out) return SSTG_E_INVAL;
  *out = M->u;
  return SSTG_OK;
}

int sstg_multi_insert_batch(sstg_multi *M, uint32_t graph, const uint32_t *src,
                            const uint32_t *dst, uint64_t n, uint32_t flags) {
  if (!M || graph >= M->k || (n && (!src || !dst))) return SSTG_E_INVAL;
  if (n == 0) return SSTG_OK;
  sstg_graph *g = M->u;
  hipSetDevice(g->device);
  const uint32_t base = M->vbase[graph];
  const uint32_t ng = M->vbase[graph + 1] - base;
  uint32_t *ds = nullptr, *dd = nullptr, *gs = nullptr, *gd = nullptr;
  SSTG_TRY(g->ws.get(kWsMulti, n, &gs));
  SSTG_TRY(g->ws.get(kWsMulti + 1, n, &gd));
  if (flags & SSTG_DEVICE_PTRS) {
    ds = (uint32_t *)src;
    dd = (uint32_t *)dst;
  } else {
    SSTG_TRY(g->ws.get(kWsMulti + 2, n, &ds));
    SSTG_TRY(g->ws.get(kWsMulti + 3, n, &dd));
    SSTG_HIP(hipMemcpyAsync(ds, src, n * 4, hipMemcpyHostToDevice, g->stream));
    SSTG_HIP(hipMemcpyAsync(dd, dst, n * 4, hipMemcpyHostToDevice, g->stream));
  }
  SSTG_LAUNCH(k_localize, blocks(n), kT, 0, g->stream, (const uint32_t *)ds,
              (const uint32_t *)dd, n, base, ng, g->n_rows, gs, gd);
  return graph_update(g, gs, gd, n, true, !(flags & SSTG_KEEP_SELF_LOOPS),
                      true);
}

int sstg_multi_insert_batches(sstg_multi *M, uint32_t nb, const uint32_t *graph,
                              const uint64_t *count, const uint32_t *src,
                              const uint32_t *dst, uint32_t flags) {
  return multi_update_batches(M, nb, graph, count, src, dst, flags, true);
}

int sstg_multi_remove_batches(sstg_multi *M, uint32_t nb, const uint32_t *graph,
                              const uint64_t *count, const uint32_t *src,
                              const uint32_t *dst, uint32_t flags) {
  return multi_update_batches(M, nb, graph, count, src, dst, flags, false);
}

int sstg_multi_run(sstg_multi *M, const uint32_t *sources, uint32_t max_levels,
                   uint32_t pr_iters, double damping, int pr_f32) {
  if (!M || !sources || max_levels == 0 || max_levels + 1 >= kMultiCounts)
    return SSTG_E_INVAL;
  sstg_graph *g = M->u;
  hipSetDevice(g->device);
  for (uint32_t i = 0; i < M->k; i++)
    if (sources[i] >= M->vbase[i + 1] - M->vbase[i]) return SSTG_E_RANGE;
  SSTG_HIP(hipMemcpyAsync(M->sources, sources, M->k * sizeof(uint32_t),
                          hipMemcpyHostToDevice, g->stream));
  if (!M->exec || M->exec_version != g->version ||
      M->exec_epoch != g->tune_epoch || M->exec_iters != pr_iters ||
      M->exec_levels != max_levels || M->exec_damping != damping ||
      M->exec_f32 != (pr_f32 ? 1 : 0))
    SSTG_TRY(capture(M, max_levels, pr_iters, damping, pr_f32 != 0));
  SSTG_HIP(hipGraphLaunch(M->exec, g->stream));
  // a BFS deeper than the recorded level budget continues level by level
  uint32_t L = max_levels;
  for (;;) {
    uint32_t c = 0;
    SSTG_HIP(hipMemcpyAsync(&c, M->counts + L, sizeof(uint32_t),
                            hipMemcpyDeviceToHost, g->stream));
    SSTG_HIP(hipStreamSynchronize(g->stream));
    if (c == 0) break;
    if (L + 1 >= kMultiCounts) return SSTG_E_UNSUPPORTED;
    DegreeBins bins;
    SSTG_TRY(degree_bins(g, &bins, kBinsPull));
    launch_level(M, bins, L, g->stream);
    SSTG_HIP(hipGetLastError());
    L++;
  }
  M->levels_run = L;
  return SSTG_OK;
}

int sstg_multi_results(sstg_multi *M, uint32_t *level, uint32_t *label,
                       double *rank64, float *rank32, uint32_t *levels_used) {
  if (!M) return SSTG_E_INVAL;
  sstg_graph *g = M->u;
  hipSetDevice(g->device);
  const uint64_t n = g->n_rows;
  if (level)
    SSTG_HIP(hipMemcpy(level, M->level, n * 4, hipMemcpyDeviceToHost));
  if (label)
    SSTG_HIP(hipMemcpy(label, M->label, n * 4, hipMemcpyDeviceToHost));
  if (rank64)
    SSTG_HIP(hipMemcpy(rank64, M->rank64, n * 8, hipMemcpyDeviceToHost));
  if (rank32)
    SSTG_HIP(hipMemcpy(rank32, M->rank32, n * 4, hipMemcpyDeviceToHost));
  if (levels_used) *levels_used = M->levels_run;
  return SSTG_OK;
}

}  // extern "C"
