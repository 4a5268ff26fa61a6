// FP64 PageRank over column segments (EdgeMapVertexMap::PR_S<double>, called
// at test.hpp:527 and 893-904).
//
// The pull x'(v) = d * sum_{u in N(v)} x(u)/deg(u) + (1-d)/n gathers one
// 8-byte contribution per edge from anywhere in an n-entry array.  Once that
// array outgrows a single XCD's 4 MB L2, most gathers miss and each miss
// moves a whole cache line.  Here the column ids are cut into segments of
// 2^sb vertices (2 MB of contributions at sb = 18) and each row into its
// runs, one run per segment the row touches.  Runs are stored grouped by
// (run-length class, segment), so one sweep over a class walks the segments
// in order and, with the blocks of a launch handed to the XCDs in contiguous
// slices, each XCD gathers from one L2-sized slice at a time:
//
//   sweep: partial[p] = sum of contrib[u] over run p   (G lanes per run)
//   combine: x'(v) = d * sum of partial over v's runs, in segment order
//
// Both steps have fixed lane strides and summation orders, so results do not
// depend on timing.  The layout is built once per graph version and kept in
// the graph's workspace.
#include <cstdlib>
#include <cstring>
#include <functional>

#include "algo.hpp"
#include "edge_map.hpp"
#include "primitives.hpp"

namespace sstg {
namespace {

constexpr int kT = 256;
constexpr int kE = 8;               // flat edges per thread in the build
constexpr uint32_t kClassMax[2] = {8, 64};  // run lengths of classes 0, 1
constexpr int kClasses = 3;         // 1 lane, 8 lanes, a wave per run

inline uint64_t blocks_for(uint64_t n) { return ceil_div(n ? n : 1, kT); }

// Walks the flat edges [kE*t, kE*t + kE) of thread t: calls
// f(e, v, col, run_start) in order.
template <class F>
__device__ __forceinline__ void walk_edges(const DevGraph &g,
                                           const uint64_t *prefix, uint64_t m,
                                           uint32_t sb, uint64_t t, F f) {
  const uint64_t e0 = t * kE;
  if (e0 >= m) return;
  const uint64_t end = e0 + kE < m ? e0 + kE : m;
  uint32_t v = (uint32_t)find_segment(prefix, g.n, e0);
  uint64_t rs = prefix[v], re = prefix[v + 1];
  const uint32_t *row = g.pool + g.off[v];
  uint32_t prev_seg = e0 > rs ? (row[e0 - rs - 1] >> sb) : 0xFFFFFFFFu;
  for (uint64_t e = e0; e < end; e++) {
    while (e >= re) {
      v++;
      rs = re;
      re = prefix[v + 1];
      row = g.pool + g.off[v];
      prev_seg = 0xFFFFFFFFu;
    }
    const uint32_t c = row[e - rs];
    const uint32_t sg = c >> sb;
    f(e, v, c, sg != prev_seg);
    prev_seg = sg;
  }
}

__global__ void __launch_bounds__(kT)
    k_runs_count(DevGraph g, const uint64_t *prefix, uint64_t m, uint32_t sb,
                 uint64_t nt, uint64_t *cnt) {
  const uint64_t t = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= nt) return;
  uint64_t c = 0;
  walk_edges(g, prefix, m, sb, t,
             [&](uint64_t, uint32_t, uint32_t, bool start) { c += start; });
  cnt[t] = c;
}

// run r: first flat edge run_e[r], row run_v[r]
__global__ void __launch_bounds__(kT)
    k_runs_emit(DevGraph g, const uint64_t *prefix, uint64_t m, uint32_t sb,
                uint64_t nt, const uint64_t *base, uint32_t *run_e,
                uint32_t *run_v) {
  const uint64_t t = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= nt) return;
  uint64_t r = base[t];
  walk_edges(g, prefix, m, sb, t,
             [&](uint64_t e, uint32_t v, uint32_t, bool start) {
               if (start) {
                 run_e[r] = (uint32_t)e;
                 run_v[r] = v;
                 r++;
               }
             });
}

// Sort key of run r: (class * segments + segment) << 32 | r; and each row's
// run range [first[v], endr[v]) in row-major run order.
__global__ void __launch_bounds__(kT)
    k_runs_key(DevGraph g, const uint64_t *prefix, uint64_t m, uint32_t sb,
               uint64_t P, uint32_t segments, const uint32_t *run_e,
               const uint32_t *run_v, uint64_t *key, uint32_t *first,
               uint32_t *endr) {
  const uint64_t r = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (r >= P) return;
  const uint32_t e = run_e[r];
  const uint64_t e1 = r + 1 < P ? run_e[r + 1] : m;
  const uint32_t len = (uint32_t)(e1 - e);
  const uint32_t v = run_v[r];
  const uint32_t sg = g.pool[g.off[v] + (e - prefix[v])] >> sb;
  const uint32_t cls = len <= kClassMax[0] ? 0 : len <= kClassMax[1] ? 1 : 2;
  key[r] = ((uint64_t)(cls * segments + sg) << 32) | r;
  if (r == 0 || run_v[r - 1] != v) first[v] = (uint32_t)r;
  if (r + 1 == P || run_v[r + 1] != v) endr[v] = (uint32_t)(r + 1);
}

// Slot p of the sorted order holds run r: pos[r] = p, len64[p] = its length.
__global__ void __launch_bounds__(kT)
    k_runs_place(const uint64_t *sorted, uint64_t P, uint64_t m,
                 const uint32_t *run_e, uint32_t *pos, uint64_t *len64) {
  const uint64_t p = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (p >= P) return;
  const uint32_t r = (uint32_t)sorted[p];
  pos[r] = (uint32_t)p;
  const uint64_t e1 = (uint64_t)r + 1 < P ? run_e[r + 1] : m;
  len64[p] = e1 - run_e[r];
  if (p == 0) len64[P] = 0;  // the exclusive scan leaves m there
}

// First slot of each class (lower bound of class * segments in the keys).
__global__ void k_class_bounds(const uint64_t *sorted, uint64_t P,
                               uint32_t segments, uint64_t *out) {
  if (threadIdx.x > kClasses) return;
  const uint64_t want = (uint64_t)(threadIdx.x * segments) << 32;
  uint64_t lo = 0, hi = P;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (sorted[mid] < want)
      lo = mid + 1;
    else
      hi = mid;
  }
  out[threadIdx.x] = lo;
}

// Copies every run's ids to its slot's place in the segmented column array.
__global__ void __launch_bounds__(kT)
    k_runs_scatter(DevGraph g, const uint64_t *prefix, uint64_t m,
                   uint32_t sb, uint64_t nt, const uint64_t *base,
                   const uint32_t *run_e, const uint32_t *pos,
                   const uint32_t *pstart, uint32_t *cols) {
  const uint64_t t = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= nt) return;
  // the run holding the thread's first edge is base[t] - 1 unless that edge
  // starts a run (edge 0 starts run 0, so base[t] >= 1 when it does not)
  uint64_t r = base[t] - 1;
  uint64_t dst = 0;
  bool have = false;
  walk_edges(g, prefix, m, sb, t,
             [&](uint64_t e, uint32_t, uint32_t c, bool start) {
               if (start) {
                 r++;
                 have = false;
               }
               if (!have) {
                 dst = (uint64_t)pstart[pos[r]] + (e - run_e[r]);
                 have = true;
               }
               cols[dst++] = c;
             });
}

__global__ void __launch_bounds__(kT)
    k_pr_seg_init(uint32_t n, const uint32_t *deg, double x0, double *rank,
                  double *contrib) {
  const uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v >= n) return;
  rank[v] = x0;
  const uint32_t d = deg[v];
  contrib[v] = d ? x0 / (double)d : 0.0;
}

// partial[p] = sum of contrib over run p, G lanes per run (G <= 64), for
// the runs [p0, p0 + count) of one class.
template <int G>
__global__ void __launch_bounds__(kT)
    k_pr_sweep(const uint32_t *cols, const uint32_t *pstart, uint64_t p0,
               uint64_t count, const double *contrib, double *partial) {
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const uint64_t gid = (uint64_t)blk * kT + threadIdx.x;
  const uint64_t grp = gid / G;
  const uint32_t lig = (uint32_t)(gid % G);
  if (grp >= count) return;  // whole groups leave together
  const uint64_t p = p0 + grp;
  const uint32_t a = pstart[p], b = pstart[p + 1];
  double acc = 0.0;
  uint32_t k = a + lig;
  for (; k + 3 * G < b; k += 4 * G) {
    const uint32_t u0 = cols[k], u1 = cols[k + G], u2 = cols[k + 2 * G],
                   u3 = cols[k + 3 * G];
    acc += contrib[u0];
    acc += contrib[u1];
    acc += contrib[u2];
    acc += contrib[u3];
  }
  for (; k < b; k += G) acc += contrib[cols[k]];
  if (G > 1)
    for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
  if (lig == 0) partial[p] = acc;
}

// x'(v) from v's runs in segment order; contrib_next for the next sweep.
__global__ void __launch_bounds__(kT)
    k_pr_combine(uint32_t n, const uint32_t *deg, const uint32_t *first,
                 const uint32_t *endr, const uint32_t *pos,
                 const double *partial, double damp, double base, double *rank,
                 double *contrib_next) {
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const uint64_t v = (uint64_t)blk * kT + threadIdx.x;
  if (v >= n) return;
  const uint32_t d = deg[v];
  double s = 0.0;
  if (d) {
    const uint32_t e = endr[v];
    for (uint32_t r = first[v]; r < e; r++) s += partial[pos[r]];
  }
  const double x = damp * s + base;
  rank[v] = x;
  contrib_next[v] = d ? x / (double)d : 0.0;
}

// workspace slots of the layout (kept) and of its build (temporary)
enum : int {
  kSlotCols = kWsPrSeg + 0,
  kSlotStart = kWsPrSeg + 1,
  kSlotPos = kWsPrSeg + 2,
  kSlotFirst = kWsPrSeg + 3,
  kSlotEnd = kWsPrSeg + 4,
  kSlotPartial = kWsPrSeg + 5,
  kSlotCnt = kWsPrSeg + 6,
  kSlotRunE = kWsPrSeg + 7,
  kSlotRunV = kWsPrSeg + 8,
  kSlotKeys = kWsPrSeg + 9,
  kSlotKeysTmp = kWsPrSeg + 10,
  kSlotContrib = kWsPrSeg + 11,
  kSlotContrib2 = kWsPrSeg + 12,
};

struct SegLayout {
  uint32_t *cols, *pstart, *pos, *first, *endr;
  uint64_t P;
  uint64_t cls[kClasses + 1];
};

int build_layout(sstg_graph *g, uint32_t sb, SegLayout *L) {
  hipStream_t s = g->stream;
  const uint64_t m = g->m;
  const uint32_t n = g->n_rows;
  Workspace &ws = g->ws;
  SSTG_TRY(ws.get(kSlotFirst, n, &L->first));
  SSTG_TRY(ws.get(kSlotEnd, n, &L->endr));
  if (g->prseg_version == g->version && g->prseg_bits == sb) {
    L->P = g->prseg_runs;
    for (int c = 0; c <= kClasses; c++) L->cls[c] = g->prseg_cls[c];
    SSTG_TRY(ws.get(kSlotCols, m, &L->cols));
    SSTG_TRY(ws.get(kSlotStart, L->P + 1, &L->pstart));
    SSTG_TRY(ws.get(kSlotPos, L->P, &L->pos));
    return SSTG_OK;
  }
  g->prseg_version = ~0ull;
  const uint64_t *prefix = nullptr;
  SSTG_TRY(graph_degree_prefix(g, &prefix));
  const DevGraph v = g->view();
  const uint64_t nt = ceil_div(m, kE);
  uint64_t *cnt = nullptr;
  SSTG_TRY(ws.get(kSlotCnt, nt + 1, &cnt));
  SSTG_LAUNCH(k_runs_count, blocks_for(nt), kT, 0, s, v, prefix, m, sb, nt,
              cnt);
  uint64_t *total = g->dev_ctr + 9;
  SSTG_TRY(scan_exclusive_u64(cnt, cnt, nt, total, ws, kWsPrim, s));
  SSTG_HIP(hipMemcpyAsync(g->host_rb + 20, total, sizeof(uint64_t),
                          hipMemcpyDeviceToHost, s));
  SSTG_HIP(hipStreamSynchronize(s));
  const uint64_t P = g->host_rb[20];
  uint32_t *run_e = nullptr, *run_v = nullptr;
  SSTG_TRY(ws.get(kSlotRunE, P, &run_e));
  SSTG_TRY(ws.get(kSlotRunV, P, &run_v));
  SSTG_LAUNCH(k_runs_emit, blocks_for(nt), kT, 0, s, v, prefix, m, sb, nt,
              (const uint64_t *)cnt, run_e, run_v);
  const uint32_t segments = (uint32_t)ceil_div(n, 1ull << sb);
  uint64_t *keys = nullptr, *tmp = nullptr;
  SSTG_TRY(ws.get(kSlotKeys, P + 1, &keys));
  SSTG_TRY(ws.get(kSlotKeysTmp, P + 1, &tmp));
  SSTG_LAUNCH(k_runs_key, blocks_for(P), kT, 0, s, v, prefix, m, sb, P,
              segments, (const uint32_t *)run_e, (const uint32_t *)run_v, keys,
              L->first, L->endr);
  int shifts[4], np = 0;
  const uint32_t kbits = bit_width_u32(kClasses * segments - 1);
  for (uint32_t b = 0; b < kbits; b += 8) shifts[np++] = 32 + (int)b;
  uint64_t *sorted = keys;
  if (np) SSTG_TRY(radix_sort_u64(keys, tmp, P, shifts, np, ws, kWsPrim, s,
                                  &sorted));
  SSTG_TRY(ws.get(kSlotPos, P, &L->pos));
  // the lengths take the buffer the sort left free
  uint64_t *len64 = sorted == keys ? tmp : keys;
  SSTG_LAUNCH(k_runs_place, blocks_for(P), kT, 0, s, (const uint64_t *)sorted,
              P, m, (const uint32_t *)run_e, L->pos, len64);
  uint64_t *bounds = g->dev_ctr + 12;
  SSTG_LAUNCH(k_class_bounds, 1, kWave, 0, s, (const uint64_t *)sorted, P,
              segments, bounds);
  SSTG_TRY(scan_exclusive_u64(len64, len64, P + 1, nullptr, ws, kWsPrim, s));
  SSTG_TRY(ws.get(kSlotStart, P + 1, &L->pstart));
  SSTG_TRY(narrow_u64_to_u32(len64, P + 1, L->pstart, s));
  SSTG_TRY(ws.get(kSlotCols, m, &L->cols));
  SSTG_LAUNCH(k_runs_scatter, blocks_for(nt), kT, 0, s, v, prefix, m, sb, nt,
              (const uint64_t *)cnt, (const uint32_t *)run_e,
              (const uint32_t *)L->pos, (const uint32_t *)L->pstart, L->cols);
  SSTG_HIP(hipMemcpyAsync(g->host_rb + 24, bounds,
                          (kClasses + 1) * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, s));
  SSTG_HIP(hipStreamSynchronize(s));
  L->P = P;
  for (int c = 0; c <= kClasses; c++) L->cls[c] = g->host_rb[24 + c];
  L->cls[kClasses] = P;
  g->prseg_runs = P;
  for (int c = 0; c <= kClasses; c++) g->prseg_cls[c] = L->cls[c];
  g->prseg_bits = sb;
  g->prseg_version = g->version;
  return SSTG_OK;
}


// One iteration: the three class sweeps, then the combine.
int pr_iteration(sstg_graph *g, const SegLayout &L, const double *contrib,
                 double *contrib_next, double *partial, double damping,
                 double *out) {
  hipStream_t s = g->stream;
  const uint32_t n = g->n_rows;
  const double base = (1.0 - damping) / (double)n;
  const uint64_t c0 = L.cls[1] - L.cls[0], c1 = L.cls[2] - L.cls[1],
                 c2 = L.cls[3] - L.cls[2];
  if (c0)
    SSTG_LAUNCH(k_pr_sweep<1>, blocks_for(c0), kT, 0, s,
                (const uint32_t *)L.cols, (const uint32_t *)L.pstart, L.cls[0],
                c0, contrib, partial);
  if (c1)
    SSTG_LAUNCH(k_pr_sweep<8>, blocks_for(c1 * 8), kT, 0, s,
                (const uint32_t *)L.cols, (const uint32_t *)L.pstart, L.cls[1],
                c1, contrib, partial);
  if (c2)
    SSTG_LAUNCH(k_pr_sweep<64>, blocks_for(c2 * 64), kT, 0, s,
                (const uint32_t *)L.cols, (const uint32_t *)L.pstart, L.cls[2],
                c2, contrib, partial);
  SSTG_LAUNCH(k_pr_combine, blocks_for(n), kT, 0, s, n,
              (const uint32_t *)g->deg, (const uint32_t *)L.first,
              (const uint32_t *)L.endr, (const uint32_t *)L.pos,
              (const double *)partial, damping, base, out, contrib_next);
  return SSTG_OK;
}

}  // namespace

// Frees the graph's recorded launch sequences (PageRank pairs and BFS).
void pagerank_release(sstg_graph *g) {
  if (g->pr_exec) hipGraphExecDestroy((hipGraphExec_t)g->pr_exec);
  if (g->pr_graph) hipGraphDestroy((hipGraph_t)g->pr_graph);
  g->pr_exec = g->pr_graph = nullptr;
  delete static_cast<IterPairKey *>(g->pr_key);
  g->pr_key = nullptr;
  bfs_release(g);
}

bool iteration_capture_enabled(const sstg_graph *g) {
  // read at every call, so a caller can switch it around one call
  const char *on = std::getenv("SSTG_GRAPH_CAPTURE");
  return !g->no_capture && on && on[0] == '1';
}

// The graph keeps one recorded pair of iterations: it is recorded again
// whenever the key (path, graph version, settings, buffers, damping)
// differs from the one it was recorded for.
int replay_iteration_pairs(sstg_graph *g, const IterPairKey &key,
                           uint32_t pairs, const std::function<int()> &record) {
  hipStream_t s = g->stream;
  const IterPairKey *have = static_cast<const IterPairKey *>(g->pr_key);
  if (!g->pr_exec || !have || std::memcmp(have, &key, sizeof(key)) != 0) {
    pagerank_release(g);
    Profiler &prof = profiler();
    const bool prof_on = prof.on;
    prof.on = false;  // no event records inside the capture
    SSTG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = record();
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    prof.on = prof_on;
    if (rc != SSTG_OK) {
      if (graph) hipGraphDestroy(graph);
      return rc;
    }
    SSTG_HIP(e);
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      hipGraphDestroy(graph);
      SSTG_HIP(ei);
    }
    g->pr_graph = graph;
    g->pr_exec = exec;
    g->pr_key = new IterPairKey(key);
  }
  Profiler &prof = profiler();
  for (uint32_t k = 0; k < pairs; k++) {
    hipEvent_t ev = nullptr;
    if (prof.on) prof.begin("pagerank_pair_replay", s, &ev);
    SSTG_HIP(hipGraphLaunch((hipGraphExec_t)g->pr_exec, s));
    if (prof.on) prof.end("pagerank_pair_replay", s, ev);
  }
  return SSTG_OK;
}

bool pagerank_segmented_applies(const sstg_graph *g) {
  return g->pr_seg_bits != 0 && g->m > 0 && g->m < (1ull << 32) &&
         (uint64_t)g->n_rows > (1ull << g->pr_seg_bits);
}

// The iterations run in pairs (contributions back in their first buffer
// after two), and a pair is recorded once into a hipGraph and replayed: one
// launch per pair instead of eight (replay_iteration_pairs).
int pagerank_segmented(sstg_graph *g, uint32_t iters, double damping,
                       double *out) {
  hipStream_t s = g->stream;
  const uint32_t n = g->n_rows;
  SegLayout L;
  SSTG_TRY(build_layout(g, g->pr_seg_bits, &L));
  double *contrib = nullptr, *contrib2 = nullptr, *partial = nullptr;
  SSTG_TRY(g->ws.get(kSlotContrib, n, &contrib));
  SSTG_TRY(g->ws.get(kSlotContrib2, n, &contrib2));
  SSTG_TRY(g->ws.get(kSlotPartial, L.P, &partial));
  const double x0 = 1.0 / (double)n;
  SSTG_LAUNCH(k_pr_seg_init, blocks_for(n), kT, 0, s, n,
              (const uint32_t *)g->deg, x0, out, contrib);
  uint32_t plain = iters;
  if (iters >= 2 && iteration_capture_enabled(g)) {
    IterPairKey key{};
    key.path = 3;  // column segments, FP64
    key.version = g->version;
    key.setting = g->pr_seg_bits;
    key.n = n;
    const void *ptrs[10] = {L.cols, L.pstart, L.pos, L.first, L.endr,
                            contrib, contrib2, partial, out, g->deg};
    for (int i = 0; i < 10; i++) key.ptr[i] = ptrs[i];
    key.damping = damping;
    SSTG_TRY(replay_iteration_pairs(g, key, iters / 2, [&]() {
      SSTG_TRY(pr_iteration(g, L, contrib, contrib2, partial, damping, out));
      return pr_iteration(g, L, contrib2, contrib, partial, damping, out);
    }));
    plain = iters & 1u;
  }
  for (uint32_t it = 0; it < plain; it++) {
    SSTG_TRY(pr_iteration(g, L, contrib, contrib2, partial, damping, out));
    std::swap(contrib, contrib2);
  }
  SSTG_HIP(hipStreamSynchronize(s));
  return SSTG_OK;
}

}  // namespace sstg
