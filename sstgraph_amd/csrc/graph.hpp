// Host-side handle of one device-resident graph.
#pragma once

#include "common.hpp"

struct sstg_graph {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t n_rows = 0, n_cols = 0;
  uint64_t m = 0;             // live edges
  uint64_t version = 0;       // bumped by every mutation
  // tinyset store (see sstg::DevGraph)
  uint64_t *off = nullptr;
  uint32_t *deg = nullptr;
  uint32_t *cap = nullptr;
  uint32_t *pool = nullptr;
  uint64_t pool_cap = 0;      // elements
  uint64_t pool_tail = 0;     // first free element
  uint64_t garbage = 0;       // elements in abandoned slots
  uint64_t rebuilds = 0;
  sstg_update_stats last{};
  // per-call scratch
  sstg::Workspace ws;
  // set for the duration of an allocation-free call: updates then skip the
  // optional pool compaction and fail (graph unchanged) when the pool tail
  // has no room, instead of allocating
  bool noalloc = false;
  // small pinned host block for read-backs
  uint64_t *host_rb = nullptr;
  // small pinned host block mapped into the device's address space
  // (fine-grained: kernels write it directly, and the host reads it after
  // its wait with no copy launched) and its device address; null where a
  // handle has none (the copy through host_rb is used then)
  uint64_t *host_ctr = nullptr;
  uint64_t *host_ctr_dev = nullptr;
  // the last value the host asked the device to write to the flag slot of
  // host_ctr (kHostFlag) once a step's counters are in place
  uint64_t host_seq = 0;
  // device-side counters block (see update.hip)
  uint64_t *dev_ctr = nullptr;
  // cached derived data: exclusive prefix of deg (n_rows + 1 entries)
  uint64_t *deg_prefix = nullptr;
  uint64_t deg_prefix_version = ~0ull;
  // cached degree bins (lists live in workspace), two sets: [kBinsPull] for
  // the vertex-pull reductions (PageRank, BC, TC orientation, the batched
  // store's BFS levels), [kBinsBfs] for the single-source BFS steps.  Bin b holds degrees in
  // (bin_hi[set][b-1], bin_hi[set][b]], the last bin the rest; tune.hip sets
  // the boundaries, tune_epoch counts the changes.  The defaults are what
  // the tuner picks on rMat at scales 22 and 24 (both sets, both seeds
  // measured)
  uint64_t bins_version[2] = {~0ull, ~0ull};
  uint64_t bins_count[2][4] = {};
  uint32_t bin_hi[2][3] = {{8, 64, 256}, {32, 128, 2048}};
  uint64_t tune_epoch = 0;
  // BFS direction switch (Beamer): top-down -> bottom-up when the frontier
  // grows and its edges exceed unexplored / alpha; back when it holds fewer
  // than n / beta vertices
  double bfs_alpha = 15.0, bfs_beta = 18.0;
  // FP64 PageRank over column segments of 2^pr_seg_bits vertices (0, the
  // default: the degree-binned vertex pull); the layout is cached for graph
  // version prseg_version
  uint32_t pr_seg_bits = 0;
  uint64_t prseg_version = ~0ull;
  uint32_t prseg_bits = 0;
  uint64_t prseg_runs = 0;
  uint64_t prseg_cls[4] = {};
  // a recorded pair of PageRank iterations (hipGraph_t / hipGraphExec_t)
  // and what it was recorded for; no_capture: never record (server views,
  // whose threads share the device with synchronous copies)
  void *pr_graph = nullptr, *pr_exec = nullptr, *pr_key = nullptr;
  // a recorded BFS and a recorded BC (one source's whole launch sequence)
  // and their keys
  void *bfs_graph = nullptr, *bfs_exec = nullptr, *bfs_key = nullptr;
  void *bc_graph = nullptr, *bc_exec = nullptr, *bc_key = nullptr;
  bool no_capture = false;
  // cached: does every stored edge have its reverse?
  uint64_t sym_version = ~0ull;
  bool symmetric = false;
  // edge weights (weights.hip): open-addressing table keyed by
  // (row << 32 | col), made by the first weighted insert
  uint64_t *wt_keys = nullptr;
  uint64_t *wt_vals = nullptr;  // (tag << 32) | weight
  uint64_t wt_cap = 0, wt_used = 0, wt_live = 0;
  uint32_t wt_tag = 1;
  uint64_t wtflat_version = ~0ull;

  sstg::DevGraph view() const {
    return sstg::DevGraph{n_rows, off, deg, cap, pool};
  }
};

namespace sstg {

// Slot capacity of a row of d ids: the next power of two of at least 4, but
// never more than the row could ever hold (n_cols) when d fits in that.
__host__ __device__ inline uint64_t slot_capacity(uint64_t d, uint32_t n_cols) {
  uint64_t c = 4;
  while (c < d) c <<= 1;
  const uint64_t lim = n_cols < 4 ? 4 : n_cols;
  if (c > lim && d <= lim) c = lim;
  return c;
}

// update_small.hip: a whole update of at most kSmallBatch pairs (device
// arrays) in one workgroup.  *fallback is set, with the graph unchanged,
// when the batch needs more pool room than is left or moves a very long
// row; the caller then runs the general pipeline.
constexpr uint32_t kSmallBatch = 4096;
int update_small(sstg_graph *g, const uint32_t *src, const uint32_t *dst,
                 uint32_t n, bool insert, bool skip_self, bool *fallback);

// update.hip
int graph_alloc(sstg_graph *g);
void graph_free(sstg_graph *g);
// host_ctr holds kHostCtrSlots words: the counter blocks from 0 and, at
// kHostFlag, the sequence number a kernel writes after them
constexpr int kHostCtrSlots = 32;
constexpr int kHostFlag = 24;

// Waits until the device has written seq to host_ctr[kHostFlag] (spinning on
// the mapped word, which arrives as soon as the kernel writes it rather than
// when the stream's completion signal is seen), then orders the host's later
// reads after it.  Falls back to a stream synchronisation when the graph has
// no mapped block or the flag has not come within 100 ms.
int wait_host_flag(sstg_graph *g, uint64_t seq);

int graph_update(sstg_graph *g, const uint32_t *src, const uint32_t *dst,
                 uint64_t n, bool insert, bool skip_self, bool device_ptrs);
int graph_rebuild(sstg_graph *g, uint64_t extra_elems);
int graph_effective_keys(sstg_graph *g, const uint32_t *src,
                         const uint32_t *dst, uint64_t n, bool insert,
                         bool skip_self, const uint64_t **out_keys,
                         uint64_t *out_count);
int graph_degree_prefix(sstg_graph *g, const uint64_t **out);
int graph_export(sstg_graph *g, uint64_t *offsets, uint32_t *cols,
                 bool device_ptrs);
int graph_has_batch(sstg_graph *g, const uint32_t *rows, const uint32_t *cols,
                    uint64_t n, uint8_t *out, bool device_ptrs);
int graph_touch_sum(sstg_graph *g, uint64_t *out);
// algo.hip
int graph_is_symmetric(sstg_graph *g, bool *out);
// pagerank.hip: drops the graph's recorded PageRank iterations
void pagerank_release(sstg_graph *g);
// tune.hip
int graph_set_thresholds(sstg_graph *g, int set, const uint32_t t[3]);
int graph_set_bfs_switch(sstg_graph *g, double alpha, double beta);
int graph_autotune(sstg_graph *g, uint32_t what, uint32_t src,
                   uint32_t pr_iters, sstg_tune_report *rep);
// weights.hip (device arrays)
int weights_upsert(sstg_graph *g, const uint32_t *src, const uint32_t *dst,
                   const uint32_t *w, uint64_t n, bool keep_self);
int weights_erase(sstg_graph *g, const uint32_t *src, const uint32_t *dst,
                  uint64_t n);
int weights_get(sstg_graph *g, const uint32_t *rows, const uint32_t *cols,
                uint64_t n, uint32_t *out);
int algo_bf_weighted(sstg_graph *g, uint32_t src, int32_t *dist);
void weights_free(sstg_graph *g);
// relabel.hip
int graph_degree_order_map(sstg_graph *g, uint32_t *map);
int graph_rewrite(sstg_graph *g, const uint32_t *map, sstg_graph *out);
int graph_statistics(sstg_graph *g, uint64_t *hist33, uint64_t *empty_rows,
                     uint64_t *max_degree, uint64_t *slot_capacity);

// rmat.hip
int er_generate(uint32_t n, uint32_t seed, uint64_t first, uint64_t count,
                uint32_t *src, uint32_t *dst, hipStream_t s);
int rmat_generate(uint32_t n, uint32_t seed, double a, double b, double c,
                  uint64_t first, uint64_t count, uint32_t *src, uint32_t *dst,
                  hipStream_t s, bool normalized = false);

// Workspace slot map of one graph (each user owns a range of slots).
enum : int {
  kWsKeysA = 0,
  kWsKeysB = 1,
  kWsFlags = 2,
  kWsBatch = 3,
  kWsSegIdx = 4,
  kWsSegV = 5,
  kWsSegP0 = 6,
  kWsSegStage = 7,
  kWsSegAlloc = 8,
  kWsSegWork = 9,
  kWsSegOldDeg = 10,
  kWsStage = 11,
  kWsInSrc = 12,
  kWsInDst = 13,
  kWsMisc = 14,
  kWsOut = 15,
  kWsOut2 = 80,
  kWsSegTile = 81,
  kWsSegIsBig = 82,
  kWsSegIsTile = 83,
  kWsBigIdx = 84,
  kWsTileIdx = 85,
  kWsBigStage = 86,
  kWsBigWork = 87,
  kWsTileLen = 88,
  kWsTileSrc = 89,
  kWsTileDst = 90,
  kWsTileAlen = 91,
  kWsTileBst = 92,
  kWsTileBlk = 93,
  kWsMulti = 94,  // 94..97
  kWsTcDag = 98,
  kWsTcDplus = 99,
  kWsTcList = 100,
  kWsTcAcc = 101,
  kWsRelKeys = 102,
  kWsRelTmp = 103,
  kWsRelSrc = 104,
  kWsRelDst = 105,
  kWsTouch = 106,
  kWsWeights = 107,
  kWsWeightFlat = 108,
  kWsWtSrc = 109,
  kWsWtDst = 110,
  kWsWtVal = 111,
  kWsAlgo = 16,      // 16..47 algorithms
  kWsPrim = 48,      // 48..63 primitives (scan recursion etc.)
  kWsPrim2 = 64,     // 64..79 second primitive range
  kWsBottomUp = 112,  // 112, 113 active lists of the bottom-up steps,
                      // 114 the BFS step counters
  kWsMultiOff = 116,  // multi store: batch offsets of a batched update
  kWsMultiBase = 117, // multi store: vertex base and size of each batch
  kWsBfsBins = 122,   // degree bin lists of the BFS set, back to back
  kWsMarkPos = 123,   // update: row position of each sorted batch key
  kWsBPos = 124,      // update: the same for the kept keys
  kWsSlots = 126,
  kWsPrSeg = 130,     // 130..142 FP64 PageRank column-segment layout
  kWsEffKeys = 143,   // the pairs an update changes (graph_effective_keys)
  kWsTcDelta = 144,   // per-block partial counts of tc_delta
  kWsSmallKeys = 145, // small updates: the plan (update_small.hip)
  kWsSmallPos = 146,
  kWsSmallRows = 147,
  kWsSmallFound = 149,  // small updates: each pair's row search
  kWsRebuild = 150,     // 150..152 pool rebuild: capacities, offsets, prefix
};

// degree threshold sets of a graph (sstg_graph::bin_hi): the reductions
// (PageRank, BC, TC orientation, the batched store's BFS levels) and the
// single-source BFS steps
constexpr int kBinsPull = 0;
constexpr int kBinsBfs = 1;

}  // namespace sstg
