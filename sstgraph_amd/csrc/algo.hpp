// EdgeMap/VertexMap algorithms on the device-resident store (algo.hip).
// All outputs are device arrays of n_rows entries.
#pragma once

#include <functional>

#include "graph.hpp"

namespace sstg {

int algo_bfs(sstg_graph *g, uint32_t src, int32_t *parent, uint32_t *level);
// FP64 PageRank; column segments (pagerank.hip) when they apply and
// allow_segments is set, the degree-binned vertex pull otherwise
int algo_pagerank(sstg_graph *g, uint32_t iters, double damping, double *out,
                  bool allow_segments = true);
bool pagerank_segmented_applies(const sstg_graph *g);
// A pair of PageRank iterations recorded into a hipGraph and replayed
// (pagerank.hip): the graph keeps one recording and its key, and records
// again when the key changes.  Opt-in, with the environment variable
// SSTG_GRAPH_CAPTURE=1 (read at each call): while a stream captures, HIP
// refuses synchronous copies on the legacy stream from any other thread,
// and the replay saves 0.2% of a scale-24 PageRank.  Never on a server
// slot's view (g->no_capture).
struct IterPairKey {
  uint32_t path;     // which PageRank: 1 vertex pull FP64, 2 FP32 + Kahan,
                     // 3 column segments
  uint32_t n;
  uint64_t version;  // graph version
  uint64_t setting;  // degree-threshold epoch or segment width
  const void *ptr[10];
  double damping;
};
bool iteration_capture_enabled(const sstg_graph *g);
// A recorded BFS (algo.hip): with the same switch, a BFS from the source
// of the last recording on the same graph version, settings, outputs and
// workspace replays that recording (hipGraph) instead of launching level by
// level with a host read-back per level.
struct RecKey {
  uint64_t version, tune_epoch, ws_generation;
  uint32_t src, n;
  const void *ptr[5];  // outputs (two), and the store's off, deg, pool
  double alpha, beta;
};
void bfs_release(sstg_graph *g);
int replay_iteration_pairs(sstg_graph *g, const IterPairKey &key,
                           uint32_t pairs, const std::function<int()> &record);
int pagerank_segmented(sstg_graph *g, uint32_t iters, double damping,
                       double *out);
int algo_pagerank_f32k(sstg_graph *g, uint32_t iters, double damping,
                       float *out);
int algo_bc(sstg_graph *g, uint32_t src, double *out);
int algo_cc(sstg_graph *g, uint32_t *out);
int algo_tc(sstg_graph *g, uint64_t *out);
// degree-oriented TC; valid only when every edge has its reverse (tc.hip)
int tc_oriented(sstg_graph *g, uint64_t *out);
// sum over u, v in L(u) of |L(u) n L(v)|, L(x) = lists[base[x] .. +len[x])
// sorted unique ids; SSTG_E_UNSUPPORTED when a list has 2^23 ids or more
int count_list_triangles(sstg_graph *g, const uint64_t *base,
                         const uint32_t *lists, const uint32_t *len,
                         uint64_t *out);
// triangles containing at least one of the changed undirected edges `keys`
// (sorted row << 32 | col pairs, both directions of each edge), each once;
// valid when every edge of g has its reverse (tc.hip)
int tc_delta(sstg_graph *g, const uint64_t *keys, uint64_t count,
             uint64_t *out);
int algo_bf(sstg_graph *g, uint32_t src, int32_t *out);
int algo_common_neighbors(sstg_graph *g, uint32_t u, uint32_t v,
                          int early_end, uint64_t *out);

}  // namespace sstg
