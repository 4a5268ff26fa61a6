/*
 * sstg.h -- C ABI of sstgraph_amd, an MI355X (gfx950) implementation of the
 * SSTGraph hot path: tinyset batch edge insert/delete and the
 * EdgeMap/VertexMap algorithms BFS, PageRank, BC, CC and TC.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, never
 * throws or exits, and returns an int status (SSTG_OK == 0, negative
 * SSTG_E_* on failure; sstg_strerror() names it).  Each declaration cites
 * the reference interface it replaces (paths relative to the SSTGraph tree;
 * SM = include/SSTGraph/SparseMatrix.hpp, TST = include/SSTGraph/internal/
 * test.hpp, TSS = include/SSTGraph/TinySet_small.hpp).
 *
 * Pointers: by default every array argument is HOST memory.  With the flag
 * SSTG_DEVICE_PTRS the arrays are device memory of the graph's device.
 *
 * Threading: one mutating call (insert/remove/rebuild) in flight per graph;
 * read-only calls on one graph may be issued from several host threads but
 * are serialised on the graph's stream.  Different graphs are independent.
 */
#ifndef SSTG_H
#define SSTG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSTG_ABI_VERSION 2

/* status codes */
#define SSTG_OK 0
#define SSTG_E_INVAL -1    /* null handle / bad argument */
#define SSTG_E_RANGE -2    /* vertex id out of range (nothing was changed) */
#define SSTG_E_NOMEM -3    /* device allocation failed or budget exceeded */
#define SSTG_E_DEVICE -4   /* HIP runtime error */
#define SSTG_E_NODEVICE -5 /* no usable gfx950 device */
#define SSTG_E_UNSUPPORTED -6
#define SSTG_E_IO -7       /* a file could not be opened, read or written */
#define SSTG_E_TOO_LARGE -8 /* beyond an allocation-free call's fixed scratch */

/* flags */
#define SSTG_DEVICE_PTRS 0x1u  /* array arguments are device pointers */
#define SSTG_SEMI_SORTED 0x2u  /* batch already grouped by row (a hint, SM:167) */
#define SSTG_KEEP_SELF_LOOPS 0x4u /* insert x==y edges (SM:536-544 single insert) */

typedef struct sstg_graph sstg_graph;

const char *sstg_strerror(int status);
int sstg_abi_version(void);
/* Number of visible HIP devices (0 on a machine without a GPU). */
int sstg_device_count(int *out);

/* ---- graph lifetime -------------------------------------------------- */
/* SparseMatrixV<true>(height, width), SM:73-87.  The graph lives on the
 * current HIP device. */
int sstg_create(uint32_t n_rows, uint32_t n_cols, uint32_t flags,
                sstg_graph **out);
/* ~SparseMatrixV, SM:141-146. */
int sstg_destroy(sstg_graph *g);

/* ---- updates ----------------------------------------------------------- */
/* SparseMatrixV<true>::insert_batch(element_type *edges, n, semi_sorted),
 * SM:556-643: contents become old U {(src[i], dst[i]) : src[i] != dst[i]}.
 * Structure of arrays instead of std::tuple; the caller's arrays are not
 * reordered.  Fails with SSTG_E_RANGE (graph unchanged) if an id is out of
 * range. */
int sstg_insert_batch(sstg_graph *g, const uint32_t *src, const uint32_t *dst,
                      uint64_t n, uint32_t flags);
/* SparseMatrixV<true>::remove_batch(tuple<el_t,el_t>*, n, semi_sorted),
 * SM:645-718: contents become old \ batch; absent edges are no-ops and
 * self-loops are not skipped. */
int sstg_remove_batch(sstg_graph *g, const uint32_t *src, const uint32_t *dst,
                      uint64_t n, uint32_t flags);
/* Allocation-free insert_batch.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory during the call (code
 * under HIP graph capture, tight serving loops, hosts that forbid it).
 * Same result as sstg_insert_batch, but it calls no hipMalloc/hipFree: its
 * kernels work in one fixed device scratch array per GPU (64 MiB, held by
 * the library for the life of the process), so
 *  - src/dst must be device arrays (flags must include SSTG_DEVICE_PTRS);
 *  - n above SSTG_NOALLOC_MAX_PAIRS is SSTG_E_TOO_LARGE, graph unchanged;
 *  - a batch whose rows need more pool room than the pool has left, or
 *    whose long rows need more staging than the scratch has left, returns
 *    SSTG_E_NOMEM with the graph unchanged (sstg_insert_batch grows the
 *    pool);
 *  - the optional compaction of abandoned slots is skipped.
 * Every call on one GPU shares that GPU's scratch: calls that overlap on
 * one GPU (several host threads, or graphs with different streams) race. */
#define SSTG_NOALLOC_MAX_PAIRS (1ull << 18)
int sstg_insert_batch_noalloc(sstg_graph *g, const uint32_t *src,
                              const uint32_t *dst, uint64_t n, uint32_t flags);
/* SparseMatrixV::insert(row, col), SM:536-544 (no self-loop skip). */
/* Weighted batch insert (SparseMatrixV<true, uint32_t>::insert_batch with
 * (row, col, value) tuples, SparseMatrix.hpp:556-643 and the valued PMA
 * insert, PMA.hpp:1095-1146): the edges as insert_batch (self-loops
 * skipped unless SSTG_KEEP_SELF_LOOPS), and every edge's weight set to its
 * value, the last occurrence in the batch winning; a present edge's weight
 * is overwritten.  Edges without a weight (inserted unweighted) weigh 1.
 * remove_batch and remove also drop weights.  Weights are kept in a device
 * hash table beside the store, made by the first weighted insert. */
int sstg_insert_batch_weighted(sstg_graph *g, const uint32_t *src,
                               const uint32_t *dst, const uint32_t *w,
                               uint64_t n, uint32_t flags);
/* value(row, col) (SparseMatrix.hpp:727-735): weight of each edge; 1 for a
 * stored edge inserted without one, 0 for an absent edge. */
int sstg_edge_weights(sstg_graph *g, const uint32_t *rows,
                      const uint32_t *cols, uint64_t n, uint32_t *out,
                      uint32_t flags);
int sstg_insert(sstg_graph *g, uint32_t row, uint32_t col);
/* SparseMatrixV::remove(row, col), SM:546-554. */
int sstg_remove(sstg_graph *g, uint32_t row, uint32_t col);

/* What the last insert_batch / remove_batch call did (summed over its
 * internal chunks).  No reference counterpart: the reference reports only
 * wall time (test.hpp:668-687); these counts give the bytes each kernel had
 * to move. */
typedef struct sstg_update_stats {
  uint64_t batch;          /* pairs passed in */
  uint64_t effective;      /* pairs that changed the graph */
  uint64_t rows;           /* rows touched */
  uint64_t tile_rows;      /* rows merged in place through LDS */
  uint64_t tile_inputs;    /* their merge inputs (moved suffix + batch run) */
  uint64_t big_rows;       /* rows merged over the flat index space */
  uint64_t staged;         /* elements copied to scratch for big rows */
  uint64_t merged;         /* flat work items of the big-row merge */
  uint64_t allocated;      /* slot elements taken at the pool tail */
  uint64_t rebuilds;       /* pool rebuilds so far on this graph */
} sstg_update_stats;
int sstg_last_update_stats(sstg_graph *g, sstg_update_stats *out);

/* ---- queries ----------------------------------------------------------- */
/* SparseMatrixV::has(row, col), SM:720-726. */
int sstg_has(sstg_graph *g, uint32_t row, uint32_t col, int *out);
/* has() for n pairs at once; out[i] = 0/1. */
int sstg_has_batch(sstg_graph *g, const uint32_t *rows, const uint32_t *cols,
                   uint64_t n, uint8_t *out, uint32_t flags);
/* num_nodes() / get_rows(), SM:56,150-155. */
int sstg_num_nodes(sstg_graph *g, uint32_t *out);
/* M(), SM:518-525. */
int sstg_num_edges(sstg_graph *g, uint64_t *out);
/* getDegree(i) for every row, SM:57-59. */
int sstg_degrees(sstg_graph *g, uint32_t *out, uint32_t flags);
/* Ascending contents of every row as CSR (offsets[n_rows+1], cols[M()]):
 * the rows' map order (TSS:295-353, map_range SM:306-315).  cols may be NULL
 * to fetch offsets only. */
int sstg_export_csr(sstg_graph *g, uint64_t *offsets, uint32_t *cols,
                    uint32_t flags);
/* get_memory_size(), SM:527-534: bytes of device memory held by the graph. */
int sstg_memory_size(sstg_graph *g, uint64_t *out);
/* touch_all_sum(), SM:436-454: sum of all stored column ids. */
int sstg_touch_all_sum(sstg_graph *g, uint64_t *out);
/* common_neighbors(u, v, early_end), SM:61-71 / TSS:517-612. */
int sstg_common_neighbors(sstg_graph *g, uint32_t u, uint32_t v,
                          int early_end, uint64_t *out);

/* ---- EdgeMap/VertexMap algorithms (the EdgeMapVertexMap:: calls at
 * TST:476-618).  Outputs are arrays of n_rows entries. -------------------- */
/* BFS(g, src), TST:486: parent[src] = src, -1 if unreached; level[v] = hop
 * depth (TST:497-509), UINT32_MAX if unreached.  Either output may be NULL. */
int sstg_bfs(sstg_graph *g, uint32_t src, int32_t *parent, uint32_t *level,
             uint32_t flags);
/* PR_S<double>(g, iters), TST:527: x0 = 1/n, x' = damping * sum over
 * neighbours u of x(u)/deg(u) + (1 - damping)/n, FP64.  The neighbours of
 * v are its row N+(v) (out-neighbours) and deg is the out-degree: on a
 * directed graph this is the transpose of Ligra's PageRank, which sums over
 * in-neighbours; on a graph whose edges all have their reverses the two
 * agree.  For the in-neighbour form pass the transposed graph. */
int sstg_pagerank_f64(sstg_graph *g, uint32_t iters, double damping,
                      double *out, uint32_t flags);
/* The same PageRank with FP32 ranks and Kahan-compensated neighbour sums. */
int sstg_pagerank_f32k(sstg_graph *g, uint32_t iters, double damping,
                       float *out, uint32_t flags);
/* BC(g, src), TST:547: single-source Brandes dependencies, FP64; out[src] =
 * |R(src)| - 1; unreached vertices 0. */
int sstg_bc(sstg_graph *g, uint32_t src, double *out, uint32_t flags);
/* CC(g), TST:571: component label = smallest vertex id in the component. */
int sstg_cc(sstg_graph *g, uint32_t *out, uint32_t flags);
/* TC(g), TST:606: number of triangles (each counted once). */
int sstg_tc(sstg_graph *g, uint64_t *out);
/* BF(g, src), TST:618 and TST:811: shortest distances from src,
 * INT32_MAX/2 for unreached vertices.  Edge weights are those set by
 * sstg_insert_batch_weighted (read as int32, 1 for an edge without one); a
 * graph that never had a weighted insert weighs every edge 1.  At most
 * n_rows relaxation rounds (a negative cycle ends there). */
int sstg_bf(sstg_graph *g, uint32_t src, int32_t *out, uint32_t flags);
/* TouchAll(g), TST:476: sum over all edges of the destination id. */
int sstg_touch_all(sstg_graph *g, uint64_t *out);

/* ---- many independent graphs processed together ------------------------
 * No single reference counterpart: the reference processes one graph per
 * call.  k graphs live in one device store (their disjoint union); one
 * replay of a captured hipGraph runs BFS (from one source per graph), CC and
 * PageRank on all of them.  Results are per vertex in graph order: entries
 * [vbase[g], vbase[g+1]) belong to graph g, with vbase the prefix sum of
 * n_per_graph; levels and labels are local to each graph. */
typedef struct sstg_multi sstg_multi;
int sstg_multi_create(uint32_t k, const uint32_t *n_per_graph,
                      sstg_multi **out);
int sstg_multi_destroy(sstg_multi *m);
/* insert_batch (SM:556-643) into graph `graph`; ids are local to it. */
int sstg_multi_insert_batch(sstg_multi *m, uint32_t graph, const uint32_t *src,
                            const uint32_t *dst, uint64_t n, uint32_t flags);
/* insert_batch (SM:556-643) / remove_batch (SM:645-718) of nb batches of
 * several graphs in ONE update pass over the union store.  Batch j holds
 * count[j] edges of graph graph[j] (ids local to it); src/dst are the batches
 * concatenated in order (host or device arrays per flags).  The reference
 * has no counterpart that takes several graphs at once: it is the batched
 * form of calling sstg_multi_insert_batch per graph, with the same result.
 * graph[j] >= k is SSTG_E_INVAL; a vertex id outside its batch's graph is
 * SSTG_E_RANGE and leaves every graph unchanged. */
int sstg_multi_insert_batches(sstg_multi *m, uint32_t nb, const uint32_t *graph,
                              const uint64_t *count, const uint32_t *src,
                              const uint32_t *dst, uint32_t flags);
int sstg_multi_remove_batches(sstg_multi *m, uint32_t nb, const uint32_t *graph,
                              const uint64_t *count, const uint32_t *src,
                              const uint32_t *dst, uint32_t flags);
/* The union as an ordinary graph handle (owned by m; do not destroy). */
int sstg_multi_graph(sstg_multi *m, sstg_graph **out);
/* BFS from sources[g] in every graph g (levels as sstg_bfs), CC labels (as
 * sstg_cc) and pr_iters PageRank iterations with each graph normalised by
 * its own vertex count (FP64, or FP32 with Kahan sums if pr_f32), all in one
 * hipGraph replay.  max_levels is the BFS depth recorded in the graph;
 * deeper BFS levels continue with plain launches. */
int sstg_multi_run(sstg_multi *m, const uint32_t *sources, uint32_t max_levels,
                   uint32_t pr_iters, double damping, int pr_f32);
/* Host copies of the last run's results (any pointer may be NULL); each
 * array has sum(n_per_graph) entries.  levels_used = BFS levels executed. */
int sstg_multi_results(sstg_multi *m, uint32_t *level, uint32_t *label,
                       double *rank64, float *rank32, uint32_t *levels_used);

/* ---- workload generator (rmat_util.h:82-131 rMat<uint32_t>) ------------ */
/* Edges i in [first, first+count) of rMat(n, seed, a, b, c) into src/dst
 * (host or device arrays per flags).  n must be a power of two.  The draw
 * is the reference's hash32 / INT32_MAX in [0, 2), so the effective quadrant
 * probabilities are a/2, b/2, c/2 and the rest; with SSTG_RMAT_NORMALIZED
 * (no reference counterpart) it is hash32 / 2^32 and they are a, b, c. */
#define SSTG_RMAT_NORMALIZED 0x10u
int sstg_rmat(uint32_t n, uint32_t seed, double a, double b, double c,
              uint64_t first, uint64_t count, uint32_t *src, uint32_t *dst,
              uint32_t flags);

/* ---- serving: one device-resident graph, a stream of requests ----------
 * No reference counterpart (the reference has no server).  Requests are
 * ordered by arrival; a query sees exactly the updates that arrived before
 * it.  Consecutive updates of one kind are applied as one device batch
 * (continuous batching, flag SSTG_SERVE_COALESCE); queries run concurrently
 * on `slots` slots (own HIP stream and workspace each), admitted in arrival
 * order while a slot is free and their device-memory estimate fits the HBM
 * budget (hbm_budget bytes, 0 = 7/8 of the free memory at creation).  With
 * SSTG_SERVE_SNAPSHOTS an update does not wait for running or queued
 * queries: the graph moves to a device copy of its store and those queries
 * run on the old arrays (held against the budget while they are needed).
 * A graph with edge weights is refused (SSTG_E_UNSUPPORTED). */
#define SSTG_SERVE_COALESCE 1u
#define SSTG_SERVE_SNAPSHOTS 2u
/* at most one PageRank or TC at a time; lighter queries pass a waiting one */
#define SSTG_SERVE_ONE_HEAVY 4u
/* the triangle count is kept through updates while every edge has its
 * reverse: each update run is logged with the triangles its changed edges
 * close (insert) or break (remove); once one TC query has counted a logged
 * version in full, a TC query on any logged version is answered from the
 * log without a device pass.  Results are those of a full count. */
#define SSTG_SERVE_MAINTAIN 8u
#define SSTG_REQ_INSERT 1u /* insert_batch (SM:556-643) */
#define SSTG_REQ_REMOVE 2u /* remove_batch (SM:645-718) */
#define SSTG_REQ_BFS 3u    /* BFS levels from `source` (TST:486) */
#define SSTG_REQ_PR 4u     /* `iters` PageRank iterations, FP64 (TST:527) */
#define SSTG_REQ_TC 5u     /* triangle count (TST:606) */
#define SSTG_REQ_CC 6u     /* component labels (TST:571) */
typedef struct sstg_request {
  uint32_t kind;
  uint32_t source;                /* BFS */
  uint32_t iters;                 /* PR */
  double damping;                 /* PR */
  const uint32_t *src, *dst;      /* updates: host arrays, copied at submit */
  uint64_t n;
} sstg_request;
typedef struct sstg_request_info {
  int status;
  double t_arrive, t_start, t_done; /* host clock, seconds */
  uint64_t scalar;                  /* TC count */
  uint64_t coalesced;               /* requests in the same device batch */
} sstg_request_info;
typedef struct sstg_server sstg_server;
int sstg_server_create(sstg_graph *g, uint32_t slots, uint64_t hbm_budget,
                       uint32_t flags, sstg_server **out);
int sstg_server_submit(sstg_server *s, const sstg_request *req,
                       uint64_t *ticket);
/* Blocks until the request is done; info may be NULL. */
int sstg_server_wait(sstg_server *s, uint64_t ticket, sstg_request_info *info);
/* Copies a finished query's result (BFS levels u32[n], PR f64[n], CC u32[n])
 * to host memory of `bytes` bytes. */
int sstg_server_fetch(sstg_server *s, uint64_t ticket, void *host_out,
                      uint64_t bytes);
/* Frees a finished request's record and device result. */
int sstg_server_release(sstg_server *s, uint64_t ticket);
/* device update batches applied, the HBM budget, store copies made */
int sstg_server_stats(sstg_server *s, uint64_t *device_batches,
                      uint64_t *budget, uint64_t *snapshots);
/* Waits for every submitted request, then stops the server. */
int sstg_server_destroy(sstg_server *s);

/* ---- relabelling, statistics, graph files ----------------------------- */
/* map[i] = new label of vertex i: the rank of (degree(i), i) in descending
 * order (SparseMatrix.hpp:737-749, degree_order_map), computed on the
 * device; map has n_rows entries (host, or device with SSTG_DEVICE_PTRS). */
int sstg_degree_order_map(sstg_graph *g, uint32_t *map, uint32_t flags);
/* A new graph holding (map[i], map[j]) for every stored edge (i, j), self-
 * loops kept (SparseMatrix.hpp:751-766, rewrite_graph).  SSTG_E_RANGE if a
 * label is out of range; SSTG_E_INVAL unless n_rows == n_cols (one map
 * relabels both).  Edge weights are not carried (the reference's
 * rewrite re-inserts two-element tuples). */
int sstg_rewrite_graph(sstg_graph *g, const uint32_t *map, uint32_t flags,
                       sstg_graph **out);
/* print_statistics (SparseMatrix.hpp:340-426): the layout-free figures of
 * the reference (elements, rows, columns, bytes, destinations by bits
 * needed) and the HBM store's own. */
typedef struct sstg_stats {
  uint64_t edges;          /* M() */
  uint32_t rows, cols;
  uint64_t memory_bytes;   /* get_memory_size() */
  uint64_t empty_rows;
  uint64_t max_degree;
  uint64_t slot_capacity;  /* sum of the rows' slot capacities (elements) */
  uint64_t pool_capacity;  /* elements allocated for all slots */
  uint64_t pool_used;      /* elements below the pool tail */
  uint64_t pool_garbage;   /* elements in abandoned slots */
  uint64_t rebuilds;       /* whole-pool rebuilds so far */
  uint64_t dest_bits[33];  /* stored edges whose destination needs x bits */
} sstg_stats;
int sstg_statistics(sstg_graph *g, sstg_stats *out);

/* ---- degree thresholds and their autotuner ------------------------------
 * No reference counterpart (the reference's tinyset thresholds are fixed
 * by its PMA layout, TinySet_small.hpp:216-293).  The vertex-pull kernels
 * give a vertex 1 lane if its degree is <= t[0], 8 lanes (4 in the BFS steps
 * and the batched store's BFS) if <= t[1], 16 lanes (a wave in TC's
 * orientation) if
 * <= t[2], else a 256-thread block (defaults 8, 64, 256 for the reductions
 * and 32, 128, 2048 for BFS; t strictly increasing).  Two sets per graph: the reductions (PageRank, BC, TC's
 * orientation, the batched store's BFS levels) and BFS (bottom-up steps).  Mutating
 * calls. */
int sstg_set_degree_thresholds(sstg_graph *g, const uint32_t t[3]);
int sstg_get_degree_thresholds(sstg_graph *g, uint32_t t[3]);
int sstg_set_bfs_degree_thresholds(sstg_graph *g, const uint32_t t[3]);
int sstg_get_bfs_degree_thresholds(sstg_graph *g, uint32_t t[3]);
/* BFS direction switch (Beamer): top-down -> bottom-up when the frontier
 * grows and its edges exceed unexplored / alpha, back when it holds fewer
 * than n / beta vertices (defaults 15 and 18; both > 0).  Per graph. */
int sstg_set_bfs_switch(sstg_graph *g, double alpha, double beta);
int sstg_get_bfs_switch(sstg_graph *g, double *alpha, double *beta);
/* Column segments of FP64 PageRank (no reference counterpart; the pull
 * itself is PR_S<double>, TST:527).  With bits in [8, 30] (default 0: off),
 * sstg_pagerank_f64 on a graph of more than 2^bits vertices cuts the
 * column ids into segments of 2^bits vertices and each row into one run per
 * segment it touches, sweeps the runs segment by segment (so the gathered
 * contributions of a sweep fit one XCD's L2) and then sums each vertex's
 * runs in segment order.  The layout is built on the first call after a
 * change of the graph and kept until the next change (about 4 bytes per
 * edge plus 16 bytes per run of device memory).  0 turns segments off (the
 * degree-binned vertex pull).  Results agree with the vertex pull within
 * FP64 rounding of the different summation order.  On rMat graphs the rows
 * split into about six runs each at 2^18-vertex segments and the per-run
 * traffic outweighs the cached gathers, so the vertex pull stays the
 * default.  Mutating calls. */
int sstg_set_pagerank_segments(sstg_graph *g, uint32_t bits);
int sstg_get_pagerank_segments(sstg_graph *g, uint32_t *bits);
#define SSTG_TUNE_PR 1u  /* objective: pr_iters FP64 PageRank iterations */
#define SSTG_TUNE_BFS 2u /* objective: one BFS from src */
typedef struct sstg_tune_report {
  uint32_t thresholds[3];           /* reductions set, chosen (now set) */
  uint32_t start_thresholds[3];     /* reductions set before tuning */
  uint32_t bfs_thresholds[3];       /* BFS set, chosen */
  uint32_t start_bfs_thresholds[3]; /* BFS set before tuning */
  double alpha, beta;               /* BFS switch, chosen */
  double start_alpha, start_beta;
  double pr_start_ms, pr_best_ms;   /* PageRank objective (0 if not tuned) */
  double bfs_start_ms, bfs_best_ms; /* BFS objective (0 if not tuned) */
  uint32_t evaluations;             /* settings timed */
} sstg_tune_report;
/* With SSTG_TUNE_PR, times pr_iters PageRank iterations (0 means 2) over a
 * grid of the reductions' thresholds; with SSTG_TUNE_BFS, one BFS from src
 * over a grid of the BFS thresholds and switch; coordinate descent each, and
 * the fastest settings are left on the graph.  Results of the algorithms are
 * unchanged except for the summation order of PageRank and the choice among
 * valid BFS parents. */
int sstg_autotune(sstg_graph *g, uint32_t what, uint32_t src,
                  uint32_t pr_iters, sstg_tune_report *out);
/* get_edges_from_file (helpers.hpp:604-631): .adj / .adj.shuf
 * (AdjacencyGraph or WeightedAdjacencyGraph; helpers.hpp:533-601) or .mtx
 * (MatrixMarket coordinate, 1-indexed, every entry returned in both
 * directions, square; helpers.hpp:382-531).  *src, *dst and, when weights
 * is not NULL, *weights (the file's integer values, 1 for unweighted files)
 * are malloc'd: free them with sstg_host_free.  Needs no device.
 * SSTG_E_IO if the file cannot be read, SSTG_E_INVAL if it is malformed. */
int sstg_edges_from_file(const char *path, uint32_t *n, uint64_t *m,
                         uint32_t **src, uint32_t **dst, uint32_t **weights);
void sstg_host_free(void *p);
/* Reads a graph file into a new n x n graph (the harness's load step,
 * test.hpp:414-449; with SSTG_LOAD_WEIGHTS the weighted load of
 * real_graph_weights, test.hpp:699-720: the file's weights, 1 for an
 * unweighted file). */
#define SSTG_LOAD_WEIGHTS 0x8u
int sstg_load_file(const char *path, uint32_t flags, sstg_graph **out);
/* Writes the graph as an AdjacencyGraph file (graph_utilities.cpp:151-183,
 * write_unweighted_graph). */
int sstg_write_adj(sstg_graph *g, const char *path);

/* Seeded Erdos-Renyi edge stream on n vertices (no reference counterpart:
 * the reference's ER tool, graph_utilities.cpp:168-212, is unseeded):
 * edge i = (hash64(seed << 32 | 2i) mod n, hash64(seed << 32 | 2i+1) mod n)
 * with rmat_util.h's hash64. */
int sstg_er(uint32_t n, uint32_t seed, uint64_t first, uint64_t count,
            uint32_t *src, uint32_t *dst, uint32_t flags);

/* ---- device memory plumbing for callers without a HIP runtime of their
 * own (tests, benchmarks, FFI hosts) ------------------------------------ */
int sstg_device_alloc(uint64_t bytes, void **out);
int sstg_device_free(void *p);
/* The copies return once the data is in place for work on any stream
 * (a graph's own stream does not wait for the null stream). */
int sstg_memcpy_h2d(void *dst, const void *src, uint64_t bytes);
int sstg_memcpy_d2h(void *dst, const void *src, uint64_t bytes);
int sstg_memcpy_d2d(void *dst, const void *src, uint64_t bytes);
int sstg_device_synchronize(void);
/* Free/total device memory in bytes (hipMemGetInfo). */
int sstg_mem_info(uint64_t *free_bytes, uint64_t *total_bytes);

/* ---- kernel timing (HIP events on the graph's stream) ------------------ */
/* When enabled, every kernel launch of the library is bracketed by HIP
 * events on the stream it is launched on. */
int sstg_profile_enable(int on);
int sstg_profile_reset(void);
/* Accumulated time (ms) and launch count of kernels whose name contains
 * `name`; forces pending events to complete. */
int sstg_profile_query(const char *name, double *total_ms, uint64_t *launches);
/* Writes "name launches total_ms" lines into buf (NUL terminated). */
int sstg_profile_dump(char *buf, uint64_t buf_bytes);

#ifdef __cplusplus
}
#endif
#endif /* SSTG_H */
