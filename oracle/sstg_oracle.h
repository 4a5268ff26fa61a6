/*
 * sstg_oracle.h -- CPU restatement of SSTGraph's hot path, used ONLY as the
 * checker (parity tests, smoke check) and as the timed CPU baseline.
 *
 * This is test infrastructure, not product code: nothing in sstgraph_amd/
 * links, imports or calls it.  Each function cites the reference file:line
 * whose behaviour it restates (paths relative to the SSTGraph tree).
 *
 * Pinning status of each function (what checks it independently of its own
 * formula) is listed in oracle/README.md.
 */
#ifndef SSTG_ORACLE_H
#define SSTG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- hashes and the RMAT generator: include/SSTGraph/internal/rmat_util.h */
uint64_t orc_hash64(uint64_t u);                 /* rmat_util.h:29-39 */
uint32_t orc_hash32(uint32_t a);                 /* rmat_util.h:42-50 */
uint64_t orc_aspen_ith_rand(uint64_t state, uint64_t i); /* rmat_util.h:62 */
uint32_t orc_log2_up(uint64_t i);                /* rmat_util.h:72-80 */
/* rMat<uint32_t>(n, seed, a, b, c)(i) for i in [first, first+count).
 * Returns 0, or -1 when n is not a power of two or a+b+c > 1
 * (rmat_util.h:92-99 abort()s there). */
/* The same recursion with r = hash32 / 2^32 (quadrant probabilities as
 * given); no reference counterpart. */
int orc_rmat_normalized(uint32_t n, uint32_t seed, double a, double b,
                        double c, uint64_t first, uint64_t count,
                        uint32_t *src, uint32_t *dst);
int orc_rmat(uint32_t n, uint32_t seed, double a, double b, double c,
             uint64_t first, uint64_t count, uint32_t *src, uint32_t *dst);

/* Seeded Erdos-Renyi edge stream (own definition, see sstg_oracle.c). */
int orc_er(uint32_t n, uint32_t seed, uint64_t first, uint64_t count,
           uint32_t *src, uint32_t *dst);

/* ---- the tinyset-semantics graph store (SparseMatrixV<true>) ---------- */
typedef struct orc_graph orc_graph;
orc_graph *orc_graph_create(uint32_t n_rows, uint32_t n_cols); /* SM:73-87 */
void orc_graph_destroy(orc_graph *g);                          /* SM:141-146 */
/* SparseMatrix.hpp:556-643: contents become old U {(x,y) in batch : x != y}.
 * Returns 0, or -2 if an id is out of range (nothing is changed then). */
int orc_insert_batch(orc_graph *g, const uint32_t *src, const uint32_t *dst,
                     uint64_t n);
/* SparseMatrix.hpp:645-718: contents become old \ batch (no self-loop skip). */
int orc_remove_batch(orc_graph *g, const uint32_t *src, const uint32_t *dst,
                     uint64_t n);
/* SparseMatrix.hpp:536-554: single-edge insert/remove (insert keeps x==y). */
int orc_insert(orc_graph *g, uint32_t row, uint32_t col);
int orc_remove(orc_graph *g, uint32_t row, uint32_t col);
int orc_has(const orc_graph *g, uint32_t row, uint32_t col); /* SM:720-726 */
uint64_t orc_num_edges(const orc_graph *g);                   /* SM:518-525 */
uint32_t orc_num_nodes(const orc_graph *g);                   /* SM:56 */
void orc_degrees(const orc_graph *g, uint32_t *out);          /* SM:57-59 */
/* Rows in ascending order (map order, TinySet_small.hpp:295-353). */
void orc_export_csr(const orc_graph *g, uint64_t *offsets, uint32_t *cols);
/* Bulk load from a sorted, duplicate-free CSR (used to seed the CPU baseline
 * from a graph built elsewhere). */
int orc_load_csr(orc_graph *g, const uint64_t *offsets, const uint32_t *cols);
uint64_t orc_touch_all_sum(const orc_graph *g);               /* SM:436-454 */

/* ---- algorithms over a CSR snapshot (offsets[n+1], cols[m]) ----------- */
/* BFS: level[v] = hop distance from src, UINT32_MAX if unreached
 * (test.hpp:497-509 derives this from the parent array); parent[src]=src,
 * parent[v] = -1 if unreached (test.hpp:488-492). Either output may be NULL. */
void orc_bfs(uint32_t n, const uint64_t *off, const uint32_t *cols,
             uint32_t src, uint32_t *level, int32_t *parent);
/* PageRank, FP64, fixed iteration count (test.hpp:527 calls PR_S<double>). */
void orc_pagerank(uint32_t n, const uint64_t *off, const uint32_t *cols,
                  uint32_t iters, double damping, double *out);
/* Single-source Brandes dependency delta_s(v), FP64 (test.hpp:547). */
void orc_bc(uint32_t n, const uint64_t *off, const uint32_t *cols,
            uint32_t src, double *out);
/* All-core CPU baselines: the same levels / dependencies with OpenMP.
 * symmetric: every edge has its reverse (< 0: checked first). */
void orc_bfs_par(uint32_t n, const uint64_t *off, const uint32_t *cols,
                 uint32_t src, int symmetric, uint32_t *level);
void orc_bc_par(uint32_t n, const uint64_t *off, const uint32_t *cols,
                uint32_t src, int symmetric, double *out);
/* Connected components, label = minimum vertex id of the component
 * (test.hpp:571-601 only uses the partition). */
void orc_cc(uint32_t n, const uint64_t *off, const uint32_t *cols,
            uint32_t *label);
/* |N(u) n N(v)|, with early_end: only w < u and w < v
 * (SparseMatrix.hpp:61-71, TinySet_small.hpp:357-407). */
uint64_t orc_common_neighbors(const uint64_t *off, const uint32_t *cols,
                              uint32_t u, uint32_t v, int early_end);
/* Triangle count, each triangle once (w < v < u) (test.hpp:606). */
uint64_t orc_tc(uint32_t n, const uint64_t *off, const uint32_t *cols);
/* Bellman-Ford hop-weighted / weighted shortest distances (test.hpp:618).
 * weights may be NULL (unit weights).  dist[v] = INT32_MAX if unreached. */
void orc_bf(uint32_t n, const uint64_t *off, const uint32_t *cols,
            const int32_t *weights, uint32_t src, int32_t *dist);

/* Set OpenMP thread count used by the timed paths (0 = library default). */
void orc_set_threads(int threads);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
