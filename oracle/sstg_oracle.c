/*
 * sstg_oracle.c -- CPU restatement of SSTGraph's hot path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  Used by tests/ as the parity checker, by
 * __graft_entry__.smoke() as the checker of one small device run, and by
 * bench.py's cpu_baseline leg as the timed CPU baseline.  The product
 * (sstgraph_amd/, include/) never links or calls this file.
 *
 * Reference citations are relative to the SSTGraph tree:
 *   SM  = include/SSTGraph/SparseMatrix.hpp
 *   TSS = include/SSTGraph/TinySet_small.hpp
 *   PMA = include/SSTGraph/PMA.hpp
 *   TST = include/SSTGraph/internal/test.hpp
 *   RMAT= include/SSTGraph/internal/rmat_util.h
 *
 * Built without fast-math (strict IEEE FP64), with OpenMP for the timed paths.
 */
#include "sstg_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 0;
void orc_set_threads(int threads) {
  g_threads = threads;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
}
int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ===================================================================== */
/* Hashes and RMAT                                                        */
/* ===================================================================== */

/* rmat_util.h:29-39 */
uint64_t orc_hash64(uint64_t u) {
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

/* rmat_util.h:42-50 */
uint32_t orc_hash32(uint32_t a) {
  a = (a + 0x7ed55d16u) + (a << 12u);
  a = (a ^ 0xc761c23cu) ^ (a >> 19u);
  a = (a + 0x165667b1u) + (a << 5u);
  a = (a + 0xd3a2646cu) ^ (a << 9u);
  a = (a + 0xfd7046c5u) + (a << 3u);
  a = (a ^ 0xb55a4f09u) ^ (a >> 16u);
  return a;
}

/* random_aspen::ith_rand, rmat_util.h:62 */
uint64_t orc_aspen_ith_rand(uint64_t state, uint64_t i) {
  return orc_hash64(i + state);
}

/* rmat_util.h:72-80 */
uint32_t orc_log2_up(uint64_t i) {
  uint32_t a = 0;
  uint64_t b = i - 1;
  while (i > 0 && b > 0) {
    b >>= 1u;
    a++;
  }
  return a;
}

/* rmat_util.h:102-105: hash32(i) / INT32_MAX in double, range [0, 2]. */
static double rmat_hash_double(uint32_t i) {
  return (double)orc_hash32(i) / (double)INT32_MAX;
}

/* The normalised draw r = hash32(i) / 2^32, in [0, 1): quadrant
 * probabilities are a, b, c, 1-a-b-c as stated, not halved.  No reference
 * counterpart; the measurement plan asks for it beside the reference draw. */
static double rmat_hash_unit(uint32_t i) {
  return (double)orc_hash32(i) / 4294967296.0;
}

/* rmat_util.h:82-131.  The recursion rMatRec(nn, start, stride) decides the
 * bit nn/2 with r = hashDouble(start) after recursing on (nn/2, start+stride);
 * unrolled here from the top level (k = 0, bit n/2) downwards. */
static int rmat_impl(uint32_t n, uint32_t seed, double a, double b, double c,
                     uint64_t first, uint64_t count, int normalized,
                     uint32_t *src, uint32_t *dst) {
  double ab = a + b, abc = a + b + c;
  if (abc > 1) return -1;
  if (n == 0 || (n & (n - 1)) != 0) return -1;
  uint32_t h = orc_hash32(seed);
  uint32_t lg = orc_log2_up(n);
#pragma omp parallel for schedule(static) if (count > 100000)
  for (int64_t t = 0; t < (int64_t)count; t++) {
    uint32_t i = (uint32_t)(first + (uint64_t)t);
    uint32_t start = orc_hash32((uint32_t)(2u * i) * h);
    uint32_t stride = orc_hash32((uint32_t)(2u * i + 1u) * h);
    uint32_t r_out = 0, c_out = 0;
    uint32_t nn = n;
    for (uint32_t k = 0; k < lg; k++) {
      double r = normalized ? rmat_hash_unit(start + k * stride)
                            : rmat_hash_double(start + k * stride);
      uint32_t half = nn >> 1;
      if (r < a) {
      } else if (r < ab) {
        c_out += half;
      } else if (r < abc) {
        r_out += half;
      } else {
        r_out += half;
        c_out += half;
      }
      nn = half;
    }
    src[t] = r_out;
    dst[t] = c_out;
  }
  return 0;
}

int orc_rmat(uint32_t n, uint32_t seed, double a, double b, double c,
             uint64_t first, uint64_t count, uint32_t *src, uint32_t *dst) {
  return rmat_impl(n, seed, a, b, c, first, count, 0, src, dst);
}

int orc_rmat_normalized(uint32_t n, uint32_t seed, double a, double b,
                        double c, uint64_t first, uint64_t count,
                        uint32_t *src, uint32_t *dst) {
  return rmat_impl(n, seed, a, b, c, first, count, 1, src, dst);
}

/* Seeded Erdos-Renyi edge stream (no reference counterpart; the reference's
 * graph_utilities.cpp ER tool is unseeded): edge i = (hash64(seed<<32|2i) mod
 * n, hash64(seed<<32|2i+1) mod n). */
int orc_er(uint32_t n, uint32_t seed, uint64_t first, uint64_t count,
           uint32_t *src, uint32_t *dst) {
  if (n == 0) return -1;
  uint64_t base = (uint64_t)seed << 32;
#pragma omp parallel for schedule(static) if (count > 100000)
  for (int64_t t = 0; t < (int64_t)count; t++) {
    uint64_t i = first + (uint64_t)t;
    src[t] = (uint32_t)(orc_hash64(base + 2 * i) % n);
    dst[t] = (uint32_t)(orc_hash64(base + 2 * i + 1) % n);
  }
  return 0;
}

/* ===================================================================== */
/* Parallel LSD radix sort of 64-bit keys (row << 32 | col)               */
/* ===================================================================== */

static uint32_t bit_width_u32(uint32_t x) {
  uint32_t w = 0;
  while (x) {
    w++;
    x >>= 1;
  }
  return w;
}

/* Sorts keys by bits [0, col_bits) and [32, 32 + row_bits), 8 bits per pass,
 * stable; the result alternates between a and tmp and is returned. */
static uint64_t *radix_sort_keys(uint64_t *a, uint64_t *tmp, size_t n,
                                 uint32_t col_bits, uint32_t row_bits) {
  int shifts[16];
  int np = 0;
  for (uint32_t s = 0; s < col_bits; s += 8) shifts[np++] = (int)s;
  for (uint32_t s = 0; s < row_bits; s += 8) shifts[np++] = (int)(32 + s);
  int T = 1;
#ifdef _OPENMP
  if (n > 65536) T = omp_get_max_threads();
#endif
  size_t *hist = (size_t *)calloc((size_t)T * 256, sizeof(size_t));
  for (int p = 0; p < np; p++) {
    int sh = shifts[p];
    memset(hist, 0, (size_t)T * 256 * sizeof(size_t));
#pragma omp parallel num_threads(T) if (T > 1)
    {
      int t = 0;
#ifdef _OPENMP
      t = omp_get_thread_num();
#endif
      size_t lo = n * (size_t)t / (size_t)T, hi = n * (size_t)(t + 1) / (size_t)T;
      size_t *h = hist + (size_t)t * 256;
      for (size_t i = lo; i < hi; i++) h[(a[i] >> sh) & 255]++;
#pragma omp barrier
#pragma omp single
      {
        size_t run = 0;
        for (int d = 0; d < 256; d++)
          for (int tt = 0; tt < T; tt++) {
            size_t c = hist[(size_t)tt * 256 + d];
            hist[(size_t)tt * 256 + d] = run;
            run += c;
          }
      }
      for (size_t i = lo; i < hi; i++) tmp[h[(a[i] >> sh) & 255]++] = a[i];
    }
    uint64_t *sw = a;
    a = tmp;
    tmp = sw;
  }
  free(hist);
  return a;
}

/* ===================================================================== */
/* Graph store: one sorted, duplicate-free uint32 row per vertex.          */
/* The reference's row is a tinyset (TSS:29-200) whose observable contents  */
/* are exactly this sorted set (map order is ascending, TSS:295-353).       */
/* ===================================================================== */

typedef struct {
  uint32_t *data;
  uint32_t len, cap;
} orc_row;

struct orc_graph {
  uint32_t n_rows, n_cols;
  orc_row *rows;
};

orc_graph *orc_graph_create(uint32_t n_rows, uint32_t n_cols) {
  orc_graph *g = (orc_graph *)calloc(1, sizeof(orc_graph));
  if (!g) return NULL;
  g->n_rows = n_rows;
  g->n_cols = n_cols;
  g->rows = (orc_row *)calloc(n_rows ? n_rows : 1, sizeof(orc_row));
  if (!g->rows) {
    free(g);
    return NULL;
  }
  return g;
}

void orc_graph_destroy(orc_graph *g) {
  if (!g) return;
  for (uint32_t i = 0; i < g->n_rows; i++) free(g->rows[i].data);
  free(g->rows);
  free(g);
}

static int row_reserve(orc_row *r, uint32_t need) {
  if (need <= r->cap) return 0;
  uint32_t nc = r->cap ? r->cap : 4;
  while (nc < need) nc = (nc > 0x7FFFFFFFu) ? need : nc * 2;
  uint32_t *p = (uint32_t *)realloc(r->data, (size_t)nc * sizeof(uint32_t));
  if (!p) return -3;
  r->data = p;
  r->cap = nc;
  return 0;
}

static uint32_t lower_bound_u32(const uint32_t *a, uint32_t len, uint32_t x) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

/* Merge the sorted, unique run b[0..k) into row r (set union).
 * Follows the semantics of repeated TSS:1036-1071 / PMA:1085-1183 inserts:
 * an element already present is a no-op (PMA:1139-1154). */
static int row_merge_insert(orc_row *r, const uint32_t *b, uint32_t k) {
  /* count new elements */
  uint32_t i = 0, j = 0, fresh = 0;
  while (j < k) {
    if (i < r->len && r->data[i] < b[j]) {
      i++;
    } else if (i < r->len && r->data[i] == b[j]) {
      i++;
      j++;
    } else {
      fresh++;
      j++;
    }
  }
  if (fresh == 0) return 0;
  if (row_reserve(r, r->len + fresh)) return -3;
  /* merge from the back, in place */
  int64_t ii = (int64_t)r->len - 1, jj = (int64_t)k - 1;
  int64_t w = (int64_t)r->len + fresh - 1;
  while (jj >= 0) {
    if (ii >= 0 && r->data[ii] > b[jj]) {
      r->data[w--] = r->data[ii--];
    } else if (ii >= 0 && r->data[ii] == b[jj]) {
      r->data[w--] = r->data[ii--];
      jj--;
    } else {
      r->data[w--] = b[jj--];
    }
  }
  r->len += fresh;
  return 0;
}

/* Remove the sorted, unique run b[0..k) from row r (set difference);
 * absent elements are no-ops (PMA:1226-1228). */
static void row_merge_remove(orc_row *r, const uint32_t *b, uint32_t k) {
  uint32_t i = 0, j = 0, w = 0;
  while (i < r->len) {
    while (j < k && b[j] < r->data[i]) j++;
    if (j < k && b[j] == r->data[i]) {
      i++;
      j++;
      continue;
    }
    r->data[w++] = r->data[i++];
  }
  r->len = w;
}

/* Shared body of the batch updates: validate, sort by (row, col), drop
 * duplicates (and self-loops when skip_self), then merge row segments in
 * parallel.  SM:556-643 sorts by row (SM:577-579), partitions at row
 * boundaries (SM:582-592) and inserts per edge in parallel over partitions
 * (SM:593-607); per-row sorted merging yields the same set. */
static int batch_update(orc_graph *g, const uint32_t *src, const uint32_t *dst,
                        uint64_t n, int insert, int skip_self) {
  for (uint64_t i = 0; i < n; i++)
    if (src[i] >= g->n_rows || dst[i] >= g->n_cols) return -2;
  if (n == 0) return 0;
  uint64_t *keys = (uint64_t *)malloc(n * sizeof(uint64_t));
  uint64_t *tmp = (uint64_t *)malloc(n * sizeof(uint64_t));
  if (!keys || !tmp) {
    free(keys);
    free(tmp);
    return -3;
  }
#pragma omp parallel for schedule(static) if (n > 100000)
  for (int64_t i = 0; i < (int64_t)n; i++)
    keys[i] = ((uint64_t)src[i] << 32) | dst[i];
  uint32_t cb = bit_width_u32(g->n_cols ? g->n_cols - 1 : 0);
  uint32_t rb = bit_width_u32(g->n_rows ? g->n_rows - 1 : 0);
  uint64_t *s = radix_sort_keys(keys, tmp, n, cb, rb);
  /* unique + self-loop filter, in place in s */
  uint64_t u = 0;
  for (uint64_t i = 0; i < n; i++) {
    uint64_t k = s[i];
    if (skip_self && (uint32_t)(k >> 32) == (uint32_t)k) continue;
    if (u > 0 && s[u - 1] == k) continue;
    s[u++] = k;
  }
  /* segment heads */
  uint64_t nseg = 0;
  for (uint64_t i = 0; i < u; i++)
    if (i == 0 || (s[i] >> 32) != (s[i - 1] >> 32)) nseg++;
  uint64_t *seg = (uint64_t *)malloc((nseg + 1) * sizeof(uint64_t));
  uint32_t *cols = (uint32_t *)malloc((u ? u : 1) * sizeof(uint32_t));
  if (!seg || !cols) {
    free(keys);
    free(tmp);
    free(seg);
    free(cols);
    return -3;
  }
  nseg = 0;
  for (uint64_t i = 0; i < u; i++) {
    if (i == 0 || (s[i] >> 32) != (s[i - 1] >> 32)) seg[nseg++] = i;
    cols[i] = (uint32_t)s[i];
  }
  seg[nseg] = u;
  int err = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : err) if (nseg > 64)
  for (int64_t t = 0; t < (int64_t)nseg; t++) {
    uint64_t b = seg[t], e = seg[t + 1];
    uint32_t row = (uint32_t)(s[b] >> 32);
    if (insert)
      err |= row_merge_insert(&g->rows[row], cols + b, (uint32_t)(e - b)) ? 1 : 0;
    else
      row_merge_remove(&g->rows[row], cols + b, (uint32_t)(e - b));
  }
  free(seg);
  free(cols);
  free(keys);
  free(tmp);
  return err ? -3 : 0;
}

int orc_insert_batch(orc_graph *g, const uint32_t *src, const uint32_t *dst,
                     uint64_t n) {
  return batch_update(g, src, dst, n, 1, 1); /* SM:570-572, 602-604 */
}

int orc_remove_batch(orc_graph *g, const uint32_t *src, const uint32_t *dst,
                     uint64_t n) {
  return batch_update(g, src, dst, n, 0, 0); /* SM:655-686: no x==y skip */
}

/* SM:536-544: single insert goes straight to the row, no self-loop skip. */
int orc_insert(orc_graph *g, uint32_t row, uint32_t col) {
  if (row >= g->n_rows || col >= g->n_cols) return -2;
  return row_merge_insert(&g->rows[row], &col, 1);
}

int orc_remove(orc_graph *g, uint32_t row, uint32_t col) {
  if (row >= g->n_rows || col >= g->n_cols) return -2;
  row_merge_remove(&g->rows[row], &col, 1);
  return 0;
}

int orc_has(const orc_graph *g, uint32_t row, uint32_t col) {
  if (row >= g->n_rows) return 0;
  const orc_row *r = &g->rows[row];
  uint32_t p = lower_bound_u32(r->data, r->len, col);
  return p < r->len && r->data[p] == col;
}

uint64_t orc_num_edges(const orc_graph *g) {
  uint64_t m = 0;
  for (uint32_t i = 0; i < g->n_rows; i++) m += g->rows[i].len;
  return m;
}

uint32_t orc_num_nodes(const orc_graph *g) { return g->n_rows; }

void orc_degrees(const orc_graph *g, uint32_t *out) {
  for (uint32_t i = 0; i < g->n_rows; i++) out[i] = g->rows[i].len;
}

void orc_export_csr(const orc_graph *g, uint64_t *offsets, uint32_t *cols) {
  uint64_t o = 0;
  for (uint32_t i = 0; i < g->n_rows; i++) {
    offsets[i] = o;
    o += g->rows[i].len;
  }
  offsets[g->n_rows] = o;
  if (!cols) return;
#pragma omp parallel for schedule(dynamic, 1024) if (g->n_rows > 4096)
  for (int64_t i = 0; i < (int64_t)g->n_rows; i++)
    if (g->rows[i].len)
      memcpy(cols + offsets[i], g->rows[i].data,
             (size_t)g->rows[i].len * sizeof(uint32_t));
}

int orc_load_csr(orc_graph *g, const uint64_t *offsets, const uint32_t *cols) {
  int err = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(| : err)
  for (int64_t i = 0; i < (int64_t)g->n_rows; i++) {
    orc_row *r = &g->rows[i];
    uint32_t len = (uint32_t)(offsets[i + 1] - offsets[i]);
    r->len = 0;
    if (len == 0) continue;
    if (row_reserve(r, len)) {
      err = 1;
      continue;
    }
    memcpy(r->data, cols + offsets[i], (size_t)len * sizeof(uint32_t));
    r->len = len;
  }
  return err ? -3 : 0;
}

/* SM:436-454 */
uint64_t orc_touch_all_sum(const orc_graph *g) {
  uint64_t sum = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : sum)
  for (int64_t i = 0; i < (int64_t)g->n_rows; i++)
    for (uint32_t j = 0; j < g->rows[i].len; j++) sum += g->rows[i].data[j];
  return sum;
}

/* ===================================================================== */
/* Algorithms on a CSR snapshot                                            */
/* ===================================================================== */

/* BFS.  Parent and depth conventions: TST:486-509 (parent[src] == src,
 * -1 = unreached, depth = UINT32_MAX for unreached). */
void orc_bfs(uint32_t n, const uint64_t *off, const uint32_t *cols,
             uint32_t src, uint32_t *level, int32_t *parent) {
  uint32_t *lv = level ? level : (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
  for (uint32_t i = 0; i < n; i++) lv[i] = UINT32_MAX;
  if (parent)
    for (uint32_t i = 0; i < n; i++) parent[i] = -1;
  if (src < n) {
    uint32_t *q = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    uint32_t qh = 0, qt = 0;
    q[qt++] = src;
    lv[src] = 0;
    if (parent) parent[src] = (int32_t)src;
    while (qh < qt) {
      uint32_t u = q[qh++];
      for (uint64_t e = off[u]; e < off[u + 1]; e++) {
        uint32_t w = cols[e];
        if (lv[w] == UINT32_MAX) {
          lv[w] = lv[u] + 1;
          if (parent) parent[w] = (int32_t)u;
          q[qt++] = w;
        }
      }
    }
    free(q);
  }
  if (!level) free(lv);
}

/* PageRank in FP64 for a fixed number of iterations (TST:527 PR_S<double>).
 * x0 = 1/n;  x'(v) = damping * sum_{u in N(v)} x(u)/deg(u) + (1-damping)/n.
 * Zero-degree vertices contribute nothing.  The sum for each vertex is taken
 * in ascending neighbour order. */
void orc_pagerank(uint32_t n, const uint64_t *off, const uint32_t *cols,
                  uint32_t iters, double damping, double *out) {
  if (n == 0) return;
  double *contrib = (double *)malloc((size_t)n * sizeof(double));
  double base = (1.0 - damping) / (double)n;
  double x0 = 1.0 / (double)n;
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < (int64_t)n; v++) out[v] = x0;
  for (uint32_t it = 0; it < iters; it++) {
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < (int64_t)n; v++) {
      uint64_t d = off[v + 1] - off[v];
      contrib[v] = d ? out[v] / (double)d : 0.0;
    }
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t v = 0; v < (int64_t)n; v++) {
      double s = 0.0;
      for (uint64_t e = off[v]; e < off[v + 1]; e++) s += contrib[cols[e]];
      out[v] = damping * s + base;
    }
  }
  free(contrib);
}

/* Single-source Brandes dependencies (TST:547 BC(g, src)).
 * sigma(s) = 1; sigma(w) = sum of sigma(v) over parents v (level(v) =
 * level(w) - 1); delta(v) = sum over children w of
 * (sigma(v) / sigma(w)) * (1 + delta(w)).  delta(s) is emitted (it equals
 * |R(s)| - 1); unreached vertices get 0. */
void orc_bc(uint32_t n, const uint64_t *off, const uint32_t *cols,
            uint32_t src, double *out) {
  for (uint32_t i = 0; i < n; i++) out[i] = 0.0;
  if (src >= n) return;
  uint32_t *lv = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
  uint32_t *q = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
  double *sigma = (double *)calloc(n, sizeof(double));
  for (uint32_t i = 0; i < n; i++) lv[i] = UINT32_MAX;
  uint32_t qh = 0, qt = 0;
  q[qt++] = src;
  lv[src] = 0;
  while (qh < qt) {
    uint32_t u = q[qh++];
    for (uint64_t e = off[u]; e < off[u + 1]; e++) {
      uint32_t w = cols[e];
      if (lv[w] == UINT32_MAX) {
        lv[w] = lv[u] + 1;
        q[qt++] = w;
      }
    }
  }
  /* sigma pushed along out-edges in BFS order (parents of w are its
   * in-neighbours one level up) */
  sigma[src] = 1.0;
  for (uint32_t i = 0; i < qt; i++) {
    uint32_t v = q[i];
    for (uint64_t e = off[v]; e < off[v + 1]; e++) {
      uint32_t w = cols[e];
      if (lv[w] == lv[v] + 1) sigma[w] += sigma[v];
    }
  }
  /* dependencies by pulling from children, in reverse BFS order */
  for (int64_t i = (int64_t)qt - 1; i >= 0; i--) {
    uint32_t v = q[i];
    double d = 0.0;
    for (uint64_t e = off[v]; e < off[v + 1]; e++) {
      uint32_t w = cols[e];
      if (lv[w] != UINT32_MAX && lv[w] == lv[v] + 1)
        d += sigma[v] / sigma[w] * (1.0 + out[w]);
    }
    out[v] = d;
  }
  free(lv);
  free(q);
  free(sigma);
}

/* ----- all-core CPU baselines of BFS and BC (timed by bench.py) -------
 * The serial functions above are the parity checker; these compute the same
 * objects with OpenMP so the baseline uses every host core, as the
 * reference's fork-join build would (its BFS and BC come from Ligra-style
 * EdgeMapVertexMap, TST:486 and TST:547).  Level-synchronous; a level is a
 * top-down step over the frontier (claims by compare-and-swap) or, on
 * symmetric input, a bottom-up step in which every unvisited vertex looks
 * for a neighbour in the frontier, taken with Beamer's switch (a growing
 * frontier whose edges exceed the unexplored edges / 15; back to top-down
 * below n / 18 frontier vertices).  Levels are unique, so they equal
 * orc_bfs's. */
static uint32_t bfs_levels_par(uint32_t n, const uint64_t *off,
                               const uint32_t *cols, uint32_t src, int sym,
                               uint32_t *lv, uint32_t *order) {
  /* a top-down step after a bottom-up one reads the frontier from order[]
   * (the bottom-up step lists it there in ascending id order) */
  for (int64_t i = 0; i < (int64_t)n; i++) lv[i] = UINT32_MAX;
  if (src >= n) return 0;
  lv[src] = 0;
  order[0] = src;
  uint64_t head = 0, tail = 1, m = off[n], explored = off[src + 1] - off[src];
  uint64_t prev_nf = 0;
  int in_bu = 0;
  uint32_t L = 0;
  int T = 1;
#ifdef _OPENMP
  T = omp_get_max_threads();
#endif
  uint64_t *cnt = (uint64_t *)calloc((size_t)T + 1, sizeof(uint64_t));
  while (tail > head) {
    uint64_t fe = 0;
#pragma omp parallel for reduction(+ : fe) schedule(static)
    for (int64_t i = (int64_t)head; i < (int64_t)tail; i++)
      fe += off[order[i] + 1] - off[order[i]];
    const uint64_t unexplored = m > explored ? m - explored : 0;
    const uint64_t nf = tail - head;
    /* Beamer's switch: bottom-up once a growing frontier's edges exceed the
     * unexplored edges / 15, back to top-down below n / 18 vertices */
    const int bottom_up =
        sym && (in_bu ? (double)nf * 18.0 >= (double)n
                      : nf > prev_nf && (double)fe * 15.0 > (double)unexplored);
    in_bu = bottom_up;
    prev_nf = nf;
    if (bottom_up) {
#pragma omp parallel for schedule(dynamic, 1024)
      for (int64_t v = 0; v < (int64_t)n; v++) {
        if (lv[v] != UINT32_MAX) continue;
        for (uint64_t e = off[v]; e < off[v + 1]; e++)
          if (lv[cols[e]] == L) {
            lv[v] = L + 1; /* read by no one at level L + 1 in this step */
            break;
          }
      }
      /* the new frontier in ascending id order, per-thread counts first */
#pragma omp parallel num_threads(T)
      {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        uint64_t lo = (uint64_t)n * t / T, hi = (uint64_t)n * (t + 1) / T, c = 0;
        for (uint64_t v = lo; v < hi; v++) c += lv[v] == L + 1;
        cnt[t + 1] = c;
#pragma omp barrier
#pragma omp single
        for (int k = 0; k < T; k++) cnt[k + 1] += cnt[k];
        uint64_t w = tail + cnt[t];
        for (uint64_t v = lo; v < hi; v++)
          if (lv[v] == L + 1) order[w++] = (uint32_t)v;
      }
      head = tail;
      tail += cnt[T];
    } else {
      uint64_t next = tail;
#pragma omp parallel
      {
        uint32_t buf[256];
        uint32_t nb = 0;
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = (int64_t)head; i < (int64_t)tail; i++) {
          const uint32_t u = order[i];
          for (uint64_t e = off[u]; e < off[u + 1]; e++) {
            const uint32_t w = cols[e];
            uint32_t want = UINT32_MAX;
            if (lv[w] == UINT32_MAX &&
                __atomic_compare_exchange_n(&lv[w], &want, L + 1, 0,
                                            __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
              buf[nb++] = w;
              if (nb == 256) {
                const uint64_t at = __atomic_fetch_add(&next, nb, __ATOMIC_RELAXED);
                memcpy(order + at, buf, nb * sizeof(uint32_t));
                nb = 0;
              }
            }
          }
        }
        if (nb) {
          const uint64_t at = __atomic_fetch_add(&next, nb, __ATOMIC_RELAXED);
          memcpy(order + at, buf, nb * sizeof(uint32_t));
        }
      }
      head = tail;
      tail = next;
    }
    uint64_t ne = 0;
#pragma omp parallel for reduction(+ : ne) schedule(static)
    for (int64_t i = (int64_t)head; i < (int64_t)tail; i++)
      ne += off[order[i] + 1] - off[order[i]];
    explored += ne;
    L++;
  }
  free(cnt);
  return (uint32_t)tail;
}

static int csr_symmetric(uint32_t n, const uint64_t *off, const uint32_t *cols) {
  int ok = 1;
#pragma omp parallel for schedule(dynamic, 1024) reduction(& : ok)
  for (int64_t v = 0; v < (int64_t)n; v++)
    for (uint64_t e = off[v]; e < off[v + 1] && ok; e++) {
      const uint32_t w = cols[e];
      const uint32_t k = lower_bound_u32(cols + off[w], (uint32_t)(off[w + 1] - off[w]),
                                         (uint32_t)v);
      if (off[w] + k >= off[w + 1] || cols[off[w] + k] != (uint32_t)v) ok = 0;
    }
  return ok;
}

/* symmetric: every edge has its reverse (< 0: check it here, outside the
 * traversal proper) */
void orc_bfs_par(uint32_t n, const uint64_t *off, const uint32_t *cols,
                 uint32_t src, int symmetric, uint32_t *level) {
  if (symmetric < 0) symmetric = csr_symmetric(n, off, cols);
  uint32_t *order = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
  bfs_levels_par(n, off, cols, src, symmetric, level, order);
  free(order);
}

/* Brandes from one source with the same conventions as orc_bc, for input
 * where every edge has its reverse: sigma and delta are pulled level by
 * level (a vertex's parents and children are its neighbours one level up
 * and down), in parallel within a level.  Directed input falls back to
 * orc_bc. */
void orc_bc_par(uint32_t n, const uint64_t *off, const uint32_t *cols,
                uint32_t src, int symmetric, double *out) {
  if (symmetric < 0) symmetric = csr_symmetric(n, off, cols);
  if (!symmetric) {
    orc_bc(n, off, cols, src, out);
    return;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) out[i] = 0.0;
  if (src >= n) return;
  uint32_t *lv = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
  uint32_t *order = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
  double *sigma = (double *)malloc((size_t)n * sizeof(double));
  const uint32_t reached = bfs_levels_par(n, off, cols, src, 1, lv, order);
  /* level boundaries in order[] (a BFS order lists levels contiguously) */
  uint32_t levels = 0;
  for (uint32_t i = 0; i < reached; i++)
    if (lv[order[i]] + 1 > levels) levels = lv[order[i]] + 1;
  uint64_t *start = (uint64_t *)calloc((size_t)levels + 1, sizeof(uint64_t));
  for (uint32_t i = 0; i < reached; i++) start[lv[order[i]] + 1]++;
  for (uint32_t k = 0; k < levels; k++) start[k + 1] += start[k];
  sigma[src] = 1.0;
  for (uint32_t L = 1; L < levels; L++) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = (int64_t)start[L]; i < (int64_t)start[L + 1]; i++) {
      const uint32_t w = order[i];
      double sg = 0.0;
      for (uint64_t e = off[w]; e < off[w + 1]; e++)
        if (lv[cols[e]] == L - 1) sg += sigma[cols[e]];
      sigma[w] = sg;
    }
  }
  for (int64_t L = (int64_t)levels - 1; L >= 0; L--) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = (int64_t)start[L]; i < (int64_t)start[L + 1]; i++) {
      const uint32_t v = order[i];
      double d = 0.0;
      for (uint64_t e = off[v]; e < off[v + 1]; e++) {
        const uint32_t w = cols[e];
        if (lv[w] == (uint32_t)L + 1) d += sigma[v] / sigma[w] * (1.0 + out[w]);
      }
      out[v] = d;
    }
  }
  free(start);
  free(lv);
  free(order);
  free(sigma);
}

static uint32_t uf_find(uint32_t *p, uint32_t x) {
  while (p[x] != x) {
    p[x] = p[p[x]];
    x = p[x];
  }
  return x;
}

/* Connected components (TST:571): label = smallest vertex id in the
 * component, edges taken as undirected. */
void orc_cc(uint32_t n, const uint64_t *off, const uint32_t *cols,
            uint32_t *label) {
  for (uint32_t i = 0; i < n; i++) label[i] = i;
  for (uint32_t u = 0; u < n; u++)
    for (uint64_t e = off[u]; e < off[u + 1]; e++) {
      uint32_t a = uf_find(label, u), b = uf_find(label, cols[e]);
      if (a == b) continue;
      if (a < b)
        label[b] = a;
      else
        label[a] = b;
    }
  for (uint32_t i = 0; i < n; i++) label[i] = uf_find(label, i);
}

/* TSS:357-407 two-pointer count with early-end bounds; SM:61-71 passes
 * (node1, node2) as bounds with early_end, line_width+1 otherwise. */
uint64_t orc_common_neighbors(const uint64_t *off, const uint32_t *cols,
                              uint32_t u, uint32_t v, int early_end) {
  uint64_t ea = early_end ? u : UINT64_MAX, eb = early_end ? v : UINT64_MAX;
  uint64_t i = off[u], ie = off[u + 1], j = off[v], je = off[v + 1];
  uint64_t cnt = 0;
  while (i < ie && j < je) {
    uint64_t a = cols[i], b = cols[j];
    if (a >= ea || b >= eb) break;
    if (a == b) {
      cnt++;
      i++;
      j++;
    } else if (a < b) {
      i++;
    } else {
      j++;
    }
  }
  return cnt;
}

/* TST:606 TC(g): for every stored edge (u, v) with v < u, count the common
 * neighbours below v (common_neighbors(u, v, true)); each triangle w<v<u is
 * counted once. */
uint64_t orc_tc(uint32_t n, const uint64_t *off, const uint32_t *cols) {
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total)
  for (int64_t u = 0; u < (int64_t)n; u++)
    for (uint64_t e = off[u]; e < off[u + 1]; e++) {
      uint32_t v = cols[e];
      if (v >= (uint32_t)u) break; /* rows are sorted */
      total += orc_common_neighbors(off, cols, (uint32_t)u, v, 1);
    }
  return total;
}

/* Bellman-Ford from src (TST:618 BF(g, src)); weights NULL means every edge
 * weighs 1 (a binary graph passes value 1, SM:189-191).  Unreached vertices
 * keep INT32_MAX / 2. */
void orc_bf(uint32_t n, const uint64_t *off, const uint32_t *cols,
            const int32_t *weights, uint32_t src, int32_t *dist) {
  const int32_t INF = INT32_MAX / 2;
  for (uint32_t i = 0; i < n; i++) dist[i] = INF;
  if (src >= n) return;
  dist[src] = 0;
  for (uint32_t round = 0; round < n; round++) {
    int changed = 0;
    for (uint32_t u = 0; u < n; u++) {
      if (dist[u] == INF) continue;
      for (uint64_t e = off[u]; e < off[u + 1]; e++) {
        int32_t w = weights ? weights[e] : 1;
        int64_t nd = (int64_t)dist[u] + w;
        if (nd < dist[cols[e]]) {
          dist[cols[e]] = (int32_t)nd;
          changed = 1;
        }
      }
    }
    if (!changed) break;
  }
}
