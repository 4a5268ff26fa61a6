// Degree-order relabelling and store statistics on the device.
//
// degree_order_map (SparseMatrix.hpp:737-749): vertex i gets the rank of the
// pair (degree(i), i) in descending order, so the highest-degree vertex is
// 0 and ties go to the larger id first.  Here the pairs are one radix sort
// of keys ((n_cols - deg) << b | (n - 1 - i)), ascending.
//
// rewrite_graph (SparseMatrix.hpp:751-766): a new graph holding
// (map[i], map[j]) for every stored (i, j).  The reference inserts edge by
// edge with insert(), which keeps self-loops; here the relabelled edge list
// is built in HBM and bulk-inserted with self-loops kept.
//
// statistics (print_statistics, SparseMatrix.hpp:340-426): the reference's
// per-tinyset byte-width tables describe its CPU layout; the layout-free
// figures are kept (elements, rows, columns, bytes, and the histogram of the
// bits each stored destination needs) and the HBM store's own are added.
#include "graph.hpp"
#include "primitives.hpp"

namespace sstg {

namespace {

constexpr int kT = 256;

__global__ void k_order_keys(const uint32_t *deg, uint32_t n, uint32_t n_cols,
                             uint32_t bits_id, uint64_t *keys) {
  const uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  keys[i] = ((uint64_t)(n_cols - deg[i]) << bits_id) | (uint64_t)(n - 1 - i);
}

__global__ void k_order_scatter(const uint64_t *sorted, uint32_t n,
                                uint32_t bits_id, uint32_t *map) {
  const uint64_t p = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (p >= n) return;
  const uint64_t mask = bits_id >= 64 ? ~0ull : ((1ull << bits_id) - 1);
  const uint32_t old = (uint32_t)(n - 1 - (sorted[p] & mask));
  map[old] = (uint32_t)p;
}

// 8 stored edges per thread; src/dst[e] = map[row], map[col] for flat e
__global__ void k_relabel_edges(DevGraph g, const uint64_t *prefix,
                                uint64_t m, const uint32_t *map, uint32_t *src,
                                uint32_t *dst) {
  const uint64_t e0 = ((uint64_t)blockIdx.x * kT + threadIdx.x) * 8;
  if (e0 >= m) return;
  const uint64_t end = e0 + 8 < m ? e0 + 8 : m;
  uint64_t u = find_segment(prefix, g.n, e0);
  for (uint64_t e = e0; e < end; e++) {
    while (e >= prefix[u + 1]) u++;
    const uint32_t c = g.pool[g.off[u] + (e - prefix[u])];
    src[e] = map[u];
    dst[e] = map[c];
  }
}

__device__ __forceinline__ uint32_t bits_needed(uint32_t x) {
  return x ? 32u - (uint32_t)__clz(x) : 0u;
}

// one wave per row: destination bit-width histogram, empty rows, largest
// degree, slot capacity total
__global__ void __launch_bounds__(kT)
    k_stats(DevGraph g, unsigned long long *out) {
  __shared__ unsigned long long hist[33];
  for (uint32_t i = threadIdx.x; i < 33; i += kT) hist[i] = 0;
  __syncthreads();
  const uint64_t row = ((uint64_t)blockIdx.x * kT + threadIdx.x) / kWave;
  const uint32_t lane = lane_id();
  if (row < g.n) {
    const uint32_t d = g.deg[row];
    const uint32_t *r = g.pool + g.off[row];
    for (uint32_t k = lane; k < d; k += kWave)
      atomicAdd(&hist[bits_needed(r[k])], 1ull);
    if (lane == 0) {
      if (d == 0) atomicAdd(&out[33], 1ull);
      atomicMax(&out[34], (unsigned long long)d);
      atomicAdd(&out[35], (unsigned long long)g.cap[row]);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < 33; i += kT)
    if (hist[i]) atomicAdd(&out[i], hist[i]);
}

__global__ void k_clear_u64(unsigned long long *p, uint32_t n) {
  const uint32_t i = blockIdx.x * kT + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

int graph_degree_order_map(sstg_graph *g, uint32_t *map) {
  hipStream_t s = g->stream;
  const uint32_t n = g->n_rows;
  if (n == 0) return SSTG_OK;
  const uint32_t bits_id = bit_width_u32(n - 1);
  const uint32_t bits_deg = bit_width_u32(g->n_cols);
  uint64_t *keys = nullptr, *tmp = nullptr;
  SSTG_TRY(g->ws.get(kWsRelKeys, n, &keys));
  SSTG_TRY(g->ws.get(kWsRelTmp, n, &tmp));
  SSTG_LAUNCH(k_order_keys, ceil_div(n, kT), kT, 0, s, (const uint32_t *)g->deg,
              n, g->n_cols, bits_id, keys);
  int shifts[8];
  int npass = 0;
  for (uint32_t b = 0; b < bits_id + bits_deg; b += 8) shifts[npass++] = (int)b;
  uint64_t *sorted = keys;
  if (npass)
    SSTG_TRY(radix_sort_u64(keys, tmp, n, shifts, npass, g->ws, kWsPrim, s,
                            &sorted));
  SSTG_LAUNCH(k_order_scatter, ceil_div(n, kT), kT, 0, s,
              (const uint64_t *)sorted, n, bits_id, map);
  return SSTG_OK;
}

int graph_rewrite(sstg_graph *g, const uint32_t *map, sstg_graph *out) {
  hipStream_t s = g->stream;
  if (g->m == 0) return SSTG_OK;
  const uint64_t *prefix = nullptr;
  SSTG_TRY(graph_degree_prefix(g, &prefix));
  uint32_t *src = nullptr, *dst = nullptr;
  SSTG_TRY(g->ws.get(kWsRelSrc, g->m, &src));
  SSTG_TRY(g->ws.get(kWsRelDst, g->m, &dst));
  SSTG_LAUNCH(k_relabel_edges, ceil_div(ceil_div(g->m, 8), kT), kT, 0, s,
              g->view(), prefix, g->m, map, src, dst);
  SSTG_HIP(hipStreamSynchronize(s));
  // out's own stream orders after the synchronisation above
  return graph_update(out, src, dst, g->m, true, false, true);
}

int graph_statistics(sstg_graph *g, uint64_t *hist33, uint64_t *empty_rows,
                     uint64_t *max_degree, uint64_t *slot_capacity) {
  hipStream_t s = g->stream;
  unsigned long long *acc = nullptr;
  SSTG_TRY(g->ws.get(kWsRelKeys, 36, &acc));
  SSTG_LAUNCH(k_clear_u64, 1, kT, 0, s, acc, 36u);
  if (g->n_rows)
    SSTG_LAUNCH(k_stats, ceil_div((uint64_t)g->n_rows * kWave, kT), kT, 0, s,
                g->view(), acc);
  uint64_t host[36];
  SSTG_HIP(hipMemcpyAsync(host, acc, sizeof(host), hipMemcpyDeviceToHost, s));
  SSTG_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < 33; i++) hist33[i] = host[i];
  *empty_rows = host[33];
  *max_degree = host[34];
  *slot_capacity = host[35];
  return SSTG_OK;
}

}  // namespace sstg
