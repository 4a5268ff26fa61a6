// EdgeMap / VertexMap on the device-resident store (Ligra-style interface of
// the reference's EdgeMapVertexMap layer, whose callers are at
// include/SSTGraph/internal/test.hpp:476-618 and whose graph-side contract is
// SparseMatrixV::map_neighbors / map_range, SparseMatrix.hpp:174-315).
//
// Functor contracts (device methods):
//   sparse push  : bool cond(d); bool update_atomic(s, d)  -> d joins the
//                  output frontier when update_atomic returns true
//   dense pull   : bool cond(d); void update(s, d)         -> d scans its
//                  neighbours until one is in the input frontier bitmap
//   vertex pull  : Acc zero(); Acc term(v, u) / add(acc, v, u);
//                  Acc merge(a, b); void store(v, acc)      -> reduction over
//                  all neighbours of v
#pragma once

#include "graph.hpp"
#include "primitives.hpp"

namespace sstg {

struct EdgeMapResult {
  uint64_t count = 0;       // vertices in the output frontier
  uint64_t degree_sum = 0;  // sum of their degrees
};

struct EdgeMapCtx {
  uint32_t *bm_cur = nullptr;   // input frontier bitmap
  uint32_t *bm_next = nullptr;  // output frontier bitmap
  uint64_t words = 0;
  uint64_t *fpre = nullptr;     // frontier degree prefix (sparse)
  uint64_t fcap = 0;
  int slot = 0;
  unsigned long long *ctr = nullptr;  // [0] count, [1] degree sum
  // bottom-up active lists: the bins' survivors back to back in one buffer
  // (bin b at act_off[b], act_n[b] entries), double-buffered; act_src < 0
  // until the first bottom-up step, whose lists are the degree bins
  uint32_t *act[2] = {};
  uint64_t act_off[4] = {0, 0, 0, 0};
  uint64_t act_n[4] = {0, 0, 0, 0};
  int act_src = -1;
};

namespace em {
namespace {  // internal linkage: this header is included by several files

constexpr int kT = 256;
constexpr int kChunk = 8;
constexpr int kBitmapRounds = 8;  // 512 vertices per wave in k_bitmap_from

__global__ void k_clear_u32(uint32_t *p, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_queue_to_bitmap(const uint32_t *q, uint64_t nq,
                                  uint32_t *bm) {
  uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i < nq) {
    uint32_t v = q[i];
    atomicOr(&bm[v >> 5], 1u << (v & 31));
  }
}

__global__ void k_frontier_degrees(DevGraph g, const uint32_t *q, uint64_t nq,
                                   uint64_t *out) {
  uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i < nq) out[i] = g.deg[q[i]];
  if (i == nq) out[nq] = 0;
}

__global__ void k_narrow(const uint64_t *in, uint64_t n, uint32_t *out) {
  uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) out[i] = (uint32_t)in[i];
}

// Wave-aggregated append of d to out; also accumulates count and degree sum.
__device__ __forceinline__ void append(bool take, uint32_t d, uint32_t dd,
                                       uint32_t *out,
                                       unsigned long long *ctr) {
  uint64_t mask = __ballot(take);
  if (!mask) return;
  uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1;
  uint32_t cnt = (uint32_t)__popcll(mask);
  unsigned long long dsum = take ? dd : 0;
  // sum degrees of the taken lanes over the active lanes
  for (int o = kWave / 2; o > 0; o >>= 1)
    dsum += __shfl_xor(dsum, o, kWave);
  unsigned long long base = 0;
  if (lane_id() == leader) {
    base = atomicAdd(&ctr[0], (unsigned long long)cnt);
    atomicAdd(&ctr[1], dsum);
  }
  base = __shfl(base, leader, kWave);
  if (take && out) out[base + mask_rank(mask)] = d;
}

// Block-wide exclusive scan of one u32 per thread (kT threads); returns the
// thread's offset and the block total in *total.
__device__ __forceinline__ uint32_t block_scan_u32(uint32_t v, uint32_t *total) {
  __shared__ uint32_t wt[kT / kWave];
  const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
  uint32_t incl = wave_inclusive_sum(v);
  if (lane == kWave - 1) wt[w] = incl;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kT / kWave; i++) {
    uint32_t x = wt[i];
    if ((uint32_t)i < w) off += x;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return off + incl - v;
}

// Adds a per-thread count and degree sum to ctr with one atomic pair per
// block.
__device__ __forceinline__ void block_count(uint32_t cnt, uint64_t dsum,
                                            unsigned long long *ctr) {
  __shared__ unsigned long long bc[2];
  if (threadIdx.x == 0) {
    bc[0] = 0;
    bc[1] = 0;
  }
  __syncthreads();
  uint64_t c = wave_sum((uint64_t)cnt), d = wave_sum(dsum);
  if (lane_id() == 0 && (c | d)) {
    atomicAdd(&bc[0], (unsigned long long)c);
    atomicAdd(&bc[1], (unsigned long long)d);
  }
  __syncthreads();
  if (threadIdx.x == 0 && (bc[0] | bc[1])) {
    atomicAdd(&ctr[0], bc[0]);
    atomicAdd(&ctr[1], bc[1]);
  }
}

template <class F>
__global__ void __launch_bounds__(kT)
    k_sparse_push(DevGraph g, const uint32_t *front, const uint64_t *fpre,
                  uint64_t nf, uint64_t mf, F f, uint32_t *next,
                  unsigned long long *ctr) {
  // each thread expands kChunk consecutive frontier edges, keeps the vertices
  // it claims in registers, and the block appends them with one atomic
  __shared__ uint32_t got_lds[kT * kChunk];
  uint32_t *got = got_lds + threadIdx.x * kChunk;
  uint64_t e = ((uint64_t)blockIdx.x * kT + threadIdx.x) * kChunk;
  uint64_t end = e + kChunk < mf ? e + kChunk : mf;
  uint32_t cnt = 0;
  uint64_t dsum = 0;
  if (e < end) {
    uint64_t i = find_segment(fpre, nf, e);
    for (; e < end; e++) {
      while (e >= fpre[i + 1]) i++;
      uint32_t s = front[i];
      uint32_t d = g.pool[g.off[s] + (e - fpre[i])];
      if (f.cond(d) && f.update_atomic(s, d)) {
        got[cnt++] = d;
        dsum += g.deg[d];
      }
    }
  }
  uint32_t total;
  uint32_t off = block_scan_u32(cnt, &total);
  __shared__ unsigned long long base;
  if (threadIdx.x == 0 && total) {
    base = atomicAdd(&ctr[0], (unsigned long long)total);
  }
  __syncthreads();
  if (total) {
    for (uint32_t k = 0; k < cnt; k++) next[base + off + k] = got[k];
  }
  block_count(0, dsum, ctr);
}

// Output frontier bitmap of a dense step: bit v = f.in_output(v); one lane
// per vertex, a ballot per 64 vertices, no atomics.
// Per-block count and degree sum stored at part[2 * block] (no atomics:
// many blocks adding to one address serialise at the L2).
__device__ __forceinline__ void block_partial(uint64_t cnt, uint64_t dsum,
                                              unsigned long long *part) {
  __shared__ unsigned long long bp[2][kT / kWave];
  cnt = wave_sum(cnt);
  dsum = wave_sum(dsum);
  if (lane_id() == 0) {
    bp[0][threadIdx.x / kWave] = cnt;
    bp[1][threadIdx.x / kWave] = dsum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long c = 0, d = 0;
    for (int w = 0; w < kT / kWave; w++) {
      c += bp[0][w];
      d += bp[1][w];
    }
    part[2 * (uint64_t)blockIdx.x] = c;
    part[2 * (uint64_t)blockIdx.x + 1] = d;
  }
}

// ctr[0..1] = the sums of the nb partial pairs (one block)
__global__ void __launch_bounds__(kT)
    k_sum_partials(const unsigned long long *part, uint64_t nb,
                   unsigned long long *ctr) {
  uint64_t c = 0, d = 0;
  for (uint64_t b = threadIdx.x; b < nb; b += kT) {
    c += part[2 * b];
    d += part[2 * b + 1];
  }
  __shared__ unsigned long long bp[2][kT / kWave];
  c = wave_sum(c);
  d = wave_sum(d);
  if (lane_id() == 0) {
    bp[0][threadIdx.x / kWave] = c;
    bp[1][threadIdx.x / kWave] = d;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tc = 0, td = 0;
    for (int w = 0; w < kT / kWave; w++) {
      tc += bp[0][w];
      td += bp[1][w];
    }
    ctr[0] = tc;
    ctr[1] = td;
  }
}

template <class F>
__global__ void __launch_bounds__(kT)
    k_bitmap_from(uint32_t n, F f, uint32_t *bm, uint64_t words, DevGraph g,
                  unsigned long long *part) {
  // kBitmapRounds words of 64 vertices per wave; the output vertices are
  // counted (and their degrees summed) per block into part
  uint64_t cnt = 0, dsum = 0;
  const uint32_t lane = lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * kT + threadIdx.x) / kWave;
  for (int r = 0; r < kBitmapRounds; r++) {
    const uint64_t v = (wave * kBitmapRounds + r) * kWave + lane;
    const bool in = v < n && f.in_output((uint32_t)v);
    const uint64_t m = __ballot(in);
    if (in) {
      cnt++;
      dsum += g.deg[v];
    }
    const uint64_t w = v >> 5;
    if ((lane & 31) == 0 && w < words)
      bm[w] = (lane == 0) ? (uint32_t)m : (uint32_t)(m >> 32);
  }
  block_partial(cnt, dsum, part);
}

constexpr uint32_t kNoVertex = 0xFFFFFFFFu;  // ids are below n_rows < 2^32

// Bottom-up step over the active vertices list[0..count) with G lanes per
// vertex (G divides 64).  A vertex whose cond holds tests one neighbour per
// lane first, then U per lane per round: the U row elements, then their U
// bitmap words, so each lane has U gathers in flight before it tests any of
// them.  The first round with a frontier member ends the scan and the lowest
// such lane's neighbour is the source.  keep[i] = 1 marks a vertex with
// neighbours that found none: it stays active for the next bottom-up step.
// No atomics: the counts come from the bitmap pass and the active list from
// a compaction of keep.
template <class F, int G, int U>
__global__ void __launch_bounds__(kT)
    k_dense_pull(DevGraph g, const uint32_t *list, uint64_t count, F f,
                 const uint32_t *bm_cur, uint8_t *keep) {
  const uint64_t gid =
      (uint64_t)xcd_block(blockIdx.x, gridDim.x) * kT + threadIdx.x;
  const uint32_t lig = (uint32_t)(gid % G);
  const uint64_t i = gid / G;
  const uint32_t glane0 = lane_id() & ~(uint32_t)(G - 1);
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << glane0);
  bool found = false, active = false;
  uint32_t dv = 0;
  uint32_t v = 0;
  if (i < count) {
    v = list[i];
    if (f.cond(v)) {
      dv = g.deg[v];
      const uint32_t *row = g.pool + g.off[v];
      // the first round tests one neighbour per lane (most unvisited
      // vertices meet the frontier among their first neighbours), later
      // rounds U per lane
      for (uint32_t base = 0, uu = 1; base < dv; base += G * uu, uu = U) {
        uint32_t u[U], w[U];
#pragma unroll
        for (int j = 0; j < U; j++) {
          const uint32_t k = base + lig + j * G;
          u[j] = ((uint32_t)j < uu && k < dv) ? row[k] : kNoVertex;
        }
#pragma unroll
        for (int j = 0; j < U; j++)
          w[j] = u[j] != kNoVertex ? bm_cur[u[j] >> 5] : 0u;
        uint32_t mine = kNoVertex;
#pragma unroll
        for (int j = U - 1; j >= 0; j--)
          if (u[j] != kNoVertex && ((w[j] >> (u[j] & 31)) & 1u)) mine = u[j];
        const uint64_t m = __ballot(mine != kNoVertex) & gmask;
        if (m) {
          const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1;
          const uint32_t su = __shfl(mine, leader, kWave);
          if (lig == 0) {
            f.update(su, v);
            found = true;
          }
          break;
        }
      }
      active = !found && dv > 0;
    }
    if (lig == 0) keep[i] = active ? 1 : 0;
  }
}

// Bottom-up step with a whole block per (high-degree) vertex: 256 x U
// neighbours per round, all loads of a round issued before their tests.
template <class F, int U>
__global__ void __launch_bounds__(kT)
    k_dense_pull_block(DevGraph g, const uint32_t *list, uint64_t count, F f,
                       const uint32_t *bm_cur, uint8_t *keep) {
  __shared__ uint32_t first_hit;
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t v = list[i];
  if (!f.cond(v)) {  // uniform across the block
    if (threadIdx.x == 0) keep[i] = 0;
    return;
  }
  const uint32_t dv = g.deg[v];
  const uint32_t *row = g.pool + g.off[v];
  if (threadIdx.x == 0) first_hit = kNoVertex;
  __syncthreads();
  for (uint32_t base = 0, uu = 1; base < dv; base += kT * uu, uu = U) {
    uint32_t u[U], w[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      const uint32_t k = base + threadIdx.x + j * kT;
      u[j] = ((uint32_t)j < uu && k < dv) ? row[k] : kNoVertex;
    }
#pragma unroll
    for (int j = 0; j < U; j++)
      w[j] = u[j] != kNoVertex ? bm_cur[u[j] >> 5] : 0u;
#pragma unroll
    for (int j = 0; j < U; j++)
      if (u[j] != kNoVertex && ((w[j] >> (u[j] & 31)) & 1u))
        atomicMin(&first_hit, base + threadIdx.x + j * kT);
    __syncthreads();
    const uint32_t h = first_hit;
    if (h != kNoVertex) {
      if (threadIdx.x == 0) {
        f.update(row[h], v);
        keep[i] = 0;
      }
      return;
    }
  }
  if (threadIdx.x == 0) keep[i] = dv > 0 ? 1 : 0;
}

// ------------------------------------------------------------ vertex pull
// Reduction over all neighbours of each vertex with G lanes per vertex
// (G <= 64 via wave shuffles; G == 256 uses the whole block).  The lane
// stride and the shuffle tree are fixed, so results do not depend on timing.
// Only vertices for which pred(v, degree) holds are reduced; pred must be the
// same for a whole group (it depends on v and its degree only).
struct AllVertices {
  __device__ __forceinline__ bool operator()(uint32_t, uint32_t) const {
    return true;
  }
};

template <class F, int G, class P>
__device__ __forceinline__ void vertex_pull_body(const DevGraph &g,
                                                 const uint32_t *list,
                                                 uint64_t count, const F &f,
                                                 const P &pred) {
  using Acc = typename F::Acc;
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const uint64_t gid = (uint64_t)blk * kT + threadIdx.x;
  const uint64_t grp = (G == 256) ? blk : gid / G;
  const uint32_t lig = (G == 256) ? threadIdx.x : (uint32_t)(gid % G);
  if (grp >= count) return;  // whole groups leave together
  const uint32_t v = list ? list[grp] : (uint32_t)grp;
  const uint32_t d = g.deg[v];
  if (!pred(v, d)) return;  // whole groups (whole blocks for G == 256)
  const uint32_t *row = g.pool + g.off[v];
  Acc acc = f.zero();
  uint32_t k = lig;
  // four neighbour ids in flight before their gathers
  for (; k + 3 * G < d; k += 4 * G) {
    uint32_t u0 = row[k], u1 = row[k + G], u2 = row[k + 2 * G],
             u3 = row[k + 3 * G];
    f.add(acc, v, u0);
    f.add(acc, v, u1);
    f.add(acc, v, u2);
    f.add(acc, v, u3);
  }
  for (; k < d; k += G) f.add(acc, v, row[k]);
  if (G > 1) {
    const int w = (G < kWave) ? G : kWave;
    for (int o = w / 2; o > 0; o >>= 1) acc = f.merge(acc, F::shfl_xor(acc, o));
  }
  if (G == 256) {
    __shared__ Acc part[kT / kWave];
    if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      Acc a = part[0];
      for (int i = 1; i < kT / kWave; i++) a = f.merge(a, part[i]);
      f.store(v, a);
    }
    return;
  }
  if (lig == 0) f.store(v, acc);
}

template <class F, int G>
__global__ void __launch_bounds__(kT)
    k_vertex_pull(DevGraph g, const uint32_t *list, uint64_t count, F f) {
  vertex_pull_body<F, G>(g, list, count, f, AllVertices{});
}

template <class F, int G, class P>
__global__ void __launch_bounds__(kT)
    k_vertex_pull_if(DevGraph g, const uint32_t *list, uint64_t count, F f,
                     P pred) {
  vertex_pull_body<F, G>(g, list, count, f, pred);
}

// predicates of vertex_pull_level
struct UpToDegree {
  uint32_t max_deg;
  __device__ __forceinline__ bool operator()(uint32_t, uint32_t d) const {
    return d <= max_deg;
  }
};
struct OnLevel {
  const uint32_t *level;
  uint32_t want;
  __device__ __forceinline__ bool operator()(uint32_t v, uint32_t) const {
    return level[v] == want;
  }
};

}  // namespace
}  // namespace em

inline int edge_map_ctx(sstg_graph *g, EdgeMapCtx *ctx, int slot) {
  ctx->words = ((uint64_t)g->n_rows + 31) / 32 + 1;
  ctx->slot = slot;
  SSTG_TRY(g->ws.get(slot + 0, ctx->words, &ctx->bm_cur));
  SSTG_TRY(g->ws.get(slot + 1, ctx->words, &ctx->bm_next));
  ctx->ctr = (unsigned long long *)(g->dev_ctr + 9);
  return SSTG_OK;
}

inline int frontier_to_bitmap(EdgeMapCtx &ctx, const uint32_t *q, uint64_t nq,
                              hipStream_t s) {
  SSTG_LAUNCH(em::k_clear_u32, ceil_div(ctx.words, em::kT), em::kT, 0, s,
              ctx.bm_cur, ctx.words);
  if (nq)
    SSTG_LAUNCH(em::k_queue_to_bitmap, ceil_div(nq, em::kT), em::kT, 0, s, q,
                nq, ctx.bm_cur);
  return SSTG_OK;
}

inline int narrow_u64_to_u32(const uint64_t *in, uint64_t n, uint32_t *out,
                             hipStream_t s) {
  if (n)
    SSTG_LAUNCH(em::k_narrow, ceil_div(n, em::kT), em::kT, 0, s, in, n, out);
  return SSTG_OK;
}

inline int read_result(sstg_graph *g, EdgeMapCtx &ctx, EdgeMapResult *res) {
  SSTG_HIP(hipMemcpyAsync(g->host_rb, ctx.ctr, 2 * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, g->stream));
  SSTG_HIP(hipStreamSynchronize(g->stream));
  res->count = g->host_rb[0];
  res->degree_sum = g->host_rb[1];
  return SSTG_OK;
}

// Sparse push over the frontier queue front[0..nf) whose degree sum is mf;
// the output frontier is appended at next.
template <class F>
int edge_map_sparse(sstg_graph *g, EdgeMapCtx &ctx, const uint32_t *front,
                    uint64_t nf, uint64_t mf, F f, uint32_t *next,
                    EdgeMapResult *res) {
  hipStream_t s = g->stream;
  SSTG_HIP(hipMemsetAsync(ctx.ctr, 0, 2 * sizeof(uint64_t), s));
  if (nf && mf) {
    SSTG_TRY(g->ws.get(ctx.slot + 2, nf + 1, &ctx.fpre));
    SSTG_LAUNCH(em::k_frontier_degrees, ceil_div(nf + 1, em::kT), em::kT, 0, s,
                g->view(), front, nf, ctx.fpre);
    SSTG_TRY(scan_exclusive_u64(ctx.fpre, ctx.fpre, nf + 1, nullptr, g->ws,
                                kWsPrim2, s));
    SSTG_LAUNCH(em::k_sparse_push<F>, ceil_div(ceil_div(mf, em::kChunk), em::kT),
                em::kT, 0, s, g->view(), front, (const uint64_t *)ctx.fpre, nf,
                mf, f, next, ctx.ctr);
  }
  return read_result(g, ctx, res);
}

// ----------------------------------------------------------- degree bins
// Vertices split by degree (cached per graph version and threshold set):
// bin 0 deg <= t[0] (one lane), bin 1 <= t[1] (8 lanes), bin 2 <= t[2] (a
// wave), bin 3 larger (a block), t = bin_hi[set]; the boundaries default to
// 8 / 64 / 2048 and are set per graph by the autotuner.
struct DegreeBins {
  uint32_t *list[4] = {nullptr, nullptr, nullptr, nullptr};
  uint64_t count[4] = {0, 0, 0, 0};
};

namespace em {
namespace {  // internal linkage: this header is included by several files
__global__ void k_bin_flags(const uint32_t *deg, uint32_t n, uint32_t lo,
                            uint32_t hi, uint8_t *flags) {
  uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v < n) {
    uint32_t d = deg[v];
    flags[v] = d >= lo && d <= hi;
  }
}
}  // namespace
}  // namespace em

// The four lists lie back to back in one buffer (list[b + 1] = list[b] +
// count[b]).
inline int degree_bins(sstg_graph *g, DegreeBins *out, int set = kBinsPull) {
  const uint32_t *t = g->bin_hi[set];
  const uint32_t lo[4] = {0, t[0] + 1, t[1] + 1, t[2] + 1};
  const uint32_t hi[4] = {t[0], t[1], t[2], 0xFFFFFFFFu};
  hipStream_t s = g->stream;
  const uint32_t n = g->n_rows;
  const int slot = set == kBinsPull ? kWsAlgo + 26 : kWsBfsBins;
  uint32_t *all = nullptr;
  SSTG_TRY(g->ws.get(slot, (uint64_t)n + 1, &all));
  if (g->bins_version[set] == g->version) {
    uint64_t off = 0;
    for (int b = 0; b < 4; b++) {
      out->count[b] = g->bins_count[set][b];
      out->list[b] = all + off;
      off += out->count[b];
    }
    return SSTG_OK;
  }
  uint8_t *flags = nullptr;
  uint64_t *tmp = nullptr;
  SSTG_TRY(g->ws.get(kWsAlgo + 24, n, &flags));
  SSTG_TRY(g->ws.get(kWsAlgo + 25, n, &tmp));
  for (int b = 0; b < 4; b++) {
    SSTG_LAUNCH(em::k_bin_flags, ceil_div(n, em::kT), em::kT, 0, s,
                (const uint32_t *)g->deg, n, lo[b], hi[b], flags);
    SSTG_TRY(select_flagged_u64(flags, nullptr, n, tmp, g->dev_ctr + 7, g->ws,
                                kWsPrim, s));
    SSTG_HIP(hipMemcpyAsync(g->host_rb + 16, g->dev_ctr + 7, sizeof(uint64_t),
                            hipMemcpyDeviceToHost, s));
    SSTG_HIP(hipStreamSynchronize(s));
    out->count[b] = g->host_rb[16];
    out->list[b] = all + (b ? (out->list[b - 1] - all) + out->count[b - 1] : 0);
    SSTG_TRY(narrow_u64_to_u32(tmp, out->count[b], out->list[b], s));
  }
  for (int b = 0; b < 4; b++) g->bins_count[set][b] = out->count[b];
  g->bins_version[set] = g->version;
  return SSTG_OK;
}

namespace em {
namespace {
// Kept flags of each bin's segment of a bottom-up step, summed: counts[b] =
// sum of keep[seg[b] .. seg[b + 1]).  Grid-stride per-block partials, then
// one block adds them (no contended atomics).
constexpr uint32_t kSegBlocks = 1024;

__global__ void __launch_bounds__(kT)
    k_segment_partials(const uint8_t *keep, const uint64_t *seg,
                       unsigned long long *part) {
  const uint64_t s1 = seg[1], s2 = seg[2], s3 = seg[3], total = seg[4];
  uint64_t c[4] = {0, 0, 0, 0};
  for (uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * kT) {
    const uint64_t k = keep[i];
    const int b = i < s1 ? 0 : i < s2 ? 1 : i < s3 ? 2 : 3;
    c[b] += k;
  }
  __shared__ unsigned long long wc[4][kT / kWave];
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const uint64_t t = wave_sum(c[b]);
    if (lane_id() == 0) wc[b][threadIdx.x / kWave] = t;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    unsigned long long t = 0;
    for (int w = 0; w < kT / kWave; w++) t += wc[threadIdx.x][w];
    part[4 * (uint64_t)blockIdx.x + threadIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kT)
    k_segment_sum(const unsigned long long *part, uint32_t nb,
                  unsigned long long *counts) {
  uint64_t c[4] = {0, 0, 0, 0};
  for (uint32_t i = threadIdx.x; i < nb; i += kT)
#pragma unroll
    for (int b = 0; b < 4; b++) c[b] += part[4 * (uint64_t)i + b];
  __shared__ unsigned long long wc[4][kT / kWave];
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const uint64_t t = wave_sum(c[b]);
    if (lane_id() == 0) wc[b][threadIdx.x / kWave] = t;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    unsigned long long t = 0;
    for (int w = 0; w < kT / kWave; w++) t += wc[threadIdx.x][w];
    counts[threadIdx.x] = t;
  }
}
}  // namespace
}  // namespace em

// Dense pull over the active vertices against ctx.bm_cur, split by degree
// bin (1, 4, 16 lanes or a block per vertex; 16 lanes rather than a wave in
// the third bin took BFS from the hub at rMat scale 24 from 2.71 to 2.52
// ms, and 4 rather than 8 in the second to 2.49 ms); the output frontier is
// ctx.bm_next, after which the two bitmaps swap roles.  The bins' active
// lists lie back to back, so one compaction of the keep flags carries the
// survivors of every bin to the next step.
template <class F>
int edge_map_dense(sstg_graph *g, EdgeMapCtx &ctx, F f, EdgeMapResult *res) {
  hipStream_t s = g->stream;
  DegreeBins bins;
  SSTG_TRY(degree_bins(g, &bins, kBinsBfs));
  const uint64_t all = bins.count[0] + bins.count[1] + bins.count[2] +
                       bins.count[3];
  const bool first = ctx.act_src < 0;
  const int dst = first ? 0 : 1 - ctx.act_src;
  SSTG_TRY(g->ws.get(kWsBottomUp + 0, all + 1, &ctx.act[0]));
  SSTG_TRY(g->ws.get(kWsBottomUp + 1, all + 1, &ctx.act[1]));
  // the first bottom-up step reads the degree bins; later ones the vertices
  // the previous step left unvisited
  const uint32_t *cur = first ? bins.list[0] : ctx.act[ctx.act_src];
  uint64_t cnt[4], off[5];
  off[0] = 0;
  for (int b = 0; b < 4; b++) {
    cnt[b] = first ? bins.count[b] : ctx.act_n[b];
    off[b + 1] = off[b] + cnt[b];
  }
  const uint64_t total = off[4];
  // ctr[0..1]: found count and degree sum; ctr[2]: survivors;
  // ctr[3 + b]: survivors of bin b
  SSTG_HIP(hipMemsetAsync(ctx.ctr, 0, 7 * sizeof(uint64_t), s));
  uint8_t *keep = nullptr;
  uint64_t *seg = nullptr;
  SSTG_TRY(g->ws.get(kWsBottomUp + 8, all + 1, &keep));
  SSTG_TRY(g->ws.get(kWsBottomUp + 2, 5, &seg));
  for (int b = 0; b < 5; b++) g->host_rb[40 + b] = off[b];  // pinned
  SSTG_HIP(hipMemcpyAsync(seg, g->host_rb + 40, sizeof(off),
                          hipMemcpyHostToDevice, s));
  DevGraph v = g->view();
  const uint32_t *bm = ctx.bm_cur;
  if (cnt[0])
    SSTG_LAUNCH((em::k_dense_pull<F, 1, 8>), ceil_div(cnt[0], em::kT), em::kT,
                0, s, v, cur + off[0], cnt[0], f, bm, keep + off[0]);
  if (cnt[1])
    SSTG_LAUNCH((em::k_dense_pull<F, 4, 8>), ceil_div(cnt[1] * 4, em::kT),
                em::kT, 0, s, v, cur + off[1], cnt[1], f, bm, keep + off[1]);
  if (cnt[2])
    SSTG_LAUNCH((em::k_dense_pull<F, 16, 8>), ceil_div(cnt[2] * 16, em::kT),
                em::kT, 0, s, v, cur + off[2], cnt[2], f, bm, keep + off[2]);
  if (cnt[3])
    SSTG_LAUNCH((em::k_dense_pull_block<F, 4>), cnt[3], em::kT, 0, s, v,
                cur + off[3], cnt[3], f, bm, keep + off[3]);
  if (total) {
    SSTG_TRY(select_flagged_u32(keep, cur, total, ctx.act[dst],
                                (uint64_t *)(ctx.ctr + 2), g->ws, kWsPrim, s));
    const uint32_t nb = (uint32_t)std::min<uint64_t>(
        ceil_div(total, (uint64_t)em::kT * 16), em::kSegBlocks);
    unsigned long long *spart = nullptr;
    SSTG_TRY(g->ws.get(kWsBottomUp + 3, 4 * (uint64_t)em::kSegBlocks, &spart));
    SSTG_LAUNCH(em::k_segment_partials, nb, em::kT, 0, s, (const uint8_t *)keep,
                (const uint64_t *)seg, spart);
    SSTG_LAUNCH(em::k_segment_sum, 1, em::kT, 0, s,
                (const unsigned long long *)spart, nb, ctx.ctr + 3);
  }
  const uint64_t bm_waves = ceil_div(ctx.words * 32, (uint64_t)kWave);
  const uint64_t bm_blocks =
      ceil_div(bm_waves, (uint64_t)(em::kT / kWave) * em::kBitmapRounds);
  unsigned long long *part = nullptr;
  SSTG_TRY(g->ws.get(kWsBottomUp + 9, 2 * bm_blocks, &part));
  SSTG_LAUNCH(em::k_bitmap_from<F>, bm_blocks, em::kT, 0, s, g->n_rows, f,
              ctx.bm_next, ctx.words, v, part);
  SSTG_LAUNCH(em::k_sum_partials, 1, em::kT, 0, s,
              (const unsigned long long *)part, bm_blocks, ctx.ctr);
  uint32_t *t = ctx.bm_cur;
  ctx.bm_cur = ctx.bm_next;
  ctx.bm_next = t;
  SSTG_HIP(hipMemcpyAsync(g->host_rb, ctx.ctr, 7 * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, s));
  SSTG_HIP(hipStreamSynchronize(s));
  res->count = g->host_rb[0];
  res->degree_sum = g->host_rb[1];
  for (int b = 0; b < 4; b++) ctx.act_n[b] = g->host_rb[3 + b];
  ctx.act_src = dst;
  return SSTG_OK;
}

// Reduction over every vertex by degree bin.  The third bin takes 16 lanes
// per vertex rather than a wave: at rMat scale 24, PageRank x20 went from
// 57.0 to 54.6 ms with its autotuned boundaries (55.5 ms with 32 lanes), as
// fewer lanes idle on the last pass over a row.
template <class F>
int vertex_pull_all(sstg_graph *g, const DegreeBins &bins, F f) {
  hipStream_t s = g->stream;
  DevGraph v = g->view();
  if (bins.count[0])
    SSTG_LAUNCH((em::k_vertex_pull<F, 1>), ceil_div(bins.count[0], em::kT),
                em::kT, 0, s, v, (const uint32_t *)bins.list[0], bins.count[0],
                f);
  if (bins.count[1])
    SSTG_LAUNCH((em::k_vertex_pull<F, 8>), ceil_div(bins.count[1] * 8, em::kT),
                em::kT, 0, s, v, (const uint32_t *)bins.list[1], bins.count[1],
                f);
  if (bins.count[2])
    SSTG_LAUNCH((em::k_vertex_pull<F, 16>),
                ceil_div(bins.count[2] * 16, em::kT), em::kT, 0, s, v,
                (const uint32_t *)bins.list[2], bins.count[2], f);
  if (bins.count[3])
    SSTG_LAUNCH((em::k_vertex_pull<F, 256>), bins.count[3], em::kT, 0, s, v,
                (const uint32_t *)bins.list[3], bins.count[3], f);
  return SSTG_OK;
}

// Vertex pull over an explicit vertex list with one wave per vertex.
template <class F>
int vertex_pull_list(sstg_graph *g, const uint32_t *list, uint64_t count, F f) {
  if (!count) return SSTG_OK;
  SSTG_LAUNCH((em::k_vertex_pull<F, 64>), ceil_div(count * 64, em::kT), em::kT,
              0, g->stream, g->view(), list, count, f);
  return SSTG_OK;
}

// Vertex pull over the vertices of one BFS level (list: that level's
// vertices, level[v] == L for each).  Large levels give their low-degree
// vertices (up to the reductions' second-bin bound) 4 lanes each and reach
// the rest through the third (16 lanes) and block degree bins filtered by
// level, instead of a wave per vertex; small levels take a wave per vertex.
// 16 lanes rather than a wave in the third bin took BC from the hub at rMat
// scale 24 from 12.5 to 11.8 ms, and 4 rather than 8 lanes for the low-degree
// vertices to 11.6 ms.
constexpr uint64_t kLevelSplitMin = 1u << 16;

template <class F>
int vertex_pull_level(sstg_graph *g, const uint32_t *list, uint64_t count,
                      const uint32_t *level, uint32_t L, F f) {
  if (!count) return SSTG_OK;
  if (count < kLevelSplitMin) return vertex_pull_list(g, list, count, f);
  DegreeBins bins;
  SSTG_TRY(degree_bins(g, &bins, kBinsPull));
  hipStream_t s = g->stream;
  const DevGraph v = g->view();
  const em::UpToDegree low{g->bin_hi[kBinsPull][1]};
  const em::OnLevel on{level, L};
  SSTG_LAUNCH((em::k_vertex_pull_if<F, 4, em::UpToDegree>),
              ceil_div(count * 4, em::kT), em::kT, 0, s, v, list, count, f, low);
  if (bins.count[2])
    SSTG_LAUNCH((em::k_vertex_pull_if<F, 16, em::OnLevel>),
                ceil_div(bins.count[2] * 16, em::kT), em::kT, 0, s, v,
                (const uint32_t *)bins.list[2], bins.count[2], f, on);
  if (bins.count[3])
    SSTG_LAUNCH((em::k_vertex_pull_if<F, 256, em::OnLevel>), bins.count[3],
                em::kT, 0, s, v, (const uint32_t *)bins.list[3], bins.count[3],
                f, on);
  return SSTG_OK;
}

// ------------------------------------------------------------ functors
// PageRank pull: rank(v) = damping * sum_u contrib(u) + base and
// contrib'(v) = rank(v) / deg(v).  With kKahan the per-lane sums carry a
// compensation term and lanes are combined with Neumaier's two-sum.
template <class Real, bool kKahan> struct PrPull;

template <class Real> struct PrPull<Real, false> {
  const Real *contrib;
  Real *contrib_next;
  Real *rank;
  const uint32_t *deg;
  Real damp, base;
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
    Real r = damp * a + base;
    rank[v] = r;
    uint32_t d = deg[v];
    contrib_next[v] = d ? r / (Real)d : (Real)0;
  }
};

struct KahanF {
  float s, c;
};

template <> struct PrPull<float, true> {
  const float *contrib;
  float *contrib_next;
  float *rank;
  const uint32_t *deg;
  float damp, base;
  using Acc = KahanF;
  __device__ __forceinline__ Acc zero() const { return KahanF{0.f, 0.f}; }
  __device__ __forceinline__ void add(Acc &a, uint32_t, uint32_t u) const {
    float y = contrib[u] - a.c;
    float t = a.s + y;
    a.c = (t - a.s) - y;
    a.s = t;
  }
  __device__ __forceinline__ Acc merge(Acc a, Acc b) const {
    // Neumaier two-sum of the partial sums; compensations add up
    float t = a.s + b.s;
    float e = (fabsf(a.s) >= fabsf(b.s)) ? (a.s - t) + b.s : (b.s - t) + a.s;
    return KahanF{t, a.c + b.c - e};
  }
  __device__ static __forceinline__ Acc shfl_xor(Acc a, int o) {
    return KahanF{__shfl_xor(a.s, o, kWave), __shfl_xor(a.c, o, kWave)};
  }
  __device__ __forceinline__ void store(uint32_t v, Acc a) const {
    float sum = a.s - a.c;
    float r = damp * sum + base;
    rank[v] = r;
    uint32_t d = deg[v];
    contrib_next[v] = d ? r / (float)d : 0.f;
  }
};

}  // namespace sstg
