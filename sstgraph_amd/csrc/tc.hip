// Triangle counting on symmetric graphs by degree orientation.
//
// The reference's TC (called at include/SSTGraph/internal/test.hpp:606)
// counts, for every stored edge (u, v) with v < u, the common neighbours of u
// and v below v (SparseMatrix.hpp:61-71 with early_end), so each triangle is
// counted once.  On a symmetric graph that count equals the number of
// triangles, which any orientation of the edges that is acyclic counts once
// as well.  This path orients each edge toward the endpoint of higher
// (degree, id) rank, so that out(u) -- the neighbours of u ranked above u --
// has at most sqrt(2m) entries, and counts |out(u) n out(v)| for every
// oriented edge u -> v: a triangle a < b < c (by rank) is found exactly once,
// at a -> b with c in both lists.
//
// The same counting kernels serve directed rows: the reference's per-edge
// count is the sum over u and v in L(u) of |L(u) n L(v)| with L(x) the ids
// of row x below x, a prefix of the sorted row, so the store itself is the
// list array there (count_list_triangles with base = off, len = lower).
//
// Layout: the oriented list of u is written at the same flat position as
// row u of the store's degree prefix (dag[prefix[u] .. + dplus[u]), sorted by
// id), so no second scan is needed.
//
// Counting stages out(u) in LDS as a linear-probing hash table (load <= 1/4)
// and streams the lists out(v), v in out(u), through it, with the lanes spread
// evenly over the concatenated lists.  Six launch tiers by |out(u)|:
//   [2, 16]    16 lanes per vertex, four vertices per wave
//   [17, 64], [65, 128], [129, 256], [257, 512]
//              one wave per vertex, 1.8 / 3.5 / 7 / 14 KB of LDS
//   > 512      one block per vertex; out(u) is hashed in tiles of 4096 ids
//              and every tile sees all of the lists out(v).
#include <algorithm>

#include "algo.hpp"
#include "edge_map.hpp"
#include "graph.hpp"
#include "primitives.hpp"

namespace sstg {

namespace {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;  // never a vertex id (ids < n_rows)
// wave tiers: |out(u)| upper bounds (powers of two, so that the table of
// 4 * MaxOut slots holds every size table_bits() picks up to MaxOut)
constexpr uint32_t kWaveTiers = 5;
constexpr uint32_t kTierOut[kWaveTiers] = {16, 64, 128, 256, 512};
constexpr uint32_t kMidOut = kTierOut[kWaveTiers - 1];
constexpr uint32_t kBlkTile = 4096;
constexpr uint32_t kBlkTable = 4 * kBlkTile;
constexpr uint32_t kBlkChunk = 256;  // lists out(v) per block step
constexpr uint32_t kAccLanes = 64;   // spread counters, 128 B apart
constexpr uint32_t kAccStride = 16;

__device__ __forceinline__ bool ranked_above(uint32_t dv, uint32_t v,
                                             uint32_t du, uint32_t u) {
  return dv > du || (dv == du && v > u);
}

__device__ __forceinline__ uint32_t tc_hash(uint32_t x, uint32_t shift) {
  return (x * 0x9E3779B1u) >> shift;
}

// log2 of the table size for k ids: smallest power of two >= 4k, >= 64
__device__ __forceinline__ uint32_t table_bits(uint32_t k) {
  uint32_t b = 6;
  while ((1u << b) < 4 * k) b++;
  return b;
}

__device__ __forceinline__ void hash_insert(uint32_t *table, uint32_t x,
                                            uint32_t bits) {
  const uint32_t mask = (1u << bits) - 1;
  uint32_t h = tc_hash(x, 32 - bits);
  while (atomicCAS(&table[h], kEmpty, x) != kEmpty) h = (h + 1) & mask;
}

__device__ __forceinline__ uint32_t hash_has(const uint32_t *table, uint32_t x,
                                             uint32_t bits) {
  const uint32_t mask = (1u << bits) - 1;
  uint32_t h = tc_hash(x, 32 - bits);
  for (;;) {
    const uint32_t e = table[h];
    if (e == x) return 1;
    if (e == kEmpty) return 0;
    h = (h + 1) & mask;
  }
}

// Streams the concatenation of the lists out(v) (list j starts at dag +
// base[j], its flat range is [pre[j], pre[j+1]), pre[k] = total) through the
// hash table, P lanes striding the flat index; four loads are issued before
// their probes so that each lane keeps four gathers in flight.  A lane keeps
// its current list's end and address shift in registers and reads the LDS
// tables only when it crosses into another list.
template <uint32_t P>
__device__ __forceinline__ uint64_t probe_lists(const uint32_t *table,
                                                uint32_t bits,
                                                const uint32_t *pre,
                                                const uint64_t *base,
                                                const uint32_t *dag,
                                                uint32_t total, uint32_t me) {
  uint64_t found = 0;
  if (me >= total) return 0;
  uint32_t seg = 0, seg_start = 0, seg_end = pre[1];
  uint64_t shift = base[0];  // dag offset of flat index t: shift + t
  for (uint32_t t0 = me; t0 < total; t0 += 4 * P) {
    uint32_t w[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const uint32_t t = t0 + r * P;
      w[r] = kEmpty;
      if (t < total) {
        if (seg_end <= t) {
          do {
            seg++;
            seg_start = seg_end;
            seg_end = pre[seg + 1];
          } while (seg_end <= t);
          shift = base[seg] - seg_start;  // wraps; shift + t is exact
        }
        w[r] = dag[shift + t];
      }
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
      if (w[r] != kEmpty) found += hash_has(table, w[r], bits);
  }
  return found;
}

// Orientation, G lanes per row (rows from one degree bin).  Kept neighbours
// are written in row order by ballot rank inside the lane group.
template <int G>
__global__ void __launch_bounds__(256)
    k_tc_orient(DevGraph g, const uint32_t *list, uint64_t count,
                const uint64_t *prefix, uint32_t *dag, uint32_t *dplus) {
  const uint64_t gid = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const uint64_t grp = gid / G;
  const uint32_t lig = (uint32_t)(gid % G);
  if (grp >= count) return;  // whole groups leave together
  const uint32_t u = list[grp];
  const uint32_t du = g.deg[u];
  const uint32_t *row = g.pool + g.off[u];
  uint32_t *out = dag + prefix[u];
  uint32_t c = 0;
  if (G == 1) {
    for (uint32_t k = 0; k < du; k++) {
      const uint32_t v = row[k];
      if (ranked_above(g.deg[v], v, du, u)) out[c++] = v;
    }
  } else {
    const uint32_t lane = lane_id();
    const uint64_t gmask =
        (G == 64) ? ~0ull : (((1ull << G) - 1) << (lane & ~(uint32_t)(G - 1)));
    for (uint32_t k0 = 0; k0 < du; k0 += G) {
      const uint32_t k = k0 + lig;
      bool keep = false;
      uint32_t v = 0;
      if (k < du) {
        v = row[k];
        keep = ranked_above(g.deg[v], v, du, u);
      }
      const uint64_t m = __ballot(keep) & gmask;
      if (keep) out[c + mask_rank(m)] = v;
      c += (uint32_t)__popcll(m);
    }
  }
  if (lig == 0) dplus[u] = c;
}

// inclusive sum over groups of L lanes (L divides 64)
template <uint32_t L>
__device__ __forceinline__ uint32_t group_inclusive_sum(uint32_t v,
                                                        uint32_t lig) {
#pragma unroll
  for (uint32_t o = 1; o < L; o <<= 1) {
    const uint32_t t = __shfl_up(v, o, L);
    if (lig >= o) v += t;
  }
  return v;
}

// L lanes per vertex u with 2 <= |out(u)| <= MaxOut, 64 / L vertices per
// block of one wave, each group with its own LDS table: short lists keep
// several vertices' setup loads in flight per wave.
template <uint32_t MaxOut, uint32_t L>
__global__ void __launch_bounds__(64)
    k_tc_count_wave(const uint64_t *lbase, const uint32_t *lists,
                    const uint32_t *llen, const uint32_t *list,
                    uint64_t count, unsigned long long *acc) {
  static_assert((MaxOut & (MaxOut - 1)) == 0, "table size must be 2^k");
  static_assert(4 * MaxOut >= 64 && 64 % L == 0, "group shape");
  constexpr uint32_t NG = 64 / L;
  __shared__ uint32_t table[NG][4 * MaxOut];
  __shared__ uint32_t pre[NG][MaxOut + 1];
  __shared__ uint64_t base[NG][MaxOut];
  const uint32_t lane = threadIdx.x, gi = lane / L, lig = lane % L;
  uint32_t *tab = table[gi];
  uint32_t *pr = pre[gi];
  uint64_t *bs = base[gi];
  uint64_t found = 0;
  for (uint64_t i0 = (uint64_t)blockIdx.x * NG; i0 < count;
       i0 += (uint64_t)gridDim.x * NG) {  // uniform across the block
    const uint64_t i = i0 + gi;
    uint32_t d = 0;
    const uint32_t *row = lists;
    if (i < count) {
      const uint32_t u = list[i];
      d = llen[u];
      row = lists + lbase[u];
    }
    const uint32_t bits = table_bits(d);
    for (uint32_t t = lig; t < (1u << bits); t += L) tab[t] = kEmpty;
    __syncthreads();
    for (uint32_t j = lig; j < d; j += L) {
      const uint32_t x = row[j];
      hash_insert(tab, x, bits);
      pr[j] = llen[x];
      bs[j] = lbase[x];
    }
    __syncthreads();
    // exclusive scan of the list lengths; pre[d] = total
    uint32_t carry = 0;
    for (uint32_t j0 = 0; j0 < d; j0 += L) {
      const uint32_t j = j0 + lig;
      const uint32_t x = j < d ? pr[j] : 0;
      const uint32_t inc = group_inclusive_sum<L>(x, lig);
      if (j < d) pr[j] = carry + inc - x;
      carry += __shfl(inc, gi * L + L - 1, 64);
    }
    if (lig == 0) pr[d] = carry;
    __syncthreads();
    found += probe_lists<L>(tab, bits, pr, bs, lists, carry, lig);
    __syncthreads();  // table, pre and base are reused by the next vertex
  }
  found = wave_sum(found);
  if (lane == 0 && found)
    atomicAdd(&acc[(blockIdx.x % kAccLanes) * kAccStride],
              (unsigned long long)found);
}

// One block of 256 per vertex u with |out(u)| > kMidOut: out(u) is hashed in
// tiles of kBlkTile ids; for each tile every list out(v), v in out(u), is
// streamed through it in steps of kBlkChunk lists.
__global__ void __launch_bounds__(256)
    k_tc_count_block(const uint64_t *lbase, const uint32_t *lists,
                     const uint32_t *llen, const uint32_t *list,
                     uint64_t count, unsigned long long *acc) {
  __shared__ uint32_t table[kBlkTable];
  __shared__ uint32_t pre[kBlkChunk + 1];
  __shared__ uint64_t base[kBlkChunk];
  const uint32_t tid = threadIdx.x;
  uint64_t found = 0;
  for (uint64_t i = blockIdx.x; i < count; i += gridDim.x) {
    const uint32_t u = list[i];
    const uint32_t d = llen[u];
    const uint32_t *row = lists + lbase[u];
    for (uint32_t t0 = 0; t0 < d; t0 += kBlkTile) {
      const uint32_t tn = (d - t0) < kBlkTile ? (d - t0) : kBlkTile;
      const uint32_t bits = table_bits(tn);
      for (uint32_t t = tid; t < (1u << bits); t += 256) table[t] = kEmpty;
      __syncthreads();
      for (uint32_t j = tid; j < tn; j += 256) hash_insert(table, row[t0 + j], bits);
      __syncthreads();
      for (uint32_t c0 = 0; c0 < d; c0 += kBlkChunk) {
        const uint32_t cn = (d - c0) < kBlkChunk ? (d - c0) : kBlkChunk;
        uint32_t len = 0;
        if (tid < cn) {
          const uint32_t x = row[c0 + tid];
          len = llen[x];
          base[tid] = lbase[x];
        }
        uint32_t total = 0;
        const uint32_t off = em::block_scan_u32(len, &total);
        if (tid < cn) pre[tid] = off;
        if (tid == 0) pre[cn] = total;
        __syncthreads();
        found += probe_lists<256>(table, bits, pre, base, lists, total, tid);
        __syncthreads();  // pre and base are rewritten by the next step
      }
    }
  }
  found = wave_sum(found);
  if ((tid & 63) == 0 && found)
    atomicAdd(&acc[(blockIdx.x % kAccLanes) * kAccStride],
              (unsigned long long)found);
}

// acc[i * kAccStride] for i < kAccLanes are the spread counters; the slot
// after them holds the largest list length
__global__ void k_tc_acc_clear(unsigned long long *acc) {
  acc[threadIdx.x * kAccStride] = 0;
  if (threadIdx.x == 0) acc[kAccLanes * kAccStride] = 0;
}

// largest len[i]: grid-stride, one atomic per block (few blocks, so the
// single target address sees few atomics)
__global__ void __launch_bounds__(256)
    k_tc_max_len(const uint32_t *len, uint32_t n, unsigned long long *acc) {
  __shared__ uint32_t wmax[4];
  uint32_t v = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    v = len[i] > v ? len[i] : v;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t t = __shfl_xor(v, o, 64);
    v = t > v ? t : v;
  }
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x / 64] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; w++) v = wmax[w] > v ? wmax[w] : v;
    if (v) atomicMax(&acc[kAccLanes * kAccStride], (unsigned long long)v);
  }
}

__global__ void k_tc_acc_sum(unsigned long long *acc) {
  uint64_t v = acc[threadIdx.x * kAccStride];
  v = wave_sum(v);
  if (threadIdx.x == 0) acc[0] = v;
}

// Triangles of a graph whose every edge has its reverse that contain at
// least one changed undirected edge, each counted once, by its smallest
// changed edge in (min << 32 | max) order.  keys: the changed pairs, sorted,
// with both directions of every edge.  One wave per changed edge {u, v},
// u < v: lanes take the ids of the shorter row and look each up in the
// longer one; a common neighbour w counts unless {u, w} or {v, w} is also
// changed with a smaller key.  Per-block partial sums, no atomics.
__device__ __forceinline__ bool sorted_has(const uint64_t *a, uint64_t n,
                                           uint64_t x) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < n && a[lo] == x;
}

__global__ void __launch_bounds__(256)
    k_tc_delta(DevGraph g, const uint64_t *keys, uint64_t count,
               unsigned long long *part) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t waves = (uint64_t)gridDim.x * 4;
  uint64_t found = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 4 + threadIdx.x / 64; i < count;
       i += waves) {
    const uint64_t kuv = keys[i];
    const uint32_t u = (uint32_t)(kuv >> 32), v = (uint32_t)kuv;
    if (u >= v) continue;  // the wave's edge: each undirected edge once
    const uint32_t du = g.deg[u], dv = g.deg[v];
    const uint32_t a = du <= dv ? u : v, b = du <= dv ? v : u;
    const uint32_t da = du <= dv ? du : dv, db = du <= dv ? dv : du;
    const uint32_t *ra = g.pool + g.off[a], *rb = g.pool + g.off[b];
    for (uint32_t j = lane; j < da; j += 64) {
      const uint32_t w = ra[j];
      if (w == u || w == v) continue;
      if (!row_has(rb, db, w)) continue;
      const uint64_t kuw = u < w ? ((uint64_t)u << 32) | w
                                 : ((uint64_t)w << 32) | u;
      const uint64_t kvw = v < w ? ((uint64_t)v << 32) | w
                                 : ((uint64_t)w << 32) | v;
      if (kuw < kuv && sorted_has(keys, count, kuw)) continue;
      if (kvw < kuv && sorted_has(keys, count, kvw)) continue;
      found++;
    }
  }
  __shared__ unsigned long long wsum[4];
  found = wave_sum(found);
  if (lane == 0) wsum[threadIdx.x / 64] = found;
  __syncthreads();
  if (threadIdx.x == 0)
    part[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ void __launch_bounds__(256)
    k_tc_delta_sum(const unsigned long long *part, uint32_t nb,
                   unsigned long long *out) {
  unsigned long long t = 0;
  for (uint32_t i = threadIdx.x; i < nb; i += 256) t += part[i];
  __shared__ unsigned long long wsum[4];
  t = wave_sum(t);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x / 64] = t;
  __syncthreads();
  if (threadIdx.x == 0) *out = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

uint32_t grid_for(uint64_t count) {
  return (uint32_t)(count < (1u << 24) ? count : (1u << 24));
}

}  // namespace

int count_list_triangles(sstg_graph *g, const uint64_t *lbase,
                         const uint32_t *lists, const uint32_t *llen,
                         uint64_t *out) {
  hipStream_t s = g->stream;
  const uint32_t n = g->n_rows;
  uint32_t *list = nullptr;
  unsigned long long *acc = nullptr;
  uint8_t *flags = nullptr;
  uint64_t *tmp = nullptr;
  SSTG_TRY(g->ws.get(kWsTcList, n, &list));
  SSTG_TRY(g->ws.get(kWsTcAcc, (kAccLanes + 1) * kAccStride, &acc));
  SSTG_TRY(g->ws.get(kWsAlgo + 24, n, &flags));
  SSTG_TRY(g->ws.get(kWsAlgo + 25, n, &tmp));
  SSTG_LAUNCH(k_tc_acc_clear, 1, kAccLanes, 0, s, acc);
  // the flat probe index of one vertex (up to 512 lists, or 256 per block
  // step) is 32-bit: lists of 2^23 ids or more go to the caller's fallback
  SSTG_LAUNCH(k_tc_max_len, (uint32_t)std::min<uint64_t>(ceil_div(n, 256), 512),
              256, 0, s, llen, n, acc);
  SSTG_HIP(hipMemcpyAsync(g->host_rb + 17, acc + kAccLanes * kAccStride,
                          sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  SSTG_HIP(hipStreamSynchronize(s));
  if (g->host_rb[17] >= (1u << 23)) return SSTG_E_UNSUPPORTED;
  for (uint32_t c = 0; c <= kWaveTiers; c++) {
    const uint32_t lo = c == 0 ? 2 : kTierOut[c - 1] + 1;
    const uint32_t hi = c < kWaveTiers ? kTierOut[c] : 0xFFFFFFFFu;
    SSTG_LAUNCH(em::k_bin_flags, ceil_div(n, em::kT), em::kT, 0, s, llen, n,
                lo, hi, flags);
    SSTG_TRY(select_flagged_u64(flags, nullptr, n, tmp, g->dev_ctr + 7, g->ws,
                                kWsPrim, s));
    SSTG_HIP(hipMemcpyAsync(g->host_rb + 16, g->dev_ctr + 7, sizeof(uint64_t),
                            hipMemcpyDeviceToHost, s));
    SSTG_HIP(hipStreamSynchronize(s));
    const uint64_t cnt = g->host_rb[16];
    if (!cnt) continue;
    SSTG_TRY(narrow_u64_to_u32(tmp, cnt, list, s));
    const uint32_t grid = grid_for(cnt);
    const uint32_t *l = list;
    switch (c) {
      case 0:
        SSTG_LAUNCH((k_tc_count_wave<kTierOut[0], 16>),
                    (uint32_t)std::min<uint64_t>(ceil_div(cnt, 4), 1u << 24), 64,
                    0, s, lbase, lists, llen, l, cnt, acc);
        break;
      case 1:
        SSTG_LAUNCH((k_tc_count_wave<kTierOut[1], 64>), grid, 64, 0, s, lbase,
                    lists, llen, l, cnt, acc);
        break;
      case 2:
        SSTG_LAUNCH((k_tc_count_wave<kTierOut[2], 64>), grid, 64, 0, s, lbase,
                    lists, llen, l, cnt, acc);
        break;
      case 3:
        SSTG_LAUNCH((k_tc_count_wave<kTierOut[3], 64>), grid, 64, 0, s, lbase,
                    lists, llen, l, cnt, acc);
        break;
      case 4:
        SSTG_LAUNCH((k_tc_count_wave<kTierOut[4], 64>), grid, 64, 0, s, lbase,
                    lists, llen, l, cnt, acc);
        break;
      default:
        SSTG_LAUNCH(k_tc_count_block, grid, 256, 0, s, lbase, lists, llen, l,
                    cnt, acc);
    }
  }
  SSTG_LAUNCH(k_tc_acc_sum, 1, kAccLanes, 0, s, acc);
  SSTG_HIP(hipMemcpyAsync(g->host_rb, acc, sizeof(uint64_t),
                          hipMemcpyDeviceToHost, s));
  SSTG_HIP(hipStreamSynchronize(s));
  *out = g->host_rb[0];
  return SSTG_OK;
}

int tc_oriented(sstg_graph *g, uint64_t *out) {
  hipStream_t s = g->stream;
  const uint32_t n = g->n_rows;
  DegreeBins bins;
  SSTG_TRY(degree_bins(g, &bins));
  const uint64_t *prefix = nullptr;
  SSTG_TRY(graph_degree_prefix(g, &prefix));
  uint32_t *dag = nullptr, *dplus = nullptr;
  SSTG_TRY(g->ws.get(kWsTcDag, g->m, &dag));
  SSTG_TRY(g->ws.get(kWsTcDplus, n, &dplus));
  const DevGraph v = g->view();
  if (bins.count[0])
    SSTG_LAUNCH(k_tc_orient<1>, ceil_div(bins.count[0], 256), 256, 0, s, v,
                (const uint32_t *)bins.list[0], bins.count[0], prefix, dag,
                dplus);
  if (bins.count[1])
    SSTG_LAUNCH(k_tc_orient<8>, ceil_div(bins.count[1] * 8, 256), 256, 0, s, v,
                (const uint32_t *)bins.list[1], bins.count[1], prefix, dag,
                dplus);
  for (int b = 2; b < 4; b++)
    if (bins.count[b])
      SSTG_LAUNCH(k_tc_orient<64>, ceil_div(bins.count[b] * 64, 256), 256, 0,
                  s, v, (const uint32_t *)bins.list[b], bins.count[b], prefix,
                  dag, dplus);
  return count_list_triangles(g, prefix, dag, dplus, out);
}

int tc_delta(sstg_graph *g, const uint64_t *keys, uint64_t count,
             uint64_t *out) {
  *out = 0;
  if (count == 0) return SSTG_OK;
  hipStream_t s = g->stream;
  const uint32_t nb = (uint32_t)std::min<uint64_t>(4096, (count + 7) / 8);
  unsigned long long *part = nullptr;
  SSTG_TRY(g->ws.get(kWsTcDelta, nb + 1, &part));
  SSTG_LAUNCH(k_tc_delta, nb, 256, 0, s, g->view(), keys, count, part);
  SSTG_LAUNCH(k_tc_delta_sum, 1, 256, 0, s, (const unsigned long long *)part,
              nb, part + nb);
  SSTG_HIP(hipMemcpyAsync(g->host_rb + 40, part + nb, sizeof(uint64_t),
                          hipMemcpyDeviceToHost, s));
  SSTG_HIP(hipStreamSynchronize(s));
  *out = g->host_rb[40];
  return SSTG_OK;
}

}  // namespace sstg
