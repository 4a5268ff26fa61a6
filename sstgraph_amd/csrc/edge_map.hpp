// EdgeMap / VertexMap on the device-resident store (Ligra-style interface of
// the reference's EdgeMapVertexMap layer, whose callers are at
// include/SSTGraph/internal/test.hpp:476-618 and whose graph-side contract is
// SparseMatrixV::map_neighbors / map_range, SparseMatrix.hpp:174-315).
//
// Functor contracts (device methods):
//   sparse push  : uint32_t peek(d), static bool open(x) -> d may be
//                  claimed; uint32_t try_claim(d) returns the value it
//                  replaced, open(that) -> claimed; void claimed(s, d)
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

// Device read-backs of a traversal.  Recording: each read is made and its
// words kept; replaying (while the stream is captured): the words come from
// the log and nothing is read, so the same launches are enqueued without a
// wait.
struct ReadLog {
  int mode = 0;  // 0 plain, 1 record, 2 replay
  std::vector<uint64_t> words;
  size_t pos = 0;
};

inline int read_back(sstg_graph *g, ReadLog *log, void *dst, const void *src,
                     size_t bytes) {
  const size_t nw = (bytes + 7) / 8;
  if (log && log->mode == 2) {
    if (log->pos + nw > log->words.size()) return SSTG_E_INVAL;
    std::memcpy(dst, log->words.data() + log->pos, bytes);
    log->pos += nw;
    return SSTG_OK;
  }
  SSTG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g->stream));
  SSTG_HIP(hipStreamSynchronize(g->stream));
  if (log && log->mode == 1) {
    const size_t at = log->words.size();
    log->words.resize(at + nw, 0);
    std::memcpy(log->words.data() + at, dst, bytes);
  }
  return SSTG_OK;
}

struct EdgeMapCtx {
  uint32_t *bm = nullptr;   // frontier bitmap of a bottom-up step
  uint32_t *vis = nullptr;  // visited filter of the top-down steps
  uint64_t words = 0;
  uint64_t *fpre = nullptr;  // frontier degree prefix (sparse)
  const uint32_t *fpre_front = nullptr;  // the frontier fpre was built for
  uint64_t fpre_n = 0;
  int slot = 0;
  // step counters, two sets of kStepCtrs (see em::kStepCtrs); a step adds
  // into set `parity` and clears the other
  unsigned long long *ctrs = nullptr;
  int parity = 0;
  // bottom-up active lists: bin b's survivors at act[.] + act_off[b] (the
  // bins' offsets in the degree-bin lists, so a bin never outgrows its
  // region), act_n[b] entries, double-buffered; act_src < 0 until the first
  // bottom-up step, whose lists are the degree bins
  uint32_t *act[2] = {};
  uint64_t act_off[4] = {0, 0, 0, 0};
  uint64_t act_n[4] = {0, 0, 0, 0};
  int act_src = -1;
  // bottom-up steps read rows from their end (set by the caller)
  int rev = 0;
  ReadLog *log = nullptr;  // read-backs, recorded or replayed (may be null)
};

namespace em {
namespace {  // internal linkage: this header is included by several files

constexpr int kT = 256;
constexpr int kChunk = 8;
constexpr int kBitmapRounds = 8;  // 512 vertices per wave in k_bits

__global__ void k_clear_u32(uint32_t *p, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_clear_u64(unsigned long long *p, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) p[i] = 0;
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

// Counters of one BFS step (u64): [0] vertices found, [1] their degree sum,
// [2 + b] survivors of degree bin b (bottom-up).  Two sets alternate: the
// kernels of a step add into one and clear the other, which the host has
// read after the step before, so no step needs a clearing call of its own.
constexpr int kStepCtrs = 8;

__device__ __forceinline__ void clear_step_ctrs(unsigned long long *c) {
  if (c && blockIdx.x == 0 && threadIdx.x < kStepCtrs) c[threadIdx.x] = 0;
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

// Per-block result of a bottom-up step, without global atomics (same-address
// atomics from ~10^5 blocks queue at one memory channel: at rMat scale 24
// they made the widest bottom-up level 5.8 ms instead of 1 ms).  The block
// writes its found vertices, compacted, to tmp_f from lbase (the list index
// of its first vertex) on, its survivors likewise to tmp_k, and
// (found << 32) | survivors to *pc.  Every thread of the block must call it.
__device__ __forceinline__ void step_partial(bool take_f, bool take_k,
                                             uint32_t v, uint64_t lbase,
                                             uint32_t *tmp_f, uint32_t *tmp_k,
                                             uint64_t *pc) {
  // at most kT of each per block: two 16-bit counts in one scan
  const uint32_t packed = (take_f ? 1u : 0u) | (take_k ? 0x10000u : 0u);
  uint32_t total;
  const uint32_t off = block_scan_u32(packed, &total);
  if (take_f) tmp_f[lbase + (off & 0xFFFFu)] = v;
  if (take_k) tmp_k[lbase + (off >> 16)] = v;
  if (threadIdx.x == 0)
    *pc = ((uint64_t)(total & 0xFFFFu) << 32) | (uint64_t)(total >> 16);
}

// Pull blocks of one bottom-up step: bin b's kernel numbers its blocks from
// pbase[b] (pbase[4] = all blocks), its block k covers list indices from
// lbase[b] + k * per[b] (lbase[b]: the bin's offset in the active lists).
struct StepBlocks {
  uint64_t pbase[5];
  uint64_t lbase[4];
  uint32_t per[4];
};

// After the exclusive scan of the packed counts pc[0..pbase[4]]: LPB lanes
// per pull block of bin b (sized to the block's at most per[b] vertices)
// move its found vertices to next (the BFS queue, in block order) and its
// survivors to the bin's region of keep.  The first block of the bin stores
// the bin's survivor count in ctr[2 + b], block 0 the found count in
// ctr[0].
constexpr int kPlaceT = 64;
template <int LPB>
__global__ void __launch_bounds__(kPlaceT)
    k_step_place(StepBlocks sb, int b, const uint64_t *pc,
                 const uint32_t *tmp_f, const uint32_t *tmp_k, uint32_t *next,
                 uint32_t *keep, unsigned long long *ctr) {
  const uint64_t gk = ((uint64_t)blockIdx.x * kPlaceT + threadIdx.x) / LPB;
  const uint32_t lig = threadIdx.x % LPB;
  if (gk >= sb.pbase[b + 1] - sb.pbase[b]) return;
  const uint64_t k = sb.pbase[b] + gk;
  const uint64_t p0 = pc[k], p1 = pc[k + 1], pb = pc[sb.pbase[b]];
  const uint32_t nfound = (uint32_t)(p1 >> 32) - (uint32_t)(p0 >> 32);
  const uint32_t nkeep = (uint32_t)p1 - (uint32_t)p0;
  const uint64_t lb = sb.lbase[b] + gk * sb.per[b];
  const uint64_t fdst = p0 >> 32;
  const uint64_t kdst = sb.lbase[b] + ((uint32_t)p0 - (uint32_t)pb);
  for (uint32_t t = lig; t < nfound; t += LPB) next[fdst + t] = tmp_f[lb + t];
  for (uint32_t t = lig; t < nkeep; t += LPB) keep[kdst + t] = tmp_k[lb + t];
  if (lig == 0) {
    if (gk == 0) ctr[2 + b] = (uint32_t)pc[sb.pbase[b + 1]] - (uint32_t)pb;
    if (k == 0) ctr[0] = pc[sb.pbase[4]] >> 32;
  }
}

constexpr uint32_t kNoVertex = 0xFFFFFFFFu;  // ids are below n_rows < 2^32

// Top-down step: thread t expands the kChunk consecutive frontier edges from
// t * kChunk.  Their row reads are issued together, then their words of the
// visited filter vis, then the cond loads, then the claims, so a lane keeps
// kChunk loads in flight instead of walking one dependent chain per edge.
// vis only filters: a bit is set only for a visited vertex (plain racy ORs
// may lose bits, never add wrong ones), so a stale or lost bit costs a
// failed claim, never a wrong result.  The block appends its claims to next
// with one atomic.
template <class F>
__global__ void __launch_bounds__(kT)
    k_sparse_push(DevGraph g, const uint32_t *front, const uint64_t *fpre,
                  uint64_t nf, uint64_t mf, F f, uint32_t *vis, uint32_t *next,
                  unsigned long long *ctr, unsigned long long *ctr_clear) {
  clear_step_ctrs(ctr_clear);
  __shared__ uint32_t got_lds[kT * kChunk];
  uint32_t *got = got_lds + threadIdx.x * kChunk;
  const uint64_t e0 = ((uint64_t)blockIdx.x * kT + threadIdx.x) * kChunk;
  uint32_t cnt = 0;
  uint64_t dsum = 0;
  // the walk below stays inside fpre[0..nf] only for edges below fpre[nf]
  if (e0 < mf) mf = min(mf, fpre[nf]);
  if (e0 < mf) {
    uint64_t i = find_segment(fpre, nf, e0);
    uint64_t seg_start = fpre[i], seg_end = fpre[i + 1];
    uint32_t s = front[i];
    uint64_t rbase = g.off[s];
    uint32_t d[kChunk], src[kChunk];
#pragma unroll
    for (int j = 0; j < kChunk; j++) {
      const uint64_t e = e0 + j;
      d[j] = kNoVertex;
      src[j] = s;
      if (e < mf) {
        while (e >= seg_end) {  // zero-degree frontier vertices skipped too
          i++;
          seg_start = seg_end;
          seg_end = fpre[i + 1];
          s = front[i];
          rbase = g.off[s];
        }
        src[j] = s;
        d[j] = g.pool[rbase + (e - seg_start)];
      }
    }
    // each phase's loads are unconditional (a non-candidate reads vertex
    // 0's entry) or predicated without use, so they are all in flight
    // before the first result is needed
    uint32_t w[kChunk], x[kChunk];
    bool c[kChunk];
#pragma unroll
    for (int j = 0; j < kChunk; j++)
      w[j] = vis[(d[j] != kNoVertex ? d[j] : 0u) >> 5];
#pragma unroll
    for (int j = 0; j < kChunk; j++)
      c[j] = d[j] != kNoVertex && !((w[j] >> (d[j] & 31)) & 1u);
#pragma unroll
    for (int j = 0; j < kChunk; j++) x[j] = f.peek(c[j] ? d[j] : 0u);
#pragma unroll
    for (int j = 0; j < kChunk; j++) {
      c[j] = c[j] && F::open(x[j]);
      x[j] = c[j] ? f.try_claim(d[j]) : 0u;
    }
    uint32_t dd[kChunk];
#pragma unroll
    for (int j = 0; j < kChunk; j++) {
      c[j] = c[j] && F::open(x[j]);  // this lane's claim won
      dd[j] = c[j] ? g.deg[d[j]] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kChunk; j++) {
      if (c[j]) {
        f.claimed(src[j], d[j]);
        got[cnt++] = d[j];
        dsum += dd[j];
        // the word as read above: a bit set meanwhile by another lane may
        // be lost, which only weakens the filter
        vis[d[j] >> 5] = w[j] | (1u << (d[j] & 31));
      }
    }
  }
  uint32_t total;
  const uint32_t off = block_scan_u32(cnt, &total);
  __shared__ unsigned long long base;
  if (threadIdx.x == 0 && total)
    base = atomicAdd(&ctr[0], (unsigned long long)total);
  __shared__ unsigned long long bsum;
  if (threadIdx.x == 0) bsum = 0;
  __syncthreads();
  const uint64_t ws = wave_sum(dsum);
  if (lane_id() == 0 && ws) atomicAdd(&bsum, (unsigned long long)ws);
  if (total)
    for (uint32_t k = 0; k < cnt; k++) next[base + off + k] = got[k];
  __syncthreads();
  if (threadIdx.x == 0 && bsum) atomicAdd(&ctr[1], bsum);
}

// bm bit v = p(v) for v < n; one lane per vertex, a ballot per 64 vertices,
// no atomics.
template <class P>
__global__ void __launch_bounds__(kT)
    k_bits(uint32_t n, P p, uint32_t *bm, uint64_t words,
           unsigned long long *ctr_clear, uint64_t *zero1) {
  clear_step_ctrs(ctr_clear);
  if (zero1 && blockIdx.x == 0 && threadIdx.x == 0) *zero1 = 0;
  if (n == 0) return;  // no vertex, no word
  const uint32_t lane = lane_id();
  const uint64_t wave = ((uint64_t)blockIdx.x * kT + threadIdx.x) / kWave;
  // every round's test issued before the first ballot (tested as they came,
  // each round waited for its load): past n a lane tests vertex n - 1, so
  // every load is in bounds and none sits behind a branch
  bool in[kBitmapRounds];
#pragma unroll
  for (int r = 0; r < kBitmapRounds; r++) {
    const uint64_t v = (wave * kBitmapRounds + r) * kWave + lane;
    const uint32_t vc = v < n ? (uint32_t)v : n - 1;
    in[r] = p(vc) & (v < n);
  }
#pragma unroll
  for (int r = 0; r < kBitmapRounds; r++) {
    const uint64_t v = (wave * kBitmapRounds + r) * kWave + lane;
    const uint64_t m = __ballot(in[r]);
    const uint64_t w = v >> 5;
    if ((lane & 31) == 0 && w < words)
      bm[w] = (lane == 0) ? (uint32_t)m : (uint32_t)(m >> 32);
  }
}

// Bottom-up step over the active vertices list[0..count) with G lanes per
// vertex (G divides 64).  A vertex whose cond holds tests one neighbour per
// lane first, then U per lane per round: the U row elements, then their U
// bitmap words, so each lane has U gathers in flight before it tests any of
// them.  The first round with a frontier member ends the scan and the lowest
// such lane's neighbour is the source.  With rev the rows are read from
// their end (largest ids first): any frontier neighbour is a valid parent,
// so the order only decides how soon one is met.  Found vertices and the
// vertices with neighbours that found none (survivors, active in the next
// bottom-up step) leave through step_partial; list[i] has list index
// lbase + i.
template <class F, int G, int U>
__global__ void __launch_bounds__(kT)
    k_dense_pull(DevGraph g, const uint32_t *list, uint64_t count, F f,
                 const uint32_t *bm_cur, int rev, uint64_t lbase,
                 uint32_t *tmp_f, uint32_t *tmp_k, uint64_t *pc) {
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const uint64_t gid = (uint64_t)blk * kT + threadIdx.x;
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
          u[j] = ((uint32_t)j < uu && k < dv) ? row[rev ? dv - 1 - k : k]
                                              : kNoVertex;
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
          if (lig == 0) f.update(su, v);
          found = true;
          break;
        }
      }
      active = !found && dv > 0;
    }
  }
  step_partial(lig == 0 && found, lig == 0 && active, v,
               lbase + (uint64_t)blk * (kT / G), tmp_f, tmp_k, pc + blk);
}

// Bottom-up step with a whole block per (high-degree) vertex: 256 x U
// neighbours per round, all loads of a round issued before their tests;
// block i's result goes to pc[i] and slot lbase + i of tmp_f or tmp_k.
// The first hit of round r goes to hit[r % 3]: thread 0 resets the slot of
// round r + 1 before round r's barrier, when every thread has read it as
// round r - 2's slot, so all threads read one value per round.
template <class F, int U>
__global__ void __launch_bounds__(kT)
    k_dense_pull_block(DevGraph g, const uint32_t *list, uint64_t count, F f,
                       const uint32_t *bm_cur, int rev, uint64_t lbase,
                       uint32_t *tmp_f, uint32_t *tmp_k, uint64_t *pc) {
  __shared__ uint32_t hit[3];
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t v = list[i];
  bool found = false, active = false;
  if (f.cond(v)) {  // uniform across the block: only this block writes v
    const uint32_t dv = g.deg[v];
    const uint32_t *row = g.pool + g.off[v];
    if (threadIdx.x < 3) hit[threadIdx.x] = kNoVertex;
    __syncthreads();
    uint32_t r = 0;
    for (uint32_t base = 0, uu = 1; base < dv; base += kT * uu, uu = U, r++) {
      uint32_t u[U], w[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const uint32_t k = base + threadIdx.x + j * kT;
        u[j] = ((uint32_t)j < uu && k < dv) ? row[rev ? dv - 1 - k : k]
                                            : kNoVertex;
      }
#pragma unroll
      for (int j = 0; j < U; j++)
        w[j] = u[j] != kNoVertex ? bm_cur[u[j] >> 5] : 0u;
#pragma unroll
      for (int j = 0; j < U; j++)
        if (u[j] != kNoVertex && ((w[j] >> (u[j] & 31)) & 1u))
          atomicMin(&hit[r % 3], base + threadIdx.x + j * kT);
      if (threadIdx.x == 0) hit[(r + 1) % 3] = kNoVertex;
      __syncthreads();
      const uint32_t h = hit[r % 3];
      if (h != kNoVertex) {
        if (threadIdx.x == 0) f.update(row[rev ? dv - 1 - h : h], v);
        found = true;
        break;
      }
    }
    active = !found && dv > 0;
  }
  if (threadIdx.x == 0) {  // one vertex: its own slot at list index lbase + i
    if (found) tmp_f[lbase + i] = v;
    if (active) tmp_k[lbase + i] = v;
    pc[i] = ((uint64_t)(found ? 1u : 0u) << 32) | (active ? 1u : 0u);
  }
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

// Buffers of a BFS: the frontier bitmap and the visited filter at slot and
// slot + 1, the frontier degree prefix at slot + 2; both counter sets (and
// kCtxExtra words after them) cleared by a kernel, which a captured launch
// sequence can hold.
constexpr int kCtxExtra = 8;
inline int edge_map_ctx(sstg_graph *g, EdgeMapCtx *ctx, int slot) {
  ctx->words = ((uint64_t)g->n_rows + 31) / 32 + 1;
  ctx->slot = slot;
  SSTG_TRY(g->ws.get(slot + 0, ctx->words, &ctx->bm));
  SSTG_TRY(g->ws.get(slot + 1, ctx->words, &ctx->vis));
  SSTG_TRY(g->ws.get(kWsBottomUp + 2, 2 * em::kStepCtrs + kCtxExtra,
                     &ctx->ctrs));
  ctx->parity = 0;
  SSTG_LAUNCH(em::k_clear_u64, 1, em::kT, 0, g->stream, ctx->ctrs,
              (uint64_t)(2 * em::kStepCtrs + kCtxExtra));
  return SSTG_OK;
}

inline int narrow_u64_to_u32(const uint64_t *in, uint64_t n, uint32_t *out,
                             hipStream_t s) {
  if (n)
    SSTG_LAUNCH(em::k_narrow, ceil_div(n, em::kT), em::kT, 0, s, in, n, out);
  return SSTG_OK;
}

// Bitmap bm of the vertices v < n for which p(v) holds (clearing the counter
// set ctr_clear when given).
template <class P>
int edge_map_bits(sstg_graph *g, EdgeMapCtx &ctx, P p, uint32_t *bm,
                  unsigned long long *ctr_clear, uint64_t *zero1 = nullptr) {
  const uint64_t waves = ceil_div(ctx.words * 32, (uint64_t)kWave);
  const uint64_t nb =
      ceil_div(waves, (uint64_t)(em::kT / kWave) * em::kBitmapRounds);
  SSTG_LAUNCH(em::k_bits<P>, nb, em::kT, 0, g->stream, g->n_rows, p, bm,
              ctx.words, ctr_clear, zero1);
  return SSTG_OK;
}

// The counters of the step just launched on the counter set ctx.parity
// (first k of them) into host_rb, one copy and one wait; then the other set
// becomes current.
inline int step_counters(sstg_graph *g, EdgeMapCtx &ctx, int k) {
  SSTG_TRY(read_back(g, ctx.log, g->host_rb,
                     ctx.ctrs + em::kStepCtrs * ctx.parity,
                     k * sizeof(uint64_t)));
  ctx.parity ^= 1;
  return SSTG_OK;
}

// Exclusive degree prefix of the frontier front[0..nf) into ctx.fpre
// (nf + 1 entries); with total given, the degree sum is read back into it
// (one copy, one wait).
inline int frontier_prefix(sstg_graph *g, EdgeMapCtx &ctx,
                           const uint32_t *front, uint64_t nf,
                           uint64_t *total) {
  hipStream_t s = g->stream;
  SSTG_TRY(g->ws.get(ctx.slot + 2, nf + 1, &ctx.fpre));
  SSTG_LAUNCH(em::k_frontier_degrees, ceil_div(nf + 1, em::kT), em::kT, 0, s,
              g->view(), front, nf, ctx.fpre);
  SSTG_TRY(scan_exclusive_u64(ctx.fpre, ctx.fpre, nf + 1, nullptr, g->ws,
                              kWsPrim2, s));
  ctx.fpre_front = front;
  ctx.fpre_n = nf;
  if (total) {
    SSTG_TRY(read_back(g, ctx.log, g->host_rb + 8, ctx.fpre + nf,
                       sizeof(uint64_t)));
    *total = g->host_rb[8];
  }
  return SSTG_OK;
}

// Sparse push over the frontier queue front[0..nf) whose degree sum is mf;
// the output frontier is appended at next, and vis takes its bits.
template <class F>
int edge_map_sparse(sstg_graph *g, EdgeMapCtx &ctx, const uint32_t *front,
                    uint64_t nf, uint64_t mf, F f, uint32_t *next,
                    EdgeMapResult *res) {
  hipStream_t s = g->stream;
  res->count = res->degree_sum = 0;
  if (!nf || !mf) return SSTG_OK;
  if (ctx.fpre_front != front || ctx.fpre_n != nf)
    SSTG_TRY(frontier_prefix(g, ctx, front, nf, nullptr));
  ctx.fpre_front = nullptr;
  unsigned long long *cur = ctx.ctrs + em::kStepCtrs * ctx.parity;
  unsigned long long *other = ctx.ctrs + em::kStepCtrs * (1 - ctx.parity);
  SSTG_LAUNCH(em::k_sparse_push<F>, ceil_div(ceil_div(mf, em::kChunk), em::kT),
              em::kT, 0, s, g->view(), front, (const uint64_t *)ctx.fpre, nf,
              mf, f, ctx.vis, next, cur, other);
  SSTG_TRY(step_counters(g, ctx, 2));
  res->count = g->host_rb[0];
  res->degree_sum = g->host_rb[1];
  return SSTG_OK;
}

// ----------------------------------------------------------- degree bins
// Vertices split by degree (cached per graph version and threshold set):
// bin 0 deg <= t[0] (one lane), bin 1 <= t[1] (8 lanes), bin 2 <= t[2] (16
// lanes), bin 3 larger (a block), t = bin_hi[set]; the boundaries default to
// 8 / 64 / 256 for the reductions and 32 / 128 / 2048 for the BFS steps and
// are set per graph by the autotuner.
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

// Dense pull over the active vertices, split by degree bin (1, 4, 16 lanes
// or a block per vertex; 16 lanes rather than a wave in the third bin took
// BFS from the hub at rMat scale 24 from 2.71 to 2.52 ms, and 4 rather than
// 8 in the second to 2.49 ms), against the bitmap of the vertices for which
// in_front holds.  Found vertices are appended at next (the BFS queue);
// each bin's survivors go to its region of the other active list.  The
// found vertices' degree sum is not computed (res->degree_sum = ~0): a
// following top-down step takes it from frontier_prefix.
template <class F, class P>
int edge_map_dense(sstg_graph *g, EdgeMapCtx &ctx, F f, P in_front,
                   uint32_t *next, EdgeMapResult *res) {
  hipStream_t s = g->stream;
  DegreeBins bins;
  SSTG_TRY(degree_bins(g, &bins, kBinsBfs));
  const uint64_t all = bins.count[0] + bins.count[1] + bins.count[2] +
                       bins.count[3];
  const bool first = ctx.act_src < 0;
  const int dst = first ? 0 : 1 - ctx.act_src;
  SSTG_TRY(g->ws.get(kWsBottomUp + 0, all + 1, &ctx.act[0]));
  SSTG_TRY(g->ws.get(kWsBottomUp + 1, all + 1, &ctx.act[1]));
  uint32_t *tmp_f = nullptr, *tmp_k = nullptr;
  SSTG_TRY(g->ws.get(kWsBottomUp + 3, all + 1, &tmp_f));
  SSTG_TRY(g->ws.get(kWsBottomUp + 8, all + 1, &tmp_k));
  static constexpr uint32_t kPer[4] = {em::kT / 1, em::kT / 4, em::kT / 16, 1};
  em::StepBlocks sb;
  uint64_t cnt[4], nblk[4];
  const uint32_t *cur[4];
  sb.pbase[0] = 0;
  for (int b = 0; b < 4; b++) {
    ctx.act_off[b] = b ? ctx.act_off[b - 1] + bins.count[b - 1] : 0;
    cnt[b] = first ? bins.count[b] : ctx.act_n[b];
    cur[b] = first ? bins.list[b] : ctx.act[ctx.act_src] + ctx.act_off[b];
    nblk[b] = ceil_div(cnt[b], (uint64_t)kPer[b]);
    sb.lbase[b] = ctx.act_off[b];
    sb.per[b] = kPer[b];
    sb.pbase[b + 1] = sb.pbase[b] + nblk[b];
  }
  const uint64_t nb_all = sb.pbase[4];
  uint64_t *pc = nullptr;
  SSTG_TRY(g->ws.get(kWsBottomUp + 9, nb_all + 1, &pc));
  unsigned long long *ctr = ctx.ctrs + em::kStepCtrs * ctx.parity;
  unsigned long long *other = ctx.ctrs + em::kStepCtrs * (1 - ctx.parity);
  // the frontier bitmap (this launch also clears the next counter set and
  // the scan's last entry)
  SSTG_TRY(edge_map_bits(g, ctx, in_front, ctx.bm, other, pc + nb_all));
  DevGraph v = g->view();
  const uint32_t *bm = ctx.bm;
  if (nblk[0])
    SSTG_LAUNCH((em::k_dense_pull<F, 1, 8>), nblk[0], em::kT, 0, s, v, cur[0],
                cnt[0], f, bm, ctx.rev, sb.lbase[0], tmp_f, tmp_k, pc + sb.pbase[0]);
  if (nblk[1])
    SSTG_LAUNCH((em::k_dense_pull<F, 4, 8>), nblk[1], em::kT, 0, s, v, cur[1],
                cnt[1], f, bm, ctx.rev, sb.lbase[1], tmp_f, tmp_k, pc + sb.pbase[1]);
  if (nblk[2])
    SSTG_LAUNCH((em::k_dense_pull<F, 16, 8>), nblk[2], em::kT, 0, s, v,
                cur[2], cnt[2], f, bm, ctx.rev, sb.lbase[2], tmp_f, tmp_k,
                pc + sb.pbase[2]);
  if (nblk[3])
    SSTG_LAUNCH((em::k_dense_pull_block<F, 4>), nblk[3], em::kT, 0, s, v,
                cur[3], cnt[3], f, bm, ctx.rev, sb.lbase[3], tmp_f, tmp_k,
                pc + sb.pbase[3]);
  if (nb_all) {
    SSTG_TRY(scan_exclusive_u64(pc, pc, nb_all + 1, nullptr, g->ws, kWsPrim,
                                s));
    const uint32_t *tf = tmp_f, *tk = tmp_k;
    uint32_t *keep = ctx.act[dst];
    const uint64_t pt = em::kPlaceT;
    if (nblk[0])
      SSTG_LAUNCH(em::k_step_place<64>, ceil_div(nblk[0] * 64, pt), pt, 0, s,
                  sb, 0, (const uint64_t *)pc, tf, tk, next, keep, ctr);
    if (nblk[1])
      SSTG_LAUNCH(em::k_step_place<16>, ceil_div(nblk[1] * 16, pt), pt, 0, s,
                  sb, 1, (const uint64_t *)pc, tf, tk, next, keep, ctr);
    if (nblk[2])
      SSTG_LAUNCH(em::k_step_place<4>, ceil_div(nblk[2] * 4, pt), pt, 0, s,
                  sb, 2, (const uint64_t *)pc, tf, tk, next, keep, ctr);
    if (nblk[3])
      SSTG_LAUNCH(em::k_step_place<1>, ceil_div(nblk[3], pt), pt, 0, s, sb, 3,
                  (const uint64_t *)pc, tf, tk, next, keep, ctr);
  }
  SSTG_TRY(step_counters(g, ctx, 6));
  res->count = g->host_rb[0];
  res->degree_sum = ~0ull;
  for (int b = 0; b < 4; b++) ctx.act_n[b] = g->host_rb[2 + b];
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
