// A whole batch update (SparseMatrixV<true>::insert_batch / remove_batch,
// SparseMatrix.hpp:556-718) of at most kSmallBatch pairs in one 1024-thread
// workgroup.
//
// The general pipeline (update.hip) sizes its grids from a read-back in the
// middle of the update and runs about fifteen launches; for a batch of a few
// thousand pairs those launches and the two host waits are the whole cost.
// Here two launches do every step: one workgroup plans the update with the
// batch held in LDS (steps 1-6), then a fixed grid of workgroups merges the
// touched rows (step 7), one row per workgroup at a time:
//
//   1. load and range-check the pairs, pack them as row << 32 | col keys;
//   2. LSD radix sort in LDS (8-bit digits over the column then row bits);
//   3. keep the first copy of each key, drop self-loops on insert, and keep
//      keys absent (insert) or present (remove) in their row, found by a
//      binary search of the row in HBM, which also gives each key's place;
//   4. compact the kept keys; the rows they touch are the runs of equal row;
//   5. per row: stays in its slot, or (insert past the slot's capacity)
//      moves to a fresh slot at the pool tail -- offsets by a block scan;
//      a batch that needs more room than the pool has left, or moves a row
//      of more than kSmMaxMove ids, stops here with nothing written;
//   6. symmetry: a symmetric graph stays symmetric iff the reverse of every
//      kept key is kept too;
//   7. merge: each row's old ids slide by the number of batch keys below
//      them (from the back on insert, from the front on remove) and the
//      batch keys drop into their places, in chunks of one workgroup whose
//      loads all finish (a barrier) before their stores.
// The plan goes to global scratch between the launches; the only host wait
// is the read-back of the counters at the end.
//
// The layout rules are the general pipeline's (slot capacities from
// slot_capacity, relocated slots left as garbage), so the two paths can
// follow each other on one graph.
#include "graph.hpp"

namespace sstg {
namespace {

constexpr int kSmT = 1024;
constexpr int kSmWaves = kSmT / kWave;
constexpr uint32_t kSmMax = kSmallBatch;
constexpr int kSmItems = kSmMax / kSmT;  // keys per thread
constexpr int kSmBins = 256;
constexpr uint64_t kSmMaxMove = 1u << 15;  // longer moves: general pipeline
constexpr int kSmMergeBlocks = 1024;      // workgroups of the merge launch

// the counter block written back (graph.dev_ctr)
enum {
  kSmErr = 0,       // an id out of range
  kSmFallback = 1,  // the general pipeline must run instead
  kSmB = 2,         // pairs that changed the graph
  kSmT_ = 3,        // rows touched
  kSmAlloc = 4,     // slot elements taken at the pool tail
  kSmGarbage = 5,   // elements of slots left behind
  kSmSym = 6,       // symmetric after the update (if it was before)
  kSmCount = 7,
};

struct SmallShifts {
  int shift[8];
  int passes;
};

// exclusive block-wide prefix of v; *total gets the block's sum
template <class T>
__device__ __forceinline__ T block_scan(T v, T *wsum, T *total) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const T incl = wave_inclusive_sum(v);
  if (lane == kWave - 1) wsum[w] = incl;
  __syncthreads();
  T off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kSmWaves; i++) {
    const T x = wsum[i];
    if (i < w) off += x;
    tot += x;
  }
  __syncthreads();  // wsum is free for the next scan
  *total = tot;
  return off + incl - v;
}

// number of a[0..k) that are <= x (a nondecreasing, in LDS)
__device__ __forceinline__ uint32_t count_le(const uint32_t *a, uint32_t k,
                                             uint64_t x) {
  uint32_t lo = 0, hi = k;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] <= x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// number of a[0..k) that are < x
__device__ __forceinline__ uint32_t count_lt(const uint32_t *a, uint32_t k,
                                             uint64_t x) {
  uint32_t lo = 0, hi = k;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ bool lds_has(const uint64_t *a, uint32_t n,
                                        uint64_t x) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < n && a[lo] == x;
}

// lower bound of x in the sorted row a[0..len) with four probes per round,
// issued together: a round trip to HBM per round, about log5(len) rounds
// instead of log2(len) for a lone key whose search is all latency
__device__ __forceinline__ uint32_t lower_bound_4(const uint32_t *a,
                                                  uint32_t len, uint32_t x) {
  uint32_t lo = 0, hi = len;  // answer in [lo, hi]
  while (hi - lo > 4) {
    const uint32_t step = (hi - lo) / 5;
    const uint32_t p1 = lo + step, p2 = p1 + step, p3 = p2 + step,
                   p4 = p3 + step;
    const uint32_t v1 = a[p1], v2 = a[p2], v3 = a[p3], v4 = a[p4];
    if (v4 < x)
      lo = p4 + 1;
    else if (v3 < x) {
      lo = p3 + 1;
      hi = p4;
    } else if (v2 < x) {
      lo = p2 + 1;
      hi = p3;
    } else if (v1 < x) {
      lo = p1 + 1;
      hi = p2;
    } else {
      hi = p1;
    }
  }
  // at most four candidates left: one more round of loads together
  const uint32_t w0 = lo < hi ? a[lo] : 0u, w1 = lo + 1 < hi ? a[lo + 1] : 0u,
                 w2 = lo + 2 < hi ? a[lo + 2] : 0u,
                 w3 = lo + 3 < hi ? a[lo + 3] : 0u;
  uint32_t below = 0;
  below += lo < hi && w0 < x;
  below += lo + 1 < hi && w1 < x;
  below += lo + 2 < hi && w2 < x;
  below += lo + 3 < hi && w3 < x;
  return lo + below;
}

// One row's merge by the whole workgroup (thread me of width): the batch
// keys keys[0..k) (low 32 bits the ids) with places pos[0..k) in the old
// row of od ids at `row`; dst is the row's new home (== row unless
// relocated).
__device__ __forceinline__ void merge_row(uint32_t *row, uint32_t *dst,
                                          uint32_t od, const uint64_t *keys,
                                          const uint32_t *pos, uint32_t k,
                                          bool insert, uint32_t me,
                                          uint32_t width) {
  const uint32_t p0 = pos[0];
  if (insert && dst != row) {
    // relocation: every old id and batch key to its merged place
    for (uint32_t j = me; j < od; j += width) dst[j + count_le(pos, k, j)] = row[j];
    for (uint32_t i = me; i < k; i += width) dst[pos[i] + i] = (uint32_t)keys[i];
    return;
  }
  if (insert) {
    // old ids at or after p0 slide right, chunks from the back: a chunk's
    // destinations lie at or above its own sources and above every source
    // of the chunks still to come
    for (int64_t e = od; e > (int64_t)p0; e -= width) {
      const int64_t j = e - (int64_t)width + me;
      const bool act = j >= (int64_t)p0;
      uint32_t val = 0, sh = 0;
      if (act) {
        val = row[j];
        sh = count_le(pos, k, (uint64_t)j);
      }
      __syncthreads();  // the chunk's loads before its stores
      if (act) row[j + sh] = val;
    }
    __syncthreads();
    for (uint32_t i = me; i < k; i += width) row[pos[i] + i] = (uint32_t)keys[i];
    return;
  }
  // remove: the ids after each removed place slide left, chunks from the
  // front; a chunk's destinations lie below its own sources
  for (uint64_t e = p0; e < od; e += width) {
    const uint64_t j = e + me;
    bool act = j < od;
    uint32_t val = 0, lt = 0;
    if (act) {
      lt = count_lt(pos, k, j);
      act = !(lt < k && pos[lt] == j);  // a removed id is not carried
      if (act) val = row[j];
    }
    __syncthreads();
    if (act) row[j - lt] = val;
  }
}

// The plan of a small update; each array holds up to kSmMax (+1) entries.
struct SmallPlan {
  uint64_t *key;     // kept keys, sorted
  uint32_t *pos;     // their places in the old rows
  uint32_t *rstart;  // first kept key of each touched row, then b
  uint64_t *aoff;    // pool offset of each relocated row's slot, then total
};

__global__ void __launch_bounds__(kSmT)
    k_update_small_plan(const uint32_t *src, const uint32_t *dst, uint32_t n,
                        DevGraph g, uint32_t n_cols, SmallShifts sh,
                        int insert, int skip_self, int track_sym,
                        uint64_t pool_tail, uint64_t pool_cap, SmallPlan plan,
                        uint64_t *ctr) {
  __shared__ uint64_t key[kSmMax];
  __shared__ uint32_t pos[kSmMax];
  __shared__ uint32_t rstart[kSmMax + 1];
  __shared__ uint32_t rdeg[kSmMax];  // touched rows' degree and capacity
  __shared__ uint32_t rcap[kSmMax];
  __shared__ uint8_t kflag[kSmMax];  // the keep decision of each sorted key
  __shared__ uint32_t wcnt[kSmWaves][kSmBins];
  __shared__ uint32_t tstart[kSmBins];
  __shared__ uint32_t wsum32[kSmWaves];
  __shared__ uint64_t wsum64[kSmWaves];
  __shared__ uint32_t s_bad, s_sym;
  const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;

  // 1. load and range-check
  if (tid == 0) {
    s_bad = 0;
    s_sym = 1;
  }
  __syncthreads();
  for (uint32_t i = tid; i < n; i += kSmT) {
    const uint32_t a = src[i], b = dst[i];
    if (a >= g.n || b >= n_cols) s_bad = 1;
    key[i] = ((uint64_t)a << 32) | b;
  }
  __syncthreads();
  if (s_bad) {
    if (tid == 0) {
      ctr[kSmErr] = 1;
      ctr[kSmFallback] = 0;
    }
    return;
  }

  // 2. LSD radix sort: each wave owns `per` consecutive 64-key items
  {
    const uint32_t per = (((n + kWave - 1) / kWave) + kSmWaves - 1) / kSmWaves;
    const uint32_t base = w * per * kWave;
    for (int p = 0; p < sh.passes; p++) {
      const int shift = sh.shift[p];
      for (int i = tid; i < kSmWaves * kSmBins; i += kSmT) (&wcnt[0][0])[i] = 0;
      __syncthreads();
      uint64_t kv[kSmItems];
      uint32_t rank[kSmItems];
#pragma unroll
      for (int j = 0; j < kSmItems; j++) {
        if ((uint32_t)j >= per) break;  // uniform across the block
        const uint32_t i = base + (uint32_t)j * kWave + lane;
        const bool valid = i < n;
        kv[j] = valid ? key[i] : 0;
        const uint32_t d = (uint32_t)(kv[j] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
          const uint64_t m = __ballot((d >> b) & 1u);
          peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const uint32_t before = wcnt[w][d];
        const uint32_t r = mask_rank(peers);
        rank[j] = before + r;
        if (valid && r == 0) wcnt[w][d] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
      }
      __syncthreads();
      uint32_t dtot = 0;
      if (tid < kSmBins) {
        uint32_t run = 0;
#pragma unroll
        for (int ww = 0; ww < kSmWaves; ww++) {
          const uint32_t c = wcnt[ww][tid];
          wcnt[ww][tid] = run;
          run += c;
        }
        dtot = run;
      }
      uint32_t tot = 0;
      const uint32_t ex = block_scan(dtot, wsum32, &tot);
      if (tid < kSmBins) tstart[tid] = ex;
      __syncthreads();
#pragma unroll
      for (int j = 0; j < kSmItems; j++) {
        if ((uint32_t)j >= per) break;
        const uint32_t i = base + (uint32_t)j * kWave + lane;
        if (i < n) {
          const uint32_t d = (uint32_t)(kv[j] >> shift) & 255u;
          key[tstart[d] + wcnt[w][d] + rank[j]] = kv[j];
        }
      }
      // every wave has placed its keys before wcnt and tstart are reset
      __syncthreads();
    }
  }

  // 3. keep: first copy, no self-loop on insert, absent (insert) or
  // present (remove).  The row searches are latency chains, so key i is
  // searched by thread i mod kSmT (one chain per thread for up to kSmT
  // keys); the decisions go through LDS to the compaction, whose threads
  // own contiguous keys [kSmItems * tid, +kSmItems)
  for (uint32_t i = tid; i < n; i += kSmT) {
    const uint64_t kk = key[i];
    bool k = i == 0 || key[i - 1] != kk;
    const uint32_t r = (uint32_t)(kk >> 32), c = (uint32_t)kk;
    if (insert && skip_self && r == c) k = false;
    uint32_t p = 0;
    if (k) {
      const uint32_t d = g.deg[r];
      const uint64_t o = g.off[r];
      bool exists = false;
      if (d) {
        const uint32_t *row = g.pool + o;
        p = lower_bound_4(row, d, c);
        exists = p < d && row[p] == c;
      }
      k = insert ? !exists : exists;
    }
    kflag[i] = k;
    pos[i] = p;
  }
  __syncthreads();
  uint64_t kv[kSmItems];
  uint32_t kp[kSmItems];
  bool keep[kSmItems];
  uint32_t cnt = 0;
#pragma unroll
  for (int j = 0; j < kSmItems; j++) {
    const uint32_t i = tid * kSmItems + j;
    keep[j] = i < n && kflag[i];
    kv[j] = i < n ? key[i] : 0;
    kp[j] = i < n ? pos[i] : 0;
    cnt += keep[j];
  }
  // 4. compact the kept keys in place (every thread holds its keys)
  uint32_t b = 0;
  uint32_t at = block_scan(cnt, wsum32, &b);
#pragma unroll
  for (int j = 0; j < kSmItems; j++)
    if (keep[j]) {
      key[at] = kv[j];
      pos[at] = kp[j];
      at++;
    }
  __syncthreads();
  if (b == 0) {
    if (tid == 0) {
      ctr[kSmErr] = 0;
      ctr[kSmFallback] = 0;
      ctr[kSmB] = 0;
      ctr[kSmT_] = 0;
      ctr[kSmAlloc] = 0;
      ctr[kSmGarbage] = 0;
      ctr[kSmSym] = 1;
    }
    return;
  }
  // rows: the runs of equal row among the kept keys
  uint32_t heads = 0;
#pragma unroll
  for (int j = 0; j < kSmItems; j++) {
    const uint32_t i = tid * kSmItems + j;
    heads += i < b && (i == 0 || (key[i] >> 32) != (key[i - 1] >> 32));
  }
  uint32_t t = 0;
  uint32_t r_at = block_scan(heads, wsum32, &t);
#pragma unroll
  for (int j = 0; j < kSmItems; j++) {
    const uint32_t i = tid * kSmItems + j;
    if (i < b && (i == 0 || (key[i] >> 32) != (key[i - 1] >> 32)))
      rstart[r_at++] = i;
  }
  if (tid == 0) rstart[t] = b;
  __syncthreads();
  // each touched row's degree and capacity, loaded row r by thread r mod kSmT
  for (uint32_t r = tid; r < t; r += kSmT) {
    const uint32_t v = (uint32_t)(key[rstart[r]] >> 32);
    rdeg[r] = g.deg[v];
    rcap[r] = g.cap[v];
  }
  __syncthreads();

  // 5. per row: slot, pool offset, garbage, length of its move
  uint64_t alloc[kSmItems];
  uint64_t asum = 0, garb = 0, worst = 0;
#pragma unroll
  for (int j = 0; j < kSmItems; j++) {
    const uint32_t r = tid * kSmItems + j;
    alloc[j] = 0;
    if (r >= t) continue;
    const uint32_t bs = rstart[r], blen = rstart[r + 1] - bs;
    const uint32_t od = rdeg[r], cp = rcap[r];
    uint64_t move;
    if (insert && (uint64_t)od + blen > cp) {
      alloc[j] = slot_capacity((uint64_t)od + blen, n_cols);
      garb += cp;
      move = od;
    } else {
      move = od - pos[bs];
    }
    asum += alloc[j];
    worst = move > worst ? move : worst;
  }
  uint64_t A = 0, G = 0;
  uint64_t aex = block_scan(asum, wsum64, &A);
#pragma unroll
  for (int j = 0; j < kSmItems; j++) {
    const uint32_t r = tid * kSmItems + j;
    if (r < t) {
      plan.aoff[r] = aex;
      aex += alloc[j];
    }
  }
  if (tid == 0) plan.aoff[t] = A;
  (void)block_scan(garb, wsum64, &G);
  // the longest move, by a max over the waves (wsum64 is free again)
  uint64_t wm = worst;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const uint64_t x = __shfl_xor(wm, o, kWave);
    wm = x > wm ? x : wm;
  }
  if (lane == 0) wsum64[w] = wm;
  __syncthreads();
  uint64_t M = 0;
#pragma unroll
  for (int i = 0; i < kSmWaves; i++) M = wsum64[i] > M ? wsum64[i] : M;
  if ((insert && pool_tail + A > pool_cap) || M > kSmMaxMove) {
    if (tid == 0) {
      ctr[kSmErr] = 0;
      ctr[kSmFallback] = 1;
    }
    return;  // nothing written
  }

  // 6. symmetry after the update.  It is only tracked from a symmetric
  // graph, where (v, u) was present exactly when (u, v) was: absent for a
  // kept insert key, present for a kept remove key.  So the graph stays
  // symmetric iff every kept key's reverse is kept too.
  if (track_sym) {
    bool ok = true;
#pragma unroll
    for (int j = 0; j < kSmItems; j++) {
      const uint32_t i = tid * kSmItems + j;
      if (i >= b) continue;
      const uint32_t u = (uint32_t)(key[i] >> 32), v = (uint32_t)key[i];
      ok &= lds_has(key, b, ((uint64_t)v << 32) | u);
    }
    if (!ok) s_sym = 0;
  }
  // the plan for the merge launch
  for (uint32_t i = tid; i < b; i += kSmT) {
    plan.key[i] = key[i];
    plan.pos[i] = pos[i];
  }
  for (uint32_t r = tid; r <= t; r += kSmT) plan.rstart[r] = rstart[r];
  __syncthreads();  // s_sym is final
  if (tid == 0) {
    ctr[kSmErr] = 0;
    ctr[kSmFallback] = 0;
    ctr[kSmB] = b;
    ctr[kSmT_] = t;
    ctr[kSmAlloc] = A;
    ctr[kSmGarbage] = G;
    ctr[kSmSym] = s_sym;
  }
}

// 7. one touched row at a time per workgroup; nothing to do when the plan
// stopped (error, fallback) or changed nothing
__global__ void __launch_bounds__(kSmT)
    k_update_small_merge(DevGraph g, int insert, uint64_t pool_tail,
                         SmallPlan plan, const uint64_t *ctr) {
  if (ctr[kSmErr] || ctr[kSmFallback]) return;
  const uint32_t t = (uint32_t)ctr[kSmT_];
  for (uint32_t r = blockIdx.x; r < t; r += gridDim.x) {
    const uint32_t bs = plan.rstart[r], k = plan.rstart[r + 1] - bs;
    const uint32_t v = (uint32_t)(plan.key[bs] >> 32);
    const uint32_t od = g.deg[v];
    const uint64_t a0 = plan.aoff[r], a1 = plan.aoff[r + 1];
    const bool reloc = insert && a1 > a0;
    uint32_t *row = g.pool + g.off[v];
    uint32_t *to = reloc ? g.pool + pool_tail + a0 : row;
    merge_row(row, to, od, plan.key + bs, plan.pos + bs, k, insert != 0,
              threadIdx.x, kSmT);
    __syncthreads();  // every thread has read deg[v] and off[v]
    if (threadIdx.x == 0) {
      g.deg[v] = insert ? od + k : od - k;
      if (reloc) {
        g.off[v] = pool_tail + a0;
        g.cap[v] = (uint32_t)(a1 - a0);
      }
    }
  }
}

}  // namespace

int update_small(sstg_graph *g, const uint32_t *src, const uint32_t *dst,
                 uint32_t n, bool insert, bool skip_self, bool *fallback) {
  *fallback = false;
  hipStream_t s = g->stream;
  SmallShifts sh{};
  const uint32_t cb = bit_width_u32(g->n_cols ? g->n_cols - 1 : 0);
  const uint32_t rb = bit_width_u32(g->n_rows ? g->n_rows - 1 : 0);
  for (uint32_t b = 0; b < cb; b += 8) sh.shift[sh.passes++] = (int)b;
  for (uint32_t b = 0; b < rb; b += 8) sh.shift[sh.passes++] = (int)(32 + b);
  const bool track = g->sym_version == g->version && g->symmetric;
  SmallPlan plan;
  SSTG_TRY(g->ws.get(kWsSmallKeys, kSmMax, &plan.key));
  SSTG_TRY(g->ws.get(kWsSmallPos, kSmMax, &plan.pos));
  SSTG_TRY(g->ws.get(kWsSmallRows, kSmMax + 1, &plan.rstart));
  SSTG_TRY(g->ws.get(kWsSmallAlloc, kSmMax + 1, &plan.aoff));
  SSTG_LAUNCH(k_update_small_plan, 1, kSmT, 0, s, src, dst, n, g->view(),
              g->n_cols, sh, insert ? 1 : 0, skip_self ? 1 : 0, track ? 1 : 0,
              g->pool_tail, g->pool_cap, plan, g->dev_ctr);
  SSTG_LAUNCH(k_update_small_merge, kSmMergeBlocks, kSmT, 0, s, g->view(),
              insert ? 1 : 0, g->pool_tail, plan,
              (const uint64_t *)g->dev_ctr);
  SSTG_HIP(hipMemcpyAsync(g->host_rb, g->dev_ctr, kSmCount * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, s));
  SSTG_HIP(hipStreamSynchronize(s));
  const uint64_t *h = g->host_rb;
  if (h[kSmErr]) return SSTG_E_RANGE;
  if (h[kSmFallback]) {
    *fallback = true;
    return SSTG_OK;
  }
  const uint64_t b = h[kSmB];
  g->last.effective += b;
  g->last.rows += h[kSmT_];
  g->last.allocated += h[kSmAlloc];
  if (b == 0) return SSTG_OK;
  // as the general pipeline: symmetry is tracked through batches that are
  // small next to the graph
  if (track && b * 8 <= g->m) {
    g->symmetric = h[kSmSym] != 0;
    g->sym_version = g->version + 1;
  } else {
    g->sym_version = ~0ull;
  }
  g->pool_tail += h[kSmAlloc];
  g->garbage += h[kSmGarbage];
  g->m = insert ? g->m + b : g->m - b;
  g->version++;
  return SSTG_OK;
}

}  // namespace sstg
