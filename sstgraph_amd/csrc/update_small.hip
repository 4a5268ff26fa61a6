// A whole batch update (SparseMatrixV<true>::insert_batch / remove_batch,
// SparseMatrix.hpp:556-718) of at most kSmallBatch pairs in one 1024-thread
// workgroup.
//
// The general pipeline (update.hip) sizes its grids from a read-back in the
// middle of the update and runs about fifteen launches; for a batch of a few
// thousand pairs those launches and the two host waits are the whole cost.
// Here three launches do every step: the row searches, spread over many
// CUs (step 0); one workgroup that plans the update with the batch held in
// LDS (steps 1-6); then a fixed grid of workgroups that merges the touched
// rows (step 7), one row per workgroup at a time:
//
//   0. (a launch of its own, one thread per pair over many CUs) search
//      each pair's row in HBM: present or not, its place, the row's degree,
//      capacity and pool offset, a 32-byte record per pair;
//   1. load and range-check the pairs, pack them as row << 32 | col keys;
//   2. sort in LDS (64-key bitonic runs merged by rank, block_sort.hpp),
//      each key carrying its pair's index;
//   3. keep the first copy of each key, drop self-loops on insert, and keep
//      keys absent (insert) or present (remove) in their row, from the
//      key's search record;
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
#include "block_sort.hpp"
#include "graph.hpp"

namespace sstg {
namespace {

constexpr int kSmT = 1024;
constexpr int kSmWaves = kSmT / kWave;
constexpr uint32_t kSmMax = kSmallBatch;
constexpr int kSmItems = kSmMax / kSmT;  // keys per thread
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

// Row searches: each round issues kFan probes per key and cuts the
// interval kFan + 1 ways, and a thread's K keys are searched side by side:
// every round issues the probes of all of them before it waits.  The value
// at the interval's upper end is kept, so the membership test needs no load
// after the last round.  A lone key's search is a chain of round trips, so
// wide rounds pay where the searches are spread over many CUs; run inside
// the plan's one workgroup, the searches were bound by that CU's memory
// pipeline (about one cache line per clock) and took 12.7-16.4 us for 1000
// keys and 42-61 us for 4000 at scale 24, whatever the fan-out.
constexpr int kFan = 8;

template <int K>
__device__ __forceinline__ void search_rows(const uint32_t *const row[K],
                                            const uint32_t len[K],
                                            const uint32_t x[K],
                                            const bool want[K], uint32_t p[K],
                                            bool found[K]) {
  // answer (first index with row[i] >= x) in [lo, hi]; row[hi] == vhi
  // when hi < len
  uint32_t lo[K], hi[K], vhi[K];
#pragma unroll
  for (int j = 0; j < K; j++) {
    lo[j] = 0;
    hi[j] = want[j] ? len[j] : 0;
    vhi[j] = 0;
  }
  for (;;) {
    uint32_t v[K][kFan];
    bool more = false;
#pragma unroll
    for (int j = 0; j < K; j++) {
      if (hi[j] - lo[j] > (uint32_t)kFan) {
        const uint32_t step = (hi[j] - lo[j]) / (kFan + 1);
#pragma unroll
        for (int m = 0; m < kFan; m++) v[j][m] = row[j][lo[j] + step * (m + 1)];
      }
    }
#pragma unroll
    for (int j = 0; j < K; j++) {
      if (hi[j] - lo[j] > (uint32_t)kFan) {
        more = true;
        const uint32_t step = (hi[j] - lo[j]) / (kFan + 1);
        uint32_t below = 0, vnext = 0;
#pragma unroll
        for (int m = 0; m < kFan; m++) below += v[j][m] < x[j];
#pragma unroll
        for (int m = 0; m < kFan; m++) vnext = below == (uint32_t)m ? v[j][m] : vnext;
        const uint32_t base = lo[j];
        if (below > 0) lo[j] = base + step * below + 1;
        if (below < (uint32_t)kFan) {
          hi[j] = base + step * (below + 1);
          vhi[j] = vnext;
        }
      }
    }
    if (!more) break;
  }
  // at most kFan candidates left: one more round of loads together
  uint32_t w[K][kFan];
#pragma unroll
  for (int j = 0; j < K; j++)
#pragma unroll
    for (int m = 0; m < kFan; m++)
      w[j][m] = lo[j] + m < hi[j] ? row[j][lo[j] + m] : 0u;
#pragma unroll
  for (int j = 0; j < K; j++) {
    uint32_t below = 0, at = 0;
#pragma unroll
    for (int m = 0; m < kFan; m++) below += lo[j] + m < hi[j] && w[j][m] < x[j];
#pragma unroll
    for (int m = 0; m < kFan; m++) at = below == (uint32_t)m ? w[j][m] : at;
    p[j] = lo[j] + below;
    found[j] = want[j] && (p[j] < hi[j] ? at == x[j]
                                        : (hi[j] < len[j] && vhi[j] == x[j]));
  }
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

// The counter block, to device memory (read by the merge launch) and, when
// hctr is not null, to mapped host memory, which the host reads after its
// wait with no copy launched.
__device__ __forceinline__ void put_counts(uint64_t *ctr, uint64_t *hctr,
                                           uint64_t err, uint64_t fallback,
                                           uint64_t b, uint64_t t, uint64_t a,
                                           uint64_t garbage, uint64_t sym) {
  const uint64_t v[kSmCount] = {err, fallback, b, t, a, garbage, sym};
#pragma unroll
  for (int k = 0; k < kSmCount; k++) {
    ctr[k] = v[k];
    if (hctr) hctr[k] = v[k];
  }
}

// Step 0: pair i's row searched for its column (one thread per pair, any
// order, duplicates and self-loops included).  found[2i] = (place, degree,
// capacity, present), found[2i + 1] = the row's pool offset (low, high
// word); pairs out of range get zeros (the plan rejects the batch).
constexpr int kSmSearchT = 64;  // small workgroups: the searches spread over CUs

__global__ void __launch_bounds__(kSmSearchT)
    k_update_small_search(const uint32_t *src, const uint32_t *dst, uint32_t n,
                          DevGraph g, uint32_t n_cols, uint4 *found) {
  const uint32_t i = blockIdx.x * kSmSearchT + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = src[i], c = dst[i];
  const uint32_t *row[1] = {g.pool};
  uint32_t len[1] = {0}, x[1] = {c}, p[1];
  bool want[1] = {r < g.n && c < n_cols}, hit[1];
  uint32_t cp = 0;
  uint64_t o = 0;
  if (want[0]) {
    len[0] = g.deg[r];
    cp = g.cap[r];
    o = g.off[r];
    row[0] = g.pool + o;
  }
  search_rows<1>(row, len, x, want, p, hit);
  found[2 * i] = make_uint4(p[0], len[0], cp, hit[0] ? 1u : 0u);
  found[2 * i + 1] = make_uint4((uint32_t)o, (uint32_t)(o >> 32), 0, 0);
}

// One touched row of a small update, as the merge launch reads it (one
// record, so a merge workgroup's first loads do not wait on one another).
struct SmallRow {
  uint64_t from, to;  // pool offset of the old slot and of the new one
  uint32_t v, od;     // row, its degree before the update
  uint32_t bs, k;     // its kept keys: plan key/pos [bs, bs + k)
  uint32_t cap, pad;  // capacity of a new slot
};

// The plan of a small update; each array holds up to kSmMax entries.
struct SmallPlan {
  uint64_t *key;     // kept keys, sorted
  uint32_t *pos;     // their places in the old rows
  SmallRow *rows;    // the touched rows
};

__global__ void __launch_bounds__(kSmT)
    k_update_small_plan(const uint32_t *src, const uint32_t *dst, uint32_t n,
                        DevGraph g, uint32_t n_cols,
                        int insert, int skip_self, int track_sym,
                        uint64_t pool_tail, uint64_t pool_cap,
                        const uint4 *found, SmallPlan plan, uint64_t *ctr,
                        uint64_t *hctr) {
  __shared__ uint64_t key[kSmMax];
  __shared__ uint32_t pos[kSmMax];
  __shared__ uint32_t rstart[kSmMax + 1];
  __shared__ uint8_t kflag[kSmMax];  // the keep decision of each sorted key
  __shared__ uint16_t kidx[kSmMax];  // each sorted key's pair index
  __shared__ uint32_t kdeg[kSmMax];  // each kept key's row: degree,
  __shared__ uint32_t kcap[kSmMax];  // capacity and pool offset
  __shared__ uint64_t koff[kSmMax];
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
  {  // a thread's pairs in flight together
    uint32_t va[kSmItems], vb[kSmItems];
#pragma unroll
    for (int j = 0; j < kSmItems; j++) {
      const uint32_t i = tid + (uint32_t)j * kSmT;
      va[j] = i < n ? src[i] : 0;
      vb[j] = i < n ? dst[i] : 0;
    }
#pragma unroll
    for (int j = 0; j < kSmItems; j++) {
      const uint32_t i = tid + (uint32_t)j * kSmT;
      if (i < n) {
        if (va[j] >= g.n || vb[j] >= n_cols) s_bad = 1;
        key[i] = ((uint64_t)va[j] << 32) | vb[j];
        kidx[i] = (uint16_t)i;
      }
    }
  }
  __syncthreads();
  if (s_bad) {
    if (tid == 0) {
      put_counts(ctr, hctr, 1, 0, 0, 0, 0, 0, 0);
    }
    return;
  }

  // 2. sort in LDS (block_sort.hpp), each key with its pair index
  block_sort_u64_u16<kSmT>(key, kidx, n);

  // 3. keep: first copy, no self-loop on insert, absent (insert) or
  // present (remove), from the key's search record (written by other CUs,
  // so read once, here); the decisions, places and the rows' degree,
  // capacity and offset go through LDS to the compaction, whose threads own
  // contiguous keys [kSmItems * tid, +kSmItems)
  for (uint32_t i = tid; i < n; i += kSmT) {
    const uint64_t kk = key[i];
    const uint32_t r = (uint32_t)(kk >> 32), c = (uint32_t)kk;
    const bool want =
        (i == 0 || key[i - 1] != kk) && !(insert && skip_self && r == c);
    uint4 f = make_uint4(0, 0, 0, 0), fo = f;
    if (want) {
      f = found[2 * kidx[i]];
      fo = found[2 * kidx[i] + 1];
    }
    kflag[i] = want && (insert ? f.w == 0 : f.w != 0);
    pos[i] = f.x;
    kdeg[i] = f.y;
    kcap[i] = f.z;
    koff[i] = ((uint64_t)fo.y << 32) | fo.x;
  }
  __syncthreads();
  uint64_t kv[kSmItems], ko[kSmItems];
  uint32_t kp[kSmItems], kd[kSmItems], kc[kSmItems];
  bool keep[kSmItems];
  uint32_t cnt = 0;
#pragma unroll
  for (int j = 0; j < kSmItems; j++) {
    const uint32_t i = tid * kSmItems + j;
    keep[j] = i < n && kflag[i];
    kv[j] = i < n ? key[i] : 0;
    kp[j] = keep[j] ? pos[i] : 0;
    kd[j] = keep[j] ? kdeg[i] : 0;
    kc[j] = keep[j] ? kcap[i] : 0;
    ko[j] = keep[j] ? koff[i] : 0;
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
      kdeg[at] = kd[j];
      kcap[at] = kc[j];
      koff[at] = ko[j];
      at++;
    }
  __syncthreads();
  if (b == 0) {
    if (tid == 0) {
      put_counts(ctr, hctr, 0, 0, 0, 0, 0, 0, 1);
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

  // 5. per row: slot, pool offset, garbage, length of its move
  // (the row's degree, capacity and pool offset: those of its first key)
  uint64_t alloc[kSmItems], from[kSmItems];
  uint32_t rdeg[kSmItems], rcap[kSmItems];
#pragma unroll
  for (int j = 0; j < kSmItems; j++) {
    const uint32_t r = tid * kSmItems + j;
    const uint32_t i = r < t ? rstart[r] : 0;
    rdeg[j] = r < t ? kdeg[i] : 0;
    rcap[j] = r < t ? kcap[i] : 0;
    from[j] = r < t ? koff[i] : 0;
  }
  uint64_t asum = 0, garb = 0, worst = 0;
#pragma unroll
  for (int j = 0; j < kSmItems; j++) {
    const uint32_t r = tid * kSmItems + j;
    alloc[j] = 0;
    if (r >= t) continue;
    const uint32_t bs = rstart[r], blen = rstart[r + 1] - bs;
    const uint32_t od = rdeg[j], cp = rcap[j];
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
      SmallRow d;
      d.from = from[j];
      d.to = alloc[j] ? pool_tail + aex : from[j];
      d.v = (uint32_t)(key[rstart[r]] >> 32);
      d.od = rdeg[j];
      d.bs = rstart[r];
      d.k = rstart[r + 1] - rstart[r];
      d.cap = (uint32_t)alloc[j];
      d.pad = 0;
      plan.rows[r] = d;
      aex += alloc[j];
    }
  }
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
      put_counts(ctr, hctr, 0, 1, 0, 0, 0, 0, 0);
    }
    return;  // nothing written
  }

  // 6. symmetry after the update.  It is only tracked from a symmetric
  // graph, where (v, u) was present exactly when (u, v) was: absent for a
  // kept insert key, present for a kept remove key.  So the graph stays
  // symmetric iff every kept key's reverse is kept too.
  // Keys tid, tid + kSmT, ... per thread (every thread busy for small b),
  // searched side by side with fixed power-of-two steps.
  if (track_sym) {
    uint64_t rev[kSmItems];
    uint32_t at[kSmItems];
#pragma unroll
    for (int j = 0; j < kSmItems; j++) {
      const uint32_t i = tid + (uint32_t)j * kSmT;
      const uint64_t kk = i < b ? key[i] : 0;
      rev[j] = (kk << 32) | (kk >> 32);
      at[j] = 0;
    }
    uint32_t top = 1;
    while (top < b) top <<= 1;
    for (uint32_t h = top >> 1; h > 0; h >>= 1) {
#pragma unroll
      for (int j = 0; j < kSmItems; j++)
        if (at[j] + h <= b && key[at[j] + h - 1] < rev[j]) at[j] += h;
    }
    bool ok = true;
#pragma unroll
    for (int j = 0; j < kSmItems; j++) {
      const uint32_t i = tid + (uint32_t)j * kSmT;
      if (i < b) ok &= at[j] < b && key[at[j]] == rev[j];
    }
    if (!ok) s_sym = 0;
  }
  // the plan for the merge launch (the rows are written above)
  for (uint32_t i = tid; i < b; i += kSmT) {
    plan.key[i] = key[i];
    plan.pos[i] = pos[i];
  }
  __syncthreads();  // s_sym is final
  if (tid == 0) {
    put_counts(ctr, hctr, 0, 0, b, t, A, G, s_sym);
  }
}

// 7. one touched row at a time per workgroup; nothing to do when the plan
// stopped (error, fallback) or changed nothing.  A workgroup's first row
// record is loaded beside the counters (the grid never exceeds kSmMax rows,
// so the load is in bounds; the record is used only if the row exists), and
// the row's key places are staged in LDS for the shift searches.
__global__ void __launch_bounds__(kSmT)
    k_update_small_merge(DevGraph g, int insert, SmallPlan plan,
                         const uint64_t *ctr) {
  __shared__ uint32_t spos[kSmMax];
  SmallRow d = plan.rows[blockIdx.x];
  if (ctr[kSmErr] || ctr[kSmFallback]) return;
  const uint32_t t = (uint32_t)ctr[kSmT_];
  for (uint32_t r = blockIdx.x; r < t; r += gridDim.x) {
    if (r != blockIdx.x) d = plan.rows[r];
    for (uint32_t i = threadIdx.x; i < d.k; i += kSmT) spos[i] = plan.pos[d.bs + i];
    __syncthreads();
    merge_row(g.pool + d.from, g.pool + d.to, d.od, plan.key + d.bs, spos, d.k,
              insert != 0, threadIdx.x, kSmT);
    __syncthreads();  // spos is read by every thread before the next row
    if (threadIdx.x == 0) {
      g.deg[d.v] = insert ? d.od + d.k : d.od - d.k;
      if (d.to != d.from) {
        g.off[d.v] = d.to;
        g.cap[d.v] = d.cap;
      }
    }
  }
}

}  // namespace

int update_small(sstg_graph *g, const uint32_t *src, const uint32_t *dst,
                 uint32_t n, bool insert, bool skip_self, bool *fallback) {
  *fallback = false;
  hipStream_t s = g->stream;
  const bool track = g->sym_version == g->version && g->symmetric;
  SmallPlan plan;
  uint4 *found = nullptr;
  SSTG_TRY(g->ws.get(kWsSmallFound, 2 * kSmMax, &found));
  SSTG_TRY(g->ws.get(kWsSmallKeys, kSmMax, &plan.key));
  SSTG_TRY(g->ws.get(kWsSmallPos, kSmMax, &plan.pos));
  SSTG_TRY(g->ws.get(kWsSmallRows, kSmMax, &plan.rows));
  if (n)
    SSTG_LAUNCH(k_update_small_search, ceil_div(n, (uint32_t)kSmSearchT),
                kSmSearchT, 0, s, src, dst, n, g->view(), g->n_cols, found);
  SSTG_LAUNCH(k_update_small_plan, 1, kSmT, 0, s, src, dst, n, g->view(),
              g->n_cols, insert ? 1 : 0, skip_self ? 1 : 0, track ? 1 : 0,
              g->pool_tail, g->pool_cap, (const uint4 *)found, plan,
              g->dev_ctr, g->host_ctr_dev);
  static_assert(kSmMergeBlocks <= kSmMax, "first row records in bounds");
  SSTG_LAUNCH(k_update_small_merge, kSmMergeBlocks, kSmT, 0, s, g->view(),
              insert ? 1 : 0, plan, (const uint64_t *)g->dev_ctr);
  const uint64_t *h = g->host_ctr;
  if (!g->host_ctr_dev) {
    SSTG_HIP(hipMemcpyAsync(g->host_rb, g->dev_ctr, kSmCount * sizeof(uint64_t),
                            hipMemcpyDeviceToHost, s));
    h = g->host_rb;
  }
  SSTG_HIP(hipStreamSynchronize(s));
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
