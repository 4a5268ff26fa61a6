// Batch edge insert / delete on the device-resident tinyset store.
//
// Reference semantics: SparseMatrixV<true>::insert_batch (SparseMatrix.hpp:
// 556-643) makes each row old U {y : (x,y) in batch, x != y}; remove_batch
// (SparseMatrix.hpp:645-718) makes it old \ batch.  Duplicate inserts and
// absent removes are no-ops (PMA.hpp:1139-1154, 1226-1228).
//
// Pipeline for one batch of n (src, dst) pairs, all on the graph's stream:
//   1. pack      keys = src << 32 | dst, range check
//   2. sort      LSD radix sort over the row and column bits
//   3. mark      keep first copy of each key; drop x==y (insert); binary
//                search the row: insert keeps absent keys, remove keeps
//                present ones
//   4. select    compact the kept keys -> B (sorted, unique, effective)
//   5. heads     row boundaries in B -> one segment per touched row
//   6. segprep   per row: new degree, in place or relocate, first moved
//                position p0, staged suffix length, output work length
//   7. scans     alloc / stage / work / tile offsets, the big and tile row
//                classes compacted (totals read back once)
//   8. stage     copy each in-place row's suffix [p0, deg) to scratch
//   9. merge     insert: merge-path merge of (suffix or whole old row) with
//                the row's batch run into the row slot (or its new slot);
//                remove: compact the staged suffix minus the batch run
//  10. finalize  deg / off / cap of the touched rows
// Steps 8 and 9 run over a flat index space (one thread per 8 outputs), so a
// hub row with a million neighbours and a thousand one-edge rows are balanced
// the same way.
#include <atomic>
#include <chrono>
#include <cstring>

#include "graph.hpp"
#include "primitives.hpp"

namespace sstg {

namespace {

constexpr int kT = 256;
constexpr int kChunk = 8;  // outputs per thread in the flat kernels

// Counter block (graph.dev_ctr):
enum {
  kCtrErr = 0,
  kCtrB = 1,        // effective batch elements
  kCtrT = 2,        // touched rows
  kCtrStage = 3,    // staged elements
  kCtrAlloc = 4,    // elements allocated at the pool tail
  kCtrWork = 5,     // flat work items of the merge step
  kCtrGarbage = 6,  // elements in slots abandoned by relocation
  kCtrScratch = 7,
  kCtrBig = 12,     // big rows
  kCtrTile = 13,    // elements (inputs) merged by the LDS tile kernel
  kCtrTiles = 14,   // tile rows
  kCtrCount = 16,
};

// Rows whose merge inputs (suffix + batch run) fit in kTileMax elements are
// merged in place by one block through LDS; larger rows are staged.  A block
// owns the rows whose inputs start in its kTile range, so its tile holds at
// most kTile + kTileMax inputs.
// slots (live elements of capacity) below which removal never compacts
constexpr uint64_t kShrinkMinElems = 1ull << 20;
constexpr int kTile = 1536;
constexpr int kTileMax = 1536;
constexpr int kTileLds = kTile + kTileMax;

__device__ __forceinline__ uint64_t grow_cap(uint64_t d, uint32_t n_cols) {
  return slot_capacity(d, n_cols);
}

// ctr[kCtrErr] = 1 if any pair of the batch is out of range (the whole-batch
// check of a batch applied in pieces)
__global__ void k_check_range(const uint32_t *src, const uint32_t *dst,
                              uint64_t n, uint32_t n_rows, uint32_t n_cols,
                              uint64_t *ctr) {
  const uint64_t stride = (uint64_t)gridDim.x * kT;
  bool bad = false;
  for (uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x; i < n; i += stride)
    bad |= (src[i] >= n_rows) | (dst[i] >= n_cols);
  if (bad) ctr[kCtrErr] = 1;
}

// pos[i] = the row position where key i sits or would go (its lower bound
// in the old row); the segment step takes each row's first moved position
// from the row's first kept key.
__global__ void k_mark(const uint64_t *keys, uint64_t n, DevGraph g,
                       int insert, int skip_self, uint8_t *flags,
                       uint32_t *pos) {
  uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  uint64_t k = keys[i];
  bool keep = (i == 0) || (keys[i - 1] != k);
  uint32_t s = (uint32_t)(k >> 32), d = (uint32_t)k;
  if (insert && skip_self && s == d) keep = false;
  uint32_t p = 0;
  if (keep) {  // degree and offset loaded together, no load after the search
    const uint32_t dg = g.deg[s];
    const uint32_t *row = g.pool + g.off[s];
    const bool exists = row_has(row, dg, d, &p);
    keep = insert ? !exists : exists;
  }
  flags[i] = keep ? 1 : 0;
  pos[i] = p;
}

__global__ void k_heads(const uint64_t *B, const uint64_t *ctr, uint64_t nmax,
                        uint8_t *flags) {
  uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= nmax) return;
  uint64_t b = ctr[kCtrB];
  flags[i] = (i < b) && (i == 0 || (B[i] >> 32) != (B[i - 1] >> 32));
}

// One thread per touched row (s < t); slot t gets zero lengths, which the
// scans read as the end of the rows.  Each row is classed as a tile row
// (merge inputs <= kTileMax, merged in place through LDS) or a big row
// (suffix staged, merged over a flat index space).  p0 is the row position
// k_mark found for the row's first kept key.
__global__ void k_segprep(const uint64_t *B, const uint32_t *B_pos,
                          const uint64_t *ctr,
                          const uint64_t *seg_idx, DevGraph g, int insert,
                          uint32_t n_cols, uint64_t nmax, uint32_t *seg_v,
                          uint32_t *seg_p0, uint32_t *seg_olddeg,
                          uint64_t *seg_alloc, uint32_t *seg_stage,
                          uint64_t *seg_work, uint32_t *seg_tile,
                          uint8_t *is_big, uint8_t *is_tile,
                          uint64_t *garbage) {
  uint64_t s = (uint64_t)blockIdx.x * kT + threadIdx.x;
  // every lane stays to the end: the abandoned slots of relocated rows are
  // summed per wave and added with one atomic per wave (one per relocated
  // row serialised on the counter when many rows move)
  uint64_t gb = 0;
  if (s <= nmax) {
    const uint64_t t = ctr[kCtrT], b = ctr[kCtrB];
    if (s >= t) {  // slot t is read as zero by the scans; later slots are not
      if (s == t) {
        seg_alloc[s] = 0;
        is_big[s] = 0;
        is_tile[s] = 0;
      }
    } else {
      uint64_t bs = seg_idx[s], be = (s + 1 < t) ? seg_idx[s + 1] : b;
      uint64_t blen = be - bs;
      uint32_t v = (uint32_t)(B[bs] >> 32);
      uint32_t od = g.deg[v], cp = g.cap[v];
      uint64_t p0 = 0, stage = 0, alloc = 0, work = 0;
      if (insert) {
        uint64_t nd = (uint64_t)od + blen;
        if (nd > cp) {  // relocate: whole row merged into a fresh slot
          alloc = grow_cap(nd, n_cols);
          work = nd;
          gb = cp;
        } else {
          p0 = od ? B_pos[bs] : 0;
          stage = od - p0;
          work = stage + blen;
        }
      } else {
        p0 = B_pos[bs];
        stage = od - p0;
        work = stage;
      }
      // merge inputs: the moved suffix (or the whole row when relocating)
      // plus the batch run
      uint64_t inputs = (alloc ? (uint64_t)od : stage) + blen;
      bool tile = inputs <= (uint64_t)kTileMax;
      seg_v[s] = v;
      seg_p0[s] = (uint32_t)p0;
      seg_olddeg[s] = od;
      seg_alloc[s] = alloc;
      seg_stage[s] = (uint32_t)stage;
      seg_work[s] = work;
      seg_tile[s] = (uint32_t)inputs;
      is_big[s] = !tile;
      is_tile[s] = tile;
    }
  }
  gb = wave_sum(gb);
  if ((threadIdx.x & (kWave - 1)) == 0 && gb)
    atomicAdd((unsigned long long *)garbage, (unsigned long long)gb);
}

// Step 7 in three launches: the six running sums of the segment slots
// (slots at or past t are zero) are scanned together -- pool elements
// allocated, big rows, tile rows, and the staged elements and merge work of
// big rows and the inputs of tile rows.  The same pass compacts the two
// classes: big row k gets big_idx[k] = its segment and its stage and work
// offsets; tile row k gets tile_idx[k] and its input offset; each offset
// array ends with its total (entry nb or nt).  Each block of a fixed grid
// (at most kSegBlocks) owns a contiguous range of slots: a reduce writes its
// six partials, one block scans them, and each block then scans its range.
constexpr int kSegSums = 6;
constexpr int kSegItems = 4;
constexpr int kSegTile = kT * kSegItems;  // slots per step of a block
constexpr uint64_t kSegBlocks = 2048;
constexpr uint64_t kSegOneBlock = 8;  // up to this many tiles: one block

struct SegScan {
  uint64_t *alloc;  // in: per slot; out: exclusive offsets
  const uint8_t *is_big, *is_tile;
  const uint32_t *stage, *tile;
  const uint64_t *work;
  uint64_t n, per;     // slots (nseg + 1), slots per block
  uint64_t *partials;  // kSegSums x blocks, sum-major; null: one block
  uint32_t blocks;
  uint64_t *big_idx, *big_stage, *big_work, *tile_idx, *tile_len;
  uint64_t *ctr;
  uint64_t *hctr;  // mapped host copy of the counter block, or null
  uint64_t seq;    // written to hctr[kHostFlag] after the counters
};

// The six running sums of the K slots s0, s0 + 1, ... below hi (allocation,
// big and tile row flags, big rows' staging and work, tile inputs; zeros
// past hi), their loads issued together: the class flags, allocations and
// tile inputs (most rows are tile rows) in one round, the big rows' staging
// and work in a second (one slot at a time, the loads of a slot depended on
// its flags).
template <int K>
__device__ __forceinline__ void seg_sums_run(const SegScan &a, uint64_t s0,
                                             uint64_t hi,
                                             uint64_t v[K][kSegSums]) {
  uint8_t big[K], til[K];
  uint64_t al[K];
  uint32_t tl[K];
#pragma unroll
  for (int j = 0; j < K; j++) {
    const bool in = s0 + j < hi;
    big[j] = in ? a.is_big[s0 + j] : 0;
    til[j] = in ? a.is_tile[s0 + j] : 0;
    al[j] = in ? a.alloc[s0 + j] : 0;
    tl[j] = in ? a.tile[s0 + j] : 0;
  }
#pragma unroll
  for (int j = 0; j < K; j++) {
    v[j][0] = al[j];
    v[j][1] = big[j] != 0;
    v[j][2] = til[j] != 0;
    v[j][3] = big[j] ? a.stage[s0 + j] : 0;
    v[j][4] = big[j] ? a.work[s0 + j] : 0;
    v[j][5] = til[j] ? tl[j] : 0;
  }
}

// End of a block's range.  Slots past t (ctr[kCtrT]) are zero and only slot
// t's alloc offset (the total) is read after the scan, so ranges stop at t.
__device__ __forceinline__ uint64_t seg_end(const SegScan &a, uint64_t lo) {
  uint64_t hi = lo + a.per < a.n ? lo + a.per : a.n;
  const uint64_t t1 = a.ctr[kCtrT] + 1;
  return hi < t1 ? hi : t1;
}

__global__ void __launch_bounds__(kT) k_seg_reduce(SegScan a) {
  __shared__ uint64_t wtot[kT / kWave][kSegSums];
  const uint64_t lo = (uint64_t)blockIdx.x * a.per;
  const uint64_t hi = seg_end(a, lo);
  uint64_t acc[kSegSums] = {0, 0, 0, 0, 0, 0};
  for (uint64_t c = lo; c < hi; c += kSegTile) {
    uint64_t v[kSegItems][kSegSums];
    seg_sums_run<kSegItems>(a, c + (uint64_t)threadIdx.x * kSegItems, hi, v);
#pragma unroll
    for (int j = 0; j < kSegItems; j++)
#pragma unroll
      for (int k = 0; k < kSegSums; k++) acc[k] += v[j][k];
  }
#pragma unroll
  for (int k = 0; k < kSegSums; k++) {
    const uint64_t w = wave_sum(acc[k]);
    if ((threadIdx.x & (kWave - 1)) == 0) wtot[threadIdx.x / kWave][k] = w;
  }
  __syncthreads();
  if (threadIdx.x < kSegSums) {
    uint64_t t = 0;
#pragma unroll
    for (int w = 0; w < kT / kWave; w++) t += wtot[w][threadIdx.x];
    a.partials[(uint64_t)threadIdx.x * a.blocks + blockIdx.x] = t;
  }
}

// One block: exclusive scan in place of each of the six partial arrays.
__global__ void __launch_bounds__(kT) k_seg_scan_partials(SegScan a) {
  constexpr int kPer = (int)(kSegBlocks / kT);
  __shared__ uint64_t wtot[kT / kWave];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const uint32_t base = threadIdx.x * kPer;
  for (int k = 0; k < kSegSums; k++) {
    uint64_t *p = a.partials + (uint64_t)k * a.blocks;
    uint64_t v[kPer], sum = 0;
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      v[j] = base + j < a.blocks ? p[base + j] : 0;
      sum += v[j];
    }
    const uint64_t incl = wave_inclusive_sum(sum);
    if (lane == kWave - 1) wtot[w] = incl;
    __syncthreads();
    uint64_t run = incl - sum;
    for (int i = 0; i < w; i++) run += wtot[i];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      if (base + j < a.blocks) p[base + j] = run;
      run += v[j];
    }
  }
}

__global__ void __launch_bounds__(kT) k_seg_scan(SegScan a) {
  __shared__ uint64_t wtot[kT / kWave][kSegSums];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const uint64_t lo = (uint64_t)blockIdx.x * a.per;
  const uint64_t hi = seg_end(a, lo);
  uint64_t carry[kSegSums];
#pragma unroll
  for (int k = 0; k < kSegSums; k++)
    carry[k] = a.partials ? a.partials[(uint64_t)k * a.blocks + blockIdx.x]
                          : 0;
  for (uint64_t c = lo; c < hi; c += kSegTile) {
    // kSegItems consecutive slots per thread: one block scan per step
    const uint64_t i0 = c + (uint64_t)threadIdx.x * kSegItems;
    uint64_t v[kSegItems][kSegSums], ex[kSegSums];
#pragma unroll
    for (int k = 0; k < kSegSums; k++) ex[k] = 0;
    seg_sums_run<kSegItems>(a, i0, hi, v);
#pragma unroll
    for (int j = 0; j < kSegItems; j++) {
#pragma unroll
      for (int k = 0; k < kSegSums; k++) ex[k] += v[j][k];
    }
#pragma unroll
    for (int k = 0; k < kSegSums; k++) {
      const uint64_t incl = wave_inclusive_sum(ex[k]);
      if (lane == kWave - 1) wtot[w][k] = incl;
      ex[k] = incl - ex[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSegSums; k++) {
      uint64_t woff = 0, tot = 0;
#pragma unroll
      for (int ww = 0; ww < kT / kWave; ww++) {
        const uint64_t x = wtot[ww][k];
        if (ww < w) woff += x;
        tot += x;
      }
      ex[k] += carry[k] + woff;
      carry[k] += tot;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSegItems; j++) {
      const uint64_t i = i0 + j;
      if (i < hi) {
        a.alloc[i] = ex[0];
        if (v[j][1]) {
          a.big_idx[ex[1]] = i;
          a.big_stage[ex[1]] = ex[3];
          a.big_work[ex[1]] = ex[4];
        }
        if (v[j][2]) {
          a.tile_idx[ex[2]] = i;
          a.tile_len[ex[2]] = ex[5];
        }
      }
#pragma unroll
      for (int k = 0; k < kSegSums; k++) ex[k] += v[j][k];
    }
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    a.ctr[kCtrAlloc] = carry[0];
    a.ctr[kCtrBig] = carry[1];
    a.ctr[kCtrTiles] = carry[2];
    a.ctr[kCtrStage] = carry[3];
    a.ctr[kCtrWork] = carry[4];
    a.ctr[kCtrTile] = carry[5];
    a.big_stage[carry[1]] = carry[3];
    a.big_work[carry[1]] = carry[4];
    a.tile_len[carry[2]] = carry[5];
    // the whole block, earlier launches' counters included, for the host
    if (a.hctr) {
      for (int k = 0; k < kCtrCount; k++) a.hctr[k] = a.ctr[k];
      __threadfence_system();  // the counters before the flag, to the host
      *(volatile uint64_t *)(a.hctr + kHostFlag) = a.seq;
    }
  }
}

struct SegArgs {
  const uint64_t *B;
  const uint64_t *seg_idx;
  const uint32_t *seg_v;
  const uint32_t *seg_p0;
  const uint32_t *seg_olddeg;
  const uint64_t *alloc_off;  // per segment, exclusive, t + 1 entries
  uint64_t t, b;
  uint64_t pool_tail;
  uint32_t *scratch;
  // big rows, compacted: segment ids and exclusive stage / work offsets
  const uint64_t *big_idx;
  const uint64_t *stage_off;
  const uint64_t *work_off;
  uint64_t nb;
  // tile rows, compacted: segment ids and exclusive input offsets
  const uint64_t *tile_idx;
  const uint64_t *tile_off;
  uint64_t nt;
};

// First index in [0, n] with a[index] >= x, found by one full wave with
// 64-way probes (about log64(n) rounds).  a is non-decreasing.
__device__ __forceinline__ uint64_t wave_lower_bound(const uint64_t *a,
                                                     uint64_t n, uint64_t x) {
  const uint32_t lane = lane_id();
  uint64_t lo = 0, hi = n;  // answer in [lo, hi]
  // the last round probes lo .. lo + 63 and must also see hi itself (as
  // "past the end"), so it needs hi - lo < 64: with hi - lo == 64 and the
  // answer at hi no lane would report it
  while (hi - lo >= (uint64_t)kWave) {
    uint64_t step = (hi - lo + kWave - 1) / kWave;
    uint64_t p = lo + lane * step;
    bool ge = p >= hi || a[p] >= x;
    uint64_t m = __ballot(ge);
    uint32_t f = m ? (uint32_t)__ffsll((unsigned long long)m) - 1 : kWave;
    if (f == 0) return lo;
    uint64_t nlo = lo + (uint64_t)(f - 1) * step + 1;
    uint64_t nhi = lo + (uint64_t)f * step;
    lo = nlo;
    hi = nhi < hi ? nhi : hi;
  }
  uint64_t p = lo + lane;
  bool ge = p >= hi || a[p] >= x;
  uint64_t m = __ballot(ge);
  return lo + (uint32_t)__ffsll((unsigned long long)m) - 1;
}

// Segment range touched by the flat range [e_lo, e_hi) of this block:
// segments [s_lo, s_hi] of the k offsets off[0..k] (wave 0 searches, the
// block reads the result from LDS).
__device__ __forceinline__ void block_segments(const uint64_t *off, uint64_t k,
                                               uint64_t e_lo, uint64_t e_hi,
                                               uint64_t *s_lo, uint64_t *s_hi) {
  __shared__ uint64_t r[2];
  if (threadIdx.x < kWave) {
    // largest s with off[s] <= e  ==  lower_bound(off, k + 1, e + 1) - 1
    uint64_t a = wave_lower_bound(off, k + 1, e_lo + 1) - 1;
    uint64_t b = wave_lower_bound(off, k + 1, e_hi) - 1;
    if (threadIdx.x == 0) {
      r[0] = a;
      r[1] = b < k ? b : k - 1;
    }
  }
  __syncthreads();
  *s_lo = r[0];
  *s_hi = r[1];
}

// largest s in [lo, hi] with off[s] <= x
__device__ __forceinline__ uint64_t find_in(const uint64_t *off, uint64_t lo,
                                            uint64_t hi, uint64_t x) {
  uint64_t h = hi + 1;
  while (h - lo > 1) {
    uint64_t m = (lo + h) >> 1;
    if (off[m] <= x)
      lo = m;
    else
      h = m;
  }
  return lo;
}

__global__ void __launch_bounds__(kT)
    k_stage(SegArgs a, DevGraph g, uint64_t total) {
  const uint64_t blo = (uint64_t)blockIdx.x * kT * kChunk;
  const uint64_t bhi = blo + (uint64_t)kT * kChunk < total
                           ? blo + (uint64_t)kT * kChunk : total;
  uint64_t s_lo, s_hi;
  block_segments(a.stage_off, a.nb, blo, bhi, &s_lo, &s_hi);
  uint64_t e = blo + (uint64_t)threadIdx.x * kChunk;
  if (e >= total) return;
  uint64_t end = e + kChunk < total ? e + kChunk : total;
  uint64_t i = find_in(a.stage_off, s_lo, s_hi, e);
  // the row's source offset once per row (big rows are long, so a thread's
  // elements nearly always share it), then every load before any store
  auto src_of = [&](uint64_t r) {
    const uint64_t s = a.big_idx[r];
    return g.off[a.seg_v[s]] + a.seg_p0[s] - a.stage_off[r];
  };
  uint64_t next = a.stage_off[i + 1], base = src_of(i);
  uint32_t val[kChunk];
#pragma unroll
  for (int k = 0; k < kChunk; k++) {
    const uint64_t ek = e + k;
    val[k] = 0;
    if (ek < end) {
      if (ek >= next) {
        i = find_in(a.stage_off, i, s_hi, ek);
        next = a.stage_off[i + 1];
        base = src_of(i);
      }
      val[k] = g.pool[base + ek];
    }
  }
#pragma unroll
  for (int k = 0; k < kChunk; k++)
    if (e + k < end) a.scratch[e + k] = val[k];
}

struct MergeSeg {
  const uint32_t *A;
  uint64_t alen;
  const uint64_t *Bk;
  uint64_t blen;
  uint32_t *dst;
  uint64_t begin, end;  // flat work range
};

// Big row i (compacted index): inputs, output and flat work range.
__device__ __forceinline__ MergeSeg load_seg(const SegArgs &a, const DevGraph &g,
                                             uint64_t i, bool insert) {
  MergeSeg m;
  uint64_t s = a.big_idx[i];
  uint64_t bs = a.seg_idx[s], be = (s + 1 < a.t) ? a.seg_idx[s + 1] : a.b;
  m.Bk = a.B + bs;
  m.blen = be - bs;
  uint32_t v = a.seg_v[s];
  bool reloc = insert && (a.alloc_off[s + 1] > a.alloc_off[s]);
  if (reloc) {
    m.A = g.pool + g.off[v];
    m.alen = a.seg_olddeg[s];
    m.dst = g.pool + a.pool_tail + a.alloc_off[s];
  } else {
    m.A = a.scratch + a.stage_off[i];
    m.alen = a.stage_off[i + 1] - a.stage_off[i];
    m.dst = g.pool + g.off[v] + a.seg_p0[s];
  }
  m.begin = a.work_off[i];
  m.end = a.work_off[i + 1];
  return m;
}

__global__ void __launch_bounds__(kT)
    k_merge_insert(SegArgs a, DevGraph g, uint64_t total) {
  const uint64_t blo = (uint64_t)blockIdx.x * kT * kChunk;
  const uint64_t bhi = blo + (uint64_t)kT * kChunk < total
                           ? blo + (uint64_t)kT * kChunk : total;
  uint64_t s_lo, s_hi;
  block_segments(a.work_off, a.nb, blo, bhi, &s_lo, &s_hi);
  uint64_t o = blo + (uint64_t)threadIdx.x * kChunk;
  if (o >= total) return;
  uint64_t end = o + kChunk < total ? o + kChunk : total;
  uint64_t i = find_in(a.work_off, s_lo, s_hi, o);
  MergeSeg m = load_seg(a, g, i, true);
  // merge-path split of output index j between A and the row's batch run
  uint64_t j = o - m.begin;
  uint64_t lo = j > m.blen ? j - m.blen : 0, hi = j < m.alen ? j : m.alen;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (m.A[mid] < (uint32_t)m.Bk[j - 1 - mid])
      lo = mid + 1;
    else
      hi = mid;
  }
  uint64_t ia = lo, ib = j - lo;
  for (; o < end; o++) {
    if (o >= m.end) {  // big rows all have work > 0: the next row starts here
      i++;
      m = load_seg(a, g, i, true);
      j = 0;
      ia = 0;
      ib = 0;
    }
    uint32_t x;
    if (ib >= m.blen || (ia < m.alen && m.A[ia] < (uint32_t)m.Bk[ib]))
      x = m.A[ia++];
    else
      x = (uint32_t)m.Bk[ib++];
    m.dst[j++] = x;
  }
}

__global__ void __launch_bounds__(kT)
    k_compact_remove(SegArgs a, DevGraph g, uint64_t total) {
  const uint64_t blo = (uint64_t)blockIdx.x * kT * kChunk;
  const uint64_t bhi = blo + (uint64_t)kT * kChunk < total
                           ? blo + (uint64_t)kT * kChunk : total;
  uint64_t s_lo, s_hi;
  block_segments(a.work_off, a.nb, blo, bhi, &s_lo, &s_hi);
  uint64_t e = blo + (uint64_t)threadIdx.x * kChunk;
  if (e >= total) return;
  uint64_t end = e + kChunk < total ? e + kChunk : total;
  uint64_t i = find_in(a.work_off, s_lo, s_hi, e);
  MergeSeg m = load_seg(a, g, i, false);
  uint64_t ia = e - m.begin;
  // batch elements below A[ia] were removed before position ia
  uint64_t rb = 0;
  {
    uint64_t lo = 0, hi = m.blen;
    uint32_t x = m.A[ia];
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if ((uint32_t)m.Bk[mid] < x)
        lo = mid + 1;
      else
        hi = mid;
    }
    rb = lo;
  }
  for (; e < end; e++) {
    if (e >= m.end) {
      i++;
      m = load_seg(a, g, i, false);
      ia = 0;
      rb = 0;
    }
    uint32_t x = m.A[ia];
    if (rb < m.blen && (uint32_t)m.Bk[rb] == x)
      rb++;
    else
      m.dst[ia - rb] = x;
    ia++;
  }
}

// In-place merge of tile rows through LDS.  Block b owns the tile rows whose
// input range starts in [b*kTile, (b+1)*kTile); it loads every input of those
// rows (old suffix or whole old row, then the batch run) into LDS,
// synchronises, and only then writes, so a row is never read after it has
// been overwritten.  Old element x (index j) lands at j + |{batch < x}|
// (insert), or at j - |{batch < x}| unless it is in the batch (remove);
// batch element y (index jb, insert) lands at jb + |{old < y}|.
// Row description of one tile row: where its inputs come from and where its
// output goes.
__device__ __forceinline__ void tile_row_info(const SegArgs &a,
                                              const DevGraph &g, uint64_t i,
                                              int insert, uint64_t *src,
                                              uint64_t *dst, uint32_t *alen,
                                              uint32_t *bstart) {
  uint64_t s = a.tile_idx[i];
  uint32_t v = a.seg_v[s];
  *bstart = (uint32_t)a.seg_idx[s];
  if (insert && a.alloc_off[s + 1] > a.alloc_off[s]) {  // relocated
    *src = g.off[v];
    *alen = a.seg_olddeg[s];
    *dst = a.pool_tail + a.alloc_off[s];
  } else {
    uint32_t p0 = a.seg_p0[s];
    *src = g.off[v] + p0;
    *alen = a.seg_olddeg[s] - p0;
    *dst = g.off[v] + p0;
  }
}

constexpr int kTileRows = 128;  // rows per block described in LDS
static_assert(kTileRows <= 256, "tile rows are numbered in bytes");
// k_merge_tile runs 256 threads over one LDS tile: the 12 KB tile, a
// 128-row table and the 3 KB row map take 18.5 KB, so eight tiles (the wave
// limit) are in flight per CU; tiles with more rows read their row
// descriptors from HBM and search them.  The block shape was chosen by
// timing the 1e7-edge insert at scale 24, before the input loads were
// issued together (input range / longest tile row:
// 1536 / 1536 merge 1.046 ms, insert 2.50 ms; 1024 / 2048 1.09 / 2.54;
// 2048 / 1024 1.015 / 2.66 and 1792 / 1280 1.01 / 2.60, whose longer rows
// go to the slower big-row path; 2048 / 1536 1.09 / 2.54 and 1536 / 2048
// 1.08 / 2.53 with a 14 KB tile, seven tiles per CU; 1.40 ms with
// 1024-thread blocks).
constexpr int kMergeThreads = 256;
constexpr int kMergeItems = kTileLds / kMergeThreads;

// Flat descriptors of the tile rows (written after any pool rebuild, so the
// row offsets they hold are current): input source, output destination,
// old-part length and start of the batch run.
__global__ void k_tile_rows(SegArgs a, DevGraph g, int insert,
                            uint64_t *t_src, uint64_t *t_dst,
                            uint32_t *t_alen, uint32_t *t_bst) {
  uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= a.nt) return;
  uint64_t src, dst;
  uint32_t alen, bst;
  tile_row_info(a, g, i, insert, &src, &dst, &alen, &bst);
  t_src[i] = src;
  t_dst[i] = dst;
  t_alen[i] = alen;
  t_bst[i] = bst;
}

// First tile row of each block and that row's input offset, side by side:
// blk[2b] = r0 = lower_bound(tile_off, b*kTile), blk[2b+1] = tile_off[r0].
// A merge block reads its range and the next one's in one 32-byte load
// instead of two dependent ones.
__global__ void k_tile_blocks(const uint64_t *tile_off, uint64_t nt,
                              uint64_t nblocks, uint64_t *blk) {
  uint64_t b = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (b > nblocks) return;
  const uint64_t r0 =
      b == nblocks ? nt : lower_bound_dev(tile_off, nt, b * (uint64_t)kTile);
  blk[2 * b] = r0;
  blk[2 * b + 1] = tile_off[r0];
}

__global__ void __launch_bounds__(kMergeThreads)
    k_merge_tile(SegArgs a, DevGraph g, int insert, const uint64_t *blk,
                 const uint64_t *t_src, const uint64_t *t_dst,
                 const uint32_t *t_alen, const uint32_t *t_bst) {
  __shared__ uint32_t buf[kTileLds];
  __shared__ uint32_t rstart[kTileRows + 1];  // input offset relative to e0
  __shared__ uint32_t ralen[kTileRows];
  __shared__ uint32_t rbstart[kTileRows];
  __shared__ uint64_t rsrc[kTileRows];
  __shared__ uint64_t rdst[kTileRows];
  __shared__ uint8_t rmap[kTileLds];  // tile row of each input (table path)
  const uint64_t *bk = blk + 2 * (uint64_t)blockIdx.x;
  const uint64_t r0 = bk[0], e0 = bk[1], r1 = bk[2], e1 = bk[3];
  if (r0 >= r1) return;
  const uint32_t W = (uint32_t)(e1 - e0);
  const uint32_t nrows = (uint32_t)(r1 - r0);
  const bool table = nrows <= (uint32_t)kTileRows;
  if (table) {
    for (uint32_t r = threadIdx.x; r < nrows; r += kMergeThreads) {
      rstart[r] = (uint32_t)(a.tile_off[r0 + r] - e0);
      ralen[r] = t_alen[r0 + r];
      rbstart[r] = t_bst[r0 + r];
      rsrc[r] = t_src[r0 + r];
      rdst[r] = t_dst[r0 + r];
    }
    if (threadIdx.x == 0) rstart[nrows] = W;
    __syncthreads();
    // each wave paints whole rows into the row map (coalesced LDS bytes), so
    // that an input finds its row with one read instead of a search
    const uint32_t ln = threadIdx.x % kWave;
    for (uint32_t r = threadIdx.x / kWave; r < nrows;
         r += kMergeThreads / kWave) {
      const uint32_t b1 = rstart[r + 1];
      for (uint32_t e = rstart[r] + ln; e < b1; e += kWave) rmap[e] = (uint8_t)r;
    }
    __syncthreads();
  }
  // Every input's address first, then all of a thread's loads, then their
  // LDS stores: a thread keeps up to kMergeItems loads in flight (with one
  // store after each load, the compiler waited for every load in turn).  A
  // batch input is the low word of its 64-bit key.  Slots past the tile
  // load the pool's first word (always there when tile rows exist) and are
  // not stored, so the loads need no branch.
  uint16_t row_of[kMergeItems];
  const uint32_t *ptr[kMergeItems];
  const uint32_t *bw = reinterpret_cast<const uint32_t *>(a.B);
  if (table) {
#pragma unroll
    for (int k = 0; k < kMergeItems; k++) {
      const uint32_t e = threadIdx.x + (uint32_t)k * kMergeThreads;
      row_of[k] = 0;
      ptr[k] = g.pool;
      if (e < W) {
        const uint32_t r = rmap[e];
        const uint32_t alen = ralen[r], j = e - rstart[r];
        row_of[k] = (uint16_t)r;
        ptr[k] = j < alen ? g.pool + rsrc[r] + j
                          : bw + 2 * ((uint64_t)rbstart[r] + (j - alen));
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < kMergeItems; k++) {
      const uint32_t e = threadIdx.x + (uint32_t)k * kMergeThreads;
      row_of[k] = 0;
      ptr[k] = g.pool;
      if (e < W) {
        const uint64_t i = find_in(a.tile_off, r0, r1 - 1, e0 + e);
        uint64_t src, dst;
        uint32_t alen, bst;
        tile_row_info(a, g, i, insert, &src, &dst, &alen, &bst);
        const uint32_t j = (uint32_t)(e0 + e - a.tile_off[i]);
        row_of[k] = (uint16_t)(i - r0);
        ptr[k] = j < alen ? g.pool + src + j : bw + 2 * ((uint64_t)bst + (j - alen));
      }
    }
  }
  uint32_t val[kMergeItems];
#pragma unroll
  for (int k = 0; k < kMergeItems; k++)  // device memory: global, not flat, loads
    val[k] = *(const __attribute__((address_space(1))) uint32_t *)ptr[k];
#pragma unroll
  for (int k = 0; k < kMergeItems; k++) {
    const uint32_t e = threadIdx.x + (uint32_t)k * kMergeThreads;
    if (e < W) buf[e] = val[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kMergeItems; k++) {
    const uint32_t e = threadIdx.x + (uint32_t)k * kMergeThreads;
    if (e >= W) continue;
    const uint32_t r = row_of[k];
    uint32_t base, alen, blen;
    uint64_t dst;
    if (table) {
      base = rstart[r];
      blen = rstart[r + 1] - base - ralen[r];
      alen = ralen[r];
      dst = rdst[r];
    } else {
      uint64_t i = r0 + r;
      uint64_t src;
      uint32_t bst;
      tile_row_info(a, g, i, insert, &src, &dst, &alen, &bst);
      base = (uint32_t)(a.tile_off[i] - e0);
      blen = (uint32_t)(a.tile_off[i + 1] - a.tile_off[i]) - alen;
    }
    const uint32_t j = e - base;
    const uint32_t *la = buf + base, *lb = buf + base + alen;
    const uint32_t x = buf[e];
    uint32_t *out = g.pool + dst;
    if (j < alen) {
      uint32_t l = 0, h = blen;  // batch elements below x
      while (l < h) {
        uint32_t m = (l + h) >> 1;
        if (lb[m] < x)
          l = m + 1;
        else
          h = m;
      }
      if (insert)
        out[j + l] = x;
      else if (!(l < blen && lb[l] == x))
        out[j - l] = x;
    } else if (insert) {
      uint32_t l = 0, h = alen;  // old elements below x
      while (l < h) {
        uint32_t m = (l + h) >> 1;
        if (la[m] < x)
          l = m + 1;
        else
          h = m;
      }
      out[(j - alen) + l] = x;
    }
  }
}

__global__ void k_finalize(SegArgs a, DevGraph g, int insert) {
  uint64_t s = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (s >= a.t) return;
  uint64_t bs = a.seg_idx[s], be = (s + 1 < a.t) ? a.seg_idx[s + 1] : a.b;
  uint32_t blen = (uint32_t)(be - bs);
  uint32_t v = a.seg_v[s];
  uint32_t od = a.seg_olddeg[s];
  g.deg[v] = insert ? od + blen : od - blen;
  if (insert && a.alloc_off[s + 1] > a.alloc_off[s]) {
    g.off[v] = a.pool_tail + a.alloc_off[s];
    g.cap[v] = (uint32_t)(a.alloc_off[s + 1] - a.alloc_off[s]);
  }
}

// After a batch on a symmetric graph: the graph stays symmetric iff every
// effective inserted edge (u, v) now has (v, u) and every removed one has
// lost it.  Clears *flag otherwise.
__global__ void k_sym_batch(const uint64_t *B, uint64_t b, DevGraph g,
                            int insert, uint32_t *flag) {
  uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= b) return;
  uint32_t u = (uint32_t)(B[i] >> 32), v = (uint32_t)B[i];
  bool has = false;
  if (v < g.n) has = row_has(g.pool + g.off[v], g.deg[v], u);
  if (has != (insert != 0)) *flag = 0;
}

// ----------------------------------------------------------- rebuild etc.
__global__ void k_deg_to_u64(const uint32_t *deg, uint64_t n, uint64_t *out) {
  uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) out[i] = deg[i];
  if (i == n) out[n] = 0;
}

__global__ void k_newcap(const uint32_t *deg, const uint32_t *cap, uint64_t n,
                         uint32_t n_cols, int recompute, uint32_t *new_cap,
                         uint64_t *cap64) {
  uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i > n) return;
  if (i == n) {
    cap64[n] = 0;
    return;
  }
  uint32_t c = cap[i];
  if (recompute) c = deg[i] ? (uint32_t)grow_cap(deg[i], n_cols) : 0;
  new_cap[i] = c;
  cap64[i] = c;
}

// The row, source slot and destination slot of each of a thread's kChunk
// consecutive edges from e0 (edges past m get slot 0 of the first edge's
// row and must not be stored).  Rows are found by prefix reads only, so the
// pool loads that follow can all be in flight at once (loaded and stored
// one edge at a time, each store made the next edge's loads wait, as the
// store might alias them).
__device__ __forceinline__ void chunk_slots(const DevGraph &g,
                                            const uint64_t *prefix, uint64_t m,
                                            const uint64_t *dst_off,
                                            uint64_t e0, uint64_t (&src)[kChunk],
                                            uint64_t (&dst)[kChunk]) {
  uint64_t v = find_segment(prefix, g.n, e0);
  uint64_t pv = prefix[v], pn = prefix[v + 1];
  uint64_t so = g.off[v], dso = dst_off ? dst_off[v] : pv;
#pragma unroll
  for (int j = 0; j < kChunk; j++) {
    const uint64_t e = e0 + (uint64_t)j;
    src[j] = so;
    dst[j] = 0;
    if (e < m) {
      while (e >= pn) {  // on to the edge's row (rows of degree 0 skipped)
        v++;
        pv = pn;
        pn = prefix[v + 1];
        so = g.off[v];
        dso = dst_off ? dst_off[v] : pv;
      }
      src[j] = so + (e - pv);
      dst[j] = dso + (e - pv);
    }
  }
}

// Copies every row (deg[v] ids) to dst_base(v) + k, flat over all edges.
// dst_off == nullptr: dense output at the row's degree prefix.
__global__ void k_gather_rows(DevGraph g, const uint64_t *prefix, uint64_t m,
                              uint32_t *out, const uint64_t *dst_off) {
  const uint64_t e0 = ((uint64_t)blockIdx.x * kT + threadIdx.x) * kChunk;
  if (e0 >= m) return;
  uint64_t src[kChunk], dst[kChunk];
  chunk_slots(g, prefix, m, dst_off, e0, src, dst);
  uint32_t val[kChunk];
#pragma unroll
  for (int j = 0; j < kChunk; j++) val[j] = g.pool[src[j]];
#pragma unroll
  for (int j = 0; j < kChunk; j++)
    if (e0 + (uint64_t)j < m) out[dst[j]] = val[j];
}

__global__ void k_has(DevGraph g, const uint32_t *rows, const uint32_t *cols,
                      uint64_t n, uint32_t n_cols, uint8_t *out) {
  uint64_t i = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  uint32_t r = rows[i], c = cols[i];
  uint8_t h = 0;
  if (r < g.n && c < n_cols) h = row_has(g.pool + g.off[r], g.deg[r], c);
  out[i] = h;
}

// per-block sums of the stored ids into part[block] (no atomics: many
// blocks adding to one address serialise)
__global__ void __launch_bounds__(kT)
    k_touch_sum(DevGraph g, const uint64_t *prefix, uint64_t m,
                uint64_t *part) {
  __shared__ uint64_t ws[kT / kWave];
  const uint64_t e0 = ((uint64_t)blockIdx.x * kT + threadIdx.x) * kChunk;
  uint64_t acc = 0;
  if (e0 < m) {
    uint64_t src[kChunk], dst[kChunk];
    chunk_slots(g, prefix, m, nullptr, e0, src, dst);
    uint32_t val[kChunk];
#pragma unroll
    for (int j = 0; j < kChunk; j++) val[j] = g.pool[src[j]];
#pragma unroll
    for (int j = 0; j < kChunk; j++)
      if (e0 + (uint64_t)j < m) acc += val[j];
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < kT / kWave; w++) t += ws[w];
    part[blockIdx.x] = t;
  }
}

// out = sum of part[0..nb) (one block)
__global__ void __launch_bounds__(kT)
    k_sum_u64(const uint64_t *part, uint64_t nb, uint64_t *out) {
  __shared__ uint64_t ws[kT / kWave];
  uint64_t t = 0;
  for (uint64_t i = threadIdx.x; i < nb; i += kT) t += part[i];
  t = wave_sum(t);
  if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t a = 0;
    for (int w = 0; w < kT / kWave; w++) a += ws[w];
    *out = a;
  }
}

inline uint64_t blocks(uint64_t n) { return ceil_div(n ? n : 1, kT); }
inline uint64_t chunk_blocks(uint64_t n) {
  return ceil_div(ceil_div(n ? n : 1, kChunk), kT);
}

}  // namespace

int graph_alloc(sstg_graph *g) {
  uint64_t n = g->n_rows ? g->n_rows : 1;
  SSTG_HIP(hipMalloc(&g->off, n * sizeof(uint64_t)));
  SSTG_HIP(hipMalloc(&g->deg, n * sizeof(uint32_t)));
  SSTG_HIP(hipMalloc(&g->cap, n * sizeof(uint32_t)));
  SSTG_HIP(hipMemsetAsync(g->off, 0, n * sizeof(uint64_t), g->stream));
  SSTG_HIP(hipMemsetAsync(g->deg, 0, n * sizeof(uint32_t), g->stream));
  SSTG_HIP(hipMemsetAsync(g->cap, 0, n * sizeof(uint32_t), g->stream));
  SSTG_HIP(hipMalloc(&g->dev_ctr, kCtrCount * sizeof(uint64_t)));
  SSTG_HIP(hipHostMalloc(&g->host_rb, 64 * sizeof(uint64_t),
                         hipHostMallocDefault));
  SSTG_HIP(hipHostMalloc(&g->host_ctr, kHostCtrSlots * sizeof(uint64_t),
                         hipHostMallocMapped | hipHostMallocCoherent));
  std::memset(g->host_ctr, 0, kHostCtrSlots * sizeof(uint64_t));
  g->host_seq = 0;
  SSTG_HIP(hipHostGetDevicePointer((void **)&g->host_ctr_dev, g->host_ctr, 0));
  SSTG_HIP(hipStreamSynchronize(g->stream));
  return SSTG_OK;
}

void graph_free(sstg_graph *g) {
  if (g->stream) hipStreamSynchronize(g->stream);
  pagerank_release(g);
  weights_free(g);
  hipFree(g->off);
  hipFree(g->deg);
  hipFree(g->cap);
  hipFree(g->pool);
  hipFree(g->dev_ctr);
  hipFree(g->deg_prefix);
  if (g->host_rb) hipHostFree(g->host_rb);
  if (g->host_ctr) hipHostFree(g->host_ctr);
  g->ws.release_all();
  g->off = nullptr;
  g->deg = g->cap = g->pool = nullptr;
  g->dev_ctr = g->deg_prefix = g->host_rb = nullptr;
  g->host_ctr = g->host_ctr_dev = nullptr;
}

// Moves every row into a fresh pool laid out in row order (the root-level
// rebalance of the packed-memory store).  recompute: reset each slot to
// grow_cap(deg) (reclaims slack left by deletes); otherwise slot sizes are
// kept.  extra: elements to leave free at the tail.
static int rebuild_impl(sstg_graph *g, uint64_t extra, bool recompute) {
  hipStream_t s = g->stream;
  const uint64_t n = g->n_rows;
  // the per-vertex temporaries stay in the workspace between rebuilds
  // rather than being allocated and freed by each (the copies into off and
  // cap below are ordered on the stream before any later use of them)
  uint32_t *new_cap = nullptr;
  uint64_t *new_off = nullptr, *prefix = nullptr;
  SSTG_TRY(g->ws.get(kWsRebuild, n + 1, &new_cap));
  SSTG_TRY(g->ws.get(kWsRebuild + 1, n + 1, &new_off));
  SSTG_LAUNCH(k_newcap, blocks(n + 1), kT, 0, s, (const uint32_t *)g->deg,
              (const uint32_t *)g->cap, n, g->n_cols, recompute ? 1 : 0,
              new_cap, new_off);
  SSTG_TRY(scan_exclusive_u64(new_off, new_off, n + 1, g->dev_ctr + kCtrScratch,
                              g->ws, kWsPrim, s));
  SSTG_HIP(hipMemcpyAsync(g->host_rb, g->dev_ctr + kCtrScratch,
                          sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  SSTG_HIP(hipStreamSynchronize(s));
  uint64_t total = g->host_rb[0];
  uint64_t want = total + extra;
  uint64_t new_cap_elems = want + want / 4 + (1u << 20);
  uint32_t *new_pool = nullptr;
  hipError_t e = hipMalloc(&new_pool, new_cap_elems * sizeof(uint32_t));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    new_cap_elems = want + 1024;
    SSTG_HIP(hipMalloc(&new_pool, new_cap_elems * sizeof(uint32_t)));
  }
  if (g->m) {
    SSTG_TRY(g->ws.get(kWsRebuild + 2, n + 1, &prefix));
    SSTG_LAUNCH(k_deg_to_u64, blocks(n + 1), kT, 0, s, (const uint32_t *)g->deg,
                n, prefix);
    SSTG_TRY(scan_exclusive_u64(prefix, prefix, n + 1, nullptr, g->ws,
                                kWsPrim, s));
    SSTG_LAUNCH(k_gather_rows, chunk_blocks(g->m), kT, 0, s, g->view(),
                (const uint64_t *)prefix, g->m, new_pool,
                (const uint64_t *)new_off);
  }
  SSTG_HIP(hipStreamSynchronize(s));
  hipFree(g->pool);
  g->pool = new_pool;
  g->pool_cap = new_cap_elems;
  g->pool_tail = total;
  g->garbage = 0;
  SSTG_HIP(hipMemcpyAsync(g->off, new_off, n * sizeof(uint64_t),
                          hipMemcpyDeviceToDevice, s));
  SSTG_HIP(hipMemcpyAsync(g->cap, new_cap, n * sizeof(uint32_t),
                          hipMemcpyDeviceToDevice, s));
  g->rebuilds++;
  return SSTG_OK;
}

int wait_host_flag(sstg_graph *g, uint64_t seq) {
  if (g->host_ctr) {
    const volatile uint64_t *f = g->host_ctr + kHostFlag;
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 1;; i++) {
      if (*f == seq) {
        std::atomic_thread_fence(std::memory_order_acquire);
        return SSTG_OK;
      }
      if ((i & 4095) == 0 && std::chrono::steady_clock::now() - t0 >
                                 std::chrono::milliseconds(100))
        break;
    }
  }
  SSTG_HIP(hipStreamSynchronize(g->stream));
  return SSTG_OK;
}

int graph_rebuild(sstg_graph *g, uint64_t extra_elems) {
  return rebuild_impl(g, extra_elems, true);
}

static int update_chunk(sstg_graph *g, const uint32_t *src, const uint32_t *dst,
                        uint64_t n, bool insert, bool skip_self) {
  hipStream_t s = g->stream;
  Workspace &ws = g->ws;
  uint64_t *ctr = g->dev_ctr;
  // density rule at the root: once abandoned slots exceed half of the used
  // pool, compact before taking more space
  if (insert && !g->noalloc && g->garbage > 0 && g->garbage * 2 > g->pool_tail)
    SSTG_TRY(rebuild_impl(g, 0, true));
  // a batch of a few thousand pairs in two launches and one host wait
  // (update_small.hip), unless it needs the pool to grow or moves a very
  // long row
  if (n <= kSmallBatch) {
    bool fallback = false;
    SSTG_TRY(update_small(g, src, dst, (uint32_t)n, insert, skip_self,
                          &fallback));
    if (!fallback) return SSTG_OK;
  }

  uint64_t *keys = nullptr, *tmp = nullptr, *B = nullptr, *seg_idx = nullptr;
  uint8_t *flags = nullptr;
  SSTG_TRY(ws.get(kWsKeysA, n, &keys));
  SSTG_TRY(ws.get(kWsKeysB, n, &tmp));
  SSTG_TRY(ws.get(kWsFlags, n + 1, &flags));
  SSTG_HIP(hipMemsetAsync(ctr, 0, kCtrCount * sizeof(uint64_t), s));
  int shifts[16], np = 0;
  uint32_t cb = bit_width_u32(g->n_cols ? g->n_cols - 1 : 0);
  uint32_t rb = bit_width_u32(g->n_rows ? g->n_rows - 1 : 0);
  for (uint32_t b = 0; b < cb; b += 8) shifts[np++] = (int)b;
  for (uint32_t b = 0; b < rb; b += 8) shifts[np++] = (int)(32 + b);
  uint64_t *sorted = nullptr;
  SSTG_TRY(pack_sort_pairs(src, dst, n, g->n_rows, g->n_cols, ctr + kCtrErr,
                           keys, tmp, shifts, np, ws, kWsPrim, s, &sorted));
  uint32_t *pos = nullptr, *B_pos = nullptr;
  SSTG_TRY(ws.get(kWsMarkPos, n, &pos));
  SSTG_TRY(ws.get(kWsBPos, n, &B_pos));
  SSTG_LAUNCH(k_mark, blocks(n), kT, 0, s, (const uint64_t *)sorted, n,
              g->view(), insert ? 1 : 0, skip_self ? 1 : 0, flags, pos);
  // B reuses the buffer that does not hold the sorted keys
  B = (sorted == keys) ? tmp : keys;
  SSTG_TRY(select_flagged_u64_u32(flags, sorted, pos, n, B, B_pos, ctr + kCtrB,
                                  ws, kWsPrim, s));
  // at most one segment per row: segment arrays hold nseg + 1 entries
  const uint64_t nseg = n < (uint64_t)g->n_rows ? n : (uint64_t)g->n_rows;
  SSTG_TRY(ws.get(kWsSegIdx, nseg + 1, &seg_idx));
  SSTG_LAUNCH(k_heads, blocks(n), kT, 0, s, (const uint64_t *)B,
              (const uint64_t *)ctr, n, flags);
  SSTG_TRY(select_flagged_u64(flags, nullptr, n, seg_idx, ctr + kCtrT, ws,
                              kWsPrim, s));
  uint32_t *seg_v = nullptr, *seg_p0 = nullptr, *seg_od = nullptr,
           *seg_stage = nullptr, *seg_tile = nullptr;
  uint64_t *seg_alloc = nullptr, *seg_work = nullptr, *big_idx = nullptr,
           *tile_idx = nullptr, *big_stage = nullptr, *big_work = nullptr,
           *tile_len = nullptr;
  uint8_t *is_big = nullptr, *is_tile = nullptr;
  SSTG_TRY(ws.get(kWsSegV, nseg + 1, &seg_v));
  SSTG_TRY(ws.get(kWsSegP0, nseg + 1, &seg_p0));
  SSTG_TRY(ws.get(kWsSegOldDeg, nseg + 1, &seg_od));
  SSTG_TRY(ws.get(kWsSegStage, nseg + 1, &seg_stage));
  SSTG_TRY(ws.get(kWsSegAlloc, nseg + 1, &seg_alloc));
  SSTG_TRY(ws.get(kWsSegWork, nseg + 1, &seg_work));
  SSTG_TRY(ws.get(kWsSegTile, nseg + 1, &seg_tile));
  SSTG_TRY(ws.get(kWsSegIsBig, nseg + 1, &is_big));
  SSTG_TRY(ws.get(kWsSegIsTile, nseg + 1, &is_tile));
  SSTG_TRY(ws.get(kWsBigIdx, nseg + 1, &big_idx));
  SSTG_TRY(ws.get(kWsTileIdx, nseg + 1, &tile_idx));
  SSTG_TRY(ws.get(kWsBigStage, nseg + 1, &big_stage));
  SSTG_TRY(ws.get(kWsBigWork, nseg + 1, &big_work));
  SSTG_TRY(ws.get(kWsTileLen, nseg + 1, &tile_len));
  SSTG_LAUNCH(k_segprep, blocks(nseg + 1), kT, 0, s, (const uint64_t *)B,
              (const uint32_t *)B_pos, (const uint64_t *)ctr, (const uint64_t *)seg_idx, g->view(),
              insert ? 1 : 0, g->n_cols, nseg, seg_v, seg_p0, seg_od,
              seg_alloc, seg_stage, seg_work, seg_tile, is_big, is_tile,
              ctr + kCtrGarbage);
  {
    SegScan sc{};
    sc.alloc = seg_alloc;
    sc.is_big = is_big;
    sc.is_tile = is_tile;
    sc.stage = seg_stage;
    sc.tile = seg_tile;
    sc.work = seg_work;
    sc.n = nseg + 1;
    sc.big_idx = big_idx;
    sc.big_stage = big_stage;
    sc.big_work = big_work;
    sc.tile_idx = tile_idx;
    sc.tile_len = tile_len;
    sc.ctr = ctr;
    sc.hctr = g->host_ctr_dev;
    sc.seq = ++g->host_seq;
    const uint64_t tiles = ceil_div(sc.n, kSegTile);
    if (tiles <= kSegOneBlock) {
      sc.per = sc.n;
      sc.blocks = 1;
      SSTG_LAUNCH(k_seg_scan, 1, kT, 0, s, sc);
    } else {
      const uint64_t tpb = ceil_div(tiles, kSegBlocks);
      sc.per = tpb * kSegTile;
      sc.blocks = (uint32_t)ceil_div(tiles, tpb);
      SSTG_TRY(ws.get(kWsPrim, (uint64_t)kSegSums * sc.blocks, &sc.partials));
      SSTG_LAUNCH(k_seg_reduce, sc.blocks, kT, 0, s, sc);
      SSTG_LAUNCH(k_seg_scan_partials, 1, kT, 0, s, sc);
      SSTG_LAUNCH(k_seg_scan, sc.blocks, kT, 0, s, sc);
    }
  }
  const uint64_t *h = g->host_ctr;
  if (!g->host_ctr_dev) {
    SSTG_HIP(hipMemcpyAsync(g->host_rb, ctr, kCtrCount * sizeof(uint64_t),
                            hipMemcpyDeviceToHost, s));
    h = g->host_rb;
    SSTG_HIP(hipStreamSynchronize(s));
  } else {
    SSTG_TRY(wait_host_flag(g, g->host_seq));
  }
  if (h[kCtrErr]) return SSTG_E_RANGE;
  const uint64_t b = h[kCtrB], t = h[kCtrT];
  const uint64_t stage_tot = h[kCtrStage], alloc_tot = h[kCtrAlloc],
                 work_tot = h[kCtrWork], garbage_add = h[kCtrGarbage],
                 tile_tot = h[kCtrTile], nb = h[kCtrBig], nt = h[kCtrTiles];
  g->last.effective += b;
  g->last.rows += t;
  g->last.tile_rows += nt;
  g->last.tile_inputs += tile_tot;
  g->last.big_rows += nb;
  g->last.staged += stage_tot;
  g->last.merged += work_tot;
  g->last.allocated += alloc_tot;
  if (b == 0) return SSTG_OK;
  if (g->pool_tail + alloc_tot > g->pool_cap) {
    // nothing has been written to the store yet: an allocation-free call
    // stops here with the graph unchanged
    if (g->noalloc) return SSTG_E_NOMEM;
    // slot sizes are kept, so the decisions made above stay valid
    SSTG_TRY(rebuild_impl(g, alloc_tot, false));
  }
  uint32_t *scratch = nullptr;
  SSTG_TRY(ws.get(kWsStage, stage_tot + 1, &scratch));
  SegArgs a;
  a.B = B;
  a.seg_idx = seg_idx;
  a.seg_v = seg_v;
  a.seg_p0 = seg_p0;
  a.seg_olddeg = seg_od;
  a.alloc_off = seg_alloc;
  a.t = t;
  a.b = b;
  a.pool_tail = g->pool_tail;
  a.scratch = scratch;
  a.big_idx = big_idx;
  a.stage_off = big_stage;
  a.work_off = big_work;
  a.nb = nb;
  a.tile_idx = tile_idx;
  a.tile_off = tile_len;
  a.nt = nt;
  if (tile_tot) {
    const uint64_t nblk = ceil_div(tile_tot, kTile);
    uint64_t *t_src = nullptr, *t_dst = nullptr, *blk = nullptr;
    uint32_t *t_alen = nullptr, *t_bst = nullptr;
    SSTG_TRY(ws.get(kWsTileSrc, nt + 1, &t_src));
    SSTG_TRY(ws.get(kWsTileDst, nt + 1, &t_dst));
    SSTG_TRY(ws.get(kWsTileAlen, nt + 1, &t_alen));
    SSTG_TRY(ws.get(kWsTileBst, nt + 1, &t_bst));
    SSTG_TRY(ws.get(kWsTileBlk, 2 * (nblk + 1), &blk));
    SSTG_LAUNCH(k_tile_rows, blocks(nt), kT, 0, s, a, g->view(),
                insert ? 1 : 0, t_src, t_dst, t_alen, t_bst);
    SSTG_LAUNCH(k_tile_blocks, blocks(nblk + 1), kT, 0, s,
                (const uint64_t *)tile_len, nt, nblk, blk);
    SSTG_LAUNCH(k_merge_tile, nblk, kMergeThreads, 0, s, a, g->view(),
                insert ? 1 : 0,
                (const uint64_t *)blk, (const uint64_t *)t_src,
                (const uint64_t *)t_dst, (const uint32_t *)t_alen,
                (const uint32_t *)t_bst);
  }
  if (stage_tot)
    SSTG_LAUNCH(k_stage, chunk_blocks(stage_tot), kT, 0, s, a, g->view(),
                stage_tot);
  if (work_tot) {
    if (insert)
      SSTG_LAUNCH(k_merge_insert, chunk_blocks(work_tot), kT, 0, s, a,
                  g->view(), work_tot);
    else
      SSTG_LAUNCH(k_compact_remove, chunk_blocks(work_tot), kT, 0, s, a,
                  g->view(), work_tot);
  }
  SSTG_LAUNCH(k_finalize, blocks(t), kT, 0, s, a, g->view(), insert ? 1 : 0);
  // track symmetry through small batches; a batch that is large next to the
  // graph leaves it unknown (checked in full when a traversal needs it)
  const bool was_symmetric = g->sym_version == g->version && g->symmetric &&
                             b * 8 <= g->m;
  uint32_t *symflag = (uint32_t *)(ctr + kCtrScratch);
  if (was_symmetric) {
    SSTG_HIP(hipMemsetAsync(symflag, 0xFF, sizeof(uint32_t), s));
    SSTG_LAUNCH(k_sym_batch, blocks(b), kT, 0, s, (const uint64_t *)B, b,
                g->view(), insert ? 1 : 0, symflag);
    SSTG_HIP(hipMemcpyAsync(g->host_rb + 32, symflag, sizeof(uint32_t),
                            hipMemcpyDeviceToHost, s));
  }
  SSTG_HIP(hipStreamSynchronize(s));
  if (was_symmetric) {
    g->symmetric = (uint32_t)g->host_rb[32] != 0;
    g->sym_version = g->version + 1;
  } else {
    g->sym_version = ~0ull;
  }
  g->pool_tail += alloc_tot;
  g->garbage += garbage_add;
  g->m = insert ? g->m + b : g->m - b;
  g->version++;
  return SSTG_OK;
}

// The pairs an update would change, without applying it: the batch packed,
// sorted and deduplicated, self-loops dropped for an insert that skips
// them, and then kept if absent (insert) or present (remove).  The sorted
// keys (row << 32 | col) land in workspace slot kWsEffKeys, which updates
// do not use, so they outlive a following graph_update.  Device arrays,
// n < 2^31.
int graph_effective_keys(sstg_graph *g, const uint32_t *src,
                         const uint32_t *dst, uint64_t n, bool insert,
                         bool skip_self, const uint64_t **out_keys,
                         uint64_t *out_count) {
  hipStream_t s = g->stream;
  Workspace &ws = g->ws;
  uint64_t *ctr = g->dev_ctr;
  *out_count = 0;
  uint64_t *out = nullptr;
  SSTG_TRY(ws.get(kWsEffKeys, n + 1, &out));
  *out_keys = out;
  if (n == 0) return SSTG_OK;
  uint64_t *keys = nullptr, *tmp = nullptr;
  uint8_t *flags = nullptr;
  uint32_t *pos = nullptr;
  SSTG_TRY(ws.get(kWsKeysA, n, &keys));
  SSTG_TRY(ws.get(kWsKeysB, n, &tmp));
  SSTG_TRY(ws.get(kWsFlags, n + 1, &flags));
  SSTG_TRY(ws.get(kWsMarkPos, n, &pos));
  SSTG_HIP(hipMemsetAsync(ctr, 0, kCtrCount * sizeof(uint64_t), s));
  int shifts[16], np = 0;
  uint32_t cb = bit_width_u32(g->n_cols ? g->n_cols - 1 : 0);
  uint32_t rb = bit_width_u32(g->n_rows ? g->n_rows - 1 : 0);
  for (uint32_t b = 0; b < cb; b += 8) shifts[np++] = (int)b;
  for (uint32_t b = 0; b < rb; b += 8) shifts[np++] = (int)(32 + b);
  uint64_t *sorted = nullptr;
  SSTG_TRY(pack_sort_pairs(src, dst, n, g->n_rows, g->n_cols, ctr + kCtrErr,
                           keys, tmp, shifts, np, ws, kWsPrim, s, &sorted));
  SSTG_LAUNCH(k_mark, blocks(n), kT, 0, s, (const uint64_t *)sorted, n,
              g->view(), insert ? 1 : 0, skip_self ? 1 : 0, flags, pos);
  SSTG_TRY(select_flagged_u64(flags, sorted, n, out, ctr + kCtrB, ws, kWsPrim,
                              s));
  SSTG_HIP(hipMemcpyAsync(g->host_rb, ctr, 2 * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, s));
  SSTG_HIP(hipStreamSynchronize(s));
  if (g->host_rb[kCtrErr]) return SSTG_E_RANGE;
  *out_count = g->host_rb[kCtrB];
  return SSTG_OK;
}

int graph_update(sstg_graph *g, const uint32_t *src, const uint32_t *dst,
                 uint64_t n, bool insert, bool skip_self, bool device_ptrs) {
  g->last = sstg_update_stats{};
  g->last.batch = n;
  if (n == 0) return SSTG_OK;
  // batches above 2^31 pairs are applied in pieces; union and difference are
  // unaffected by the split.  Every id is checked before the first piece, so
  // an out-of-range id leaves the graph unchanged however large the batch.
  const uint64_t kMaxChunk = 1ull << 31;
  if (n > kMaxChunk) {
    bool bad = false;
    if (device_ptrs) {
      SSTG_HIP(hipMemsetAsync(g->dev_ctr + kCtrErr, 0, sizeof(uint64_t),
                              g->stream));
      SSTG_LAUNCH(k_check_range, 4096, kT, 0, g->stream, src, dst, n,
                  g->n_rows, g->n_cols, g->dev_ctr);
      SSTG_HIP(hipMemcpyAsync(g->host_rb, g->dev_ctr + kCtrErr,
                              sizeof(uint64_t), hipMemcpyDeviceToHost,
                              g->stream));
      SSTG_HIP(hipStreamSynchronize(g->stream));
      bad = g->host_rb[0] != 0;
    } else {
      for (uint64_t i = 0; i < n && !bad; i++)
        bad = src[i] >= g->n_rows || dst[i] >= g->n_cols;
    }
    if (bad) return SSTG_E_RANGE;
  }
  for (uint64_t base = 0; base < n; base += kMaxChunk) {
    uint64_t len = n - base < kMaxChunk ? n - base : kMaxChunk;
    const uint32_t *s = src + base, *d = dst + base;
    if (!device_ptrs) {
      uint32_t *ds = nullptr, *dd = nullptr;
      SSTG_TRY(g->ws.get(kWsInSrc, len, &ds));
      SSTG_TRY(g->ws.get(kWsInDst, len, &dd));
      SSTG_HIP(hipMemcpyAsync(ds, s, len * sizeof(uint32_t),
                              hipMemcpyHostToDevice, g->stream));
      SSTG_HIP(hipMemcpyAsync(dd, d, len * sizeof(uint32_t),
                              hipMemcpyHostToDevice, g->stream));
      s = ds;
      d = dd;
    }
    SSTG_TRY(update_chunk(g, s, d, len, insert, skip_self));
  }
  // density rule on removal (the packed-memory array's lower bound, as
  // PMA.hpp:686-702 halves a list): once the live ids fill less than a
  // quarter of the live slots, compact the pool with every slot reset to the
  // capacity of its row, which gives the memory back
  if (!insert && !g->noalloc && g->pool_tail - g->garbage > kShrinkMinElems &&
      g->m * 4 < g->pool_tail - g->garbage)
    SSTG_TRY(rebuild_impl(g, 0, true));
  g->last.rebuilds = g->rebuilds;
  return SSTG_OK;
}

int graph_degree_prefix(sstg_graph *g, const uint64_t **out) {
  if (!g->deg_prefix) {
    SSTG_HIP(hipMalloc(&g->deg_prefix, ((uint64_t)g->n_rows + 1) *
                                           sizeof(uint64_t)));
    g->deg_prefix_version = ~0ull;
  }
  if (g->deg_prefix_version != g->version) {
    uint64_t n = g->n_rows;
    SSTG_LAUNCH(k_deg_to_u64, blocks(n + 1), kT, 0, g->stream,
                (const uint32_t *)g->deg, n, g->deg_prefix);
    SSTG_TRY(scan_exclusive_u64(g->deg_prefix, g->deg_prefix, n + 1, nullptr,
                                g->ws, kWsPrim, g->stream));
    g->deg_prefix_version = g->version;
  }
  *out = g->deg_prefix;
  return SSTG_OK;
}

int graph_export(sstg_graph *g, uint64_t *offsets, uint32_t *cols,
                 bool device_ptrs) {
  const uint64_t *prefix = nullptr;
  SSTG_TRY(graph_degree_prefix(g, &prefix));
  hipStream_t s = g->stream;
  uint64_t n = g->n_rows;
  if (offsets) {
    SSTG_HIP(hipMemcpyAsync(offsets, prefix, (n + 1) * sizeof(uint64_t),
                            device_ptrs ? hipMemcpyDeviceToDevice
                                        : hipMemcpyDeviceToHost,
                            s));
  }
  if (cols && g->m) {
    uint32_t *dcols = cols;
    if (!device_ptrs) SSTG_TRY(g->ws.get(kWsMisc, g->m, &dcols));
    SSTG_LAUNCH(k_gather_rows, chunk_blocks(g->m), kT, 0, s, g->view(), prefix,
                g->m, dcols, (const uint64_t *)nullptr);
    if (!device_ptrs)
      SSTG_HIP(hipMemcpyAsync(cols, dcols, g->m * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, s));
  }
  SSTG_HIP(hipStreamSynchronize(s));
  return SSTG_OK;
}

int graph_has_batch(sstg_graph *g, const uint32_t *rows, const uint32_t *cols,
                    uint64_t n, uint8_t *out, bool device_ptrs) {
  if (n == 0) return SSTG_OK;
  hipStream_t s = g->stream;
  const uint32_t *r = rows, *c = cols;
  uint8_t *o = out;
  if (!device_ptrs) {
    uint32_t *dr = nullptr, *dc = nullptr;
    SSTG_TRY(g->ws.get(kWsInSrc, n, &dr));
    SSTG_TRY(g->ws.get(kWsInDst, n, &dc));
    SSTG_TRY(g->ws.get(kWsFlags, n, &o));
    SSTG_HIP(hipMemcpyAsync(dr, rows, n * 4, hipMemcpyHostToDevice, s));
    SSTG_HIP(hipMemcpyAsync(dc, cols, n * 4, hipMemcpyHostToDevice, s));
    r = dr;
    c = dc;
  }
  SSTG_LAUNCH(k_has, blocks(n), kT, 0, s, g->view(), r, c, n, g->n_cols, o);
  if (!device_ptrs)
    SSTG_HIP(hipMemcpyAsync(out, o, n, hipMemcpyDeviceToHost, s));
  SSTG_HIP(hipStreamSynchronize(s));
  return SSTG_OK;
}

int graph_touch_sum(sstg_graph *g, uint64_t *out) {
  const uint64_t *prefix = nullptr;
  SSTG_TRY(graph_degree_prefix(g, &prefix));
  hipStream_t s = g->stream;
  uint64_t *acc = g->dev_ctr + kCtrScratch;
  const uint64_t nb = chunk_blocks(g->m);
  uint64_t *part = nullptr;
  SSTG_TRY(g->ws.get(kWsTouch, nb, &part));
  if (g->m) {
    SSTG_LAUNCH(k_touch_sum, nb, kT, 0, s, g->view(), prefix, g->m, part);
    SSTG_LAUNCH(k_sum_u64, 1, kT, 0, s, (const uint64_t *)part, nb, acc);
  } else {
    SSTG_HIP(hipMemsetAsync(acc, 0, sizeof(uint64_t), s));
  }
  SSTG_HIP(hipMemcpyAsync(g->host_rb, acc, sizeof(uint64_t),
                          hipMemcpyDeviceToHost, s));
  SSTG_HIP(hipStreamSynchronize(s));
  *out = g->host_rb[0];
  return SSTG_OK;
}

}  // namespace sstg
