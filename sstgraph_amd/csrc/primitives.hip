// Scans, stream compaction and LSD radix sort for gfx950 (wave64).
#include "primitives.hpp"
#include "block_sort.hpp"
#include "graph.hpp"

namespace sstg {

namespace {

// Scans and selects run in at most three launches whatever their length:
// each block of a fixed grid (at most kScanBlocks) reduces one contiguous
// range of whole tiles, one block scans those per-block partials, and each
// block then scans its range tile by tile from its partial's offset.  Up to
// kScanOneBlock tiles one block does the whole job in one launch.  A tile is
// kScanItems values per thread (512), small so that short scans still spread
// over many blocks; the cap of 2048 blocks (8 per CU) keeps the partials to
// one block's worth and the machine full.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 2;
constexpr int kScanTile = kScanThreads * kScanItems;  // 512
constexpr uint64_t kScanBlocks = 2048;
constexpr int kPartItems = (int)(kScanBlocks / kScanThreads);  // 8
constexpr uint64_t kScanOneBlock = 8;

struct Ranges {
  uint64_t per;     // values per block, a multiple of kScanTile
  uint32_t blocks;  // ceil(n / per) <= kScanBlocks
};

Ranges ranges_of(uint64_t n) {
  const uint64_t tiles = ceil_div(n, kScanTile);
  const uint64_t tpb = ceil_div(tiles, kScanBlocks);
  return Ranges{tpb * kScanTile, (uint32_t)ceil_div(tiles, tpb)};
}

// Block-wide exclusive scan of one value per thread (256 threads = 4 waves).
// Returns the exclusive prefix; *total gets the block sum.
template <class T>
__device__ __forceinline__ T block_exclusive(T v, T *wtot, T *total) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  T incl = wave_inclusive_sum(v);
  if (lane == kWave - 1) wtot[w] = incl;
  __syncthreads();
  T woff = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kScanThreads / kWave; i++) {
    T x = wtot[i];
    if (i < w) woff += x;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return woff + incl - v;
}

template <class T>
__device__ __forceinline__ T block_sum(T v, T *wtot) {
  v = wave_sum(v);
  if ((threadIdx.x & (kWave - 1)) == 0) wtot[threadIdx.x / kWave] = v;
  __syncthreads();
  T t = 0;
#pragma unroll
  for (int i = 0; i < kScanThreads / kWave; i++) t += wtot[i];
  return t;
}

// Values the scans read: a plain array, or a byte flag counted as 0 / 1.
template <class T> struct ArrayIn {
  const T *p;
  __device__ T operator()(uint64_t i) const { return p[i]; }
};
struct FlagIn {
  const uint8_t *p;
  __device__ uint64_t operator()(uint64_t i) const { return p[i] ? 1 : 0; }
};

template <class T, class In>
__global__ void __launch_bounds__(kScanThreads)
    k_range_reduce(In in, uint64_t n, uint64_t per, T *partials) {
  __shared__ T wtot[kScanThreads / kWave];
  const uint64_t lo = (uint64_t)blockIdx.x * per;
  const uint64_t hi = lo + per < n ? lo + per : n;
  T acc = 0;
  for (uint64_t c = lo; c < hi; c += kScanTile) {
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
      uint64_t i = c + (uint64_t)j * kScanThreads + threadIdx.x;
      if (i < hi) acc += in(i);
    }
  }
  T t = block_sum(acc, wtot);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// Exclusive scan, in place, of nb <= kScanBlocks per-block partials (one
// block, kPartItems consecutive values per thread).
template <class T>
__global__ void __launch_bounds__(kScanThreads)
    k_scan_partials(T *p, uint32_t nb) {
  __shared__ T wtot[kScanThreads / kWave];
  const uint32_t base = threadIdx.x * kPartItems;
  T v[kPartItems], sum = 0;
#pragma unroll
  for (int j = 0; j < kPartItems; j++) {
    v[j] = base + j < nb ? p[base + j] : (T)0;
    sum += v[j];
  }
  T tot;
  T run = block_exclusive(sum, wtot, &tot);
#pragma unroll
  for (int j = 0; j < kPartItems; j++) {
    if (base + j < nb) p[base + j] = run;
    run += v[j];
  }
}

// Exclusive scan of each block's range from its offset (offsets null: one
// block, offset 0); the last block writes the grand total.
template <class T, class In>
__global__ void __launch_bounds__(kScanThreads)
    k_range_scan(In in, T *out, uint64_t n, uint64_t per, const T *offsets,
                 T *d_total) {
  __shared__ T wtot[kScanThreads / kWave];
  const uint64_t lo = (uint64_t)blockIdx.x * per;
  const uint64_t hi = lo + per < n ? lo + per : n;
  T carry = offsets ? offsets[blockIdx.x] : (T)0;
  for (uint64_t c = lo; c < hi; c += kScanTile) {
    T vals[kScanItems];
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
      uint64_t i = c + (uint64_t)j * kScanThreads + threadIdx.x;
      vals[j] = (i < hi) ? in(i) : (T)0;
    }
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
      uint64_t i = c + (uint64_t)j * kScanThreads + threadIdx.x;
      T tot;
      T ex = block_exclusive(vals[j], wtot, &tot);
      if (i < hi) out[i] = carry + ex;
      carry += tot;
    }
  }
  if (d_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
    *d_total = carry;
}

template <class T>
int scan_impl(const T *in, T *out, uint64_t n, T *d_total, Workspace &ws,
              int slot, hipStream_t s) {
  if (n == 0) {
    if (d_total) SSTG_HIP(hipMemsetAsync(d_total, 0, sizeof(T), s));
    return SSTG_OK;
  }
  const ArrayIn<T> src{in};
  if (ceil_div(n, kScanTile) <= kScanOneBlock) {
    SSTG_LAUNCH((k_range_scan<T, ArrayIn<T>>), 1, kScanThreads, 0, s, src, out,
                n, n, (const T *)nullptr, d_total);
    return SSTG_OK;
  }
  const Ranges r = ranges_of(n);
  T *partials = nullptr;
  SSTG_TRY(ws.get(slot, r.blocks, &partials));
  SSTG_LAUNCH((k_range_reduce<T, ArrayIn<T>>), r.blocks, kScanThreads, 0, s,
              src, n, r.per, partials);
  SSTG_LAUNCH(k_scan_partials<T>, 1, kScanThreads, 0, s, partials, r.blocks);
  SSTG_LAUNCH((k_range_scan<T, ArrayIn<T>>), r.blocks, kScanThreads, 0, s, src,
              out, n, r.per, (const T *)partials, d_total);
  return SSTG_OK;
}

// ------------------------------------------------------------- select
// Writes the flagged values of each block's range in order from its offset
// (offsets null: one block, offset 0), and values2 beside them into out2
// when given; the last block writes the count.
template <class V>
__global__ void __launch_bounds__(kScanThreads)
    k_select_range(const uint8_t *flags, const V *values, uint64_t n,
                   uint64_t per, const uint64_t *offsets, V *out,
                   uint64_t *d_count, const uint32_t *values2,
                   uint32_t *out2) {
  __shared__ uint64_t wtot[kScanThreads / kWave];
  const uint64_t lo = (uint64_t)blockIdx.x * per;
  const uint64_t hi = lo + per < n ? lo + per : n;
  uint64_t carry = offsets ? offsets[blockIdx.x] : 0;
  for (uint64_t c = lo; c < hi; c += kScanTile) {
    // flags and values loaded together, before the scans (a value loaded
    // after its scan cost a second round trip per step)
    uint8_t f[kScanItems];
    V val[kScanItems];
    uint32_t val2[kScanItems];
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
      uint64_t i = c + (uint64_t)j * kScanThreads + threadIdx.x;
      const bool in = i < hi;
      f[j] = (in && flags[i]) ? 1 : 0;
      val[j] = (in && values) ? values[i] : (V)i;
      val2[j] = (in && out2) ? values2[i] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
      uint64_t tot;
      uint64_t ex = block_exclusive((uint64_t)f[j], wtot, &tot);
      if (f[j]) {
        out[carry + ex] = val[j];
        if (out2) out2[carry + ex] = val2[j];
      }
      carry += tot;
    }
  }
  if (d_count && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
    *d_count = carry;
}

// --------------------------------------------------------- radix sort
constexpr int kRsThreads = 256;
constexpr int kRsWaves = kRsThreads / kWave;
constexpr int kRsItems = 8;
constexpr int kRsTile = kRsThreads * kRsItems;  // 2048 keys
constexpr int kRsBins = 256;

// Digit histogram of each tile; hist is digit-major: hist[d * tiles + t].
__global__ void __launch_bounds__(kRsThreads)
    k_rs_hist(const uint64_t *keys, uint64_t n, int shift, uint32_t *hist,
              uint32_t tiles) {
  __shared__ uint32_t h[kRsWaves][kRsBins];
  const int w = threadIdx.x / kWave;
  // consecutive tiles on one XCD: the counts of 16 consecutive tiles share a
  // cache line of each digit's row, which then fills in one L2
  const uint32_t tile = xcd_block(blockIdx.x, gridDim.x);
  for (int i = threadIdx.x; i < kRsWaves * kRsBins; i += kRsThreads)
    (&h[0][0])[i] = 0;
  __syncthreads();
  uint64_t base = (uint64_t)tile * kRsTile;
  // all of a thread's keys in flight before the first count
  uint64_t kv[kRsItems];
#pragma unroll
  for (int j = 0; j < kRsItems; j++) {
    uint64_t i = base + (uint64_t)j * kRsThreads + threadIdx.x;
    kv[j] = i < n ? keys[i] : 0;
  }
#pragma unroll
  for (int j = 0; j < kRsItems; j++) {
    uint64_t i = base + (uint64_t)j * kRsThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[w][(uint32_t)(kv[j] >> shift) & 255u], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRsBins; d += kRsThreads) {
    uint32_t c = 0;
#pragma unroll
    for (int ww = 0; ww < kRsWaves; ww++) c += h[ww][d];
    hist[(uint64_t)d * tiles + tile] = c;
  }
}

// keys[i] = src[i] << 32 | dst[i]; a pair out of range sets *err and packs
// as (0, 0).
__global__ void k_pack_pairs(const uint32_t *src, const uint32_t *dst,
                             uint64_t n, uint32_t n_rows, uint32_t n_cols,
                             uint64_t *keys, uint64_t *err) {
  const uint64_t i = (uint64_t)blockIdx.x * kRsThreads + threadIdx.x;
  if (i >= n) return;
  uint32_t a = src[i], b = dst[i];
  if (a >= n_rows || b >= n_cols) {
    *err = 1;
    a = 0;
    b = 0;
  }
  keys[i] = ((uint64_t)a << 32) | b;
}

// k_pack_pairs over k_rs_hist's tiles (the same tile order and layout),
// counting each packed key's digit at shift as k_rs_hist would.
__global__ void __launch_bounds__(kRsThreads)
    k_rs_pack_hist(const uint32_t *src, const uint32_t *dst, uint64_t n,
                   uint32_t n_rows, uint32_t n_cols, uint64_t *keys,
                   uint64_t *err, int shift, uint32_t *hist, uint32_t tiles) {
  __shared__ uint32_t h[kRsWaves][kRsBins];
  const int w = threadIdx.x / kWave;
  const uint32_t tile = xcd_block(blockIdx.x, gridDim.x);
  for (int i = threadIdx.x; i < kRsWaves * kRsBins; i += kRsThreads)
    (&h[0][0])[i] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)tile * kRsTile;
  uint32_t a[kRsItems], b[kRsItems];
#pragma unroll
  for (int j = 0; j < kRsItems; j++) {
    const uint64_t i = base + (uint64_t)j * kRsThreads + threadIdx.x;
    a[j] = i < n ? src[i] : 0u;
    b[j] = i < n ? dst[i] : 0u;
  }
  bool bad = false;
#pragma unroll
  for (int j = 0; j < kRsItems; j++) {
    const uint64_t i = base + (uint64_t)j * kRsThreads + threadIdx.x;
    if (i < n) {
      if (a[j] >= n_rows || b[j] >= n_cols) {
        bad = true;
        a[j] = 0;
        b[j] = 0;
      }
      const uint64_t k = ((uint64_t)a[j] << 32) | b[j];
      keys[i] = k;
      atomicAdd(&h[w][(uint32_t)(k >> shift) & 255u], 1u);
    }
  }
  if (bad) *err = 1;
  __syncthreads();
  for (int d = threadIdx.x; d < kRsBins; d += kRsThreads) {
    uint32_t c = 0;
#pragma unroll
    for (int ww = 0; ww < kRsWaves; ww++) c += h[ww][d];
    hist[(uint64_t)d * tiles + tile] = c;
  }
}

// Block d: the exclusive prefix of digit d's tile counts hist[d * tiles + t]
// over the tiles, in place, and the digit's total in dtot[d].  The scatter
// adds the totals of the smaller digits itself, so one launch replaces a
// three-launch scan of the whole table.
__global__ void __launch_bounds__(kScanThreads)
    k_rs_digit_scan(uint32_t *hist, uint32_t tiles, uint32_t *dtot) {
  __shared__ uint32_t wtot[kScanThreads / kWave];
  constexpr uint32_t kItems = 8;   // consecutive counts per thread and round
  constexpr uint32_t kRounds = 4;  // rounds whose loads are issued together
  constexpr uint32_t kRound = kScanThreads * kItems;
  uint32_t *h = hist + (uint64_t)blockIdx.x * tiles;
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < tiles; c0 += kRound * kRounds) {
    uint32_t v[kRounds][kItems];
#pragma unroll
    for (uint32_t r = 0; r < kRounds; r++)
#pragma unroll
      for (uint32_t j = 0; j < kItems; j++) {
        const uint32_t i = c0 + r * kRound + threadIdx.x * kItems + j;
        v[r][j] = i < tiles ? h[i] : 0u;
      }
#pragma unroll
    for (uint32_t r = 0; r < kRounds; r++) {
      const uint32_t i0 = c0 + r * kRound + threadIdx.x * kItems;
      uint32_t sum = 0;
#pragma unroll
      for (uint32_t j = 0; j < kItems; j++) sum += v[r][j];
      uint32_t tot;
      uint32_t run = carry + block_exclusive(sum, wtot, &tot);
#pragma unroll
      for (uint32_t j = 0; j < kItems; j++) {
        if (i0 + j < tiles) h[i0 + j] = run;
        run += v[r][j];
      }
      carry += tot;
    }
  }
  if (threadIdx.x == 0) dtot[blockIdx.x] = carry;
}

// Stable scatter of one tile.  Keys are read wave-striped (wave w, item j,
// lane l -> tile offset w*64*ITEMS + j*64 + l), which is also their order for
// stability.  Lanes with equal digits are found with 8 ballots; ranks are
// counted per wave in LDS, turned into tile-local positions, the tile is
// sorted in LDS and then written out digit run by digit run (coalesced).
__global__ void __launch_bounds__(kRsThreads)
    k_rs_scatter(const uint64_t *in, uint64_t *out, uint64_t n, int shift,
                 const uint32_t *offs, const uint32_t *digit_tot,
                 uint32_t tiles) {
  static_assert(kRsThreads == kRsBins, "one thread per digit");
  __shared__ uint64_t stage[kRsTile];
  __shared__ uint32_t wcnt[kRsWaves][kRsBins];
  __shared__ uint32_t tstart[kRsBins];  // tile-local start of each digit
  __shared__ uint32_t gstart[kRsBins];  // global start of each digit
  __shared__ uint32_t wtot[kRsWaves];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  // consecutive tiles on one XCD: the digit runs of tiles t and t + 1 meet
  // inside one cache line, which then fills in one L2 (with tiles dealt
  // round-robin over the XCDs, two L2s wrote each such line in parts)
  const uint32_t tile = xcd_block(blockIdx.x, gridDim.x);
  for (int i = threadIdx.x; i < kRsWaves * kRsBins; i += kRsThreads)
    (&wcnt[0][0])[i] = 0;
  {  // global start of digit d: the tile's prefix within the digit plus the
     // totals of the smaller digits
    const uint32_t d = threadIdx.x;
    const uint32_t own = offs[(uint64_t)d * tiles + tile], t = digit_tot[d];
    const uint32_t incl = wave_inclusive_sum(t);
    if (lane == kWave - 1) wtot[w] = incl;
    __syncthreads();
    uint32_t woff = 0;
#pragma unroll
    for (int ww = 0; ww < kRsWaves; ww++)
      if (ww < w) woff += wtot[ww];
    gstart[d] = own + woff + incl - t;
  }
  __syncthreads();

  const uint64_t base = (uint64_t)tile * kRsTile +
                        (uint64_t)w * (kWave * kRsItems);
  uint64_t key[kRsItems];
  uint32_t rank[kRsItems];
  // all of a lane's keys in flight before the ranking (ranking each key as
  // it arrived waited for every load in turn)
#pragma unroll
  for (int j = 0; j < kRsItems; j++) {
    uint64_t i = base + (uint64_t)j * kWave + lane;
    key[j] = i < n ? in[i] : 0;
  }
#pragma unroll
  for (int j = 0; j < kRsItems; j++) {
    uint64_t i = base + (uint64_t)j * kWave + lane;
    bool valid = i < n;
    uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    uint32_t before = wcnt[w][d];
    uint32_t r = mask_rank(peers);
    rank[j] = before + r;
    // the lowest lane of each peer group advances the wave's digit count
    if (valid && r == 0) wcnt[w][d] = before + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // per digit: exclusive offsets across waves, tile total
  uint32_t dtot = 0;
  if (threadIdx.x < kRsBins) {
    const int d = threadIdx.x;
    uint32_t run = 0;
#pragma unroll
    for (int ww = 0; ww < kRsWaves; ww++) {
      uint32_t c = wcnt[ww][d];
      wcnt[ww][d] = run;
      run += c;
    }
    dtot = run;
  }
  // exclusive scan of digit totals over 256 digits (one per thread)
  {
    uint32_t incl = wave_inclusive_sum(dtot);
    if (lane == kWave - 1) wtot[w] = incl;
    __syncthreads();
    uint32_t woff = 0;
#pragma unroll
    for (int ww = 0; ww < kRsWaves; ww++)
      if (ww < w) woff += wtot[ww];
    if (threadIdx.x < kRsBins) tstart[threadIdx.x] = woff + incl - dtot;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kRsItems; j++) {
    uint64_t i = base + (uint64_t)j * kWave + lane;
    if (i < n) {
      uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
      stage[tstart[d] + wcnt[w][d] + rank[j]] = key[j];
    }
  }
  __syncthreads();
  const uint64_t tbase = (uint64_t)tile * kRsTile;
  const uint32_t valid_n =
      (uint32_t)((n - tbase) < (uint64_t)kRsTile ? (n - tbase) : kRsTile);
  for (uint32_t i = threadIdx.x; i < valid_n; i += kRsThreads) {
    uint64_t k = stage[i];
    uint32_t d = (uint32_t)(k >> shift) & 255u;
    out[(uint64_t)gstart[d] + (i - tstart[d])] = k;
  }
}

// One-block LSD radix sort of up to kSsMax keys.  The keys stay in LDS
// across all passes (each pass ranks, scans and scatters in place), so a
// short batch costs one launch instead of three per pass.  Ranking is the
// same wave-striped stable ballot ranking as k_rs_scatter, with 16 waves.
constexpr int kSsThreads = 1024;
constexpr int kSsWaves = kSsThreads / kWave;
constexpr int kSsItems = 16;
constexpr uint32_t kSsMax = kSsThreads * kSsItems;  // 16384 keys, 128 KB

struct SortShifts {
  int shift[16];
  int passes;
};

// Block b sorts keys[b * chunk, min(n, (b + 1) * chunk)), chunk <= kSsMax.
__global__ void __launch_bounds__(kSsThreads)
    k_sort_small(uint64_t *keys, uint32_t n, SortShifts sh, uint32_t chunk) {
  __shared__ uint64_t stage[kSsMax];
  keys += (uint64_t)blockIdx.x * chunk;
  n = min(n - blockIdx.x * chunk, chunk);
  __shared__ uint32_t wcnt[kSsWaves][kRsBins];
  __shared__ uint32_t tstart[kRsBins];
  __shared__ uint32_t wtot[kSsWaves];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  // each wave owns a contiguous run of `per` 64-key items (stability needs
  // wave order = key order), sized to n so that short batches use all waves
  const uint32_t per = (((n + kWave - 1) / kWave) + kSsWaves - 1) / kSsWaves;
  const uint32_t base = (uint32_t)w * per * kWave;
  for (uint32_t i = threadIdx.x; i < n; i += kSsThreads) stage[i] = keys[i];
  for (int p = 0; p < sh.passes; p++) {
    const int shift = sh.shift[p];
    for (int i = threadIdx.x; i < kSsWaves * kRsBins; i += kSsThreads)
      (&wcnt[0][0])[i] = 0;
    __syncthreads();
    uint64_t key[kSsItems];
    uint32_t rank[kSsItems];
#pragma unroll
    for (int j = 0; j < kSsItems; j++) {
      if ((uint32_t)j >= per) break;  // uniform across the block
      const uint32_t i = base + (uint32_t)j * kWave + lane;
      const bool valid = i < n;
      key[j] = valid ? stage[i] : 0;
      const uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
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
    if (threadIdx.x < kRsBins) {
      const int d = threadIdx.x;
      uint32_t run = 0;
#pragma unroll
      for (int ww = 0; ww < kSsWaves; ww++) {
        const uint32_t c = wcnt[ww][d];
        wcnt[ww][d] = run;
        run += c;
      }
      dtot = run;
    }
    {
      const uint32_t incl = wave_inclusive_sum(dtot);
      if (lane == kWave - 1) wtot[w] = incl;
      __syncthreads();
      uint32_t woff = 0;
#pragma unroll
      for (int ww = 0; ww < kRsBins / kWave; ww++)
        if (ww < w) woff += wtot[ww];
      if (threadIdx.x < kRsBins) tstart[threadIdx.x] = woff + incl - dtot;
    }
    __syncthreads();
    // every key of the pass is in registers: scatter in place
#pragma unroll
    for (int j = 0; j < kSsItems; j++) {
      if ((uint32_t)j >= per) break;
      const uint32_t i = base + (uint32_t)j * kWave + lane;
      if (i < n) {
        const uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
        stage[tstart[d] + wcnt[w][d] + rank[j]] = key[j];
      }
    }
    // the scatter reads wcnt and tstart: no wave may clear them for the
    // next pass before every wave has placed its keys
    __syncthreads();
  }
  for (uint32_t i = threadIdx.x; i < n; i += kSsThreads) keys[i] = stage[i];
}

// ---------------------------------------------------------- merge sort
constexpr uint32_t kMsChunk = kBlockSortMax;

// Block b sorts keys[b * kMsChunk, +kMsChunk) (the last chunk may be short)
// by whole-key comparison (block_sort.hpp).
__global__ void __launch_bounds__(kSsThreads)
    k_sort_chunks(uint64_t *keys, uint64_t n) {
  __shared__ uint64_t s[kMsChunk];
  keys += (uint64_t)blockIdx.x * kMsChunk;
  const uint32_t m = (uint32_t)min(n - (uint64_t)blockIdx.x * kMsChunk,
                                   (uint64_t)kMsChunk);
  {  // the chunk's loads in flight together
    constexpr int kPer = kMsChunk / kSsThreads;
    uint64_t v[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const uint32_t i = threadIdx.x + (uint32_t)j * kSsThreads;
      v[j] = i < m ? keys[i] : 0;
    }
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const uint32_t i = threadIdx.x + (uint32_t)j * kSsThreads;
      if (i < m) s[i] = v[j];
    }
  }
  block_sort_u64<kSsThreads>(s, m);  // begins and ends with a barrier
  for (uint32_t i = threadIdx.x; i < m; i += kSsThreads) keys[i] = s[i];
}

// Keys of n <= kMergeSortMax: k_sort_chunks sorts runs of kMsChunk keys
// (one block each, all in parallel), then each k_merge_pass doubles the
// sorted runs in one launch.  A 10^5-key batch takes 6 launches where the
// radix sort takes 17 (6 digit passes of histogram, count scan and
// scatter, the first histogram counted while packing).
constexpr int kMpThreads = 256;
constexpr int kMpItems = 8;
constexpr uint32_t kMpOut = kMpThreads * kMpItems;  // outputs per block
// above this the radix sort wins (insert at scale 24, radix against merge:
// 1e6 keys 0.670 against 0.703 ms, 7e5 0.565 against 0.587, 2^19 0.473
// against 0.483, 3.5e5 0.384 either way, 2^18 + 1 0.332 against 0.329)
constexpr uint64_t kMergeSortMax = 1ull << 18;

// Number of a[0..la) among the first d outputs of the merge of a and b
// (ties from a first): the smallest i with a[i] > b[d - 1 - i], or the
// largest possible i.  One wave tests up to 64 candidates per round.
__device__ uint64_t mp_split_wave(const uint64_t *a, uint64_t la,
                                  const uint64_t *b, uint64_t lb, uint64_t d) {
  uint64_t lo = d > lb ? d - lb : 0, hi = d < la ? d : la;  // answer in [lo, hi]
  const uint32_t lane = lane_id();
  while (hi > lo) {  // uniform across the wave
    const uint64_t span = hi - lo;
    const uint64_t q = span <= (uint64_t)kWave
                           ? lo + lane
                           : lo + ((span * (lane + 1)) / (kWave + 1));
    const bool valid = q < hi;
    // the candidates a[q] > b[d-1-q] do not all lie below the answer
    const bool below = valid && a[q] <= b[d - 1 - q];
    const uint64_t mb = __ballot(below), mv = __ballot(valid);
    const uint32_t t = (uint32_t)__popcll(mb);  // a prefix of the valid lanes
    const uint32_t nv = (uint32_t)__popcll(mv);
    uint64_t nlo = lo, nhi = hi;
    if (t > 0) nlo = __shfl(q, t - 1, kWave) + 1;
    if (t < nv) nhi = __shfl(q, t, kWave);
    lo = nlo;
    hi = nhi;
  }
  return lo;
}

// src holds sorted runs of `run` keys (the last may be shorter); dst gets
// sorted runs of 2 * run.  bpp blocks per pair of runs: block k of a pair
// merges the pair's outputs [k * kMpOut, (k + 1) * kMpOut).  Waves 0 and 1
// find where that range starts and ends in the two runs, the block stages
// both pieces in LDS, and each thread merges kMpItems outputs from its own
// split.
__global__ void __launch_bounds__(kMpThreads)
    k_merge_pass(const uint64_t *src, uint64_t *dst, uint64_t n, uint64_t run,
                 uint32_t bpp) {
  __shared__ uint64_t sa[kMpOut], sb[kMpOut];
  __shared__ uint64_t split[2];
  const uint64_t pair = blockIdx.x / bpp;
  const uint64_t pb = (uint64_t)(blockIdx.x % bpp) * kMpOut;
  const uint64_t a0 = pair * 2 * run;
  if (a0 >= n) return;  // uniform across the block
  const uint64_t la = min(run, n - a0);
  const uint64_t b0 = a0 + la;
  const uint64_t lb = b0 < n ? min(run, n - b0) : 0;
  const uint64_t tot = la + lb;
  if (pb >= tot) return;  // uniform
  const uint64_t pe = min(pb + (uint64_t)kMpOut, tot);
  const uint32_t w = threadIdx.x / kWave;
  if (w < 2) {
    const uint64_t i = mp_split_wave(src + a0, la, src + b0, lb, w ? pe : pb);
    if (lane_id() == 0) split[w] = i;
  }
  __syncthreads();
  const uint64_t ia0 = split[0], ia1 = split[1];
  const uint32_t na = (uint32_t)(ia1 - ia0);
  const uint32_t nb = (uint32_t)((pe - ia1) - (pb - ia0));
  const uint64_t ib0 = pb - ia0;
  {  // both pieces (na + nb <= kMpOut) in flight before the LDS stores
    uint64_t va[kMpItems], vb[kMpItems];
#pragma unroll
    for (int j = 0; j < kMpItems; j++) {
      const uint32_t i = threadIdx.x + (uint32_t)j * kMpThreads;
      va[j] = i < na ? src[a0 + ia0 + i] : 0;
      vb[j] = i < nb ? src[b0 + ib0 + i] : 0;
    }
#pragma unroll
    for (int j = 0; j < kMpItems; j++) {
      const uint32_t i = threadIdx.x + (uint32_t)j * kMpThreads;
      if (i < na) sa[i] = va[j];
      if (i < nb) sb[i] = vb[j];
    }
  }
  __syncthreads();
  const uint32_t cnt = (uint32_t)(pe - pb);
  const uint32_t d = threadIdx.x * kMpItems;
  if (d >= cnt) return;
  uint32_t lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
  while (lo < hi) {
    const uint32_t m = (lo + hi) >> 1;
    if (sa[m] <= sb[d - 1 - m])
      lo = m + 1;
    else
      hi = m;
  }
  uint32_t i = lo, j = d - lo;
  uint64_t *out = dst + a0 + pb + d;
#pragma unroll
  for (int k = 0; k < kMpItems; k++) {
    if (d + k >= cnt) break;
    const bool from_a = j >= nb || (i < na && sa[i] <= sb[j]);
    out[k] = from_a ? sa[i++] : sb[j++];
  }
}

}  // namespace

int scan_exclusive_u64(const uint64_t *in, uint64_t *out, uint64_t n,
                       uint64_t *d_total, Workspace &ws, int ws_base,
                       hipStream_t s) {
  return scan_impl<uint64_t>(in, out, n, d_total, ws, ws_base, s);
}

int scan_exclusive_u32(const uint32_t *in, uint32_t *out, uint64_t n,
                       uint32_t *d_total, Workspace &ws, int ws_base,
                       hipStream_t s) {
  return scan_impl<uint32_t>(in, out, n, d_total, ws, ws_base, s);
}

template <class V>
static int select_flagged(const uint8_t *flags, const V *values, uint64_t n,
                          V *out, uint64_t *d_count, Workspace &ws,
                          int ws_base, hipStream_t s,
                          const uint32_t *values2 = nullptr,
                          uint32_t *out2 = nullptr) {
  if (n == 0) {
    SSTG_HIP(hipMemsetAsync(d_count, 0, sizeof(uint64_t), s));
    return SSTG_OK;
  }
  if (ceil_div(n, kScanTile) <= kScanOneBlock) {
    SSTG_LAUNCH(k_select_range<V>, 1, kScanThreads, 0, s, flags, values, n, n,
                (const uint64_t *)nullptr, out, d_count, values2, out2);
    return SSTG_OK;
  }
  const Ranges r = ranges_of(n);
  uint64_t *counts = nullptr;
  SSTG_TRY(ws.get(ws_base, r.blocks, &counts));
  SSTG_LAUNCH((k_range_reduce<uint64_t, FlagIn>), r.blocks, kScanThreads, 0, s,
              FlagIn{flags}, n, r.per, counts);
  SSTG_LAUNCH(k_scan_partials<uint64_t>, 1, kScanThreads, 0, s, counts,
              r.blocks);
  SSTG_LAUNCH(k_select_range<V>, r.blocks, kScanThreads, 0, s, flags, values,
              n, r.per, (const uint64_t *)counts, out, d_count, values2, out2);
  return SSTG_OK;
}

int select_flagged_u64(const uint8_t *flags, const uint64_t *values,
                       uint64_t n, uint64_t *out, uint64_t *d_count,
                       Workspace &ws, int ws_base, hipStream_t s) {
  return select_flagged<uint64_t>(flags, values, n, out, d_count, ws, ws_base,
                                  s);
}

int select_flagged_u64_u32(const uint8_t *flags, const uint64_t *values,
                           const uint32_t *values2, uint64_t n, uint64_t *out,
                           uint32_t *out2, uint64_t *d_count, Workspace &ws,
                           int ws_base, hipStream_t s) {
  return select_flagged<uint64_t>(flags, values, n, out, d_count, ws, ws_base,
                                  s, values2, out2);
}

int select_flagged_u32(const uint8_t *flags, const uint32_t *values,
                       uint64_t n, uint32_t *out, uint64_t *d_count,
                       Workspace &ws, int ws_base, hipStream_t s) {
  return select_flagged<uint32_t>(flags, values, n, out, d_count, ws, ws_base,
                                  s);
}

// The radix passes; with pairs, pass 0's histogram comes from the packing
// kernel, which writes the keys.
struct PackedPairs {
  const uint32_t *src, *dst;
  uint32_t n_rows, n_cols;
  uint64_t *err;
};

static int radix_passes(uint64_t *keys, uint64_t *tmp, uint64_t n,
                        const int *shifts, int npass, Workspace &ws,
                        int ws_base, hipStream_t s, uint64_t **result,
                        const PackedPairs *pairs) {
  uint32_t tiles = (uint32_t)ceil_div(n, kRsTile);
  uint32_t *hist = nullptr;
  SSTG_TRY(ws.get(ws_base, (uint64_t)tiles * kRsBins, &hist));
  uint32_t *dtot = nullptr;
  SSTG_TRY(ws.get(ws_base + 1, kRsBins, &dtot));
  uint64_t *src = keys, *dst = tmp;
  for (int p = 0; p < npass; p++) {
    if (p == 0 && pairs)
      SSTG_LAUNCH(k_rs_pack_hist, tiles, kRsThreads, 0, s, pairs->src,
                  pairs->dst, n, pairs->n_rows, pairs->n_cols, keys,
                  pairs->err, shifts[0], hist, tiles);
    else
      SSTG_LAUNCH(k_rs_hist, tiles, kRsThreads, 0, s, (const uint64_t *)src,
                  n, shifts[p], hist, tiles);
    SSTG_LAUNCH(k_rs_digit_scan, kRsBins, kScanThreads, 0, s, hist, tiles,
                dtot);
    SSTG_LAUNCH(k_rs_scatter, tiles, kRsThreads, 0, s, (const uint64_t *)src,
                dst, n, shifts[p], (const uint32_t *)hist,
                (const uint32_t *)dtot, tiles);
    uint64_t *t = src;
    src = dst;
    dst = t;
  }
  *result = src;
  return SSTG_OK;
}

int radix_sort_u64(uint64_t *keys, uint64_t *tmp, uint64_t n,
                   const int *shifts, int npass, Workspace &ws, int ws_base,
                   hipStream_t s, uint64_t **result) {
  *result = keys;
  if (n <= 1 || npass == 0) return SSTG_OK;
  if (n >= (1ull << 32) || npass > 16) return SSTG_E_INVAL;
  if (n <= kSsMax) {
    SortShifts sh{};
    for (int p = 0; p < npass; p++) sh.shift[p] = shifts[p];
    sh.passes = npass;
    SSTG_LAUNCH(k_sort_small, 1, kSsThreads, 0, s, keys, (uint32_t)n, sh,
                kSsMax);
    return SSTG_OK;
  }
  return radix_passes(keys, tmp, n, shifts, npass, ws, ws_base, s, result,
                      nullptr);
}

int sort_keys_u64(uint64_t *keys, uint64_t *tmp, uint64_t n, const int *shifts,
                  int npass, Workspace &ws, int ws_base, hipStream_t s,
                  uint64_t **result) {
  *result = keys;
  if (n <= 1 || npass == 0) return SSTG_OK;
  if (n > kMergeSortMax)
    return radix_sort_u64(keys, tmp, n, shifts, npass, ws, ws_base, s, result);
  SSTG_LAUNCH(k_sort_chunks, (uint32_t)ceil_div(n, (uint64_t)kMsChunk),
              kSsThreads, 0, s, keys, n);
  uint64_t *src = keys, *dst = tmp;
  for (uint64_t run = kMsChunk; run < n; run *= 2) {
    const uint64_t pairs = ceil_div(n, 2 * run);
    const uint32_t bpp = (uint32_t)ceil_div(2 * run, (uint64_t)kMpOut);
    SSTG_LAUNCH(k_merge_pass, pairs * bpp, kMpThreads, 0, s,
                (const uint64_t *)src, dst, n, run, bpp);
    uint64_t *t = src;
    src = dst;
    dst = t;
  }
  *result = src;
  return SSTG_OK;
}

int pack_sort_pairs(const uint32_t *src, const uint32_t *dst, uint64_t n,
                    uint32_t n_rows, uint32_t n_cols, uint64_t *err,
                    uint64_t *keys, uint64_t *tmp, const int *shifts,
                    int npass, Workspace &ws, int ws_base, hipStream_t s,
                    uint64_t **result) {
  *result = keys;
  if (n == 0) return SSTG_OK;
  if (n > kMergeSortMax && npass > 0) {
    if (n >= (1ull << 32) || npass > 16) return SSTG_E_INVAL;
    const PackedPairs pairs{src, dst, n_rows, n_cols, err};
    return radix_passes(keys, tmp, n, shifts, npass, ws, ws_base, s, result,
                        &pairs);
  }
  SSTG_LAUNCH(k_pack_pairs, ceil_div(n, (uint64_t)kRsThreads), kRsThreads, 0,
              s, src, dst, n, n_rows, n_cols, keys, err);
  return sort_keys_u64(keys, tmp, n, shifts, npass, ws, ws_base, s, result);
}

}  // namespace sstg
