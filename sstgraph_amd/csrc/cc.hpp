// Connected components by union-find on the device store, shared by the
// single-graph CC (algo.hip) and the many-graph pass (multi.hip).
//
// The reference's harness only uses the partition CC returns
// (internal/test.hpp:571-601); labels here are the smallest vertex id of each
// component, kept so by always hooking the larger root under the smaller
// (p[x] <= x throughout, so every root is its component's minimum).
//
// Symmetric graphs take an Afforest-style schedule: each vertex first links
// to kSampleLinks neighbours at hashed row positions (the smallest
// neighbours would funnel whole graphs onto a few roots and their CAS), the
// forest is compressed, the most common root among 64 samples per graph is
// taken as that graph's large component, and only vertices outside it link
// all of their edges.  A skipped edge (v, w) with v inside the large
// component is still linked from w's side, because w's row holds v; if w is
// inside too, the edge joins nothing new.  Rows without their reverses link
// every stored edge.
#pragma once

#include "graph.hpp"

namespace sstg {
namespace cc {
namespace {  // internal linkage: included by several files

constexpr int kT = 256;
constexpr uint32_t kSampleLinks = 2;
constexpr uint32_t kInfRoot = 0xFFFFFFFFu;  // no vertex: ids are below n

__device__ __forceinline__ uint32_t find(uint32_t *p, uint32_t x) {
  uint32_t y = p[x];
  while (y != x) {
    uint32_t z = p[y];
    if (z != y) p[x] = z;  // path halving; only ever lowers p[x]
    x = y;
    y = z;
  }
  return x;
}

// union of the sets of u and w: the larger root is hooked under the smaller
__device__ __forceinline__ void link(uint32_t *p, uint32_t u, uint32_t w) {
  uint32_t ru = find(p, u), rw = find(p, w);
  while (ru != rw) {
    const uint32_t hi = ru > rw ? ru : rw, lo = ru > rw ? rw : ru;
    const uint32_t old = atomicCAS(&p[hi], hi, lo);
    if (old == hi) break;
    ru = find(p, old);
    rw = find(p, lo);
  }
}

__device__ __forceinline__ uint32_t sample_hash(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

__global__ void k_identity(uint32_t n, uint32_t *p) {
  const uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v < n) p[v] = (uint32_t)v;
}

// p[v] = min(v, first neighbour): rows are sorted, so this is a link
// without atomics (the start of the every-edge schedule)
__global__ void k_init(DevGraph g, uint32_t *p) {
  const uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v >= g.n) return;
  uint32_t m = (uint32_t)v;
  if (g.deg[v]) {
    const uint32_t first = g.pool[g.off[v]];
    if (first < m) m = first;
  }
  p[v] = m;
}

// kSampleLinks neighbours of every vertex, at hashed row positions
__global__ void k_link_sample(DevGraph g, uint32_t *p) {
  const uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v >= g.n) return;
  const uint32_t d = g.deg[v];
  if (d == 0) return;
  const uint32_t *row = g.pool + g.off[v];
  const uint32_t h = sample_hash((uint32_t)v);
  for (uint32_t k = 0; k < kSampleLinks && k < d; k++)
    link(p, (uint32_t)v, row[(h + k) % d]);
}

__global__ void k_compress(uint32_t n, uint32_t *p) {
  const uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v >= n) return;
  p[v] = find(p, (uint32_t)v);
}

// giant[i] = the most common root among 64 sampled vertices of graph i
// (vertices [vbase[i], vbase[i+1]), or [0, n) when vbase is null); one wave
// per graph, ties to the lowest sampled lane.  Positions are hashed: evenly
// spaced ids share their low bits, and in rMat graphs ids with many zero bits
// are mostly isolated.
__global__ void k_mode(const uint32_t *vbase, uint32_t graphs, uint32_t n,
                       uint32_t *p, uint32_t *giant) {
  const uint64_t gi = ((uint64_t)blockIdx.x * kT + threadIdx.x) / kWave;
  const uint32_t lane = lane_id();
  if (gi >= graphs) return;  // whole waves leave together
  const uint32_t b = vbase ? vbase[gi] : 0u;
  const uint32_t e = vbase ? vbase[gi + 1] : n;
  const uint32_t len = e - b;
  uint32_t r = kInfRoot;
  if (len)
    r = find(p, b + sample_hash((uint32_t)gi * kWave + lane + 1u) % len);
  uint32_t best = 0, best_r = kInfRoot;
  for (int l = 0; l < kWave; l++) {
    const uint32_t x = __shfl(r, l, kWave);
    uint32_t c = (uint32_t)__popcll(__ballot(x != kInfRoot && r == x));
    if (c > best) {
      best = c;
      best_r = x;
    }
  }
  if (lane == 0) giant[gi] = best_r;
}

// every edge of the vertices outside their graph's large component
__global__ void k_link_rest(DevGraph g, uint32_t *p, const uint32_t *giant,
                            const uint32_t *graph_of) {
  const uint64_t v = (uint64_t)blockIdx.x * kT + threadIdx.x;
  if (v >= g.n) return;
  const uint32_t d = g.deg[v];
  if (d <= kSampleLinks) return;  // all of its edges were sampled
  // inside the large component now: its edges are linked from the other
  // side or join nothing new (compression races can leave p[v] one hop
  // short of the root, so roots are compared, not p[v])
  const uint32_t c = giant[graph_of ? graph_of[v] : 0u];
  if (find(p, (uint32_t)v) == find(p, c)) return;
  const uint32_t *row = g.pool + g.off[v];
  for (uint32_t k = 0; k < d; k++) link(p, (uint32_t)v, row[k]);
}

// every stored edge, flat over edges (8 per thread)
__global__ void k_hook_all(DevGraph g, const uint64_t *prefix, uint64_t m,
                           uint32_t *p) {
  uint64_t e = ((uint64_t)blockIdx.x * kT + threadIdx.x) * 8;
  if (e >= m) return;
  const uint64_t end = e + 8 < m ? e + 8 : m;
  uint64_t v = find_segment(prefix, g.n, e);
  for (; e < end; e++) {
    while (e >= prefix[v + 1]) v++;
    link(p, (uint32_t)v, g.pool[g.off[v] + (e - prefix[v])]);
  }
}

}  // namespace
}  // namespace cc
}  // namespace sstg
