"""Seeded workload definitions shared by bench.py and the tests.

These are the harness side of the hot path (test.hpp:396-697): the update
batches of the dynamic-update loop, the 512-graph batched workload and the
served request trace.  Everything here is host-side numpy; edge streams come
from the rMat generator (rmat_util.h:82-131) through a caller-supplied
function so the same definition serves the device generator and any other
generator that is bit-identical to it.
"""
import numpy as np


def hash64_np(u):
    """rmat_util.h:29-39 (hash64) on uint64 numpy arrays, wrapping."""
    u = np.asarray(u).astype(np.uint64)
    with np.errstate(over="ignore"):
        v = u * np.uint64(3935559000370003845) + np.uint64(2691343689449507681)
        v ^= v >> np.uint64(21)
        v ^= v << np.uint64(37)
        v ^= v >> np.uint64(4)
        v *= np.uint64(4768777513237032717)
        v ^= v << np.uint64(20)
        v ^= v >> np.uint64(41)
        v ^= v << np.uint64(5)
    return v


UPDATE_SIZES = (1000, 10000, 100000, 1000000, 10000000)


def update_batch(rmat, n, off, cols, size, seed):
    """One batch of the dynamic-update workload (test.hpp:636-690 shape):
    80% fresh rMat(n, seed) edges followed by 20% edges of the graph given as
    CSR (off, cols), edge index hash64(seed << 32 | j) mod m for j = 0, 1, ..
    rmat(n, seed, count) -> (src, dst) host arrays.  Returns (src, dst)."""
    fresh = int(size * 0.8)
    fs, fd = rmat(n, seed, fresh)
    m = int(off[-1])
    j = np.arange(size - fresh, dtype=np.uint64) + np.uint64(seed) * np.uint64(1 << 32)
    if m == 0 or size == fresh:
        return np.ascontiguousarray(fs, np.uint32), np.ascontiguousarray(fd, np.uint32)
    pick = (hash64_np(j) % np.uint64(m)).astype(np.int64)
    rows = (np.searchsorted(off, pick, side="right") - 1).astype(np.uint32)
    return (np.concatenate([fs, rows]).astype(np.uint32),
            np.concatenate([fd, cols[pick]]).astype(np.uint32))


def c4_specs(k=512, seed=3):
    """Graph specs of the batched workload: scale = 8 + floor(11 U) in
    [8, 18]; even graphs rMat with edge factor 8..32, odd graphs Erdos-Renyi
    with average degree 4..64 (uniforms from hash64 of the seed).  Each spec
    is (kind, scale, param, generator seed)."""
    specs = []
    for g in range(k):
        u = hash64_np(np.array([(seed << 32) | (3 * g + j) for j in range(3)],
                               dtype=np.uint64)).astype(np.float64) / 2.0**64
        scale = min(18, 8 + int(11 * u[0]))
        if g % 2 == 0:
            specs.append(("rmat", scale, 8 + int(25 * u[1]), 1000 + g))
        else:
            specs.append(("er", scale, 4 + int(61 * u[2]), 2000 + g))
    return specs


def c4_edge_count(spec):
    """Generated (one-directional) edges of a batched-workload graph."""
    kind, scale, param, _ = spec
    n = 1 << scale
    return param * n if kind == "rmat" else param * n // 2


def c4_host_edges(spec, rmat, er):
    """Symmetrised edge stream of one batched-workload graph as host arrays:
    the generated edges (u, v) followed by their reverses."""
    kind, scale, param, seed = spec
    n = 1 << scale
    count = c4_edge_count(spec)
    s, d = rmat(n, seed, count) if kind == "rmat" else er(n, seed, count)
    return n, np.concatenate([s, d]), np.concatenate([d, s])


def serve_trace(n, deg, count, rate, seed=4):
    """Seeded request trace: Poisson arrivals at `rate` per second; 55% update
    batches (70% inserts of fresh rMat edges, k log-uniform in [1e2, 1e6],
    30% removes of an earlier inserted batch), 25% BFS from a random
    non-isolated vertex, 12% PageRank (10 iterations), 8% TC.  Returns
    (arrival offsets, requests); an update is ("insert"|"remove", (k, seed))."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    nz = np.nonzero(deg)[0]
    reqs, inserted = [], []
    for i in range(count):
        r = rng.random()
        if r < 0.55:
            if inserted and rng.random() < 0.3:
                reqs.append(("remove", inserted[int(rng.integers(len(inserted)))]))
            else:
                size = int(10 ** rng.uniform(2, 6))
                reqs.append(("insert", (size, 100 + i)))
                inserted.append((size, 100 + i))
        elif r < 0.80:
            reqs.append(("bfs", int(nz[int(rng.integers(len(nz)))])))
        elif r < 0.92:
            reqs.append(("pr", 10))
        else:
            reqs.append(("tc", None))
    return t, reqs


def serve_batches(rmat, n, reqs):
    """Host arrays of every update batch of a trace: k rMat(n, seed) edges
    applied with their reverses, so the served graph stays undirected."""
    batches = {}
    for kind, arg in reqs:
        if kind == "insert" and arg not in batches:
            bs, bd = rmat(n, arg[1], arg[0])
            batches[arg] = (np.concatenate([bs, bd]), np.concatenate([bd, bs]))
    return batches
