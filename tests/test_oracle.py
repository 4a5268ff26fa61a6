"""Pins the CPU oracle against the reference's own outputs (golden vectors
from the reference's rmat_util.h), networkx, closed forms and brute force.
CPU only."""
import json
import os
from collections import deque

import numpy as np
import pytest

from oracle import oracle as O
from tests.graphs import GOLDEN, csr_from_edges, random_graph, read_mtx


@pytest.fixture(scope="module", autouse=True)
def _built():
    if not os.path.exists(O.LIB_PATH):
        O.build()


# --------------------------------------------------------------- RMAT, hashes
def test_hash_kats_match_reference_build():
    kats = json.load(open(os.path.join(GOLDEN, "hash_kats.json")))
    for x, v in kats["h32"]:
        assert O.hash32(x) == v
    for x, v in kats["h64"]:
        assert O.hash64(x) == v
    for i, v in kats["aspen"]:
        assert int(O.lib().orc_aspen_ith_rand(0, i)) == v
    for x, v in kats["log2_up"]:
        assert int(O.lib().orc_log2_up(x)) == v


def test_rmat_matches_reference_build():
    cases = json.load(open(os.path.join(GOLDEN, "rmat_vectors.json")))
    assert len(cases) >= 10
    for c in cases:
        s, d = O.rmat(c["n"], c["seed"], c["count"], first=c["first"],
                      a=c["a"], b=c["b"], c=c["c"])
        assert s.tolist() == c["src"], (c["n"], c["seed"], c["first"])
        assert d.tolist() == c["dst"], (c["n"], c["seed"], c["first"])


def test_rmat_effective_quadrants():
    # hashDouble spans [0, 2): effective probabilities are a/2, b/2, c/2,
    # 1-(a+b+c)/2 (rmat_util.h:102-105).
    n = 1 << 20
    s, d = O.rmat(n, 9, 400_000)
    top_r, top_c = s >= n // 2, d >= n // 2
    q = [np.mean(~top_r & ~top_c), np.mean(~top_r & top_c),
         np.mean(top_r & ~top_c), np.mean(top_r & top_c)]
    np.testing.assert_allclose(q, [0.25, 0.05, 0.05, 0.65], atol=0.005)
    assert s.max() < n and d.max() < n


def test_rmat_normalized_quadrants_and_draws():
    """The normalised draw r = hash32 / 2^32 gives the stated probabilities;
    a few edges recomputed bit by bit from hash32 (itself pinned by the
    reference build's vectors) fix the recursion order and the scaling."""
    n = 1 << 20
    s, d = O.rmat(n, 9, 400_000, normalized=True)
    top_r, top_c = s >= n // 2, d >= n // 2
    q = [np.mean(~top_r & ~top_c), np.mean(~top_r & top_c),
         np.mean(top_r & ~top_c), np.mean(top_r & top_c)]
    np.testing.assert_allclose(q, [0.5, 0.1, 0.1, 0.3], atol=0.005)
    m32 = 0xFFFFFFFF
    h = O.hash32(9)
    for i in (0, 1, 12345, 399_999):
        start = O.hash32((2 * i * h) & m32)
        stride = O.hash32(((2 * i + 1) * h) & m32)
        r = c = 0
        for k in range(20):
            half = n >> (k + 1)
            x = O.hash32((start + k * stride) & m32) / 2.0**32
            if x < 0.5:
                pass
            elif x < 0.6:
                c += half
            elif x < 0.7:
                r += half
            else:
                r += half
                c += half
        assert (int(s[i]), int(d[i])) == (r, c), i
    s2, d2 = O.rmat(n, 9, 1000)
    assert not np.array_equal(s2, s[:1000])


def test_rmat_rejects_non_power_of_two():
    with pytest.raises(ValueError):
        O.rmat(1000, 0, 4)


# ---------------------------------------------------------- tinyset semantics
def _rows(g):
    off, cols = g.csr()
    return {i: cols[off[i]:off[i + 1]].tolist() for i in range(len(off) - 1)
            if off[i + 1] > off[i]}


def test_insert_batch_is_set_union_without_self_loops():
    rng = np.random.default_rng(1)
    n = 300
    g = O.Graph(n)
    ref = set()
    for step in range(6):
        k = int(rng.integers(1, 3000))
        s = rng.integers(0, n, k).astype(np.uint32)
        d = rng.integers(0, n, k).astype(np.uint32)
        g.insert_batch(s, d)
        ref |= {(a, b) for a, b in zip(s.tolist(), d.tolist()) if a != b}
        rows = _rows(g)
        got = {(r, c) for r, cs in rows.items() for c in cs}
        assert got == ref
        for cs in rows.values():  # ascending map order (Tinyset_test.cpp:153-196)
            assert cs == sorted(set(cs))
        assert g.num_edges() == len(ref)


def test_remove_batch_is_set_difference_including_self_loops():
    rng = np.random.default_rng(2)
    n = 200
    g = O.Graph(n)
    s = rng.integers(0, n, 5000).astype(np.uint32)
    d = rng.integers(0, n, 5000).astype(np.uint32)
    g.insert_batch(s, d)
    g.insert(5, 5)  # single insert keeps self-loops (SparseMatrix.hpp:536-544)
    ref = {(a, b) for a, b in zip(s.tolist(), d.tolist()) if a != b} | {(5, 5)}
    assert g.has(5, 5)
    rs = np.concatenate([rng.integers(0, n, 3000), [5]]).astype(np.uint32)
    rd = np.concatenate([rng.integers(0, n, 3000), [5]]).astype(np.uint32)
    g.remove_batch(rs, rd)
    ref -= set(zip(rs.tolist(), rd.tolist()))
    got = {(r, c) for r, cs in _rows(g).items() for c in cs}
    assert got == ref and not g.has(5, 5)


def test_insert_then_remove_same_batch_drops_preexisting_edges():
    # test.hpp:668-677 inserts then removes the same array: edges of the batch
    # that were present before are removed as well.
    rng = np.random.default_rng(3)
    n = 128
    g = O.Graph(n)
    s = rng.integers(0, n, 2000).astype(np.uint32)
    d = rng.integers(0, n, 2000).astype(np.uint32)
    g.insert_batch(s, d)
    before = {(r, c) for r, cs in _rows(g).items() for c in cs}
    bs = rng.integers(0, n, 500).astype(np.uint32)
    bd = rng.integers(0, n, 500).astype(np.uint32)
    g.insert_batch(bs, bd)
    g.remove_batch(bs, bd)
    after = {(r, c) for r, cs in _rows(g).items() for c in cs}
    assert after == before - set(zip(bs.tolist(), bd.tolist()))


def test_zero_and_0xffff_multiples_stay_present():
    # Tinyset_test.cpp:198-240: keys with (i & 0xFFFF) == 0, including 0.
    maxv = 1 << 20
    g = O.Graph(2, maxv)
    keys = np.arange(0, maxv, 0x10000, dtype=np.uint32)
    rng = np.random.default_rng(4)
    other = rng.integers(0, maxv, 20000).astype(np.uint32)
    allk = np.concatenate([keys, other])
    rng.shuffle(allk)
    g.insert_batch(np.ones(len(allk), np.uint32), allk)
    for k in keys.tolist():
        assert g.has(1, k)
    assert g.has(1, 0)


def test_matrix_cellwise_and_touch_all_sum():
    # SparseMatrix_test.cpp:8-118 (binary variant): cell-by-cell membership,
    # touch_all_sum equals the sum of keys, M() == 0 after removing all.
    rng = np.random.default_rng(5)
    rows_n = 100_000
    r = rng.integers(0, rows_n - 1, 20000).astype(np.uint32)
    c = rng.integers(0, rows_n - 1, 20000).astype(np.uint32)
    g = O.Graph(rows_n)
    g.insert_batch(r, c)
    ref = {(a, b) for a, b in zip(r.tolist(), c.tolist()) if a != b}
    for a, b in list(ref)[:2000]:
        assert g.has(a, b)
    assert g.touch_all_sum() == sum(b for _, b in ref)
    rr = np.array([a for a, _ in ref], np.uint32)
    cc = np.array([b for _, b in ref], np.uint32)
    g.remove_batch(rr, cc)
    assert g.num_edges() == 0


def test_out_of_range_ids_rejected_and_graph_unchanged():
    g = O.Graph(10)
    g.insert_batch(np.array([1, 2], np.uint32), np.array([3, 4], np.uint32))
    with pytest.raises(ValueError):
        g.insert_batch(np.array([1, 10], np.uint32), np.array([2, 3], np.uint32))
    assert g.num_edges() == 2


# ---------------------------------------------------------------- algorithms
def _golden_graph(key):
    gg = json.load(open(os.path.join(GOLDEN, "graph_goldens.json")))[key]
    n, s, d = O.rmat_graph_edges(gg["scale"], gg["edge_factor"], gg["seed"])
    g = O.Graph(n)
    g.insert_batch(s, d)
    return gg, g


@pytest.mark.parametrize("key", ["rmat_s10_ef16_seed0", "rmat_s12_ef8_seed7",
                                 "rmat_s16_ef16_seed0"])
def test_bfs_cc_tc_match_networkx(key):
    import hashlib
    gg, g = _golden_graph(key)
    off, cols = g.csr()
    assert len(cols) == gg["m_directed"]
    lev, par = O.bfs(off, cols, 0)
    lab = O.cc(off, cols)
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    assert sha(lev) == gg["sha_levels"]
    assert sha(lab) == gg["sha_cc_labels"]
    assert O.tc(off, cols) == gg["triangles"]
    if "levels" in gg:
        assert lev.tolist() == gg["levels"]
        assert lab.tolist() == gg["cc_labels"]
    # parent array: parent[src]==src, parents one level up, -1 iff unreached
    reached = lev != 0xFFFFFFFF
    assert par[0] == 0 and np.all((par >= 0) == reached)
    v = np.nonzero(reached & (np.arange(len(lev)) != 0))[0]
    assert np.all(lev[par[v]] + 1 == lev[v])


def test_examples_closed_forms():
    # johnson8-2-4: Kneser K(8,2), 15-regular, 28 vertices.
    n, s, d = read_mtx(os.path.join(GOLDEN, "johnson8-2-4.mtx"))
    off, cols = csr_from_edges(n, s, d)
    assert len(cols) == 420
    lev, _ = O.bfs(off, cols, 0)
    assert np.bincount(lev).tolist() == [1, 15, 12]
    assert O.tc(off, cols) == 420
    assert len(set(O.cc(off, cols).tolist())) == 1
    np.testing.assert_allclose(O.pagerank(off, cols, 10), 1 / 28, rtol=1e-14)
    np.testing.assert_allclose(O.pagerank(off, cols, 37), 1 / 28, rtol=1e-14)
    b = O.bc(off, cols, 0)
    assert b[0] == pytest.approx(27.0)
    np.testing.assert_allclose(b[lev == 1], 0.8, rtol=1e-12)
    assert np.all(b[lev == 2] == 0)
    # star.mtx: 20 vertices, one triangle 1-7-14 (1-indexed)
    n, s, d = read_mtx(os.path.join(GOLDEN, "star.mtx"))
    off, cols = csr_from_edges(n, s, d)
    assert len(cols) == 40
    lev, _ = O.bfs(off, cols, 0)
    assert np.bincount(lev).tolist() == [1, 7, 12]
    assert O.tc(off, cols) == 1
    b = O.bc(off, cols, 0)
    assert b[0] == pytest.approx(19.0)
    expect = np.zeros(n)
    expect[[6, 13]] = 6.0
    expect[0] = 19.0
    np.testing.assert_allclose(b, expect, atol=1e-12)


def _dense_pagerank(off, cols, iters, damping):
    n = len(off) - 1
    A = np.zeros((n, n))
    for v in range(n):
        A[v, cols[off[v]:off[v + 1]]] = 1.0   # row v lists N(v)
    deg = A.sum(axis=1)
    inv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    x = np.full(n, 1.0 / n)
    for _ in range(iters):
        x = damping * (A @ (x * inv)) + (1 - damping) / n
    return x


@pytest.mark.parametrize("symmetric", [True, False])
def test_pagerank_matches_dense_power_iteration(symmetric):
    off, cols = random_graph(300, 1500, 11, symmetric=symmetric)
    for iters, damp in [(1, 0.85), (10, 0.85), (20, 0.5)]:
        got = O.pagerank(off, cols, iters, damp)
        np.testing.assert_allclose(got, _dense_pagerank(off, cols, iters, damp),
                                   rtol=1e-12, atol=0)


def _brute_dependency(n, adj, s):
    # delta_s(v) = sum_{t != s} sigma_st(v) / sigma_st, from all-pairs BFS.
    def bfs_counts(x):
        dist = [-1] * n
        sig = [0] * n
        dist[x], sig[x] = 0, 1
        q = deque([x])
        while q:
            u = q.popleft()
            for w in adj[u]:
                if dist[w] < 0:
                    dist[w] = dist[u] + 1
                    q.append(w)
                if dist[w] == dist[u] + 1:
                    sig[w] += sig[u]
        return dist, sig
    ds, ss = bfs_counts(s)
    allp = [bfs_counts(v) for v in range(n)]
    out = np.zeros(n)
    for v in range(n):
        if ds[v] < 0:
            continue
        dv, sv = allp[v]
        tot = 0.0
        for t in range(n):
            if t == s or t == v or ds[t] < 0 or dv[t] < 0:
                continue
            if ds[v] + dv[t] == ds[t]:
                tot += ss[v] * sv[t] / ss[t]
        out[v] = tot
    return out


@pytest.mark.parametrize("seed,symmetric", [(21, True), (22, True), (23, True),
                                            (24, False), (25, False)])
def test_bc_matches_brute_force_path_counting(seed, symmetric):
    n = 40
    off, cols = random_graph(n, 70 if symmetric else 160, seed, symmetric=symmetric)
    adj = [cols[off[v]:off[v + 1]].tolist() for v in range(n)]
    for s in [0, 7]:
        got = O.bc(off, cols, s)
        np.testing.assert_allclose(got, _brute_dependency(n, adj, s),
                                   rtol=1e-12, atol=1e-12)
        lev, _ = O.bfs(off, cols, s)
        reached = lev != 0xFFFFFFFF
        assert got[s] == pytest.approx(reached.sum() - 1)
        # sum_{v != s} delta = sum_{t reached, t != s} (d(s,t) - 1)
        lt = lev[reached & (np.arange(n) != s)].astype(np.int64)
        assert got[reached].sum() - got[s] == pytest.approx((lt - 1).sum())


def test_common_neighbors_early_end_brute_force():
    off, cols = random_graph(60, 400, 31)
    for u in range(0, 60, 7):
        for v in range(0, 60, 5):
            nu = set(cols[off[u]:off[u + 1]].tolist())
            nv = set(cols[off[v]:off[v + 1]].tolist())
            assert O.common_neighbors(off, cols, u, v, False) == len(nu & nv)
            assert O.common_neighbors(off, cols, u, v, True) == \
                len({w for w in nu & nv if w < u and w < v})


def test_tc_brute_force_small():
    import itertools
    off, cols = random_graph(45, 250, 41)
    adj = [set(cols[off[v]:off[v + 1]].tolist()) for v in range(45)]
    brute = sum(1 for a, b, c in itertools.combinations(range(45), 3)
                if b in adj[a] and c in adj[a] and c in adj[b])
    assert O.tc(off, cols) == brute


def test_bf_unit_weights_equals_bfs_levels():
    off, cols = random_graph(500, 900, 51)
    lev, _ = O.bfs(off, cols, 3)
    dist = O.bf(off, cols, 3)
    unreached = lev == 0xFFFFFFFF
    assert np.all(dist[unreached] == (2**31 - 1) // 2)
    assert np.all(dist[~unreached] == lev[~unreached])


def test_bf_weighted_matches_dijkstra():
    """Positive weights: Bellman-Ford distances are Dijkstra's
    (scipy.sparse.csgraph), a routine independent of the oracle."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import dijkstra
    n = 400
    off, cols = random_graph(n, 1500, 61, symmetric=False)
    rng = np.random.default_rng(5)
    w = rng.integers(1, 50, cols.size).astype(np.int32)
    a = csr_matrix((w.astype(np.float64), cols.astype(np.int64),
                    off.astype(np.int64)), shape=(n, n))
    for src in (0, 7, 123):
        got = O.bf(off, cols, src, w)
        ref = dijkstra(a, indices=src)
        inf = ~np.isfinite(ref)
        assert inf.any() and (~inf).sum() > 100
        assert np.all(got[inf] == (2**31 - 1) // 2)
        assert np.array_equal(got[~inf], ref[~inf].astype(np.int64))


def test_bf_negative_weights_on_dag():
    """Negative weights without a negative cycle (edges only from lower to
    higher ids): the distances are the DAG's exact shortest paths, computed
    by relaxing vertices in id order."""
    n = 300
    rng = np.random.default_rng(9)
    s = rng.integers(0, n, 2000)
    d = rng.integers(0, n, 2000)
    lo, hi = np.minimum(s, d), np.maximum(s, d)
    off, cols = csr_from_edges(n, lo, hi)
    w = rng.integers(-20, 30, cols.size).astype(np.int32)
    got = O.bf(off, cols, 0, w)
    inf = (2**31 - 1) // 2
    dist = [inf] * n
    dist[0] = 0
    for u in range(n):
        if dist[u] == inf:
            continue
        for e in range(int(off[u]), int(off[u + 1])):
            v = int(cols[e])
            dist[v] = min(dist[v], dist[u] + int(w[e]))
    assert got.tolist() == dist


def test_er_generator_is_uniform_and_seeded():
    n = 1000
    s, d = O.er(n, 11, 200000)
    assert s.max() < n and d.max() < n
    counts = np.bincount(np.concatenate([s, d]), minlength=n)
    # 400 draws per vertex on average; a uniform stream stays within 6 sigma
    assert abs(counts.mean() - 400) < 1e-9 and counts.std() < 6 * 20
    s2, _ = O.er(n, 11, 10, first=5)
    assert np.array_equal(s2, s[5:15])
    s3, _ = O.er(n, 12, 10)
    assert not np.array_equal(s3, s[:10])


@pytest.mark.parametrize("symmetric", [True, False])
def test_all_core_bfs_and_bc_equal_the_serial_oracle(symmetric):
    """The all-core baselines compute the same levels (bit-exact, both step
    directions taken on the symmetric rMat graph) and the same dependencies
    (within FP64 reordering) as the serial checker."""
    n, s, d = O.rmat_graph_edges(14, 16, 5)
    if not symmetric:
        s, d = s[: len(s) // 2], d[: len(d) // 2]
    off, cols = csr_from_edges(n, s, d)
    deg = np.diff(off)
    for src in (0, int(np.argmax(deg)), int(np.nonzero(deg)[0][-1])):
        lev, _ = O.bfs(off, cols, src)
        assert np.array_equal(O.bfs_par(off, cols, src), lev)
        want = O.bc(off, cols, src)
        got = O.bc_par(off, cols, src)
        tol = 1e-9 * np.abs(want) + 1e-12
        assert (np.abs(got - want) <= tol).all()
    assert O.bfs_par(off, cols, n - 1).dtype == np.uint32
