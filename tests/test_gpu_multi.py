"""Many graphs in one device store (sstg_multi_*): BFS, CC and PageRank for
all graphs from one captured hipGraph replay must equal each graph's own
oracle results (levels and labels bit-exact, FP64 PR within 1e-9, FP32+Kahan
PR within 1e-5)."""
import numpy as np
import pytest

import sstgraph_amd as S
from oracle import oracle as O
from tests.gpu_common import need_gpu, rel_close, say

pytestmark = pytest.mark.gpu


def _make(k, seed, symmetric=True):
    rng = np.random.default_rng(seed)
    graphs = []
    for g in range(k):
        scale = int(rng.integers(4, 12))
        n = 1 << scale
        if g % 2 == 0:
            s, d = O.rmat(n, 1000 + g, int(rng.integers(8, 17)) * n)
        else:
            s, d = O.er(n, 2000 + g, n * int(rng.integers(2, 9)))
        if symmetric:
            s, d = np.concatenate([s, d]), np.concatenate([d, s])
        og = O.Graph(n)
        og.insert_batch(s, d)
        off, cols = og.csr()
        src = int(np.argmax(np.diff(off)))
        graphs.append((n, s, d, off, cols, src))
    graphs.append((1, np.zeros(0, np.uint32), np.zeros(0, np.uint32),
                   np.zeros(2, np.uint64), np.zeros(0, np.uint32), 0))
    return graphs


@pytest.mark.parametrize("max_levels", [64, 2])
def test_multi_matches_per_graph_oracle(max_levels):
    need_gpu()
    graphs = _make(23, 5)
    M = S.MultiGraph([g[0] for g in graphs])
    for i, (n, s, d, off, cols, src) in enumerate(graphs):
        if len(s):
            M.insert_batch(i, s, d)
    sources = [g[5] for g in graphs]
    for f32 in (False, True):
        say("multi run", max_levels, "f32" if f32 else "f64")
        M.run(sources, max_levels=max_levels, pr_iters=10, f32=f32)
        r = M.results(f32=f32)
        for i, (n, s, d, off, cols, src) in enumerate(graphs):
            lev, _ = O.bfs(off, cols, src)
            assert np.array_equal(M.part(r["level"], i), lev), i
            assert np.array_equal(M.part(r["label"], i), O.cc(off, cols)), i
            want = O.pagerank(off, cols, 10, 0.85)
            ok, w = rel_close(M.part(r["rank"], i), want, 1e-5 if f32 else 1e-9)
            assert ok, (i, w)


def test_multi_directed_graphs():
    """Rows without their reverses: BFS stays top-down and CC links every
    stored edge (no large-component skip)."""
    need_gpu()
    graphs = _make(9, 8, symmetric=False)
    M = S.MultiGraph([g[0] for g in graphs])
    for i, (n, s, d, off, cols, src) in enumerate(graphs):
        if len(s):
            M.insert_batch(i, s, d)
    M.run([g[5] for g in graphs], max_levels=64, pr_iters=10)
    r = M.results()
    for i, (n, s, d, off, cols, src) in enumerate(graphs):
        assert np.array_equal(M.part(r["level"], i), O.bfs(off, cols, src)[0]), i
        assert np.array_equal(M.part(r["label"], i), O.cc(off, cols)), i
        ok, w = rel_close(M.part(r["rank"], i), O.pagerank(off, cols, 10, 0.85), 1e-9)
        assert ok, (i, w)


def test_multi_equals_single_graph_runs():
    """Each graph's result is the same alone (its own SparseMatrixV) and in
    the batch."""
    need_gpu()
    graphs = _make(6, 9)
    M = S.MultiGraph([g[0] for g in graphs])
    singles = []
    for i, (n, s, d, off, cols, src) in enumerate(graphs):
        if len(s):
            M.insert_batch(i, s, d)
        g = S.SparseMatrixV(n)
        if len(s):
            g.insert_batch(s, d)
        singles.append(g)
    M.run([g[5] for g in graphs], pr_iters=7)
    r = M.results()
    for i, g in enumerate(singles):
        _, lev = S.EdgeMapVertexMap.BFS(g, graphs[i][5], levels=True)
        assert np.array_equal(M.part(r["level"], i), lev)
        assert np.array_equal(M.part(r["label"], i), S.EdgeMapVertexMap.CC(g))
        ok, w = rel_close(M.part(r["rank"], i), S.EdgeMapVertexMap.PR_S(g, 7), 1e-12)
        assert ok, w


def test_batched_updates_match_per_graph_oracles():
    """insert_batches / remove_batches: uneven batches of many graphs (some
    empty, one graph twice in a call) in one update pass leave every graph
    as its own oracle's after the same per-graph updates, row for row; the
    captured run then matches the oracles too."""
    need_gpu()
    graphs = _make(17, 12)
    sizes = [g[0] for g in graphs]
    M = S.MultiGraph(sizes)
    oracles = [O.Graph(n) for n in sizes]
    rng = np.random.default_rng(13)
    # the base edges of every graph in one call; graph 3's in two batches
    batches = []
    for i, (n, s, d, off, cols, src) in enumerate(graphs):
        if i == 3:
            h = len(s) // 2
            batches += [(i, s[:h], d[:h]), (i, s[h:], d[h:])]
        else:
            batches.append((i, s, d))
        oracles[i].insert_batch(s, d)
    M.insert_batches(batches)

    def check(what):
        for i, og in enumerate(oracles):
            off, cols = M.export_csr(i)
            ooff, ocols = og.csr()
            assert np.array_equal(off, ooff.astype(np.uint64)), (what, i)
            assert np.array_equal(cols, ocols), (what, i)

    check("base")
    for step in range(3):
        ins, rem = [], []
        for i, n in enumerate(sizes):
            k = int(rng.integers(0, 3 * n)) if i % 4 else 0
            s = rng.integers(0, n, k).astype(np.uint32)
            d = rng.integers(0, n, k).astype(np.uint32)
            ins.append((i, np.concatenate([s, d]), np.concatenate([d, s])))
            oracles[i].insert_batch(ins[-1][1], ins[-1][2])
        say("batched insert step", step)
        M.insert_batches(ins[::-1])  # any batch order
        check(f"insert {step}")
        for i, n in enumerate(sizes):
            off, cols = oracles[i].csr()
            m = len(cols)
            pick = rng.integers(0, max(m, 1), m // 3) if m else np.zeros(0, int)
            rows = (np.searchsorted(off, pick, side="right") - 1).astype(np.uint32)
            s, d = rows, cols[pick].astype(np.uint32)
            rem.append((i, np.concatenate([s, d]), np.concatenate([d, s])))
        for i, s, d in rem:
            oracles[i].remove_batch(s, d)
        M.remove_batches(rem)
        check(f"remove {step}")
    M.run([g[5] for g in graphs], pr_iters=10)
    r = M.results()
    for i, og in enumerate(oracles):
        off, cols = og.csr()
        src = graphs[i][5]
        assert np.array_equal(M.part(r["level"], i), O.bfs(off, cols, src)[0]), i
        assert np.array_equal(M.part(r["label"], i), O.cc(off, cols)), i
        ok, w = rel_close(M.part(r["rank"], i), O.pagerank(off, cols, 10, 0.85),
                          1e-9)
        assert ok, (i, w)
    # device batches, back to back
    s = np.array([1, 2, 0, 5], np.uint32)
    d = np.array([2, 1, 5, 0], np.uint32)
    M.insert_batches_device([0, 2], [2, 2], S.DeviceArray.from_host(s),
                            S.DeviceArray.from_host(d))
    oracles[0].insert_batch(s[:2], d[:2])
    oracles[2].insert_batch(s[2:], d[2:])
    check("device batches")


def test_batched_updates_reject_bad_ids_atomically():
    need_gpu()
    M = S.MultiGraph([10, 20, 5])
    M.insert_batches([(0, [1, 2], [2, 1]), (1, [3], [4])])
    before = [M.export_csr(i) for i in range(3)]
    # vertex 5 is outside graph 2: nothing of the call is applied
    with pytest.raises(S.IdOutOfRange):
        M.insert_batches([(1, [7], [8]), (2, [1, 5], [2, 0])])
    with pytest.raises(S.SSTGError):
        M.insert_batches([(3, [0], [1])])  # no graph 3
    after = [M.export_csr(i) for i in range(3)]
    for (a, b), (c, e) in zip(before, after):
        assert np.array_equal(a, c) and np.array_equal(b, e)
    M.insert_batches([])  # nothing to do
    M.remove_batches([(0, [1], [2]), (1, [], [])])
    off, cols = M.export_csr(0)
    assert cols.tolist() == [1] and off.tolist()[:4] == [0, 0, 0, 1]


def test_multi_rejects_out_of_range_local_ids():
    need_gpu()
    M = S.MultiGraph([10, 20])
    with pytest.raises(S.IdOutOfRange):
        M.insert_batch(0, np.array([1, 10], np.uint32), np.array([2, 3], np.uint32))
    M.insert_batch(1, np.array([1, 19], np.uint32), np.array([19, 1], np.uint32))
    M.run([0, 1], pr_iters=1)
    r = M.results()
    assert M.part(r["level"], 1)[19] == 1


def test_er_generator_device_matches_oracle():
    need_gpu()
    for n, seed, cnt, first in [(1000, 3, 5000, 0), (1 << 20, 7, 1 << 16, 12345)]:
        s, d = S.er(n, seed, cnt, first=first)
        os_, od = O.er(n, seed, cnt, first=first)
        assert np.array_equal(s, os_) and np.array_equal(d, od)
