"""Edge weights (SparseMatrixV<true, uint32_t>) and weighted Bellman-Ford on
the device.

The model is a Python dict: a weighted insert sets each edge's weight, the
last occurrence in a batch winning (the valued PMA overwrites a present key,
PMA.hpp:1095-1146); an edge inserted without a weight weighs 1; removing an
edge drops its weight; value() of an absent edge is 0 (PMA.hpp:1397-1420).
Weighted BF must equal the oracle's Bellman-Ford (pinned against scipy's
Dijkstra and an exact DAG relaxation) run with the model's weights aligned
to the exported CSR, bit for bit."""
import numpy as np
import pytest

import sstgraph_amd as S
from oracle import oracle as O
from tests.gpu_common import need_gpu, say

pytestmark = pytest.mark.gpu
EMV = S.EdgeMapVertexMap
INF = (2**31 - 1) // 2


@pytest.fixture(autouse=True)
def _gpu():
    need_gpu()


class Model:
    def __init__(self, n):
        self.n = n
        self.w = {}  # edge -> weight

    def insert_weighted(self, s, d, w, keep_self=False):
        for a, b, x in zip(s.tolist(), d.tolist(), w.tolist()):
            if a != b or keep_self:
                self.w[(a, b)] = x

    def insert(self, s, d):
        for a, b in zip(s.tolist(), d.tolist()):
            if a != b:
                self.w.setdefault((a, b), 1)

    def remove(self, s, d):
        for a, b in zip(s.tolist(), d.tolist()):
            self.w.pop((a, b), None)

    def aligned(self, off, cols):
        rows = np.repeat(np.arange(self.n), np.diff(off).astype(np.int64))
        return np.array([self.w[(int(r), int(c))] for r, c in zip(rows, cols)],
                        dtype=np.uint32)


def _check(g, model, sources, what):
    off, cols = g.export_csr()
    rows = np.repeat(np.arange(model.n), np.diff(off).astype(np.int64))
    assert sorted(zip(rows.tolist(), cols.tolist())) == sorted(model.w), \
        f"{what}: edge set"
    w = model.aligned(off, cols)
    got = g.values(rows, cols)
    assert np.array_equal(got, w), f"{what}: weights"
    for src in sources:
        say(what, "BF from", src)
        dist = EMV.BF(g, src)
        want = O.bf(off, cols, src, w.view(np.int32))
        assert np.array_equal(dist, want), f"{what}: BF from {src}"


def test_weighted_batches_last_write_wins_and_bf():
    n = 3000
    rng = np.random.default_rng(1)
    g = S.SparseMatrixV(n)
    m = Model(n)
    for b in range(3):
        # duplicates inside a batch (20000 draws over a small id range)
        s = rng.integers(0, n, 20000).astype(np.uint32)
        d = rng.integers(0, n, 20000).astype(np.uint32)
        s[:50] = d[:50]  # self-loops are skipped
        w = rng.integers(1, 1000, 20000).astype(np.uint32)
        g.insert_batch_weighted(s, d, w)
        m.insert_weighted(s, d, w)
        say("batch", b, "edges", len(m.w))
        _check(g, m, [0, int(s[100])], f"batch {b}")


def test_overwrite_remove_and_unweighted_edges():
    n = 1500
    rng = np.random.default_rng(2)
    g = S.SparseMatrixV(n)
    m = Model(n)
    s = rng.integers(0, n, 8000).astype(np.uint32)
    d = rng.integers(0, n, 8000).astype(np.uint32)
    w = rng.integers(1, 100, 8000).astype(np.uint32)
    g.insert_batch_weighted(s, d, w)
    m.insert_weighted(s, d, w)
    # overwrite half of them
    w2 = rng.integers(100, 200, 4000).astype(np.uint32)
    g.insert_batch_weighted(s[:4000], d[:4000], w2)
    m.insert_weighted(s[:4000], d[:4000], w2)
    # unweighted inserts: new edges weigh 1, present ones keep their weight
    s3 = rng.integers(0, n, 3000).astype(np.uint32)
    d3 = rng.integers(0, n, 3000).astype(np.uint32)
    s3[:500], d3[:500] = s[:500], d[:500]
    g.insert_batch(s3, d3)
    m.insert(s3, d3)
    _check(g, m, [1], "after overwrite")
    # removals drop weights; re-inserted unweighted they weigh 1
    g.remove_batch(s[2000:5000], d[2000:5000])
    m.remove(s[2000:5000], d[2000:5000])
    g.remove(int(s[0]), int(d[0]))
    m.remove(s[:1], d[:1])
    _check(g, m, [1], "after remove")
    g.insert_batch(s[2000:2500], d[2000:2500])
    m.insert(s[2000:2500], d[2000:2500])
    _check(g, m, [1, 77], "after reinsert")
    # absent edges and rows out of range read 0
    assert g.value(int(s[3000]), int(d[3000])) == m.w.get(
        (int(s[3000]), int(d[3000])), 0)
    assert g.values([n + 5, 0], [0, 0]).tolist() == [0, 0]


def test_keep_self_loops_and_device_batches():
    n = 600
    rng = np.random.default_rng(3)
    g = S.SparseMatrixV(n)
    m = Model(n)
    s = rng.integers(0, n, 5000).astype(np.uint32)
    d = rng.integers(0, n, 5000).astype(np.uint32)
    s[:20] = d[:20]
    w = rng.integers(1, 9, 5000).astype(np.uint32)
    ds, dd, dw = (S.DeviceArray.from_host(x) for x in (s, d, w))
    g.insert_batch_weighted(ds, dd, dw)
    m.insert_weighted(s, d, w)
    _check(g, m, [5], "device batch")
    # single inserts keep self-loops (SparseMatrix.hpp:536-554)
    lib = S.lib()
    r = np.array([7], np.uint32)
    x = np.array([42], np.uint32)
    rc = lib.sstg_insert_batch_weighted(g.handle, r.ctypes.data, r.ctypes.data,
                                        x.ctypes.data, 1, S.SSTG_KEEP_SELF_LOOPS)
    assert rc == 0
    m.insert_weighted(r, r, x, keep_self=True)
    assert g.value(7, 7) == 42
    _check(g, m, [7], "self-loop")


def test_negative_weights_without_negative_cycles():
    """int32 reading of the weights (BF's dist is int32, test.hpp:611-633):
    negative weights on a DAG (edges from lower to higher ids)."""
    n = 2000
    rng = np.random.default_rng(4)
    s = rng.integers(0, n, 12000)
    d = rng.integers(0, n, 12000)
    lo = np.minimum(s, d).astype(np.uint32)
    hi = np.maximum(s, d).astype(np.uint32)
    w = rng.integers(-40, 60, 12000).astype(np.int32).view(np.uint32)
    g = S.SparseMatrixV(n)
    m = Model(n)
    g.insert_batch_weighted(lo, hi, w)
    m.insert_weighted(lo, hi, w)
    _check(g, m, [0, 3, 100], "negative weights")


def test_weighted_rmat_scale16_growing_table():
    """rMat scale 16, edge factor 8, both directions, inserted in 4 weighted
    batches (the table rehashes as it grows), then BF from 3 sources."""
    n = 1 << 16
    s, d = O.rmat(n, 7, 8 * n)
    s, d = np.concatenate([s, d]), np.concatenate([d, s])
    rng = np.random.default_rng(6)
    w = rng.integers(1, 255, s.size).astype(np.uint32)
    w = np.concatenate([w[: s.size // 2], w[: s.size // 2]])  # symmetric
    g = S.SparseMatrixV(n)
    wt = {}
    for part in np.array_split(np.arange(s.size), 4):
        g.insert_batch_weighted(s[part], d[part], w[part])
    for a, b, x in zip(s.tolist(), d.tolist(), w.tolist()):
        if a != b:
            wt[(a, b)] = x
    off, cols = g.export_csr()
    assert int(off[-1]) == len(wt)
    rows = np.repeat(np.arange(n), np.diff(off).astype(np.int64))
    want_w = np.array([wt[(int(a), int(b))] for a, b in zip(rows, cols)],
                      dtype=np.uint32)
    assert np.array_equal(g.values(rows, cols), want_w)
    deg = np.diff(off)
    for src in (0, int(np.argmax(deg)), int(rows[len(rows) // 3])):
        say("rmat16 BF from", src)
        dist = EMV.BF(g, src)
        assert np.array_equal(dist, O.bf(off, cols, src, want_w.view(np.int32)))
    # unweighted BF on the same edges is the hop distance
    g2 = S.SparseMatrixV(n)
    g2.insert_batch(s, d)
    lev, _ = O.bfs(off, cols, 0)
    hop = EMV.BF(g2, 0)
    assert np.array_equal(hop[lev != S.UNREACHED], lev[lev != S.UNREACHED])
    assert np.all(hop[lev == S.UNREACHED] == INF)


def test_weighted_file_load(tmp_path):
    n = 50
    rng = np.random.default_rng(8)
    s = rng.integers(0, n, 300)
    d = rng.integers(0, n, 300)
    key = np.unique(s * n + d)
    s, d = key // n, key % n
    w = rng.integers(1, 1000, key.size)
    off = np.searchsorted(s, np.arange(n))
    p = tmp_path / "w.adj"
    p.write_text("WeightedAdjacencyGraph\n%d\n%d\n" % (n, key.size) +
                 "".join(f"{x}\n" for x in off) + "".join(f"{x}\n" for x in d) +
                 "".join(f"{x}\n" for x in w))
    g = S.SparseMatrixV.from_file(str(p), weights=True)
    keep = s != d
    assert np.array_equal(g.values(s[keep], d[keep]), w[keep].astype(np.uint32))
    g1 = S.SparseMatrixV.from_file(str(p))
    assert np.all(g1.values(s[keep], d[keep]) == 1)
    # the server's store copies carry no weights: a weighted graph is refused
    with pytest.raises(S.SSTGError) as e:
        S.Server(g, slots=2)
    assert e.value.code == S.SSTG_E_UNSUPPORTED
