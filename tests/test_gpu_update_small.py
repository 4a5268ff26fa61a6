"""Batches of at most 4096 pairs take a two-launch path (update_small.hip):
one workgroup plans the update in LDS, a grid merges the rows.  Its result
must equal the oracle's row for row in every case the general pipeline
handles: in-place and relocated rows, hub rows moved by the whole
workgroup, removes of many ids of one row, duplicates and self-loops, and
the cases it hands back to the general pipeline (no pool room, a row
moving more than 2^15 ids).  Symmetry tracked through small batches keeps
BFS bottom-up steps exact."""
import numpy as np
import pytest

import sstgraph_amd as S
from oracle import oracle as O
from tests.gpu_common import assert_same_csr, need_gpu, say

pytestmark = pytest.mark.gpu
EMV = S.EdgeMapVertexMap


@pytest.fixture(autouse=True)
def _gpu():
    need_gpu()


def _pair(n, seed, m, hubs=()):
    g, og = S.SparseMatrixV(n), O.Graph(n)
    rng = np.random.default_rng(seed)
    s = rng.integers(0, n, m).astype(np.uint32)
    d = rng.integers(0, n, m).astype(np.uint32)
    for h, k in hubs:
        hs = np.full(k, h, np.uint32)
        hd = rng.choice(n, k, replace=False).astype(np.uint32)
        s, d = np.concatenate([s, hs]), np.concatenate([d, hd])
    g.insert_batch(s, d)
    og.insert_batch(s, d)
    return g, og, rng


def test_small_batches_every_size_insert_and_remove():
    n = 1 << 14
    g, og, rng = _pair(n, 1, 200000, hubs=[(7, 9000), (8, 3000)])
    for k in (1, 2, 63, 64, 65, 700, 1023, 1024, 1025, 4000, 4095, 4096):
        say("small batch", k)
        s = rng.integers(0, n, k).astype(np.uint32)
        d = rng.integers(0, n, k).astype(np.uint32)
        s[: k // 3] = 7  # a hub row: its suffix moves through the whole block
        g.insert_batch(s, d)
        og.insert_batch(s, d)
        assert_same_csr(g, og, f"insert {k}")
        r = rng.permutation(k)[: max(1, k // 2)]
        g.remove_batch(s[r], d[r])
        og.remove_batch(s[r], d[r])
        assert_same_csr(g, og, f"remove {k}")


def test_small_batch_relocations_duplicates_and_self_loops():
    n = 5000
    g, og, rng = _pair(n, 2, 20000)
    for step in range(12):
        k = int(rng.integers(100, 4096))
        rows = rng.integers(0, 40, k).astype(np.uint32)  # few rows: they outgrow
        cols = rng.integers(0, n, k).astype(np.uint32)   # their slots
        cols[::7] = rows[::7]                            # self-loops
        rows = np.concatenate([rows, rows[: k // 10]])   # duplicates
        cols = np.concatenate([cols, cols[: k // 10]])
        rows, cols = rows[:4096], cols[:4096]
        g.insert_batch(rows, cols)
        og.insert_batch(rows, cols)
        assert_same_csr(g, og, f"insert step {step}")
        st = g.last_update_stats()
        say("step", step, "allocated", st["allocated"], "rows", st["rows"])
    # removes of nearly whole rows, absent ids and repeats
    off, cols = og.csr()
    for v in range(0, 40, 3):
        row = cols[off[v]:off[v + 1]]
        rs = np.full(len(row) + 20, v, np.uint32)
        rc = np.concatenate([row, rng.integers(0, n, 20)]).astype(np.uint32)
        rs, rc = rs[:4096], rc[:4096]
        g.remove_batch(rs, rc)
        og.remove_batch(rs, rc)
    assert_same_csr(g, og, "removes")
    g.insert(3, 3)  # single insert keeps a self-loop
    og.insert(3, 3)
    g.remove_batch(np.array([3, 3], np.uint32), np.array([3, 3], np.uint32))
    og.remove_batch(np.array([3, 3], np.uint32), np.array([3, 3], np.uint32))
    assert_same_csr(g, og, "self-loop remove")


def test_small_batch_hands_back_to_the_general_pipeline():
    # an empty graph has no pool: the first small insert needs room
    n = 1 << 16
    g, og = S.SparseMatrixV(n), O.Graph(n)
    s = np.arange(1000, dtype=np.uint32)
    d = (s * 7 + 1).astype(np.uint32) % n
    g.insert_batch(s, d)
    og.insert_batch(s, d)
    assert_same_csr(g, og, "first insert")
    # a row of 40000 ids and a key below all of them: a 40000-id move
    hs = np.full(40000, 9, np.uint32)
    hd = np.arange(100, 40100, dtype=np.uint32)
    g.insert_batch(hs, hd)
    og.insert_batch(hs, hd)
    g.insert_batch(np.array([9, 10], np.uint32), np.array([5, 6], np.uint32))
    og.insert_batch(np.array([9, 10], np.uint32), np.array([5, 6], np.uint32))
    assert_same_csr(g, og, "long move")
    g.remove_batch(np.array([9], np.uint32), np.array([5], np.uint32))
    og.remove_batch(np.array([9], np.uint32), np.array([5], np.uint32))
    assert_same_csr(g, og, "long remove")
    with pytest.raises(S.IdOutOfRange):
        g.insert_batch(np.array([1, n], np.uint32), np.array([2, 3], np.uint32))
    assert_same_csr(g, og, "out of range leaves the graph unchanged")


def test_small_symmetric_batches_keep_bfs_exact():
    n, s, d = O.rmat_graph_edges(14, 8, 3)
    g, og = S.SparseMatrixV(n), O.Graph(n)
    g.insert_batch(s, d)
    og.insert_batch(s, d)
    rng = np.random.default_rng(4)
    for step in range(10):
        es, ed = O.rmat(n, 300 + step, int(rng.integers(10, 2000)))
        bs, bd = np.concatenate([es, ed]), np.concatenate([ed, es])
        if step % 4 == 3:
            bs, bd = bs[: len(bs) // 2], bd[: len(bd) // 2]  # one direction
        if step % 2:
            g.remove_batch(bs, bd)
            og.remove_batch(bs, bd)
        else:
            g.insert_batch(bs, bd)
            og.insert_batch(bs, bd)
        off, cols = og.csr()
        src = int(np.argmax(np.diff(off)))
        _, lev = EMV.BFS(g, src, levels=True)
        assert np.array_equal(lev, O.bfs(off, cols, src)[0]), step
    assert_same_csr(g, og, "end")
