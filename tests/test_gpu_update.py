"""Batch insert / delete on the device store vs the CPU oracle: contents must
be bit-exact after every batch.  Calls go through the C ABI (libsstg.so)."""
import json
import os

import numpy as np
import pytest

import sstgraph_amd as S
from oracle import oracle as O
from tests.gpu_common import assert_same_csr, need_gpu, say
from tests.graphs import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu():
    need_gpu()


def test_device_rmat_matches_reference_vectors():
    cases = json.load(open(os.path.join(GOLDEN, "rmat_vectors.json")))
    for c in cases:
        say("rmat", c["n"], c["seed"], c["first"])
        s, d = S.rmat(c["n"], c["seed"], c["count"], first=c["first"],
                      a=c["a"], b=c["b"], c=c["c"])
        assert s.tolist() == c["src"] and d.tolist() == c["dst"]
    # a long run against the oracle restatement (pinned above)
    s, d = S.rmat(1 << 20, 12, 1 << 20)
    os_, od = O.rmat(1 << 20, 12, 1 << 20)
    assert np.array_equal(s, os_) and np.array_equal(d, od)
    # the normalised draw (no reference counterpart) against the oracle's
    s, d = S.rmat(1 << 20, 12, 1 << 20, first=77, normalized=True)
    os_, od = O.rmat(1 << 20, 12, 1 << 20, first=77, normalized=True)
    assert np.array_equal(s, os_) and np.array_equal(d, od)
    for a, b, c in ((0.57, 0.19, 0.19), (0.25, 0.25, 0.25), (1.0, 0.0, 0.0)):
        s, d = S.rmat(1 << 12, 3, 50000, a=a, b=b, c=c, normalized=True)
        os_, od = O.rmat(1 << 12, 3, 50000, a=a, b=b, c=c, normalized=True)
        assert np.array_equal(s, os_) and np.array_equal(d, od), (a, b, c)


@pytest.mark.parametrize("n,sizes", [(1000, [1, 7, 100, 101, 1000, 20000]),
                                     (70000, [5000, 300000, 2])])
def test_insert_batches_match_oracle(n, sizes):
    rng = np.random.default_rng(n)
    g, og = S.SparseMatrixV(n), O.Graph(n)
    for k in sizes:
        say("insert", n, k)
        s = rng.integers(0, n, k).astype(np.uint32)
        d = rng.integers(0, n, k).astype(np.uint32)
        g.insert_batch(s, d)
        og.insert_batch(s, d)
        assert_same_csr(g, og, f"insert n={n} k={k}")


def test_batch_sizes_around_the_sort_and_scan_paths():
    """Batches on both sides of the one-block LDS sort (4096 keys for an
    update, 16384 for the radix sort's own one-block path), of the one-block
    scans (4096 values) and of the merge sort (4K-key block sorts and merge
    passes up to 2^18 keys, the radix sort beyond), inserted and then half
    removed; ids below 2^17 so every batch sorts on three column and three
    row digits."""
    n = 1 << 17
    rng = np.random.default_rng(11)
    g, og = S.SparseMatrixV(n), O.Graph(n)
    for k in (2, 63, 64, 65, 4095, 4096, 4097, 8192, 12289, 16383, 16384,
              16385, 32768, 40000, 49153, 1 << 18, (1 << 18) + 1, 1 << 20,
              (1 << 20) + 1, 1 << 21):
        say("batch", k)
        s = rng.integers(0, n, k).astype(np.uint32)
        d = rng.integers(0, n, k).astype(np.uint32)
        g.insert_batch(s, d)
        og.insert_batch(s, d)
        assert_same_csr(g, og, f"insert k={k}")
        g.remove_batch(s[::2], d[::2])
        og.remove_batch(s[::2], d[::2])
        assert_same_csr(g, og, f"remove k={k}")


@pytest.mark.parametrize("rows", [4095, 4096, 4159])
def test_many_big_rows_bit_exact(rows):
    """Batches whose rows are all merged by the big-row path (each row of
    2100 ids gets one id below all of them, so the whole row moves): the
    flat merge kernels find each block's rows with a 64-way wave search over
    rows + 1 offsets, which must be exact at every size, including 4097 ..
    4160 offsets, where the search's last round covers exactly 64 entries."""
    n, deg = 1 << 16, 2100
    r = np.repeat(np.arange(rows, dtype=np.int64), deg)
    k = np.tile(np.arange(deg, dtype=np.int64), rows)
    c = 1000 + (r * 7919 + k * 31) % 64536  # distinct per row, all >= 1000
    g, og = S.SparseMatrixV(n), O.Graph(n)
    g.insert_batch(r.astype(np.uint32), c.astype(np.uint32))
    og.insert_batch(r.astype(np.uint32), c.astype(np.uint32))
    bs = np.arange(rows, dtype=np.uint32)
    bd = ((bs + 1) % 1000).astype(np.uint32)  # below every stored id, never bs
    # each pair twice: the batch exceeds the small-batch path's 4096 pairs,
    # so the general pipeline's big-row merge runs for every row count
    bs, bd = np.concatenate([bs, bs]), np.concatenate([bd, bd])
    say("big rows", rows)
    g.insert_batch(bs, bd)
    og.insert_batch(bs, bd)
    assert g.last_update_stats()["big_rows"] == rows
    assert_same_csr(g, og, f"insert into {rows} big rows")
    g.remove_batch(bs, bd)
    og.remove_batch(bs, bd)
    assert_same_csr(g, og, f"remove from {rows} big rows")


def test_remove_batches_match_oracle():
    rng = np.random.default_rng(7)
    n = 5000
    g, og = S.SparseMatrixV(n), O.Graph(n)
    s = rng.integers(0, n, 200000).astype(np.uint32)
    d = rng.integers(0, n, 200000).astype(np.uint32)
    g.insert_batch(s, d)
    og.insert_batch(s, d)
    for step, k in enumerate([1, 50, 3000, 60000, 150000]):
        say("remove", k)
        # half existing edges (some repeated), half random (mostly absent)
        pick = rng.integers(0, len(s), k // 2 + 1)
        rs = np.concatenate([s[pick], rng.integers(0, n, k // 2)]).astype(np.uint32)
        rd = np.concatenate([d[pick], rng.integers(0, n, k // 2)]).astype(np.uint32)
        g.remove_batch(rs, rd)
        og.remove_batch(rs, rd)
        assert_same_csr(g, og, f"remove step {step}")
        # re-insert part so rows regrow into their slots
        g.insert_batch(rs[: k // 3], rd[: k // 3])
        og.insert_batch(rs[: k // 3], rd[: k // 3])
        assert_same_csr(g, og, f"reinsert step {step}")


def test_self_loop_and_duplicate_rules():
    g, og = S.SparseMatrixV(50), O.Graph(50)
    s = np.array([1, 1, 2, 2, 2, 3, 3], np.uint32)
    d = np.array([1, 4, 2, 5, 5, 4, 4], np.uint32)
    g.insert_batch(s, d)
    og.insert_batch(s, d)
    assert_same_csr(g, og, "batch")
    assert not g.has(1, 1) and not g.has(2, 2) and g.has(2, 5)
    g.insert(7, 7)  # single insert keeps the self-loop (SparseMatrix.hpp:536-544)
    og.insert(7, 7)
    assert g.has(7, 7)
    assert_same_csr(g, og, "single insert")
    g.remove_batch(np.array([7, 9], np.uint32), np.array([7, 9], np.uint32))
    og.remove_batch(np.array([7, 9], np.uint32), np.array([7, 9], np.uint32))
    assert not g.has(7, 7)
    assert_same_csr(g, og, "remove self-loop")


def test_out_of_range_ids_leave_graph_unchanged():
    g = S.SparseMatrixV(100)
    g.insert_batch(np.array([1, 2], np.uint32), np.array([3, 4], np.uint32))
    with pytest.raises(S.IdOutOfRange):
        g.insert_batch(np.array([1, 100], np.uint32), np.array([5, 6], np.uint32))
    with pytest.raises(S.IdOutOfRange):
        g.remove_batch(np.array([1, 2], np.uint32), np.array([3, 100], np.uint32))
    assert g.M() == 2 and g.has(1, 3) and g.has(2, 4)
    g.insert_batch(np.array([], np.uint32), np.array([], np.uint32))
    assert g.M() == 2
    # long batches on both sides of the merge-sort limit (2^18 keys): above
    # it the pairs are packed by the kernel that counts the first radix digit
    rng = np.random.default_rng(5)
    for n in (1 << 18, (1 << 18) + 3):
        for at in (0, n - 1):
            say("out of range", n, at)
            s = rng.integers(0, 100, n, dtype=np.uint32)
            d = rng.integers(0, 100, n, dtype=np.uint32)
            d[at] = 100
            with pytest.raises(S.IdOutOfRange):
                g.insert_batch(s, d)
            s[at], d[at] = 100, 0
            with pytest.raises(S.IdOutOfRange):
                g.remove_batch(s, d)
            assert g.M() == 2 and g.has(1, 3) and g.has(2, 4)


def test_rectangular_and_zero_key():
    # keys with (i & 0xFFFF) == 0, including 0, stay present
    # (Tinyset_test.cpp:198-240)
    maxv = 1 << 20
    g = S.SparseMatrixV(3, maxv)
    keys = np.arange(0, maxv, 0x10000, dtype=np.uint32)
    other = np.random.default_rng(4).integers(0, maxv, 50000).astype(np.uint32)
    allk = np.concatenate([keys, other])
    g.insert_batch(np.ones(len(allk), np.uint32), allk)
    assert g.has_batch(np.ones(len(keys), np.uint32), keys).all()
    off, cols = g.export_csr()
    assert cols[off[1]:off[2]].tolist() == sorted(set(allk.tolist()))


def test_hub_row_grows_through_relocations():
    rng = np.random.default_rng(9)
    n = 1 << 18
    g, og = S.SparseMatrixV(n), O.Graph(n)
    for k in [10, 1000, 100000, 400000, 3]:
        say("hub", k)
        s = np.full(k, 5, np.uint32)
        s[: k // 10] = rng.integers(0, n, k // 10)
        d = rng.integers(0, n, k).astype(np.uint32)
        g.insert_batch(s, d)
        og.insert_batch(s, d)
        assert_same_csr(g, og, f"hub k={k}")
    assert g.touch_all_sum() == og.touch_all_sum()
    assert np.array_equal(g.degrees(), og.degrees())


def _c1():
    n, s, d = O.rmat_graph_edges(16, 16, 0)
    return n, s, d


def test_c1_build_then_update_batches_bit_exact():
    """Scale-16 RMAT (seed 0) built by batch insert, then insert+remove
    batches of 1e3, 1e4, 1e5 RMAT edges (seeds 100..) with 20% existing
    edges; contents compared after every batch."""
    n, s, d = _c1()
    gold = json.load(open(os.path.join(GOLDEN, "graph_goldens.json")))
    gold = gold["rmat_s16_ef16_seed0"]
    g, og = S.SparseMatrixV(n), O.Graph(n)
    g.insert_batch(s, d)
    og.insert_batch(s, d)
    assert g.M() == gold["m_directed"]
    assert_same_csr(g, og, "C1 build")
    off, cols = og.csr()
    for k, bsz in enumerate([1000, 10000, 100000]):
        say("C1 batch", bsz)
        fs, fd = O.rmat(n, 100 + k, int(bsz * 0.8))
        pick = (np.array([O.hash64(100 + k + 7919 * j) for j in range(bsz - len(fs))],
                         dtype=np.uint64) % np.uint64(len(cols))).astype(np.int64)
        rows = np.searchsorted(off, pick, side="right") - 1
        bs = np.concatenate([fs, rows.astype(np.uint32)])
        bd = np.concatenate([fd, cols[pick]])
        g.insert_batch(bs, bd)
        og.insert_batch(bs, bd)
        assert_same_csr(g, og, f"C1 insert {bsz}")
        g.remove_batch(bs, bd)
        og.remove_batch(bs, bd)
        assert_same_csr(g, og, f"C1 remove {bsz}")


def test_scale22_update_batches_bit_exact():
    """The update workload's shape at scale 22: rMat(2^22, seed 2, 16n)
    symmetrised, then insert+remove batches of 1e5, 1e6 and 1e7 edges (80%
    fresh rMat edges, 20% present ones), contents compared with the oracle
    after every batch.  The large batches take the relocation and big-row
    paths, which the statistics confirm."""
    n, s, d = O.rmat_graph_edges(22, 16, 2)
    g, og = S.SparseMatrixV(n), O.Graph(n)
    say("scale 22 build", len(s))
    g.insert_batch(s, d)
    og.insert_batch(s, d)
    assert_same_csr(g, og, "scale 22 build")
    off, cols = og.csr()
    seen = {"big_rows": 0, "allocated": 0, "tile_rows": 0}
    for k, bsz in enumerate([100_000, 1_000_000, 10_000_000]):
        say("scale 22 batch", bsz)
        fs, fd = O.rmat(n, 100 + k, int(bsz * 0.8))
        pick = np.random.default_rng(100 + k).integers(0, len(cols), bsz - len(fs))
        rows = np.searchsorted(off, pick, side="right") - 1
        bs = np.concatenate([fs, rows.astype(np.uint32)])
        bd = np.concatenate([fd, cols[pick]])
        g.insert_batch(bs, bd)
        og.insert_batch(bs, bd)
        st = g.last_update_stats()
        for key in seen:
            seen[key] += st[key]
        assert_same_csr(g, og, f"scale 22 insert {bsz}")
        g.remove_batch(bs, bd)
        og.remove_batch(bs, bd)
        assert_same_csr(g, og, f"scale 22 remove {bsz}")
    say("paths taken", seen)
    assert seen["big_rows"] > 0 and seen["allocated"] > 0 and seen["tile_rows"] > 0


def test_heavy_removal_compacts_the_pool():
    """Removing most of a graph's edges (live ids below a quarter of the
    live slots) compacts the pool and shrinks every slot to its row, as the
    packed-memory array's lower density bound does; contents stay exact and
    later inserts grow rows again."""
    n = 1 << 16
    rng = np.random.default_rng(21)
    s = rng.integers(0, n, 3_000_000).astype(np.uint32)
    d = rng.integers(0, n, 3_000_000).astype(np.uint32)
    g, og = S.SparseMatrixV(n), O.Graph(n)
    g.insert_batch(s, d)
    og.insert_batch(s, d)
    before = g.statistics()
    keep = rng.random(len(s)) < 0.1
    g.remove_batch(s[~keep], d[~keep])
    og.remove_batch(s[~keep], d[~keep])
    after = g.statistics()
    say("pool capacity", before["pool_capacity"], "->", after["pool_capacity"],
        "rebuilds", before["rebuilds"], "->", after["rebuilds"])
    assert after["rebuilds"] > before["rebuilds"]
    assert after["pool_capacity"] < before["pool_capacity"] // 2
    assert after["slot_capacity"] <= 4 * max(after["edges"], 1) + 4 * n
    assert_same_csr(g, og, "after compaction")
    g.insert_batch(s[:200000], d[:200000])
    og.insert_batch(s[:200000], d[:200000])
    assert_same_csr(g, og, "regrown")
