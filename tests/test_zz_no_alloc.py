"""The allocation-free insert (sstg_insert_batch_noalloc) through the C ABI:
its result must equal sstg_insert_batch's bit for bit, call after call (the
second call smaller than the first, so it runs over the first call's
leftover scratch), and it must leave the graph's own device memory as it
was.  Oversized batches and host pointers are refused with the graph
unchanged."""
import ctypes

import numpy as np
import pytest

import sstgraph_amd as S
from oracle import oracle as O
from tests.gpu_common import need_gpu, say

pytestmark = pytest.mark.gpu

MAX_PAIRS = 1 << 18  # SSTG_NOALLOC_MAX_PAIRS
SSTG_E_TOO_LARGE = -8


def _noalloc(g, ds, dd, n, flags=S.SSTG_DEVICE_PTRS):
    L = S.lib()
    L.sstg_insert_batch_noalloc.restype = ctypes.c_int
    L.sstg_insert_batch_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_uint64,
                                            ctypes.c_uint32]
    return L.sstg_insert_batch_noalloc(g.handle, ds, dd, n, flags)


def _same(a, b):
    oa, ca = a.export_csr()
    ob, cb = b.export_csr()
    return np.array_equal(oa, ob) and np.array_equal(ca, cb)


def _batch(rng, n, k, hub=None, hub_k=0):
    s = rng.integers(0, n, k).astype(np.uint32)
    d = rng.integers(0, n, k).astype(np.uint32)
    if hub is not None:
        s[:hub_k] = hub  # one long row: the big-row merge and its staging
    return s, d


def test_noalloc_insert_equals_insert_batch_twice():
    need_gpu()
    n, s, d = O.rmat_graph_edges(16, 16, 0)
    a, b = S.SparseMatrixV(n), S.SparseMatrixV(n)
    a.insert_batch(s, d)
    b.insert_batch(s, d)
    og = O.Graph(n)
    og.insert_batch(s, d)
    hub = int(np.argmax(a.degrees()))
    rng = np.random.default_rng(5)
    for call, (k, hub_k) in enumerate([(MAX_PAIRS, 3000), (20000, 0)]):
        bs, bd = _batch(rng, n, k, hub, hub_k)
        ds, dd = S.DeviceArray.from_host(bs), S.DeviceArray.from_host(bd)
        mem = a.get_memory_size()
        free0 = S.mem_info()[0]
        say("noalloc call", call, "pairs", k)
        rc = _noalloc(a, ds.ptr, dd.ptr, k)
        assert rc == 0, S.strerror(rc)
        assert a.get_memory_size() == mem, "the call grew the graph's store"
        if call:
            # the first call on the device may load the library's scratch;
            # after it no call takes device memory
            assert S.mem_info()[0] == free0, "the call allocated device memory"
        st = a.last_update_stats()
        b.insert_batch(ds, dd)
        og.insert_batch(bs, bd)
        assert _same(a, b), f"call {call}: noalloc differs from insert_batch"
        off, cols = a.export_csr()
        ooff, ocols = og.csr()
        assert np.array_equal(off, ooff) and np.array_equal(cols, ocols)
        say("  effective", st["effective"], "tile rows", st["tile_rows"],
            "big rows", st["big_rows"])
        if hub_k:
            assert st["big_rows"] >= 1
        ds.free()
        dd.free()


def test_noalloc_refusals_leave_graph_unchanged():
    need_gpu()
    n = 1 << 12
    g = S.SparseMatrixV(n)
    rng = np.random.default_rng(3)
    s, d = _batch(rng, n, 50000)
    g.insert_batch(s, d)
    off0, cols0 = g.export_csr()
    # larger than the fixed scratch
    k = MAX_PAIRS + 1
    bs, bd = _batch(rng, n, k)
    ds, dd = S.DeviceArray.from_host(bs), S.DeviceArray.from_host(bd)
    assert _noalloc(g, ds.ptr, dd.ptr, k) == SSTG_E_TOO_LARGE
    # host pointers are refused (the call stages nothing)
    assert _noalloc(g, S._ptr(bs), S._ptr(bd), 10, 0) == S.SSTG_E_INVAL
    # an out-of-range id
    bad = np.array([1, n], np.uint32)
    db = S.DeviceArray.from_host(bad)
    assert _noalloc(g, db.ptr, db.ptr, 2) == S.SSTG_E_RANGE
    off, cols = g.export_csr()
    assert np.array_equal(off, off0) and np.array_equal(cols, cols0)
    # a graph with no pool yet has no room: refused, nothing allocated
    e = S.SparseMatrixV(n)
    mem = e.get_memory_size()
    assert _noalloc(e, ds.ptr, dd.ptr, 1000) == S.SSTG_E_NOMEM
    assert e.M() == 0 and e.get_memory_size() == mem
    # and the same small batch is accepted where the pool has room
    assert _noalloc(g, ds.ptr, dd.ptr, 1000) == 0
    ref = S.SparseMatrixV(n)
    ref.insert_batch(_rows(off0), cols0)
    ref.insert_batch(bs[:1000], bd[:1000])
    assert _same(g, ref)


def _rows(off):
    return np.repeat(np.arange(len(off) - 1, dtype=np.uint32),
                     np.diff(off).astype(np.int64))
