"""sstgraph_amd -- SSTGraph's hot path on MI355X (gfx950).

Python mirror of the reference's operator interface over the C ABI in
include/sstg.h (libsstg.so next to this file):

  SparseMatrixV(height, width)           SparseMatrix.hpp:73-87
    .insert_batch(src, dst)               SparseMatrix.hpp:556-643
    .remove_batch(src, dst)               SparseMatrix.hpp:645-718
    .insert / .remove / .has              SparseMatrix.hpp:536-554, 720-726
    .num_nodes / .getDegree / .M          SparseMatrix.hpp:56-59, 518-525
    .get_memory_size / .touch_all_sum     SparseMatrix.hpp:527-534, 436-454
    .common_neighbors(u, v, early_end)    SparseMatrix.hpp:61-71
  EdgeMapVertexMap.BFS / PR_S / BC / CC / TC / BF / TouchAll
                                          test.hpp:476-618 call sites

All compute runs in the HIP library; there is no CPU fallback.  Importing
works without a GPU (for inspection), but creating a graph raises
DeviceUnavailable when no gfx950 device is present.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsstg.so")

SSTG_OK = 0
SSTG_E_INVAL = -1
SSTG_E_RANGE = -2
SSTG_E_NOMEM = -3
SSTG_E_DEVICE = -4
SSTG_E_NODEVICE = -5
SSTG_E_UNSUPPORTED = -6
SSTG_E_IO = -7

SSTG_DEVICE_PTRS = 0x1
SSTG_SERVE_COALESCE = 0x1
SSTG_SERVE_SNAPSHOTS = 0x2
SSTG_SERVE_ONE_HEAVY = 0x4
SSTG_SERVE_MAINTAIN = 0x8
SSTG_SEMI_SORTED = 0x2
SSTG_KEEP_SELF_LOOPS = 0x4
SSTG_LOAD_WEIGHTS = 0x8
SSTG_RMAT_NORMALIZED = 0x10
SSTG_TUNE_PR = 0x1
SSTG_TUNE_BFS = 0x2

UNREACHED = 0xFFFFFFFF


class SSTGError(RuntimeError):
    def __init__(self, code, what):
        self.code = code
        super().__init__(f"{what}: {strerror(code)} ({code})")


class DeviceUnavailable(SSTGError):
    pass


class IdOutOfRange(SSTGError, ValueError):
    pass


class Stats(ctypes.Structure):
    """sstg_stats (print_statistics figures)."""
    _fields_ = [("edges", ctypes.c_uint64), ("rows", ctypes.c_uint32),
                ("cols", ctypes.c_uint32), ("memory_bytes", ctypes.c_uint64),
                ("empty_rows", ctypes.c_uint64), ("max_degree", ctypes.c_uint64),
                ("slot_capacity", ctypes.c_uint64),
                ("pool_capacity", ctypes.c_uint64), ("pool_used", ctypes.c_uint64),
                ("pool_garbage", ctypes.c_uint64), ("rebuilds", ctypes.c_uint64),
                ("dest_bits", ctypes.c_uint64 * 33)]


class TuneReport(ctypes.Structure):
    """sstg_tune_report (the degree-threshold autotuner's outcome)."""
    _fields_ = [("thresholds", ctypes.c_uint32 * 3),
                ("start_thresholds", ctypes.c_uint32 * 3),
                ("bfs_thresholds", ctypes.c_uint32 * 3),
                ("start_bfs_thresholds", ctypes.c_uint32 * 3),
                ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
                ("start_alpha", ctypes.c_double), ("start_beta", ctypes.c_double),
                ("pr_start_ms", ctypes.c_double), ("pr_best_ms", ctypes.c_double),
                ("bfs_start_ms", ctypes.c_double), ("bfs_best_ms", ctypes.c_double),
                ("evaluations", ctypes.c_uint32)]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        for k in ("thresholds", "start_thresholds", "bfs_thresholds",
                  "start_bfs_thresholds"):
            d[k] = list(d[k])
        d["evaluations"] = int(d["evaluations"])
        return d


def _autotune(handle, pr, bfs_source, pr_iters):
    what = (SSTG_TUNE_PR if pr else 0) | (SSTG_TUNE_BFS if bfs_source is not None else 0)
    rep = TuneReport()
    _check(lib().sstg_autotune(handle, what, 0 if bfs_source is None else bfs_source,
                               pr_iters, ctypes.byref(rep)), "autotune")
    return rep.as_dict()


class UpdateStats(ctypes.Structure):
    """sstg_update_stats: what the last batch update did."""
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "batch", "effective", "rows", "tile_rows", "tile_inputs", "big_rows",
        "staged", "merged", "allocated", "rebuilds")]

    def as_dict(self):
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


REQ_INSERT, REQ_REMOVE, REQ_BFS, REQ_PR, REQ_TC, REQ_CC = 1, 2, 3, 4, 5, 6


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("source", ctypes.c_uint32),
                ("iters", ctypes.c_uint32), ("damping", ctypes.c_double),
                ("src", ctypes.c_void_p), ("dst", ctypes.c_void_p),
                ("n", ctypes.c_uint64)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("t_arrive", ctypes.c_double),
                ("t_start", ctypes.c_double), ("t_done", ctypes.c_double),
                ("scalar", ctypes.c_uint64), ("coalesced", ctypes.c_uint64)]


_lib = None


def lib():
    """The loaded libsstg.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"sstgraph_amd: {LIB_PATH} is missing; build it with "
            "`python -m sstgraph_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    P = ctypes.POINTER
    sig = {
        "sstg_strerror": (ctypes.c_char_p, [i32]),
        "sstg_abi_version": (i32, []),
        "sstg_device_count": (i32, [P(i32)]),
        "sstg_create": (i32, [u32, u32, u32, P(vp)]),
        "sstg_destroy": (i32, [vp]),
        "sstg_insert_batch": (i32, [vp, vp, vp, u64, u32]),
        "sstg_remove_batch": (i32, [vp, vp, vp, u64, u32]),
        "sstg_insert_batch_weighted": (i32, [vp, vp, vp, vp, u64, u32]),
        "sstg_edge_weights": (i32, [vp, vp, vp, u64, vp, u32]),
        "sstg_insert": (i32, [vp, u32, u32]),
        "sstg_remove": (i32, [vp, u32, u32]),
        "sstg_has": (i32, [vp, u32, u32, P(i32)]),
        "sstg_last_update_stats": (i32, [vp, P(UpdateStats)]),
        "sstg_has_batch": (i32, [vp, vp, vp, u64, vp, u32]),
        "sstg_num_nodes": (i32, [vp, P(u32)]),
        "sstg_num_edges": (i32, [vp, P(u64)]),
        "sstg_degrees": (i32, [vp, vp, u32]),
        "sstg_export_csr": (i32, [vp, vp, vp, u32]),
        "sstg_memory_size": (i32, [vp, P(u64)]),
        "sstg_touch_all_sum": (i32, [vp, P(u64)]),
        "sstg_touch_all": (i32, [vp, P(u64)]),
        "sstg_common_neighbors": (i32, [vp, u32, u32, i32, P(u64)]),
        "sstg_bfs": (i32, [vp, u32, vp, vp, u32]),
        "sstg_pagerank_f64": (i32, [vp, u32, ctypes.c_double, vp, u32]),
        "sstg_pagerank_f32k": (i32, [vp, u32, ctypes.c_double, vp, u32]),
        "sstg_bc": (i32, [vp, u32, vp, u32]),
        "sstg_cc": (i32, [vp, vp, u32]),
        "sstg_tc": (i32, [vp, P(u64)]),
        "sstg_bf": (i32, [vp, u32, vp, u32]),
        "sstg_rmat": (i32, [u32, u32, ctypes.c_double, ctypes.c_double,
                            ctypes.c_double, u64, u64, vp, vp, u32]),
        "sstg_er": (i32, [u32, u32, u64, u64, vp, vp, u32]),
        "sstg_multi_create": (i32, [u32, vp, P(vp)]),
        "sstg_multi_destroy": (i32, [vp]),
        "sstg_multi_insert_batch": (i32, [vp, u32, vp, vp, u64, u32]),
        "sstg_multi_insert_batches": (i32, [vp, u32, vp, vp, vp, vp, u32]),
        "sstg_multi_remove_batches": (i32, [vp, u32, vp, vp, vp, vp, u32]),
        "sstg_multi_graph": (i32, [vp, P(vp)]),
        "sstg_multi_run": (i32, [vp, vp, u32, u32, ctypes.c_double, i32]),
        "sstg_multi_results": (i32, [vp, vp, vp, vp, vp, P(u32)]),
        "sstg_server_create": (i32, [vp, u32, u64, u32, P(vp)]),
        "sstg_server_submit": (i32, [vp, P(Request), P(u64)]),
        "sstg_server_wait": (i32, [vp, u64, P(RequestInfo)]),
        "sstg_server_fetch": (i32, [vp, u64, vp, u64]),
        "sstg_server_release": (i32, [vp, u64]),
        "sstg_server_stats": (i32, [vp, P(u64), P(u64), P(u64)]),
        "sstg_server_destroy": (i32, [vp]),
        "sstg_degree_order_map": (i32, [vp, vp, u32]),
        "sstg_rewrite_graph": (i32, [vp, vp, u32, P(vp)]),
        "sstg_statistics": (i32, [vp, P(Stats)]),
        "sstg_set_degree_thresholds": (i32, [vp, vp]),
        "sstg_get_degree_thresholds": (i32, [vp, vp]),
        "sstg_autotune": (i32, [vp, u32, u32, u32, P(TuneReport)]),
        "sstg_set_bfs_switch": (i32, [vp, ctypes.c_double, ctypes.c_double]),
        "sstg_set_pagerank_segments": (i32, [vp, u32]),
        "sstg_get_pagerank_segments": (i32, [vp, P(u32)]),
        "sstg_set_bfs_degree_thresholds": (i32, [vp, vp]),
        "sstg_get_bfs_degree_thresholds": (i32, [vp, vp]),
        "sstg_get_bfs_switch": (i32, [vp, P(ctypes.c_double), P(ctypes.c_double)]),
        "sstg_edges_from_file": (i32, [ctypes.c_char_p, P(u32), P(u64), P(vp),
                                       P(vp), P(vp)]),
        "sstg_host_free": (None, [vp]),
        "sstg_load_file": (i32, [ctypes.c_char_p, u32, P(vp)]),
        "sstg_write_adj": (i32, [vp, ctypes.c_char_p]),
        "sstg_device_alloc": (i32, [u64, P(vp)]),
        "sstg_device_free": (i32, [vp]),
        "sstg_memcpy_h2d": (i32, [vp, vp, u64]),
        "sstg_memcpy_d2h": (i32, [vp, vp, u64]),
        "sstg_memcpy_d2d": (i32, [vp, vp, u64]),
        "sstg_device_synchronize": (i32, []),
        "sstg_mem_info": (i32, [P(u64), P(u64)]),
        "sstg_profile_enable": (i32, [i32]),
        "sstg_profile_reset": (i32, []),
        "sstg_profile_query": (i32, [ctypes.c_char_p, P(ctypes.c_double), P(u64)]),
        "sstg_profile_dump": (i32, [ctypes.c_char_p, u64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def strerror(code):
    try:
        return lib().sstg_strerror(code).decode()
    except Exception:  # pragma: no cover - library missing
        return "error"


def _check(rc, what):
    if rc == SSTG_OK:
        return
    if rc == SSTG_E_NODEVICE:
        raise DeviceUnavailable(rc, what)
    if rc == SSTG_E_RANGE:
        raise IdOutOfRange(rc, what)
    if rc == SSTG_E_IO:
        raise OSError(f"{what}: {lib().sstg_strerror(rc).decode()}")
    raise SSTGError(rc, what)


def device_count():
    n = ctypes.c_int(0)
    _check(lib().sstg_device_count(ctypes.byref(n)), "sstg_device_count")
    return n.value


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _u32(a):
    """Vertex ids as a contiguous uint32 array.  Ids of another integer type
    must fit in [0, 2^32): a negative or wider id would otherwise wrap onto
    a valid vertex."""
    a = np.asarray(a)
    if a.dtype != np.uint32 and a.size:
        if not np.issubdtype(a.dtype, np.integer):
            raise TypeError(f"vertex ids must be integers, not {a.dtype}")
        if int(a.min()) < 0 or int(a.max()) >= 1 << 32:
            raise IdOutOfRange(SSTG_E_RANGE, "vertex id outside [0, 2^32)")
    return np.ascontiguousarray(a, dtype=np.uint32)


class DeviceArray:
    """A device allocation owned by Python (plumbing for callers that keep
    batches or results in HBM)."""

    def __init__(self, count, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(lib().sstg_device_alloc(max(self.nbytes, 1), ctypes.byref(p)),
               "sstg_device_alloc")
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if a.nbytes:
            _check(lib().sstg_memcpy_h2d(d.ptr, _ptr(a), a.nbytes), "h2d")
        return d

    def to_host(self):
        out = np.empty(self.count, dtype=self.dtype)
        if self.nbytes:
            _check(lib().sstg_memcpy_d2h(_ptr(out), self.ptr, self.nbytes), "d2h")
        return out

    def free(self):
        if self.ptr is not None and _lib is not None:
            lib().sstg_device_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def rmat(n, seed, count, first=0, a=0.5, b=0.1, c=0.1, device=False,
         normalized=False):
    """Edges [first, first+count) of rMat<uint32_t>(n, seed, a, b, c)
    (rmat_util.h:82-131), generated on the GPU.  device=True returns a pair
    of DeviceArray; normalized=True draws in [0, 1) (quadrant probabilities
    a, b, c as given) instead of the reference's [0, 2)."""
    nf = SSTG_RMAT_NORMALIZED if normalized else 0
    if device:
        s, d = DeviceArray(count, np.uint32), DeviceArray(count, np.uint32)
        _check(lib().sstg_rmat(n, seed, a, b, c, first, count, s.ptr, d.ptr,
                               SSTG_DEVICE_PTRS | nf), "sstg_rmat")
        return s, d
    s = np.empty(count, np.uint32)
    d = np.empty(count, np.uint32)
    _check(lib().sstg_rmat(n, seed, a, b, c, first, count, _ptr(s), _ptr(d), nf),
           "sstg_rmat")
    return s, d


def er(n, seed, count, first=0, device=False):
    """Seeded Erdos-Renyi edge stream (see sstg.h sstg_er), on the GPU."""
    if device:
        s, d = DeviceArray(count, np.uint32), DeviceArray(count, np.uint32)
        _check(lib().sstg_er(n, seed, first, count, s.ptr, d.ptr,
                             SSTG_DEVICE_PTRS), "sstg_er")
        return s, d
    s = np.empty(count, np.uint32)
    d = np.empty(count, np.uint32)
    _check(lib().sstg_er(n, seed, first, count, _ptr(s), _ptr(d), 0), "sstg_er")
    return s, d


class MultiGraph:
    """k independent graphs in one device store; BFS + CC + PageRank on all
    of them in one captured hipGraph replay (sstg_multi_*)."""

    def __init__(self, sizes):
        self.sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
        self.k = len(self.sizes)
        self.vbase = np.concatenate([[0], np.cumsum(self.sizes.astype(np.int64))])
        h = ctypes.c_void_p()
        _check(lib().sstg_multi_create(self.k, _ptr(self.sizes), ctypes.byref(h)),
               "sstg_multi_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().sstg_multi_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def insert_batch(self, graph, src, dst):
        if isinstance(src, DeviceArray):
            _check(lib().sstg_multi_insert_batch(self._h, graph, src.ptr, dst.ptr,
                                                 min(src.count, dst.count),
                                                 SSTG_DEVICE_PTRS), "multi insert")
            return
        s, d = _u32(src), _u32(dst)
        _check(lib().sstg_multi_insert_batch(self._h, graph, _ptr(s), _ptr(d),
                                             s.size, 0), "multi insert")

    def _batches(self, fn, batches, what):
        g = np.array([b[0] for b in batches], dtype=np.uint32)
        srcs = [_u32(b[1]) for b in batches]
        dsts = [_u32(b[2]) for b in batches]
        c = np.array([min(a.size, b.size) for a, b in zip(srcs, dsts)],
                     dtype=np.uint64)
        s = np.ascontiguousarray(np.concatenate(
            [a[:k] for a, k in zip(srcs, c)] + [np.empty(0, np.uint32)]))
        d = np.ascontiguousarray(np.concatenate(
            [a[:k] for a, k in zip(dsts, c)] + [np.empty(0, np.uint32)]))
        _check(fn(self._h, g.size, _ptr(g), _ptr(c), _ptr(s), _ptr(d), 0),
               what)

    def _batches_device(self, fn, graphs, counts, src, dst, what):
        g = _u32(graphs)
        c = np.ascontiguousarray(counts, dtype=np.uint64)
        if int(c.sum()) > min(src.count, dst.count):
            raise ValueError("counts exceed the device arrays")
        _check(fn(self._h, g.size, _ptr(g), _ptr(c), src.ptr, dst.ptr,
                  SSTG_DEVICE_PTRS), what)

    def insert_batches(self, batches):
        """insert_batch of several graphs in one update pass: batches is a
        list of (graph, src, dst) with ids local to each graph."""
        self._batches(lib().sstg_multi_insert_batches, batches,
                      "multi insert batches")

    def remove_batches(self, batches):
        """remove_batch of several graphs in one update pass."""
        self._batches(lib().sstg_multi_remove_batches, batches,
                      "multi remove batches")

    def insert_batches_device(self, graphs, counts, src, dst):
        """insert_batches with the batches back to back in DeviceArrays."""
        self._batches_device(lib().sstg_multi_insert_batches, graphs, counts,
                             src, dst, "multi insert batches")

    def remove_batches_device(self, graphs, counts, src, dst):
        self._batches_device(lib().sstg_multi_remove_batches, graphs, counts,
                             src, dst, "multi remove batches")

    def edges(self):
        """Directed edges stored over all graphs."""
        h = ctypes.c_void_p()
        _check(lib().sstg_multi_graph(self._h, ctypes.byref(h)), "multi graph")
        out = ctypes.c_uint64(0)
        _check(lib().sstg_num_edges(h, ctypes.byref(out)), "M")
        return out.value

    def export_csr(self, graph):
        """Graph `graph`'s rows as (offsets, columns) with local ids."""
        h = ctypes.c_void_p()
        _check(lib().sstg_multi_graph(self._h, ctypes.byref(h)), "multi graph")
        n = int(self.vbase[-1])
        off = np.empty(n + 1, np.uint64)
        _check(lib().sstg_export_csr(h, _ptr(off), None, 0), "export_csr")
        cols = np.empty(max(int(off[-1]), 1), np.uint32)
        _check(lib().sstg_export_csr(h, _ptr(off), _ptr(cols), 0), "export_csr")
        lo, hi = int(self.vbase[graph]), int(self.vbase[graph + 1])
        part = off[lo:hi + 1]
        c = cols[int(part[0]):int(part[-1])] - np.uint32(lo)
        return part - part[0], c

    def autotune(self, pr_iters=2):
        """Tunes the degree thresholds of the union store on PageRank
        iterations over all graphs; the next run re-captures with them."""
        h = ctypes.c_void_p()
        _check(lib().sstg_multi_graph(self._h, ctypes.byref(h)), "multi graph")
        return _autotune(h, True, None, pr_iters)

    def run(self, sources, max_levels=64, pr_iters=10, damping=0.85, f32=False):
        src = _u32(sources)
        _check(lib().sstg_multi_run(self._h, _ptr(src), max_levels, pr_iters,
                                    damping, 1 if f32 else 0), "multi run")

    def results(self, f32=False):
        n = int(self.vbase[-1])
        level = np.empty(n, np.uint32)
        label = np.empty(n, np.uint32)
        rank = np.empty(n, np.float32 if f32 else np.float64)
        used = ctypes.c_uint32(0)
        _check(lib().sstg_multi_results(self._h, _ptr(level), _ptr(label),
                                        None if f32 else _ptr(rank),
                                        _ptr(rank) if f32 else None,
                                        ctypes.byref(used)), "multi results")
        return dict(level=level, label=label, rank=rank, levels_used=used.value)

    def part(self, arr, g):
        return arr[self.vbase[g]:self.vbase[g + 1]]


def edges_from_file(path, weights=False):
    """get_edges_from_file (helpers.hpp:604-631) for .adj / .adj.shuf / .mtx:
    returns (n, src, dst) or (n, src, dst, weights) as numpy uint32 arrays.
    Runs on the host; no device needed."""
    n, m = ctypes.c_uint32(0), ctypes.c_uint64(0)
    ps, pd, pw = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    _check(lib().sstg_edges_from_file(os.fsencode(path), ctypes.byref(n),
                                      ctypes.byref(m), ctypes.byref(ps),
                                      ctypes.byref(pd),
                                      ctypes.byref(pw) if weights else None),
           f"read {path}")
    out = []
    try:
        for p in (ps, pd) + ((pw,) if weights else ()):
            buf = (ctypes.c_uint32 * max(m.value, 1)).from_address(p.value)
            out.append(np.frombuffer(buf, np.uint32, count=m.value).copy())
    finally:
        for p in (ps, pd, pw):
            if p.value:
                lib().sstg_host_free(p)
    return (n.value, *out)


class SparseMatrixV:
    """Device-resident SparseMatrixV<true> (CSR, binary): one sorted set of
    column ids per row."""

    def __init__(self, height, width=None, _handle=None):
        self.n_rows = int(height)
        self.n_cols = int(height if width is None else width)
        if _handle is not None:
            self._h = _handle
            return
        h = ctypes.c_void_p()
        _check(lib().sstg_create(self.n_rows, self.n_cols, 0, ctypes.byref(h)),
               "sstg_create")
        self._h = h

    @classmethod
    def from_file(cls, path, weights=False):
        """Reads a .adj / .mtx file into a new graph (sstg_load_file); with
        weights=True the file's weights are set (1 for an unweighted file)."""
        h = ctypes.c_void_p()
        _check(lib().sstg_load_file(os.fsencode(path),
                                    SSTG_LOAD_WEIGHTS if weights else 0,
                                    ctypes.byref(h)),
               f"load {path}")
        n = ctypes.c_uint32(0)
        _check(lib().sstg_num_nodes(h, ctypes.byref(n)), "num_nodes")
        return cls(n.value, n.value, _handle=h)

    def write_adj(self, path):
        """Writes an AdjacencyGraph file (graph_utilities.cpp:151-183)."""
        _check(lib().sstg_write_adj(self._h, os.fsencode(path)), f"write {path}")

    def degree_order_map(self):
        """SparseMatrix.hpp:737-749: new label of each vertex by descending
        (degree, id), computed on the device."""
        out = np.empty(self.n_rows, np.uint32)
        _check(lib().sstg_degree_order_map(self._h, _ptr(out), 0), "degree_order_map")
        return out

    def rewrite_graph(self, new_label_map):
        """SparseMatrix.hpp:751-766: a new graph with every edge relabelled."""
        m = _u32(new_label_map)
        if m.size != self.n_rows:
            raise ValueError("label map needs one entry per row")
        h = ctypes.c_void_p()
        _check(lib().sstg_rewrite_graph(self._h, m.ctypes.data, 0, ctypes.byref(h)),
               "rewrite_graph")
        return SparseMatrixV(self.n_rows, self.n_cols, _handle=h)

    def degree_thresholds(self, bfs=False):
        """Degree boundaries of the 1-lane / 8-lane / wave / block groups of
        the reductions (PageRank, BC, TC), or of BFS with bfs=True."""
        t = np.zeros(3, np.uint32)
        fn = (lib().sstg_get_bfs_degree_thresholds if bfs
              else lib().sstg_get_degree_thresholds)
        _check(fn(self._h, _ptr(t)), "thresholds")
        return t.tolist()

    def set_degree_thresholds(self, t, bfs=False):
        t = _u32(t)
        if t.size != 3:
            raise ValueError("three thresholds")
        fn = (lib().sstg_set_bfs_degree_thresholds if bfs
              else lib().sstg_set_degree_thresholds)
        _check(fn(self._h, _ptr(t)), "thresholds")

    def bfs_switch(self):
        """(alpha, beta) of the BFS direction switch."""
        a, b = ctypes.c_double(0), ctypes.c_double(0)
        _check(lib().sstg_get_bfs_switch(self._h, ctypes.byref(a), ctypes.byref(b)),
               "bfs_switch")
        return a.value, b.value

    def set_bfs_switch(self, alpha, beta):
        _check(lib().sstg_set_bfs_switch(self._h, alpha, beta), "bfs_switch")

    def pagerank_segments(self):
        """Column-segment width (log2 vertices) of FP64 PageRank; 0 = off."""
        b = ctypes.c_uint32(0)
        _check(lib().sstg_get_pagerank_segments(self._h, ctypes.byref(b)),
               "pagerank_segments")
        return b.value

    def set_pagerank_segments(self, bits):
        _check(lib().sstg_set_pagerank_segments(self._h, int(bits)),
               "pagerank_segments")

    def autotune(self, pr=True, bfs_source=None, pr_iters=2):
        """Picks the degree thresholds by timing PageRank iterations and/or a
        BFS from bfs_source on this graph (sstg_autotune); returns the
        report and leaves the chosen thresholds set."""
        return _autotune(self._h, pr, bfs_source, pr_iters)

    def statistics(self):
        st = Stats()
        _check(lib().sstg_statistics(self._h, ctypes.byref(st)), "statistics")
        d = {k: getattr(st, k) for k, _ in Stats._fields_ if k != "dest_bits"}
        d["dest_bits"] = list(st.dest_bits)
        return d

    def print_statistics(self, file=None):
        """print_statistics (SparseMatrix.hpp:340-426): the layout-free lines
        of the reference, then the HBM store's own figures."""
        st = self.statistics()
        say = lambda *a: print(*a, file=file)
        say("matrix stored in csr")
        say(f"there are {st['edges']} elements")
        say(f"there are {st['rows']} rows and {st['cols']} columns")
        say(f"The graph uses {st['memory_bytes']} bytes of memory")
        say(f"{st['empty_rows']} rows are empty; the largest row has "
            f"{st['max_degree']} elements")
        say(f"slots hold {st['slot_capacity']} elements; the pool has "
            f"{st['pool_capacity']} ({st['pool_used']} below the tail, "
            f"{st['pool_garbage']} abandoned), {st['rebuilds']} rebuilds")
        say("count of outgoing edges which uses x bits")
        total = 0
        for i, c in enumerate(st["dest_bits"]):
            total += c
            say(f"{i}, {c}, total so far = {total}")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().sstg_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # ---- updates
    def _batch(self, fn, src, dst, flags, what):
        if isinstance(src, DeviceArray):
            n = min(src.count, dst.count)
            _check(fn(self._h, src.ptr, dst.ptr, n, flags | SSTG_DEVICE_PTRS), what)
            return
        if dst is None:  # array of (row, col) pairs
            e = np.asarray(src)
            src, dst = e[:, 0], e[:, 1]
        s, d = _u32(src), _u32(dst)
        if s.shape != d.shape:
            raise ValueError("src and dst differ in length")
        _check(fn(self._h, _ptr(s), _ptr(d), s.size, flags), what)

    def insert_batch(self, src, dst=None, semi_sorted=False):
        self._batch(lib().sstg_insert_batch, src, dst,
                    SSTG_SEMI_SORTED if semi_sorted else 0, "insert_batch")

    def remove_batch(self, src, dst=None, semi_sorted=False):
        self._batch(lib().sstg_remove_batch, src, dst,
                    SSTG_SEMI_SORTED if semi_sorted else 0, "remove_batch")

    def insert_batch_weighted(self, src, dst=None, weights=None):
        """insert_batch of (row, col, value) tuples on a SparseMatrixV<true,
        uint32_t> (SparseMatrix.hpp:556-643): the edges are inserted and each
        one's weight set, the last occurrence in the batch winning.  Takes
        three arrays, or one array of (row, col, value) rows."""
        if isinstance(src, DeviceArray):
            n = min(src.count, dst.count, weights.count)
            _check(lib().sstg_insert_batch_weighted(
                self._h, src.ptr, dst.ptr, weights.ptr, n, SSTG_DEVICE_PTRS),
                "insert_batch_weighted")
            return
        if dst is None:
            e = np.asarray(src)
            src, dst, weights = e[:, 0], e[:, 1], e[:, 2]
        s, d = _u32(src), _u32(dst)
        w = np.asarray(weights)
        if w.dtype != np.uint32 and w.size and (
                int(w.min()) < -(1 << 31) or int(w.max()) >= 1 << 32):
            raise ValueError("weights must fit in 32 bits")
        # negative weights keep their two's complement bits (read as int32)
        w = np.ascontiguousarray(w.astype(np.int64) & 0xFFFFFFFF, dtype=np.uint32) \
            if w.dtype != np.uint32 else np.ascontiguousarray(w)
        if s.shape != d.shape or s.shape != w.shape:
            raise ValueError("src, dst and weights differ in length")
        _check(lib().sstg_insert_batch_weighted(self._h, _ptr(s), _ptr(d), _ptr(w),
                                                s.size, 0),
               "insert_batch_weighted")

    def value(self, row, col):
        """SparseMatrix.hpp:727-735: the edge's weight (1 if it was inserted
        without one), 0 if the edge is absent."""
        return int(self.values([row], [col])[0])

    def values(self, rows, cols):
        r, c = _u32(rows), _u32(cols)
        if r.shape != c.shape:
            raise ValueError("rows and cols differ in length")
        out = np.empty(r.size, np.uint32)
        _check(lib().sstg_edge_weights(self._h, _ptr(r), _ptr(c), r.size, _ptr(out), 0),
               "edge_weights")
        return out

    def last_update_stats(self):
        st = UpdateStats()
        _check(lib().sstg_last_update_stats(self._h, ctypes.byref(st)),
               "last_update_stats")
        return st.as_dict()

    def insert(self, row, col):
        _check(lib().sstg_insert(self._h, row, col), "insert")

    def remove(self, row, col):
        _check(lib().sstg_remove(self._h, row, col), "remove")

    # ---- queries
    def has(self, row, col):
        out = ctypes.c_int(0)
        _check(lib().sstg_has(self._h, row, col, ctypes.byref(out)), "has")
        return bool(out.value)

    def has_batch(self, rows, cols):
        r, c = _u32(rows), _u32(cols)
        out = np.empty(r.size, np.uint8)
        _check(lib().sstg_has_batch(self._h, _ptr(r), _ptr(c), r.size, _ptr(out), 0),
               "has_batch")
        return out.astype(bool)

    def num_nodes(self):
        return self.n_rows

    def M(self):
        out = ctypes.c_uint64(0)
        _check(lib().sstg_num_edges(self._h, ctypes.byref(out)), "M")
        return out.value

    num_edges = M

    def degrees(self):
        out = np.empty(self.n_rows, np.uint32)
        _check(lib().sstg_degrees(self._h, _ptr(out), 0), "degrees")
        return out

    def getDegree(self, node):
        return int(self.degrees()[node])

    def export_csr(self):
        off = np.empty(self.n_rows + 1, np.uint64)
        _check(lib().sstg_export_csr(self._h, _ptr(off), None, 0), "export_csr")
        cols = np.empty(max(int(off[-1]), 1), np.uint32)
        _check(lib().sstg_export_csr(self._h, _ptr(off), _ptr(cols), 0),
               "export_csr")
        return off, cols[: int(off[-1])]

    def get_memory_size(self):
        out = ctypes.c_uint64(0)
        _check(lib().sstg_memory_size(self._h, ctypes.byref(out)), "memory_size")
        return out.value

    def touch_all_sum(self):
        out = ctypes.c_uint64(0)
        _check(lib().sstg_touch_all_sum(self._h, ctypes.byref(out)), "touch_all_sum")
        return out.value

    def common_neighbors(self, u, v, early_end=False):
        out = ctypes.c_uint64(0)
        _check(lib().sstg_common_neighbors(self._h, u, v, 1 if early_end else 0,
                                           ctypes.byref(out)), "common_neighbors")
        return out.value

    def map_neighbors(self, i, f):
        """Host-side ordered neighbour scan of row i (SparseMatrix.hpp:218-
        253): calls f(i, nbr) in ascending order; f returning True stops the
        scan (early exit, PMA.hpp:420-426)."""
        off, cols = self.export_csr()
        for c in cols[off[i]:off[i + 1]].tolist():
            if f(i, c):
                break


class EdgeMapVertexMap:
    """The algorithm entry points the reference harness calls
    (test.hpp:476-618), returning numpy arrays instead of malloc'd memory."""

    @staticmethod
    def BFS(g, src, levels=False, out=None):
        """Parent array (and levels).  out=(parent, level) DeviceArrays keeps
        the results in HBM (either may be None)."""
        n = g.n_rows
        if out is not None:
            p, l = out
            _check(lib().sstg_bfs(g.handle, src, p.ptr if p is not None else None,
                                  l.ptr if l is not None else None,
                                  SSTG_DEVICE_PTRS), "BFS")
            return out
        parent = np.empty(n, np.int32)
        level = np.empty(n, np.uint32)
        _check(lib().sstg_bfs(g.handle, src, _ptr(parent), _ptr(level), 0), "BFS")
        return (parent, level) if levels else parent

    @staticmethod
    def PR_S(g, iters=10, damping=0.85, out=None):
        """PageRank in FP64; out=DeviceArray keeps the ranks in HBM."""
        if out is not None:
            _check(lib().sstg_pagerank_f64(g.handle, iters, damping, out.ptr,
                                           SSTG_DEVICE_PTRS), "PR_S")
            return out
        out = np.empty(g.n_rows, np.float64)
        _check(lib().sstg_pagerank_f64(g.handle, iters, damping, _ptr(out), 0),
               "PR_S")
        return out

    @staticmethod
    def PR_F32K(g, iters=10, damping=0.85):
        out = np.empty(g.n_rows, np.float32)
        _check(lib().sstg_pagerank_f32k(g.handle, iters, damping, _ptr(out), 0),
               "PR_F32K")
        return out

    @staticmethod
    def BC(g, src):
        out = np.empty(g.n_rows, np.float64)
        _check(lib().sstg_bc(g.handle, src, _ptr(out), 0), "BC")
        return out

    @staticmethod
    def CC(g):
        out = np.empty(g.n_rows, np.uint32)
        _check(lib().sstg_cc(g.handle, _ptr(out), 0), "CC")
        return out

    @staticmethod
    def TC(g):
        out = ctypes.c_uint64(0)
        _check(lib().sstg_tc(g.handle, ctypes.byref(out)), "TC")
        return out.value

    @staticmethod
    def BF(g, src):
        """Bellman-Ford distances (weights from insert_batch_weighted, else
        1 per edge), INT32_MAX // 2 where unreached."""
        out = np.empty(g.n_rows, np.int32)
        _check(lib().sstg_bf(g.handle, src, _ptr(out), 0), "BF")
        return out

    @staticmethod
    def TouchAll(g):
        out = ctypes.c_uint64(0)
        _check(lib().sstg_touch_all(g.handle, ctypes.byref(out)), "TouchAll")
        return out.value


class Profile:
    """Kernel timings recorded by the library with HIP events on the stream
    each kernel runs on."""

    @staticmethod
    def enable(on=True):
        lib().sstg_profile_enable(1 if on else 0)

    @staticmethod
    def reset():
        lib().sstg_profile_reset()

    @staticmethod
    def query(name):
        ms = ctypes.c_double(0)
        n = ctypes.c_uint64(0)
        lib().sstg_profile_query(name.encode(), ctypes.byref(ms), ctypes.byref(n))
        return ms.value, n.value

    @staticmethod
    def dump():
        buf = ctypes.create_string_buffer(1 << 16)
        lib().sstg_profile_dump(buf, len(buf))
        rows = []
        for line in buf.value.decode().splitlines():
            name, n, ms = line.rsplit(" ", 2)
            rows.append((name, int(n), float(ms)))
        return rows


def mem_info():
    f, t = ctypes.c_uint64(0), ctypes.c_uint64(0)
    _check(lib().sstg_mem_info(ctypes.byref(f), ctypes.byref(t)), "mem_info")
    return f.value, t.value


def synchronize():
    _check(lib().sstg_device_synchronize(), "synchronize")


class Server:
    """Serving loop over one device-resident graph (sstg_server_*): updates
    are applied in arrival order, consecutive same-kind updates as one device
    batch when coalesce is set; queries run concurrently on `slots` slots
    under an HBM budget; with snapshots, updates proceed on a copy of the
    store while earlier queries finish on the old one."""

    def __init__(self, g, slots=4, hbm_budget=0, coalesce=True, snapshots=False,
                 one_heavy=False, maintain=False):
        self.g = g
        h = ctypes.c_void_p()
        flags = (SSTG_SERVE_COALESCE if coalesce else 0) | \
            (SSTG_SERVE_SNAPSHOTS if snapshots else 0) | \
            (SSTG_SERVE_ONE_HEAVY if one_heavy else 0) | \
            (SSTG_SERVE_MAINTAIN if maintain else 0)
        _check(lib().sstg_server_create(g.handle, slots, hbm_budget, flags,
                                        ctypes.byref(h)),
               "sstg_server_create")
        self._h = h
        self._keep = {}

    def _submit(self, req, keep=None):
        t = ctypes.c_uint64(0)
        _check(lib().sstg_server_submit(self._h, ctypes.byref(req), ctypes.byref(t)),
               "submit")
        return t.value

    def insert(self, src, dst):
        s, d = _u32(src), _u32(dst)
        return self._submit(Request(REQ_INSERT, 0, 0, 0.0, s.ctypes.data,
                                    d.ctypes.data, s.size))

    def remove(self, src, dst):
        s, d = _u32(src), _u32(dst)
        return self._submit(Request(REQ_REMOVE, 0, 0, 0.0, s.ctypes.data,
                                    d.ctypes.data, s.size))

    def bfs(self, source):
        return self._submit(Request(REQ_BFS, source, 0, 0.0, None, None, 0))

    def pagerank(self, iters=10, damping=0.85):
        return self._submit(Request(REQ_PR, 0, iters, damping, None, None, 0))

    def tc(self):
        return self._submit(Request(REQ_TC, 0, 0, 0.0, None, None, 0))

    def cc(self):
        return self._submit(Request(REQ_CC, 0, 0, 0.0, None, None, 0))

    def wait(self, ticket):
        info = RequestInfo()
        _check(lib().sstg_server_wait(self._h, ticket, ctypes.byref(info)), "wait")
        return {k: getattr(info, k) for k, _ in RequestInfo._fields_}

    def fetch(self, ticket, dtype):
        out = np.empty(self.g.n_rows, dtype=dtype)
        _check(lib().sstg_server_fetch(self._h, ticket, _ptr(out), out.nbytes), "fetch")
        return out

    def release(self, ticket):
        _check(lib().sstg_server_release(self._h, ticket), "release")

    def stats(self):
        b, bud, sn = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(lib().sstg_server_stats(self._h, ctypes.byref(b), ctypes.byref(bud),
                                       ctypes.byref(sn)), "stats")
        return {"device_batches": b.value, "budget": bud.value, "snapshots": sn.value}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _check(lib().sstg_server_destroy(self._h), "server destroy")
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
