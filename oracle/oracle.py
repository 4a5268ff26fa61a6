"""ctypes wrapper of the CPU oracle (oracle/sstg_oracle.c).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product package
(sstgraph_amd) never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libsstg_oracle.so")
REF_RMAT = os.path.join(_HERE, "_ref", "rmat_ref")

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")

_lib = None


def build():
    """Compile the oracle (and the reference RMAT tool where the reference
    tree is present) with the Makefile beside this file."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    u32, u64, i32, f64 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_double
    sig = {
        "orc_hash64": (u64, [u64]),
        "orc_hash32": (u32, [u32]),
        "orc_aspen_ith_rand": (u64, [u64, u64]),
        "orc_log2_up": (u32, [u64]),
        "orc_rmat": (ctypes.c_int, [u32, u32, f64, f64, f64, u64, u64, _u32p, _u32p]),
        "orc_rmat_normalized": (ctypes.c_int,
                                [u32, u32, f64, f64, f64, u64, u64, _u32p, _u32p]),
        "orc_er": (ctypes.c_int, [u32, u32, u64, u64, _u32p, _u32p]),
        "orc_graph_create": (vp, [u32, u32]),
        "orc_graph_destroy": (None, [vp]),
        "orc_insert_batch": (ctypes.c_int, [vp, _u32p, _u32p, u64]),
        "orc_remove_batch": (ctypes.c_int, [vp, _u32p, _u32p, u64]),
        "orc_insert": (ctypes.c_int, [vp, u32, u32]),
        "orc_remove": (ctypes.c_int, [vp, u32, u32]),
        "orc_has": (ctypes.c_int, [vp, u32, u32]),
        "orc_num_edges": (u64, [vp]),
        "orc_num_nodes": (u32, [vp]),
        "orc_degrees": (None, [vp, _u32p]),
        "orc_export_csr": (None, [vp, _u64p, vp]),
        "orc_load_csr": (ctypes.c_int, [vp, _u64p, _u32p]),
        "orc_touch_all_sum": (u64, [vp]),
        "orc_bfs": (None, [u32, _u64p, _u32p, u32, _u32p, _i32p]),
        "orc_pagerank": (None, [u32, _u64p, _u32p, u32, f64, _f64p]),
        "orc_bc": (None, [u32, _u64p, _u32p, u32, _f64p]),
        "orc_bfs_par": (None, [u32, _u64p, _u32p, u32, ctypes.c_int, _u32p]),
        "orc_bc_par": (None, [u32, _u64p, _u32p, u32, ctypes.c_int, _f64p]),
        "orc_cc": (None, [u32, _u64p, _u32p, _u32p]),
        "orc_common_neighbors": (u64, [_u64p, _u32p, u32, u32, ctypes.c_int]),
        "orc_tc": (u64, [u32, _u64p, _u32p]),
        "orc_bf": (None, [u32, _u64p, _u32p, vp, u32, _i32p]),
        "orc_set_threads": (None, [ctypes.c_int]),
        "orc_get_threads": (ctypes.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def hash32(x):
    return int(lib().orc_hash32(x))


def hash64(x):
    return int(lib().orc_hash64(x))


def rmat(n, seed, count, first=0, a=0.5, b=0.1, c=0.1, normalized=False):
    src = np.empty(count, dtype=np.uint32)
    dst = np.empty(count, dtype=np.uint32)
    fn = lib().orc_rmat_normalized if normalized else lib().orc_rmat
    rc = fn(n, seed, a, b, c, first, count, src, dst)
    if rc != 0:
        raise ValueError("rMat: n must be a power of two and a+b+c <= 1")
    return src, dst


def er(n, seed, count, first=0):
    src = np.empty(count, dtype=np.uint32)
    dst = np.empty(count, dtype=np.uint32)
    if lib().orc_er(n, seed, first, count, src, dst) != 0:
        raise ValueError("er: n must be positive")
    return src, dst


def set_threads(t):
    lib().orc_set_threads(int(t))


def threads():
    return int(lib().orc_get_threads())


class Graph:
    """Oracle tinyset graph (sorted unique rows)."""

    def __init__(self, n_rows, n_cols=None):
        self.n_rows = int(n_rows)
        self.n_cols = int(n_rows if n_cols is None else n_cols)
        self._h = lib().orc_graph_create(self.n_rows, self.n_cols)
        if not self._h:
            raise MemoryError("orc_graph_create")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_graph_destroy(h)
            self._h = None

    def insert_batch(self, src, dst):
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        rc = lib().orc_insert_batch(self._h, src, dst, len(src))
        if rc:
            raise ValueError(f"orc_insert_batch rc={rc}")

    def remove_batch(self, src, dst):
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        rc = lib().orc_remove_batch(self._h, src, dst, len(src))
        if rc:
            raise ValueError(f"orc_remove_batch rc={rc}")

    def insert(self, r, c):
        if lib().orc_insert(self._h, r, c):
            raise ValueError("orc_insert")

    def remove(self, r, c):
        if lib().orc_remove(self._h, r, c):
            raise ValueError("orc_remove")

    def has(self, r, c):
        return bool(lib().orc_has(self._h, r, c))

    def num_edges(self):
        return int(lib().orc_num_edges(self._h))

    def degrees(self):
        out = np.empty(self.n_rows, dtype=np.uint32)
        lib().orc_degrees(self._h, out)
        return out

    def csr(self):
        off = np.empty(self.n_rows + 1, dtype=np.uint64)
        lib().orc_export_csr(self._h, off, None)
        cols = np.empty(max(int(off[-1]), 1), dtype=np.uint32)
        lib().orc_export_csr(self._h, off, cols.ctypes.data)
        return off, cols[: int(off[-1])]

    def load_csr(self, off, cols):
        off = np.ascontiguousarray(off, dtype=np.uint64)
        cols = np.ascontiguousarray(cols, dtype=np.uint32)
        if cols.size == 0:
            cols = np.zeros(1, dtype=np.uint32)
        if lib().orc_load_csr(self._h, off, cols):
            raise MemoryError("orc_load_csr")

    def touch_all_sum(self):
        return int(lib().orc_touch_all_sum(self._h))


def _csr_args(off, cols):
    off = np.ascontiguousarray(off, dtype=np.uint64)
    cols = np.ascontiguousarray(cols, dtype=np.uint32)
    if cols.size == 0:
        cols = np.zeros(1, dtype=np.uint32)
    return off, cols


def bfs(off, cols, src):
    off, cols = _csr_args(off, cols)
    n = len(off) - 1
    level = np.empty(n, dtype=np.uint32)
    parent = np.empty(n, dtype=np.int32)
    lib().orc_bfs(n, off, cols, src, level, parent)
    return level, parent


def bfs_par(off, cols, src, symmetric=None):
    """BFS levels on every host core (the timed baseline; equals bfs).
    symmetric: whether every edge has its reverse (None: checked first)."""
    off, cols = _csr_args(off, cols)
    n = len(off) - 1
    level = np.empty(n, dtype=np.uint32)
    lib().orc_bfs_par(n, off, cols, src, -1 if symmetric is None else int(symmetric),
                      level)
    return level


def bc_par(off, cols, src, symmetric=None):
    """Brandes dependencies on every host core (the timed baseline)."""
    off, cols = _csr_args(off, cols)
    n = len(off) - 1
    out = np.empty(n, dtype=np.float64)
    lib().orc_bc_par(n, off, cols, src, -1 if symmetric is None else int(symmetric),
                     out)
    return out


def pagerank(off, cols, iters, damping=0.85):
    off, cols = _csr_args(off, cols)
    n = len(off) - 1
    out = np.empty(n, dtype=np.float64)
    lib().orc_pagerank(n, off, cols, iters, damping, out)
    return out


def bc(off, cols, src):
    off, cols = _csr_args(off, cols)
    n = len(off) - 1
    out = np.empty(n, dtype=np.float64)
    lib().orc_bc(n, off, cols, src, out)
    return out


def cc(off, cols):
    off, cols = _csr_args(off, cols)
    n = len(off) - 1
    out = np.empty(n, dtype=np.uint32)
    lib().orc_cc(n, off, cols, out)
    return out


def tc(off, cols):
    off, cols = _csr_args(off, cols)
    return int(lib().orc_tc(len(off) - 1, off, cols))


def common_neighbors(off, cols, u, v, early_end):
    off, cols = _csr_args(off, cols)
    return int(lib().orc_common_neighbors(off, cols, u, v, 1 if early_end else 0))


def bf(off, cols, src, weights=None):
    off, cols = _csr_args(off, cols)
    n = len(off) - 1
    out = np.empty(n, dtype=np.int32)
    wp = None
    if weights is not None:
        weights = np.ascontiguousarray(weights, dtype=np.int32)
        wp = weights.ctypes.data
    lib().orc_bf(n, off, cols, wp, src, out)
    return out


def rmat_graph_edges(scale, edge_factor, seed, a=0.5, b=0.1, c=0.1):
    """RMAT base-graph edge list as in the workloads: 2^scale vertices,
    edge_factor * 2^scale generated edges, symmetrised (each (u,v) also adds
    (v,u)).  Self-loops and duplicates are left in; insert_batch drops them."""
    n = 1 << scale
    s, d = rmat(n, seed, edge_factor * n, a=a, b=b, c=c)
    return n, np.concatenate([s, d]), np.concatenate([d, s])
