"""Small graph helpers shared by the tests (pure Python / numpy)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_mtx(path):
    """MatrixMarket coordinate reader: 1-indexed, and every entry is taken in
    both directions whether the file says symmetric or general -- the
    convention of the reference's loader (helpers.hpp:505-521: each line
    (s, d) yields (s-1, d-1) and (d-1, s-1))."""
    with open(path) as f:
        f.readline()
        lines = [ln for ln in f if not ln.startswith("%")]
    r, c, m = (int(x) for x in lines[0].split())
    pairs = np.array([[int(x) for x in ln.split()[:2]] for ln in lines[1:1 + m]],
                     dtype=np.int64).reshape(-1, 2) - 1
    src, dst = pairs[:, 0], pairs[:, 1]
    src, dst = np.concatenate([src, dst]), np.concatenate([dst, src])
    return r, src.astype(np.uint32), dst.astype(np.uint32)


def csr_from_edges(n, src, dst, drop_self=True):
    src = np.asarray(src, dtype=np.int64)
    dst = np.asarray(dst, dtype=np.int64)
    if drop_self:
        keep = src != dst
        src, dst = src[keep], dst[keep]
    key = np.unique(src * n + dst)
    s, d = key // n, key % n
    off = np.zeros(n + 1, dtype=np.uint64)
    np.add.at(off, s + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    return off, d.astype(np.uint32)


def random_graph(n, m, seed, symmetric=True):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, n, m)
    d = rng.integers(0, n, m)
    if symmetric:
        s, d = np.concatenate([s, d]), np.concatenate([d, s])
    return csr_from_edges(n, s, d)
