"""Triangle counting on the device: the degree-oriented LDS-hash path (used on
symmetric graphs) against closed forms and the CPU oracle, across its launch
tiers (|out(u)| <= 64, 128, 256, 512, and > 512 with 4096-id hash tiles), with
self-loops, and under relabelling.  Counts are bit-exact."""
from math import comb

import numpy as np
import pytest

import sstgraph_amd as S
from oracle import oracle as O
from tests.gpu_common import need_gpu, say
from tests.graphs import csr_from_edges, random_graph

pytestmark = pytest.mark.gpu
EMV = S.EdgeMapVertexMap


@pytest.fixture(autouse=True)
def _gpu():
    need_gpu()


def clique_edges(k, base=0):
    a, b = np.triu_indices(k, 1)
    a = a.astype(np.uint32) + base
    b = b.astype(np.uint32) + base
    return np.concatenate([a, b]), np.concatenate([b, a])


def _graph(n, s, d):
    g = S.SparseMatrixV(n)
    g.insert_batch(s, d)
    return g


@pytest.mark.parametrize("k", [3, 4, 40, 65, 129, 200, 257, 300, 513, 1200])
def test_clique_closed_form(k):
    """K_k has C(k,3) triangles; equal degrees make the orientation follow
    ids, so |out(u)| = k-1-u covers every tier boundary."""
    say("clique", k)
    s, d = clique_edges(k)
    g = _graph(k, s, d)
    assert EMV.TC(g) == comb(k, 3)


def test_clique_two_hash_tiles():
    """|out(u)| > 4096 for the first vertices of K_4500: out(u) is hashed in
    two tiles, each seeing every list out(v)."""
    k = 4500
    s, d = clique_edges(k)
    g = _graph(k, s, d)
    assert EMV.TC(g) == comb(k, 3)


def test_mixed_hubs_and_clique_vs_oracle():
    rng = np.random.default_rng(5)
    n = 30000
    parts = [clique_edges(700, base=1000)]
    hub = np.uint32(7)
    leaves = rng.choice(n, 5000, replace=False).astype(np.uint32)
    leaves = leaves[leaves != hub]
    parts.append((np.full(len(leaves), hub, np.uint32), leaves))
    s = rng.integers(0, n, 200000).astype(np.uint32)
    d = rng.integers(0, n, 200000).astype(np.uint32)
    parts.append((s, d))
    S_ = np.concatenate([p[0] for p in parts])
    D_ = np.concatenate([p[1] for p in parts])
    S2, D2 = np.concatenate([S_, D_]), np.concatenate([D_, S_])
    off, cols = csr_from_edges(n, S2, D2)
    g = _graph(n, S2, D2)
    want = O.tc(off, cols)
    say("mixed triangles", want)
    assert EMV.TC(g) == want


def test_self_loops_do_not_count():
    """insert() keeps self-loops (SparseMatrix.hpp:536-554); they never close
    a triangle, and the graph stays symmetric."""
    n = 2000
    off, cols = random_graph(n, 40000, 11, symmetric=True)
    rows = np.repeat(np.arange(n, dtype=np.uint32), np.diff(off).astype(np.int64))
    g = _graph(n, rows, cols)
    og = O.Graph(n)
    og.load_csr(off, cols)
    for v in (0, 5, 17, 1999):
        g.insert(v, v)
        og.insert(v, v)
    o_off, o_cols = og.csr()
    assert EMV.TC(g) == O.tc(o_off, o_cols)


def test_relabelling_invariant_at_scale():
    """The triangle count of an rMat scale-18 graph is unchanged by a random
    relabelling of its vertices, and equals the oracle's."""
    n, s, d = O.rmat_graph_edges(18, 16, 9)
    og = O.Graph(n)
    og.insert_batch(s, d)
    off, cols = og.csr()
    want = O.tc(off, cols)
    g = _graph(n, s, d)
    assert EMV.TC(g) == want
    perm = np.random.default_rng(3).permutation(n).astype(np.uint32)
    g2 = _graph(n, perm[s], perm[d])
    assert EMV.TC(g2) == want
    say("scale 18 triangles", want)


def test_directed_graph_keeps_per_edge_count():
    """Rows without their reverses: the count follows the reference's
    per-edge rule, not the triangle count."""
    n = 3000
    off, cols = random_graph(n, 60000, 21, symmetric=False)
    rows = np.repeat(np.arange(n, dtype=np.uint32), np.diff(off).astype(np.int64))
    g = _graph(n, rows, cols)
    assert EMV.TC(g) == O.tc(off, cols)


def test_directed_dense_rows_block_tier():
    """Directed rows with more than 512 ids below the row: the per-edge rule
    runs on the store's own rows through the block tier."""
    rng = np.random.default_rng(23)
    n = 1500
    keep = rng.random((n, n)) < 0.7
    np.fill_diagonal(keep, False)
    s, d = np.nonzero(keep)
    s, d = s.astype(np.uint32), d.astype(np.uint32)
    off, cols = csr_from_edges(n, s, d)
    g = _graph(n, s, d)
    want = O.tc(off, cols)
    say("directed dense count", want)
    assert EMV.TC(g) == want

