"""Degree-order relabelling (degree_order_map / rewrite_graph), statistics
and the graph-file round trip on the device.  The map must equal the
reference's definition (descending (degree, id) pairs, SparseMatrix.hpp:
737-749) computed in numpy; the rewritten graph must hold exactly the
relabelled edges and keep the triangle count, component structure and BFS
level histogram."""
import os

import numpy as np
import pytest

import sstgraph_amd as S
from oracle import oracle as O
from tests.gpu_common import need_gpu, say
from tests.graphs import GOLDEN, csr_from_edges, random_graph

pytestmark = pytest.mark.gpu
EMV = S.EdgeMapVertexMap


@pytest.fixture(autouse=True)
def _gpu():
    need_gpu()


def numpy_degree_order_map(deg):
    n = len(deg)
    order = sorted(range(n), key=lambda i: (int(deg[i]), i), reverse=True)
    m = np.empty(n, np.uint32)
    m[np.array(order, dtype=np.int64)] = np.arange(n, dtype=np.uint32)
    return m


def _graph_from_csr(n, off, cols):
    g = S.SparseMatrixV(n)
    rows = np.repeat(np.arange(n, dtype=np.uint32), np.diff(off).astype(np.int64))
    g.insert_batch(rows, cols)
    return g


@pytest.mark.parametrize("seed,n,m", [(1, 1, 0), (2, 50, 40), (3, 4000, 30000)])
def test_degree_order_map_matches_definition(seed, n, m):
    off, cols = random_graph(n, m, seed)
    g = _graph_from_csr(n, off, cols)
    got = g.degree_order_map()
    assert np.array_equal(got, numpy_degree_order_map(np.diff(off)))


def test_rewrite_graph_relabels_every_edge_and_keeps_structure():
    n, s, d = O.rmat_graph_edges(14, 8, 5)
    og = O.Graph(n)
    og.insert_batch(s, d)
    off, cols = og.csr()
    g = S.SparseMatrixV(n)
    g.insert_batch(s, d)
    g.insert(3, 3)  # a self-loop survives the rewrite, as insert() keeps it
    mp = g.degree_order_map()
    assert np.array_equal(mp, numpy_degree_order_map(g.degrees()))
    h = g.rewrite_graph(mp)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(off).astype(np.int64))
    w_off, w_cols = csr_from_edges(n, np.concatenate([mp[rows], [mp[3]]]),
                                   np.concatenate([mp[cols], [mp[3]]]),
                                   drop_self=False)
    h_off, h_cols = h.export_csr()
    assert np.array_equal(h_off, w_off) and np.array_equal(h_cols, w_cols)
    assert EMV.TC(h) == EMV.TC(g) == O.tc(off, cols)
    # components map onto components; BFS level histograms agree
    lab_g, lab_h = EMV.CC(g), EMV.CC(h)
    assert len(np.unique(lab_g)) == len(np.unique(lab_h))
    assert np.array_equal(lab_h[mp] == lab_h[mp][0], lab_g == lab_g[0])
    src = int(np.argmax(np.diff(off)))
    _, lg = EMV.BFS(g, src, levels=True)
    _, lh = EMV.BFS(h, int(mp[src]), levels=True)
    assert np.array_equal(np.sort(lg), np.sort(lh))
    assert np.array_equal(lh[mp], lg)
    # the hub is label 0 after the rewrite
    assert int(np.argmax(h.degrees())) == 0


def test_rewrite_graph_refuses_out_of_range_labels():
    g = S.SparseMatrixV(10)
    g.insert_batch(np.array([1, 2], np.uint32), np.array([2, 1], np.uint32))
    mp = np.arange(10, dtype=np.uint32)
    mp[2] = 10
    with pytest.raises(S.IdOutOfRange):
        g.rewrite_graph(mp)


def test_rewrite_graph_refuses_rectangular_graphs():
    """One map relabels rows and columns, so a graph with more columns than
    rows is refused instead of reading past the map."""
    g = S.SparseMatrixV(4, 100)
    g.insert_batch(np.array([1, 2], np.uint32), np.array([90, 3], np.uint32))
    with pytest.raises(S.SSTGError) as e:
        g.rewrite_graph(np.arange(4, dtype=np.uint32))
    assert e.value.code == S.SSTG_E_INVAL


def test_statistics_match_numpy():
    n = 3000
    off, cols = random_graph(n, 20000, 9)
    g = _graph_from_csr(n, off, cols)
    st = g.statistics()
    deg = np.diff(off).astype(np.int64)
    assert st["edges"] == len(cols) and st["rows"] == n and st["cols"] == n
    assert st["empty_rows"] == int((deg == 0).sum())
    assert st["max_degree"] == int(deg.max())
    bits = np.array([int(c).bit_length() for c in cols])
    assert st["dest_bits"] == np.bincount(bits, minlength=33).tolist()
    assert st["memory_bytes"] == g.get_memory_size()
    assert st["slot_capacity"] >= len(cols)
    assert st["pool_used"] <= st["pool_capacity"]


def test_file_round_trip(tmp_path):
    g = S.SparseMatrixV.from_file(os.path.join(GOLDEN, "johnson8-2-4.mtx"))
    assert g.n_rows == 28 and g.M() == 420 and EMV.TC(g) == 420
    p = str(tmp_path / "j.adj")
    g.write_adj(p)
    h = S.SparseMatrixV.from_file(p)
    a, b = g.export_csr(), h.export_csr()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    say("round trip", h.M())
