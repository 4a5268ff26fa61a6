"""Graph files on the host (sstg_edges_from_file): .mtx and .adj read as the
reference's get_edges_from_file reads them (helpers.hpp:363-631), malformed
files refused with an error instead of an exit.  No device needed."""
import os

import numpy as np
import pytest

import sstgraph_amd as S
from tests.graphs import GOLDEN, read_mtx


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


@pytest.mark.parametrize("name", ["johnson8-2-4.mtx", "star.mtx"])
def test_mtx_examples_match_reference_convention(name):
    n, s, d = S.edges_from_file(os.path.join(GOLDEN, name))
    rn, rs, rd = read_mtx(os.path.join(GOLDEN, name))
    assert n == rn
    # entry i yields (s-1, d-1) then (d-1, s-1)
    k = len(rs) // 2
    want_s = np.empty(2 * k, np.uint32)
    want_d = np.empty(2 * k, np.uint32)
    want_s[0::2], want_d[0::2] = rs[:k], rd[:k]
    want_s[1::2], want_d[1::2] = rd[:k], rs[:k]
    assert np.array_equal(s, want_s) and np.array_equal(d, want_d)


def test_mtx_general_is_symmetrised_and_weights(tmp_path):
    p = _write(tmp_path, "g.mtx",
               "%%MatrixMarket matrix coordinate integer general\n"
               "% a comment\n%another\n3 3 2\n1 2 7\n3 1 9\n")
    n, s, d, w = S.edges_from_file(p, weights=True)
    assert n == 3
    assert s.tolist() == [0, 1, 2, 0] and d.tolist() == [1, 0, 0, 2]
    assert w.tolist() == [7, 7, 9, 9]
    p = _write(tmp_path, "r.mtx",
               "%%MatrixMarket matrix coordinate real symmetric\n2 2 1\n2 1 3.75\n")
    n, s, d, w = S.edges_from_file(p, weights=True)
    assert (s.tolist(), d.tolist(), w.tolist()) == ([1, 0], [0, 1], [3, 3])
    p = _write(tmp_path, "p.mtx",
               "%%MatrixMarket matrix coordinate pattern symmetric\n2 2 1\n2 1\n")
    assert S.edges_from_file(p, weights=True)[3].tolist() == [1, 1]


def test_adj_unweighted_and_weighted(tmp_path):
    # rows: 0 -> {1, 2}, 1 -> {}, 2 -> {0}
    p = _write(tmp_path, "a.adj", "AdjacencyGraph\n3\n3\n0\n2\n2\n1\n2\n0\n")
    n, s, d = S.edges_from_file(p)
    assert n == 3 and s.tolist() == [0, 0, 2] and d.tolist() == [1, 2, 0]
    assert S.edges_from_file(p, weights=True)[3].tolist() == [1, 1, 1]
    p = _write(tmp_path, "w.adj.shuf",
               "WeightedAdjacencyGraph 3 3 0 2 2 1 2 0 5 6 7")
    n, s, d, w = S.edges_from_file(p, weights=True)
    assert s.tolist() == [0, 0, 2] and w.tolist() == [5, 6, 7]


def test_large_adj_parsed_in_parallel(tmp_path):
    rng = np.random.default_rng(1)
    n = 50000
    deg = rng.integers(0, 40, n)
    off = np.concatenate([[0], np.cumsum(deg)[:-1]])
    m = int(deg.sum())
    dst = rng.integers(0, n, m)
    text = "AdjacencyGraph\n%d\n%d\n" % (n, m) + "\n".join(map(str, off)) + "\n" + \
        "\n".join(map(str, dst)) + "\n"
    p = _write(tmp_path, "big.adj", text)
    rn, s, d = S.edges_from_file(p)
    assert rn == n
    assert np.array_equal(s, np.repeat(np.arange(n), deg).astype(np.uint32))
    assert np.array_equal(d, dst.astype(np.uint32))


@pytest.mark.parametrize("name,text", [
    ("bad.mtx", "%%MatrixMarket matrix array real general\n2 2 1\n1 1 1\n"),
    ("short.mtx", "%%MatrixMarket matrix coordinate pattern general\n2 2 2\n1 2\n"),
    ("zero.mtx", "%%MatrixMarket matrix coordinate pattern general\n2 2 1\n0 1\n"),
    ("rect.mtx", "%%MatrixMarket matrix coordinate pattern general\n2 3 1\n1 3\n"),
    ("hdr.adj", "Graph\n1\n1\n0\n0\n"),
    ("len.adj", "AdjacencyGraph\n2\n2\n0\n1\n1\n"),
    ("order.adj", "AdjacencyGraph\n2\n2\n2\n0\n1\n1\n"),
    ("x.txt", "AdjacencyGraph\n1\n1\n0\n0\n"),
])
def test_malformed_files_are_refused(tmp_path, name, text):
    p = _write(tmp_path, name, text)
    with pytest.raises(S.SSTGError) as e:
        S.edges_from_file(p)
    assert e.value.code == S.SSTG_E_INVAL


def test_missing_file_is_an_io_error(tmp_path):
    with pytest.raises(OSError):
        S.edges_from_file(str(tmp_path / "nope.adj"))
