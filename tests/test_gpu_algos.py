"""BFS, PageRank, BC, CC, TC and BF on the device vs the CPU oracle.
BFS levels, the CC labels and the triangle count are bit-exact; FP64 PR and
BC agree within 1e-9 relative (atol 1e-12 for exact zeros)."""
import json
import os

import numpy as np
import pytest

import sstgraph_amd as S
from oracle import oracle as O
from tests.gpu_common import need_gpu, rel_close, say
from tests.graphs import GOLDEN, csr_from_edges, random_graph, read_mtx

pytestmark = pytest.mark.gpu
EMV = S.EdgeMapVertexMap
RTOL = 1e-9


@pytest.fixture(autouse=True)
def _gpu():
    need_gpu()


def _device_graph(n, off, cols):
    g = S.SparseMatrixV(n)
    rows = np.repeat(np.arange(n, dtype=np.uint32), np.diff(off).astype(np.int64))
    g.insert_batch(rows, cols)
    return g


@pytest.fixture(scope="module")
def c1():
    need_gpu()
    n, s, d = O.rmat_graph_edges(16, 16, 0)
    og = O.Graph(n)
    og.insert_batch(s, d)
    off, cols = og.csr()
    g = S.SparseMatrixV(n)
    g.insert_batch(s, d)
    return n, g, off, cols


def _check_bfs(g, off, cols, src):
    parent, level = EMV.BFS(g, src, levels=True)
    olev, _ = O.bfs(off, cols, src)
    assert np.array_equal(level, olev), f"BFS levels from {src}"
    reached = olev != S.UNREACHED
    assert np.array_equal(parent >= 0, reached)
    assert parent[src] == src
    v = np.nonzero(reached)[0]
    v = v[v != src]
    p = parent[v]
    assert np.all(level[p] + 1 == level[v])
    # the tree edge runs parent -> v
    for x, px in list(zip(v.tolist(), p.tolist()))[:: max(1, len(v) // 2000)]:
        assert x in cols[off[px]:off[px + 1]]


def test_c1_bfs_levels_match(c1):
    n, g, off, cols = c1
    gold = json.load(open(os.path.join(GOLDEN, "graph_goldens.json")))
    gold = gold["rmat_s16_ef16_seed0"]
    for src in [0, 1, 12345, n - 1]:
        say("C1 bfs", src)
        _check_bfs(g, off, cols, src)
    import hashlib
    _, level = EMV.BFS(g, 0, levels=True)
    assert hashlib.sha256(level.tobytes()).hexdigest() == gold["sha_levels"]


def test_c1_cc_tc_match(c1):
    n, g, off, cols = c1
    gold = json.load(open(os.path.join(GOLDEN, "graph_goldens.json")))
    gold = gold["rmat_s16_ef16_seed0"]
    say("C1 cc")
    lab = EMV.CC(g)
    assert np.array_equal(lab, O.cc(off, cols))
    say("C1 tc")
    assert EMV.TC(g) == O.tc(off, cols) == gold["triangles"]


def test_c1_pagerank_within_1e9(c1):
    n, g, off, cols = c1
    for iters in [1, 10, 20]:
        say("C1 pr", iters)
        got = EMV.PR_S(g, iters, 0.85)
        want = O.pagerank(off, cols, iters, 0.85)
        ok, worst = rel_close(got, want, RTOL)
        assert ok, f"PR iters={iters} worst rel err {worst}"


def test_c1_bc_within_1e9(c1):
    n, g, off, cols = c1
    for src in [0, 777]:
        say("C1 bc", src)
        got = EMV.BC(g, src)
        want = O.bc(off, cols, src)
        ok, worst = rel_close(got, want, RTOL, atol_floor=1e-12)
        assert ok, f"BC src={src} worst rel err {worst}"


def test_c1_bf_and_touchall(c1):
    n, g, off, cols = c1
    dist = EMV.BF(g, 0)
    assert np.array_equal(dist, O.bf(off, cols, 0))
    assert EMV.TouchAll(g) == int(cols.astype(np.uint64).sum())


def test_common_neighbors(c1):
    n, g, off, cols = c1
    rng = np.random.default_rng(3)
    for u, v in rng.integers(0, n, (40, 2)).tolist() + [(n - 1, n - 2), (5, 5)]:
        for ee in (False, True):
            assert g.common_neighbors(u, v, ee) == \
                O.common_neighbors(off, cols, u, v, ee)


def test_examples_closed_forms():
    for name, tc, hist, bc0 in [("johnson8-2-4.mtx", 420, [1, 15, 12], 27.0),
                                ("star.mtx", 1, [1, 7, 12], 19.0)]:
        n, s, d = read_mtx(os.path.join(GOLDEN, name))
        g = S.SparseMatrixV(n)
        g.insert_batch(s, d)
        off, cols = csr_from_edges(n, s, d)
        assert EMV.TC(g) == tc
        _, lev = EMV.BFS(g, 0, levels=True)
        assert np.bincount(lev).tolist() == hist
        b = EMV.BC(g, 0)
        assert b[0] == pytest.approx(bc0, rel=1e-12)
        ok, _ = rel_close(b, O.bc(off, cols, 0), RTOL, 1e-12)
        assert ok
    n, s, d = read_mtx(os.path.join(GOLDEN, "johnson8-2-4.mtx"))
    g = S.SparseMatrixV(n)
    g.insert_batch(s, d)
    np.testing.assert_allclose(EMV.PR_S(g, 10), 1 / 28, rtol=1e-13)


@pytest.mark.parametrize("symmetric", [True, False])
def test_random_graphs_all_algorithms(symmetric):
    for seed, n, m in [(1, 300, 900), (2, 5000, 60000), (3, 2, 1), (4, 1, 0)]:
        say("random", seed, n, m, symmetric)
        off, cols = random_graph(n, m, seed, symmetric=symmetric)
        g = _device_graph(n, off, cols)
        _check_bfs(g, off, cols, 0)
        assert np.array_equal(EMV.CC(g), O.cc(off, cols))
        assert EMV.TC(g) == O.tc(off, cols)
        ok, w = rel_close(EMV.PR_S(g, 7, 0.85), O.pagerank(off, cols, 7, 0.85), RTOL)
        assert ok, w
        ok, w = rel_close(EMV.BC(g, 0), O.bc(off, cols, 0), RTOL, 1e-12)
        assert ok, w


def test_pagerank_f32_kahan_error_vs_f64(c1):
    n, g, off, cols = c1
    want = O.pagerank(off, cols, 10, 0.85)
    got = EMV.PR_F32K(g, 10, 0.85).astype(np.float64)
    rel = np.abs(got - want) / want
    say("PR f32+Kahan max rel err", rel.max())
    # FP32 ranks carry ~2^-24 relative rounding per operation
    assert rel.max() < 1e-5


def test_symmetry_tracking_through_updates():
    """Bottom-up steps are only valid on symmetric graphs; the library tracks
    symmetry through small batches.  BFS/BC must stay exact as the graph goes
    from symmetric to directed and back."""
    rng = np.random.default_rng(17)
    n = 20000
    off, cols = random_graph(n, 300000, 19, symmetric=True)
    g = _device_graph(n, off, cols)
    og = O.Graph(n)
    og.load_csr(off, cols)

    def check(what):
        say("symmetry", what)
        o_off, o_cols = og.csr()
        _check_bfs(g, o_off, o_cols, 1)
        ok, w = rel_close(EMV.BC(g, 1), O.bc(o_off, o_cols, 1), RTOL, 1e-12)
        assert ok, (what, w)

    check("symmetric base")
    s = rng.integers(0, n, 500).astype(np.uint32)
    d = rng.integers(0, n, 500).astype(np.uint32)
    g.insert_batch(np.concatenate([s, d]), np.concatenate([d, s]))
    og.insert_batch(np.concatenate([s, d]), np.concatenate([d, s]))
    check("symmetric batch")
    g.insert_batch(s[:50], d[:50] ^ 1)   # one direction only
    og.insert_batch(s[:50], d[:50] ^ 1)
    check("directed batch")
    g.insert_batch(d[:50] ^ 1, s[:50])   # the reverses
    og.insert_batch(d[:50] ^ 1, s[:50])
    check("reverses added")
    g.remove_batch(np.concatenate([s[:100], d[:100]]), np.concatenate([d[:100], s[:100]]))
    og.remove_batch(np.concatenate([s[:100], d[:100]]), np.concatenate([d[:100], s[:100]]))
    check("symmetric removal")


def test_cc_many_components():
    """A large rMat component next to cliques, paths, stars, isolated
    vertices and self-loops: the sampled large component is skipped after
    the first links, every other vertex links all of its edges, and the
    labels (smallest id per component) equal the oracle's."""
    rng = np.random.default_rng(31)
    n = 1 << 16
    s, d = O.rmat(1 << 15, 77, 8 << 15)      # ids below 2^15
    parts_s, parts_d = [s], [d]
    base = 1 << 15
    for k in range(200):                     # small components above 2^15
        size = int(rng.integers(2, 40))
        ids = base + np.arange(size, dtype=np.uint32)
        base += size
        if k % 3 == 0:                       # clique
            a, b = np.triu_indices(size, 1)
            parts_s.append(ids[a]); parts_d.append(ids[b])
        elif k % 3 == 1:                     # path
            parts_s.append(ids[:-1]); parts_d.append(ids[1:])
        else:                                # star
            parts_s.append(np.full(size - 1, ids[0], np.uint32)); parts_d.append(ids[1:])
    S_ = np.concatenate(parts_s).astype(np.uint32)
    D_ = np.concatenate(parts_d).astype(np.uint32)
    S2, D2 = np.concatenate([S_, D_]), np.concatenate([D_, S_])
    off, cols = csr_from_edges(n, S2, D2)
    g = _device_graph(n, off, cols)
    g.insert(n - 1, n - 1)                   # a self-loop on an isolated vertex
    og = O.Graph(n)
    og.load_csr(off, cols)
    og.insert(n - 1, n - 1)
    o_off, o_cols = og.csr()
    assert np.array_equal(EMV.CC(g), O.cc(o_off, o_cols))



@pytest.mark.parametrize("bits", [8, 10, 12, 14])
def test_pagerank_column_segments_match_oracle(c1, bits):
    """FP64 PageRank over column segments of 2^bits vertices equals the
    oracle within 1e-9 and the vertex pull within FP64 reordering, for 1, 10
    and 20 iterations; the layout cached for one graph version is rebuilt
    after an update."""
    n, g0, off, cols = c1
    s = np.repeat(np.arange(n, dtype=np.uint32), np.diff(off).astype(np.int64))
    g = S.SparseMatrixV(n)
    g.insert_batch(s, cols)
    g.set_pagerank_segments(0)
    plain = {it: EMV.PR_S(g, it) for it in (1, 10, 20)}
    g.set_pagerank_segments(bits)
    assert g.pagerank_segments() == bits
    for it in (1, 10, 20, 10):  # the repeat reuses the cached layout
        say("segments", bits, "iters", it)
        got = EMV.PR_S(g, it)
        ok, w = rel_close(got, O.pagerank(off, cols, it, 0.85), RTOL)
        assert ok, (bits, it, w)
        ok, w = rel_close(got, plain[it], 1e-12)
        assert ok, (bits, it, w)
    # an update changes the graph version: the next call rebuilds
    rng = np.random.default_rng(bits)
    bs = rng.integers(0, n, 5000).astype(np.uint32)
    bd = rng.integers(0, n, 5000).astype(np.uint32)
    g.insert_batch(bs, bd)
    og = O.Graph(n)
    og.load_csr(off, cols)
    og.insert_batch(bs, bd)
    o_off, o_cols = og.csr()
    ok, w = rel_close(EMV.PR_S(g, 10), O.pagerank(o_off, o_cols, 10, 0.85), RTOL)
    assert ok, w


def test_pagerank_segments_long_runs_and_directed_rows():
    """Runs of every length class (a hub whose ids fill one segment, mid
    rows, single ids), empty rows, rows without their reverses and a
    self-loop, on segments of 2^8 vertices, against the oracle."""
    n = 1 << 12
    rng = np.random.default_rng(23)
    s = [np.full(3000, 7, np.int64), np.full(200, 9, np.int64),
         rng.integers(0, n, 20000)]
    d = [np.arange(3000) % 256 + 512, rng.integers(1024, 1300, 200),
         rng.integers(0, n, 20000)]
    S_ = np.concatenate(s).astype(np.uint32)
    D_ = np.concatenate(d).astype(np.uint32)
    g = S.SparseMatrixV(n)
    g.insert_batch(S_, D_)           # directed: no reverses added
    g.insert(11, 11)                  # a self-loop (single insert keeps it)
    og = O.Graph(n)
    og.insert_batch(S_, D_)
    og.insert(11, 11)
    off, cols = og.csr()
    assert (np.diff(off) == 0).any()
    g.set_pagerank_segments(8)
    for it in (1, 5, 20):
        ok, w = rel_close(EMV.PR_S(g, it), O.pagerank(off, cols, it, 0.85), RTOL)
        assert ok, (it, w)


@pytest.fixture
def graph_capture():
    """Recorded PageRank iteration pairs are opt-in (SSTG_GRAPH_CAPTURE=1,
    read by the library at each call)."""
    old = os.environ.get("SSTG_GRAPH_CAPTURE")
    os.environ["SSTG_GRAPH_CAPTURE"] = "1"
    yield
    if old is None:
        del os.environ["SSTG_GRAPH_CAPTURE"]
    else:
        os.environ["SSTG_GRAPH_CAPTURE"] = old


def test_pagerank_replayed_pairs_stay_exact(c1, graph_capture):
    """PageRank iterations run as replays of one recorded pair: calls on an
    unchanged graph repeat bit for bit, with other device work (BFS, CC)
    and synchronisations between them, and a call after an update or a
    threshold change (which records again) still matches the oracle; odd
    iteration counts finish with one plain iteration."""
    n, g0, off, cols = c1
    s = np.repeat(np.arange(n, dtype=np.uint32), np.diff(off).astype(np.int64))
    g = S.SparseMatrixV(n)
    g.insert_batch(s, cols)
    first = EMV.PR_S(g, 10)
    ok, w = rel_close(first, O.pagerank(off, cols, 10, 0.85), RTOL)
    assert ok, w
    for k in range(4):
        EMV.BFS(g, k, levels=True)
        EMV.CC(g)
        S.synchronize()
        assert np.array_equal(EMV.PR_S(g, 10), first), k
    for it in (2, 3, 9):
        ok, w = rel_close(EMV.PR_S(g, it), O.pagerank(off, cols, it, 0.85), RTOL)
        assert ok, (it, w)
    g.set_degree_thresholds([4, 32, 1024])
    ok, w = rel_close(EMV.PR_S(g, 10), O.pagerank(off, cols, 10, 0.85), RTOL)
    assert ok, w
    rng = np.random.default_rng(5)
    bs = rng.integers(0, n, 3000).astype(np.uint32)
    bd = rng.integers(0, n, 3000).astype(np.uint32)
    g.insert_batch(bs, bd)
    og = O.Graph(n)
    og.load_csr(off, cols)
    og.insert_batch(bs, bd)
    o_off, o_cols = og.csr()
    for it in (10, 11):
        ok, w = rel_close(EMV.PR_S(g, it), O.pagerank(o_off, o_cols, it, 0.85), RTOL)
        assert ok, (it, w)
    ok, w = rel_close(EMV.PR_F32K(g, 10).astype(np.float64),
                      O.pagerank(o_off, o_cols, 10, 0.85), 1e-5)
    assert ok, w


def test_bfs_replayed_runs_match_oracle(c1, graph_capture):
    """With recording on, the first BFS from a source records its launch
    sequence and later calls with the same source, outputs and graph
    version replay it.  Every call, replayed or not, gives the oracle's
    levels and a valid parent tree: repeated calls with other device work
    between them, a change of source (records again), a threshold or
    direction-switch change, and an update (a new graph version)."""
    n, _, off, cols = c1
    s = np.repeat(np.arange(n, dtype=np.uint32), np.diff(off).astype(np.int64))
    g = S.SparseMatrixV(n)
    g.insert_batch(s, cols)
    hub = int(np.argmax(np.diff(off)))
    for src in (hub, hub, 0, 12345, 12345, hub):
        say("replayed bfs", src)
        _check_bfs(g, off, cols, src)
    for k in range(3):
        EMV.PR_S(g, 2)
        EMV.CC(g)
        S.synchronize()
        _check_bfs(g, off, cols, hub)
    g.set_bfs_switch(4.0, 40.0)
    _check_bfs(g, off, cols, hub)
    _check_bfs(g, off, cols, hub)
    g.set_bfs_switch(15.0, 18.0)
    g.set_degree_thresholds([8, 64, 512], bfs=True)
    _check_bfs(g, off, cols, hub)
    _check_bfs(g, off, cols, hub)
    rng = np.random.default_rng(9)
    bs = rng.integers(0, n, 4000).astype(np.uint32)
    bd = rng.integers(0, n, 4000).astype(np.uint32)
    g.insert_batch(np.concatenate([bs, bd]), np.concatenate([bd, bs]))
    og = O.Graph(n)
    og.load_csr(off, cols)
    og.insert_batch(np.concatenate([bs, bd]), np.concatenate([bd, bs]))
    o_off, o_cols = og.csr()
    for src in (hub, hub, 7, 7):
        _check_bfs(g, o_off, o_cols, src)


def test_bc_replayed_runs_match_oracle(c1, graph_capture):
    """With recording on, BC from a source records the whole launch sequence
    (its BFS, the grouping by level and the sigma and delta pulls) and later
    calls with the same source, output and graph version replay it: every
    call agrees with the oracle within 1e-9, across repeated calls with
    other device work between them, source changes, a threshold change and
    an update."""
    n, _, off, cols = c1
    s = np.repeat(np.arange(n, dtype=np.uint32), np.diff(off).astype(np.int64))
    g = S.SparseMatrixV(n)
    g.insert_batch(s, cols)
    hub = int(np.argmax(np.diff(off)))
    want = {v: O.bc(off, cols, v) for v in (hub, 12345)}
    for src in (hub, hub, 12345, 12345, hub):
        say("replayed bc", src)
        ok, w = rel_close(EMV.BC(g, src), want[src], RTOL, atol_floor=1e-12)
        assert ok, (src, w)
    for k in range(2):
        EMV.BFS(g, k, levels=True)
        EMV.PR_S(g, 2)
        S.synchronize()
        ok, w = rel_close(EMV.BC(g, hub), want[hub], RTOL, atol_floor=1e-12)
        assert ok, (k, w)
    g.set_degree_thresholds([4, 32, 1024])
    for _ in range(2):
        ok, w = rel_close(EMV.BC(g, hub), want[hub], RTOL, atol_floor=1e-12)
        assert ok, w
    rng = np.random.default_rng(11)
    bs = rng.integers(0, n, 4000).astype(np.uint32)
    bd = rng.integers(0, n, 4000).astype(np.uint32)
    g.insert_batch(np.concatenate([bs, bd]), np.concatenate([bd, bs]))
    og = O.Graph(n)
    og.load_csr(off, cols)
    og.insert_batch(np.concatenate([bs, bd]), np.concatenate([bd, bs]))
    o_off, o_cols = og.csr()
    w2 = O.bc(o_off, o_cols, hub)
    for _ in range(2):
        ok, w = rel_close(EMV.BC(g, hub), w2, RTOL, atol_floor=1e-12)
        assert ok, w
