"""Degree thresholds and their autotuner.  The thresholds only choose how
many lanes work on a vertex, so under any valid setting BFS levels, CC
labels and the triangle count stay bit-exact against the oracle and FP64
PageRank within 1e-9 relative; the tuner must leave a valid setting that is
no slower than where it started, and the batched store must re-capture its
hipGraph with the new bins."""
import numpy as np
import pytest

import sstgraph_amd as S
from oracle import oracle as O
from tests.gpu_common import need_gpu, rel_close, say

pytestmark = pytest.mark.gpu
EMV = S.EdgeMapVertexMap


@pytest.fixture(autouse=True)
def _gpu():
    need_gpu()


@pytest.fixture(scope="module")
def g14():
    need_gpu()
    n = 1 << 14
    s, d = O.rmat(n, 3, 16 * n)
    s, d = np.concatenate([s, d]), np.concatenate([d, s])
    g = S.SparseMatrixV(n)
    g.insert_batch(s, d)
    off, cols = g.export_csr()
    return g, off, cols


def _check_all(g, off, cols, what):
    deg = np.diff(off)
    hub = int(np.argmax(deg))
    for src in (0, hub):
        _, lev = EMV.BFS(g, src, levels=True)
        olev, _ = O.bfs(off, cols, src)
        assert np.array_equal(lev, olev), f"{what}: BFS from {src}"
    assert np.array_equal(EMV.CC(g), O.cc(off, cols)), f"{what}: CC"
    assert EMV.TC(g) == O.tc(off, cols), f"{what}: TC"
    ok, err = rel_close(EMV.PR_S(g, 10), O.pagerank(off, cols, 10), 1e-9)
    assert ok, f"{what}: PR rel err {err}"


# the library's defaults (the tuner's picks on rMat at scales 22 and 24)
PULL_DEFAULT = [8, 64, 256]
BFS_DEFAULT = [32, 128, 2048]


def test_threshold_settings_keep_results_exact(g14):
    g, off, cols = g14
    assert g.degree_thresholds() == PULL_DEFAULT
    assert g.degree_thresholds(bfs=True) == BFS_DEFAULT
    sets = ([1, 16, 256], [32, 512, 65536], [4, 32, 1024], [8, 64, 2048])
    for t, tb in zip(sets, sets[::-1]):
        say("thresholds", t, "bfs", tb)
        g.set_degree_thresholds(t)
        g.set_degree_thresholds(tb, bfs=True)
        assert g.degree_thresholds() == t
        assert g.degree_thresholds(bfs=True) == tb
        _check_all(g, off, cols, f"{t} / bfs {tb}")
    g.set_degree_thresholds(PULL_DEFAULT)
    g.set_degree_thresholds(BFS_DEFAULT, bfs=True)


def test_bfs_switch_settings_keep_levels_exact(g14):
    g, off, cols = g14
    assert g.bfs_switch() == (15.0, 18.0)
    hub = int(np.argmax(np.diff(off)))
    olev, _ = O.bfs(off, cols, hub)
    for a, b in ((1e-3, 1e-3), (1e6, 1e6), (4, 6), (60, 36), (15, 18)):
        g.set_bfs_switch(a, b)
        assert g.bfs_switch() == (a, b)
        parent, lev = EMV.BFS(g, hub, levels=True)
        assert np.array_equal(lev, olev), (a, b)
        # parents are one level up and adjacent
        reached = np.nonzero(lev != S.UNREACHED)[0]
        for v in reached[:: max(1, len(reached) // 500)]:
            p = int(parent[v])
            if v == hub:
                assert p == hub
                continue
            assert lev[p] + 1 == lev[v]
            assert v in set(cols[off[p]:off[p + 1]].tolist())
    for a, b in ((0, 18), (15, -1), (float("nan"), 18)):
        with pytest.raises(S.SSTGError):
            g.set_bfs_switch(a, b)
    assert g.bfs_switch() == (15.0, 18.0)


def test_invalid_thresholds_rejected(g14):
    g, _, _ = g14
    for t in ([8, 8, 2048], [64, 8, 2048], [1, 2, 2]):
        with pytest.raises(S.SSTGError):
            g.set_degree_thresholds(t)
        with pytest.raises(S.SSTGError):
            g.set_degree_thresholds(t, bfs=True)
    assert g.degree_thresholds() == PULL_DEFAULT
    assert g.degree_thresholds(bfs=True) == BFS_DEFAULT
    with pytest.raises(S.SSTGError):
        g.autotune(pr=False, bfs_source=None)
    with pytest.raises(S.IdOutOfRange):
        g.autotune(pr=True, bfs_source=g.n_rows)


def test_autotune_picks_a_valid_faster_setting(g14):
    g, off, cols = g14
    hub = int(np.argmax(np.diff(off)))
    rep = g.autotune(pr=True, bfs_source=hub)
    say("autotune", rep)
    t, tb = rep["thresholds"], rep["bfs_thresholds"]
    assert t[0] < t[1] < t[2] and tb[0] < tb[1] < tb[2]
    assert g.degree_thresholds() == t
    assert g.degree_thresholds(bfs=True) == tb
    assert rep["start_thresholds"] == PULL_DEFAULT
    assert rep["start_bfs_thresholds"] == BFS_DEFAULT
    assert (rep["start_alpha"], rep["start_beta"]) == (15.0, 18.0)
    assert g.bfs_switch() == (rep["alpha"], rep["beta"])
    assert rep["evaluations"] >= 20
    assert 0 < rep["pr_best_ms"] <= rep["pr_start_ms"]
    assert 0 < rep["bfs_best_ms"] <= rep["bfs_start_ms"]
    _check_all(g, off, cols, f"tuned {t} / bfs {tb}")
    # a PageRank-only tuning leaves the BFS settings alone
    rep2 = g.autotune(pr=True, bfs_source=None)
    assert rep2["bfs_thresholds"] == tb and rep2["bfs_start_ms"] == 0
    g.set_degree_thresholds(PULL_DEFAULT)
    g.set_degree_thresholds(BFS_DEFAULT, bfs=True)
    g.set_bfs_switch(15, 18)


def test_multigraph_recaptures_after_tuning():
    sizes = [1 << 10, 1 << 12, 1 << 9, 1 << 13]
    M = S.MultiGraph(sizes)
    plain = []
    for i, n in enumerate(sizes):
        s, d = O.rmat(n, 20 + i, 12 * n)
        s, d = np.concatenate([s, d]), np.concatenate([d, s])
        M.insert_batch(i, s, d)
        pg = S.SparseMatrixV(n)
        pg.insert_batch(s, d)
        plain.append(pg)
    sources = [int(np.argmax(pg.degrees())) for pg in plain]
    M.run(sources, pr_iters=10)
    before = M.results()
    rep = M.autotune()
    say("multi autotune", rep)
    M.run(sources, pr_iters=10)
    after = M.results()
    assert np.array_equal(before["level"], after["level"])
    assert np.array_equal(before["label"], after["label"])
    ok, err = rel_close(after["rank"], before["rank"], 1e-9)
    assert ok, err
    for i, pg in enumerate(plain):
        off, cols = pg.export_csr()
        olev, _ = O.bfs(off, cols, sources[i])
        assert np.array_equal(M.part(after["level"], i), olev)
