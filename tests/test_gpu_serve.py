"""Serving loop (sstg_server_*): every query result in a mixed stream must be
the result the query would get alone on the graph as it stood at its arrival
(BFS levels, CC labels and TC exact; PR within 1e-9), whether requests are
coalesced and run concurrently (with or without store snapshots, with or
without one heavy query at a time) or served one by one."""
import numpy as np
import pytest

import sstgraph_amd as S
from oracle import oracle as O
from tests.gpu_common import need_gpu, rel_close, say

pytestmark = pytest.mark.gpu


def _stream(n, rng, count):
    reqs = []
    for i in range(count):
        r = rng.random()
        if r < 0.45:
            k = int(10 ** rng.uniform(1, 3.5))
            s, d = O.rmat(n, 500 + i, k)
            kind = "insert" if rng.random() < 0.7 else "remove"
            # keep the graph symmetric most of the time; every 7th update is
            # one-directional so bottom-up steps switch off and on
            if i % 7:
                s, d = np.concatenate([s, d]), np.concatenate([d, s])
            reqs.append((kind, s, d))
        elif r < 0.7:
            reqs.append(("bfs", int(rng.integers(0, n)), None))
        elif r < 0.82:
            reqs.append(("pr", 5, None))
        elif r < 0.92:
            reqs.append(("cc", None, None))
        else:
            reqs.append(("tc", None, None))
    return reqs


def _expected(n, base_s, base_d, reqs):
    og = O.Graph(n)
    og.insert_batch(base_s, base_d)
    exp = []
    for kind, a, b in reqs:
        if kind == "insert":
            og.insert_batch(a, b)
            exp.append(None)
        elif kind == "remove":
            og.remove_batch(a, b)
            exp.append(None)
        else:
            off, cols = og.csr()
            if kind == "bfs":
                exp.append(O.bfs(off, cols, a)[0])
            elif kind == "pr":
                exp.append(O.pagerank(off, cols, a, 0.85))
            elif kind == "cc":
                exp.append(O.cc(off, cols))
            else:
                exp.append(O.tc(off, cols))
    return exp


def _serve(n, base_s, base_d, reqs, slots, coalesce, snapshots=False,
           one_heavy=False, maintain=False):
    g = S.SparseMatrixV(n)
    g.insert_batch(base_s, base_d)
    srv = S.Server(g, slots=slots, coalesce=coalesce, snapshots=snapshots,
                   one_heavy=one_heavy, maintain=maintain)
    tickets = []
    for kind, a, b in reqs:  # submitted back to back: they queue and overlap
        if kind == "insert":
            tickets.append(srv.insert(a, b))
        elif kind == "remove":
            tickets.append(srv.remove(a, b))
        elif kind == "bfs":
            tickets.append(srv.bfs(a))
        elif kind == "pr":
            tickets.append(srv.pagerank(a))
        elif kind == "cc":
            tickets.append(srv.cc())
        else:
            tickets.append(srv.tc())
    out = []
    coalesced = 0
    for (kind, a, b), t in zip(reqs, tickets):
        info = srv.wait(t)
        assert info["status"] == 0, (kind, info)
        assert info["t_done"] >= info["t_start"] >= info["t_arrive"]
        coalesced = max(coalesced, info["coalesced"])
        if kind == "bfs":
            out.append(srv.fetch(t, np.uint32))
        elif kind == "pr":
            out.append(srv.fetch(t, np.float64))
        elif kind == "cc":
            out.append(srv.fetch(t, np.uint32))
        elif kind == "tc":
            out.append(info["scalar"])
        else:
            out.append(None)
        srv.release(t)
    stats = srv.stats()
    srv.close()
    return out, coalesced, stats


def _compare(reqs, got, exp):
    for i, ((kind, a, b), x, e) in enumerate(zip(reqs, got, exp)):
        if kind in ("bfs", "cc"):
            assert np.array_equal(x, e), (i, kind)
        elif kind == "pr":
            ok, w = rel_close(x, e, 1e-9)
            assert ok, (i, w)
        elif kind == "tc":
            assert x == e, (i, x, e)


def test_served_results_equal_alone_results():
    need_gpu()
    rng = np.random.default_rng(42)
    n, s, d = O.rmat_graph_edges(13, 8, 4)
    reqs = _stream(n, rng, 60)
    exp = _expected(n, s, d, reqs)
    say("serve concurrent+coalesced")
    got, coalesced, stats = _serve(n, s, d, reqs, slots=4, coalesce=True)
    _compare(reqs, got, exp)
    n_updates = sum(1 for r in reqs if r[0] in ("insert", "remove"))
    say("updates", n_updates, "device batches", stats["device_batches"],
        "max coalesced", coalesced)
    assert stats["device_batches"] <= n_updates
    say("serve concurrent+coalesced+snapshots")
    got2, _, stats2 = _serve(n, s, d, reqs, slots=4, coalesce=True, snapshots=True)
    _compare(reqs, got2, exp)
    say("store snapshots", stats2["snapshots"])
    assert stats2["snapshots"] >= 1
    say("serve one heavy at a time, lighter requests passing, snapshots")
    got3, _, stats3 = _serve(n, s, d, reqs, slots=4, coalesce=True,
                             snapshots=True, one_heavy=True)
    _compare(reqs, got3, exp)
    say("store snapshots", stats3["snapshots"])
    say("serve with the triangle count carried through updates")
    got4, _, _ = _serve(n, s, d, reqs, slots=4, coalesce=True, snapshots=True,
                        one_heavy=True, maintain=True)
    _compare(reqs, got4, exp)
    say("serve serialised")
    got1, coalesced1, stats1 = _serve(n, s, d, reqs, slots=1, coalesce=False)
    _compare(reqs, got1, exp)
    assert coalesced1 == 1 and stats1["device_batches"] == n_updates


def test_admission_by_hbm_budget():
    need_gpu()
    n, s, d = O.rmat_graph_edges(12, 8, 6)
    g = S.SparseMatrixV(n)
    g.insert_batch(s, d)
    # a budget below one PageRank's estimate: PR is refused at submit, BFS
    # (smaller estimate) is admitted one at a time
    srv = S.Server(g, slots=4, hbm_budget=n * 30, coalesce=True)
    with pytest.raises(S.SSTGError) as e:
        srv.pagerank(3)
    assert e.value.code == S.SSTG_E_NOMEM
    ts = [srv.bfs(v) for v in range(0, n, n // 8)]
    off, cols = O.Graph(n), None
    og = O.Graph(n)
    og.insert_batch(s, d)
    o_off, o_cols = og.csr()
    for v, t in zip(range(0, n, n // 8), ts):
        assert srv.wait(t)["status"] == 0
        assert np.array_equal(srv.fetch(t, np.uint32), O.bfs(o_off, o_cols, v)[0])
        srv.release(t)
    srv.close()


def test_triangle_count_carried_through_updates():
    """With the count maintained, TC queries between symmetric inserts and
    removes (some of them closing or breaking many triangles at once: a
    clique's edges inserted and removed in one batch) equal the oracle's
    count at their position; one-directional updates stop the carrying and a
    later TC counts in full again."""
    need_gpu()
    rng = np.random.default_rng(8)
    n, s, d = O.rmat_graph_edges(12, 8, 9)
    clique = np.arange(100, 140, dtype=np.uint32)
    a, b = np.triu_indices(len(clique), 1)
    cs, cd = clique[a], clique[b]
    reqs = [("tc", None, None)]
    for i in range(24):
        if i == 5:
            reqs.append(("insert", np.concatenate([cs, cd]), np.concatenate([cd, cs])))
        elif i == 11:
            reqs.append(("remove", np.concatenate([cs[::2], cd[::2]]),
                         np.concatenate([cd[::2], cs[::2]])))
        elif i == 17:  # one direction only: the graph turns directed
            reqs.append(("insert", rng.integers(0, n, 50).astype(np.uint32),
                         rng.integers(0, n, 50).astype(np.uint32)))
        else:
            k = int(rng.integers(10, 3000))
            es, ed = O.rmat(n, 900 + i, k)
            kind = "insert" if i % 3 else "remove"
            reqs.append((kind, np.concatenate([es, ed]), np.concatenate([ed, es])))
        reqs.append(("tc", None, None))
    exp = _expected(n, s, d, reqs)
    for one_heavy in (True, False):
        say("carried triangle count, one heavy" if one_heavy else "carried, concurrent")
        got, _, _ = _serve(n, s, d, reqs, slots=4, coalesce=True, snapshots=True,
                           one_heavy=one_heavy, maintain=True)
        _compare(reqs, got, exp)
