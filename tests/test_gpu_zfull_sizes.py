"""Parity at the benchmark configurations' full sizes.

- rMat scale 22 (seed 1, 16n, symmetrised): BFS from vertex 0 and from the
  largest hub, 20 FP64 PageRank iterations, CC, BC from 0 and from the hub,
  and TC against the oracle on the same graph (test.hpp:486-606 calls).
- rMat scale 24 (seed 2) built by batch insert, then the update loop of
  test.hpp:636-690: insert and remove of batches of 1e3 .. 1e7 edges (seeds
  100.., 20% existing edges), every row compared after every batch.
- The 512-graph batched workload (scales 8..18, seed 3): one captured replay
  of BFS, CC and 10 PageRank iterations, every graph against its own oracle.
- rMat scale 25 (seed 4) served: every query result of a served request
  stream equals the same query run alone on the same graph version.

BFS levels, CC labels, TC counts and rows are bit-exact; FP64 PageRank and
BC agree within 1e-9 relative; FP32 PageRank with Kahan sums within 1e-5 of
the FP64 oracle (the measured error is printed).
"""
import hashlib

import numpy as np
import pytest

import sstgraph_amd as S
from oracle import oracle as O
from sstgraph_amd import workloads as W
from tests.gpu_common import assert_same_csr, need_gpu, rel_close, say

pytestmark = pytest.mark.gpu
EMV = S.EdgeMapVertexMap
RTOL = 1e-9


def _sym_rmat(scale, seed, edge_factor=16):
    """Symmetrised rMat base-graph stream from the device generator (pinned
    against the reference generator by test_gpu_update)."""
    n = 1 << scale
    s, d = S.rmat(n, seed, edge_factor * n)
    return n, np.concatenate([s, d]), np.concatenate([d, s])


@pytest.fixture(scope="module")
def c2():
    need_gpu()
    n, s, d = _sym_rmat(22, 1)
    say("scale 22 build", len(s))
    g = S.SparseMatrixV(n)
    g.insert_batch(s, d)
    og = O.Graph(n)
    og.insert_batch(s, d)
    del s, d
    off, cols = og.csr()
    del og
    assert g.M() == len(cols)
    return n, g, off, cols


@pytest.mark.timeout(900)
def test_c2_scale22_algorithms_match_oracle(c2):
    n, g, off, cols = c2
    goff, gcols = g.export_csr()
    assert np.array_equal(goff, off) and np.array_equal(gcols, cols)
    del goff, gcols
    deg = np.diff(off)
    hub = int(np.argmax(deg))
    for src in (0, hub):
        say("scale 22 bfs", src)
        parent, level = EMV.BFS(g, src, levels=True)
        olev, _ = O.bfs(off, cols, src)
        assert np.array_equal(level, olev), f"BFS levels from {src}"
        reached = olev != S.UNREACHED
        assert np.array_equal(parent >= 0, reached) and parent[src] == src
        v = np.nonzero(reached)[0]
        v = v[v != src]
        assert np.all(level[parent[v]] + 1 == level[v])
    say("scale 22 pagerank x20")
    want = O.pagerank(off, cols, 20, 0.85)
    ok, worst = rel_close(EMV.PR_S(g, 20), want, RTOL)
    say("  FP64 max rel err", worst)
    assert ok, worst
    ok32, worst32 = rel_close(EMV.PR_F32K(g, 20), want, 1e-5)
    say("  FP32+Kahan max rel err vs FP64 oracle", worst32)
    assert ok32, worst32
    say("scale 22 cc")
    assert np.array_equal(EMV.CC(g), O.cc(off, cols))
    for src in (0, hub):
        say("scale 22 bc", src)
        ok, worst = rel_close(EMV.BC(g, src), O.bc(off, cols, src), RTOL, 1e-12)
        assert ok, (src, worst)
    say("scale 22 tc")
    assert EMV.TC(g) == O.tc(off, cols)


@pytest.mark.timeout(1200)
def test_c3_scale24_update_chain_bit_exact():
    need_gpu()
    n, s, d = _sym_rmat(24, 2)
    say("scale 24 build", len(s))
    g = S.SparseMatrixV(n)
    g.insert_batch(s, d)
    og = O.Graph(n)
    og.insert_batch(s, d)
    del s, d
    assert_same_csr(g, og, "scale 24 build")
    off, cols = og.csr()
    gen = lambda nn, sd, c: S.rmat(nn, sd, c)  # noqa: E731
    for k, size in enumerate(W.UPDATE_SIZES):
        bs, bd = W.update_batch(gen, n, off, cols, size, 100 + k)
        say("scale 24 insert", size)
        g.insert_batch(bs, bd)
        og.insert_batch(bs, bd)
        assert_same_csr(g, og, f"scale 24 insert {size}")
        say("scale 24 remove", size)
        g.remove_batch(bs, bd)
        og.remove_batch(bs, bd)
        assert_same_csr(g, og, f"scale 24 remove {size}")
    say("scale 24 edges after the chain", g.M())


@pytest.mark.timeout(1200)
def test_c4_512_graphs_one_replay_match_oracles():
    need_gpu()
    specs = W.c4_specs(512, 3)
    sizes = [1 << sp[1] for sp in specs]
    M = S.MultiGraph(sizes)
    gen_r = lambda nn, sd, c: S.rmat(nn, sd, c)  # noqa: E731
    gen_e = lambda nn, sd, c: S.er(nn, sd, c)  # noqa: E731
    sources, want = [], []
    m_total = 0
    for i, sp in enumerate(specs):
        n, s, d = W.c4_host_edges(sp, gen_r, gen_e)
        M.insert_batch(i, s, d)
        og = O.Graph(n)
        og.insert_batch(s, d)
        off, cols = og.csr()
        del og, s, d
        m_total += len(cols)
        src = int(np.argmax(np.diff(off)))
        sources.append(src)
        want.append((O.bfs(off, cols, src)[0], O.cc(off, cols),
                     O.pagerank(off, cols, 10, 0.85)))
        if i % 64 == 63:
            say("batched workload: oracle done for", i + 1, "graphs")
    say("batched workload", len(specs), "graphs", sum(sizes), "vertices",
        m_total, "directed edges")
    assert M.edges() == m_total
    worst32 = 0.0
    for f32 in (False, True):
        say("batched replay", "f32+kahan" if f32 else "f64")
        M.run(sources, pr_iters=10, f32=f32)
        r = M.results(f32=f32)
        for i, (lev, lab, pr) in enumerate(want):
            if not f32:
                assert np.array_equal(M.part(r["level"], i), lev), i
                assert np.array_equal(M.part(r["label"], i), lab), i
            ok, w = rel_close(M.part(r["rank"], i), pr, 1e-5 if f32 else RTOL)
            assert ok, (i, f32, w)
            if f32:
                worst32 = max(worst32, w)
    say("FP32+Kahan PageRank max rel err vs FP64 oracle over 512 graphs", worst32)
    M.close()


def _fingerprint(a):
    return hashlib.sha1(np.ascontiguousarray(a).tobytes()).hexdigest()[:12]


@pytest.mark.timeout(1200)
def test_c5_scale25_served_equals_alone():
    """A served stream (4 slots, coalesced updates, store snapshots, one
    heavy query at a time, the triangle count carried through updates) on
    rMat scale 25: each query's result equals the same query run alone, in
    arrival order, on a second copy of the graph that receives the same
    updates."""
    need_gpu()
    n = 1 << 25
    src, dst = S.rmat(n, 4, 16 * n, device=True)
    L = S.lib()
    sym_s = S.DeviceArray(2 * 16 * n, np.uint32)
    sym_d = S.DeviceArray(2 * 16 * n, np.uint32)
    half = 4 * 16 * n
    import ctypes
    for a, b in ((sym_s, src), (sym_d, dst)):
        S._check(L.sstg_memcpy_d2d(a.ptr, b.ptr, half), "d2d")
    for a, b in ((sym_s, dst), (sym_d, src)):
        S._check(L.sstg_memcpy_d2d(ctypes.c_void_p(a.ptr.value + half), b.ptr, half), "d2d")
    src.free()
    dst.free()
    say("scale 25 build", 2 * 16 * n)
    served_g = S.SparseMatrixV(n)
    served_g.insert_batch(sym_s, sym_d)
    alone_g = S.SparseMatrixV(n)
    alone_g.insert_batch(sym_s, sym_d)
    sym_s.free()
    sym_d.free()
    assert served_g.M() == alone_g.M()
    say("scale 25 edges", served_g.M())
    _, reqs = W.serve_trace(n, served_g.degrees(), 40, 1.0)
    reqs.append(("cc", None))
    batches = W.serve_batches(lambda nn, sd, c: S.rmat(nn, sd, c), n, reqs)
    srv = S.Server(served_g, slots=4, coalesce=True, snapshots=True, one_heavy=True,
                   maintain=True)
    tickets = []
    for kind, arg in reqs:
        if kind in ("insert", "remove"):
            bs, bd = batches[arg]
            tickets.append(srv.insert(bs, bd) if kind == "insert" else srv.remove(bs, bd))
        elif kind == "bfs":
            tickets.append(srv.bfs(arg))
        elif kind == "pr":
            tickets.append(srv.pagerank(arg))
        elif kind == "cc":
            tickets.append(srv.cc())
        else:
            tickets.append(srv.tc())
    served = []
    for (kind, arg), t in zip(reqs, tickets):
        info = srv.wait(t)
        assert info["status"] == 0, (kind, info)
        if kind in ("bfs", "cc"):
            served.append(srv.fetch(t, np.uint32))
        elif kind == "pr":
            served.append(srv.fetch(t, np.float64))
        elif kind == "tc":
            served.append(info["scalar"])
        else:
            served.append(None)
        srv.release(t)
    stats = srv.stats()
    srv.close()
    say("served", len(reqs), "requests;", stats)
    worst = 0.0
    for i, ((kind, arg), got) in enumerate(zip(reqs, served)):
        if kind == "insert":
            alone_g.insert_batch(*batches[arg])
        elif kind == "remove":
            alone_g.remove_batch(*batches[arg])
        elif kind == "bfs":
            _, lev = EMV.BFS(alone_g, arg, levels=True)
            assert np.array_equal(got, lev), (i, "bfs", arg)
            say(i, "bfs", arg, "levels", _fingerprint(lev))
        elif kind == "cc":
            lab = EMV.CC(alone_g)
            assert np.array_equal(got, lab), (i, "cc")
            say(i, "cc components", int(np.unique(lab).size))
        elif kind == "pr":
            pr = EMV.PR_S(alone_g, arg)
            ok, w = rel_close(got, pr, RTOL)
            worst = max(worst, w)
            assert ok, (i, "pr", w)
            say(i, "pr", "max rel diff", w)
        else:
            tc = EMV.TC(alone_g)
            assert got == tc, (i, "tc", got, tc)
            say(i, "tc", tc)
    assert served_g.M() == alone_g.M()
    say("PageRank served vs alone max rel diff", worst)
