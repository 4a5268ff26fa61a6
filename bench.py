"""Benchmark of the SSTGraph hot path on one MI355X.

Workload (dynamic updates on RMAT scale 24): base graph rMat(n=2^24, seed 2,
a/b/c = 0.5/0.1/0.1) with 16n generated edges, symmetrised, built by batch
insert into the device tinyset store; then update batches of 1e3 .. 1e7
edges (80% fresh rMat edges with seeds 100.., 20% edges already present).
A step is the reference's update loop body (test.hpp:668-677): insert_batch
of the batch, then remove_batch of the same batch.  The timed batch is the
largest one (1e7 edges); the other sizes are reported under "extra".

value = batch edges / mean insert_batch time (the reference's throughput
definition, test.hpp:681-687: every batch edge counts), inputs resident in
HBM.  BFS (source 0) and 20 PageRank iterations on the same graph are
reported as GTEPS under "extra".

    python bench.py [--steps K] [--warmup W] [--scale 24] [--batch 1e7]
                    [--no-cpu] [--gpus 1]
    python bench.py --workload batched   # 512 graphs, one hipGraph replay
    python bench.py --workload serve     # served request stream, p50/p99
    python bench.py --workload algos     # BFS, PR, CC, BC, TC at scale 22
    python bench.py --workload pagerank  # (or bfs, bc) one algorithm, for profiles
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

from sstgraph_amd.workloads import (c4_specs, serve_batches, serve_trace,  # noqa: E402
                                    update_batch)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


RMAT_FLAGS = 0  # SSTG_RMAT_NORMALIZED with --rmat-normalized
RMAT_NOTE = ("hashDouble in [0,2): effective 0.25/0.05/0.05/0.65 (reference generator)",
             "normalised draws in [0,1): 0.5/0.1/0.1/0.3 (--rmat-normalized)")


def device_symmetric_rmat(S, n, seed, count):
    """2*count pairs on the device: rMat edges (u,v) followed by (v,u)."""
    L = S.lib()
    src = S.DeviceArray(2 * count, np.uint32)
    dst = S.DeviceArray(2 * count, np.uint32)
    S._check(L.sstg_rmat(n, seed, 0.5, 0.1, 0.1, 0, count, src.ptr, dst.ptr,
                         S.SSTG_DEVICE_PTRS | RMAT_FLAGS), "rmat")
    half = ctypes.c_void_p(src.ptr.value + 4 * count)
    halfd = ctypes.c_void_p(dst.ptr.value + 4 * count)
    S._check(L.sstg_memcpy_d2d(half, dst.ptr, 4 * count), "d2d")
    S._check(L.sstg_memcpy_d2d(halfd, src.ptr, 4 * count), "d2d")
    return src, dst


def make_batch(S, n, off, cols, size, seed):
    """80% fresh rMat(n, seed) edges + 20% edges present in the base graph
    (edge index hash64(seed, j) mod m), as device and host arrays."""
    bs, bd = update_batch(
        lambda nn, sd, c: S.rmat(nn, sd, c, normalized=RMAT_FLAGS != 0),
        n, off, cols, size, seed)
    return S.DeviceArray.from_host(bs), S.DeviceArray.from_host(bd), bs, bd


def timed(fn):
    t0 = time.perf_counter()
    fn()
    return time.perf_counter() - t0


def cpu_baseline(off, cols, n, batches, headline, budget_s=24.0):
    """The oracle's batch insert (per-row sorted merge, OpenMP) timed on
    this host on the same base graph and the same batches, each batch
    inserted and then removed again as in the reference's update loop
    (test.hpp:668-677): on all host threads and on one, at every batch size
    (median of up to 5 runs, within a time budget per size)."""
    from oracle import oracle as O  # CPU baseline leg only
    if not os.path.exists(O.LIB_PATH):
        O.build()
    threads = O.threads()
    og = O.Graph(n)
    og.load_csr(off, cols)
    per = {}
    for size, (bs, bd) in batches.items():
        rec = {}
        for t in (threads, 1):
            O.set_threads(t)
            times = []
            t_start = time.perf_counter()
            lim = budget_s / len(batches) / (2 if t == 1 else 1)
            while len(times) < 5 and (len(times) < 1 or time.perf_counter() - t_start < lim):
                times.append(timed(lambda: og.insert_batch(bs, bd)))
                og.remove_batch(bs, bd)
            med = float(np.median(times))
            rec["all_cores" if t == threads else "one_core"] = dict(
                inserts_per_s=size / med, insert_ms=1e3 * med, runs=len(times))
        per[str(size)] = rec
    O.set_threads(threads)
    h = per[str(headline)]
    return dict(value=h["all_cores"]["inserts_per_s"], unit="edges/s", cores=threads,
                kind="port",
                sample=f"oracle insert_batch of the {headline}-edge batch into the "
                       f"scale-24 base graph, median of {h['all_cores']['runs']} runs; "
                       f"every batch size and one core under per_batch",
                single_core=dict(value=h["one_core"]["inserts_per_s"], unit="edges/s",
                                 cores=1, runs=h["one_core"]["runs"]),
                per_batch=per, host=host_cpu())


def host_cpu():
    """The host's logical CPU count and model name (from /proc/cpuinfo)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(logical_cpus=os.cpu_count(), model=model)


def cpu_algos(off, cols, hub, iters=2):
    """The oracle's BFS and BC from the hub and FP64 PageRank iterations on
    the same base graph, on all host threads (OpenMP, level-synchronous BFS
    and Brandes) and on one thread."""
    from oracle import oracle as O  # CPU baseline leg only
    threads = O.threads()
    m = int(off[-1])
    out = {"threads": threads}
    for t in (threads, 1):
        O.set_threads(t)
        key = "all_cores" if t == threads else "one_core"
        # the base graph is symmetrised by construction
        t_bfs = timed(lambda: O.bfs_par(off, cols, hub, symmetric=True))
        t_bc = timed(lambda: O.bc_par(off, cols, hub, symmetric=True))
        t_pr = timed(lambda: O.pagerank(off, cols, iters))
        out[key] = dict(cores=t, bfs_hub_ms=1e3 * t_bfs, bfs_gteps=m / t_bfs / 1e9,
                        bc_hub_ms=1e3 * t_bc, pagerank_ms_per_iter=1e3 * t_pr / iters,
                        pagerank_gteps=m * iters / t_pr / 1e9)
    O.set_threads(threads)
    return out


def op_model_bytes(n, deg, hs, hd, skip_self=True):
    """Row-merge model of one insert batch: 8 B per batch pair + 4 B per old
    and per new element of every touched row + 16 B per touched row."""
    keys = np.unique(hs.astype(np.int64) * n + hd.astype(np.int64))
    ks = keys // n
    if skip_self:
        ks = ks[ks != keys % n]
    trow, tcount = np.unique(ks, return_counts=True)
    d_old = int(deg[trow].sum())
    # new rows hold at most the old ones plus the batch's distinct ids
    return 8 * len(hs) + 4 * (2 * d_old + int(tcount.sum())) + 16 * len(trow)


def c4_edges(S, spec):
    """Symmetrised edge stream of one batched-workload graph, on the device."""
    kind, scale, param, seed = spec
    n = 1 << scale
    count = param * n if kind == "rmat" else param * n // 2
    L = S.lib()
    src = S.DeviceArray(2 * count, np.uint32)
    dst = S.DeviceArray(2 * count, np.uint32)
    if kind == "rmat":
        S._check(L.sstg_rmat(n, seed, 0.5, 0.1, 0.1, 0, count, src.ptr, dst.ptr,
                             S.SSTG_DEVICE_PTRS | RMAT_FLAGS), "rmat")
    else:
        S._check(L.sstg_er(n, seed, 0, count, src.ptr, dst.ptr,
                           S.SSTG_DEVICE_PTRS), "er")
    S._check(L.sstg_memcpy_d2d(ctypes.c_void_p(src.ptr.value + 4 * count),
                               dst.ptr, 4 * count), "d2d")
    S._check(L.sstg_memcpy_d2d(ctypes.c_void_p(dst.ptr.value + 4 * count),
                               src.ptr, 4 * count), "d2d")
    return n, src, dst


def c4_update_batches(S, specs, frac=100, seed_shift=7919):
    """One fresh update batch per batched-workload graph, of uneven sizes
    (about 1/frac of the graph's generated edges, both directions), back to
    back in two device arrays; returns (graphs, counts, src, dst)."""
    L = S.lib()
    counts = []
    for kind, scale, param, seed in specs:
        n = 1 << scale
        counts.append(max(8, (param * n if kind == "rmat" else param * n // 2) // frac))
    total = 2 * sum(counts)
    src = S.DeviceArray(total, np.uint32)
    dst = S.DeviceArray(total, np.uint32)
    off = 0
    for (kind, scale, param, seed), c in zip(specs, counts):
        n = 1 << scale
        ps = ctypes.c_void_p(src.ptr.value + 4 * off)
        pd = ctypes.c_void_p(dst.ptr.value + 4 * off)
        if kind == "rmat":
            S._check(L.sstg_rmat(n, seed + seed_shift, 0.5, 0.1, 0.1, 0, c, ps, pd,
                                 S.SSTG_DEVICE_PTRS | RMAT_FLAGS), "rmat")
        else:
            S._check(L.sstg_er(n, seed + seed_shift, 0, c, ps, pd,
                               S.SSTG_DEVICE_PTRS), "er")
        S._check(L.sstg_memcpy_d2d(ctypes.c_void_p(ps.value + 4 * c), pd, 4 * c), "d2d")
        S._check(L.sstg_memcpy_d2d(ctypes.c_void_p(pd.value + 4 * c), ps, 4 * c), "d2d")
        off += 2 * c
    return list(range(len(specs))), [2 * c for c in counts], src, dst


def run_batched(args):
    """512 graphs: BFS (from each graph's largest hub), CC and 10 PageRank
    iterations on all of them from one hipGraph replay, against per-graph
    plain launches of the same three algorithms."""
    import sstgraph_amd as S
    specs = c4_specs(args.graphs)
    sizes = [1 << sp[1] for sp in specs]
    M = S.MultiGraph(sizes)
    plain = []
    m_total = 0
    t0 = time.perf_counter()
    for g, sp in enumerate(specs):
        n, src, dst = c4_edges(S, sp)
        M.insert_batch(g, src, dst)
        pg = S.SparseMatrixV(n)
        pg.insert_batch(src, dst)
        m_total += pg.M()
        plain.append(pg)
        src.free()
        dst.free()
    log(f"built {len(specs)} graphs ({sum(sizes)} vertices, {m_total} directed "
        f"edges) in {time.perf_counter() - t0:.1f}s")
    sources = [int(np.argmax(pg.degrees())) for pg in plain]
    iters = 10
    tune = None
    if not args.no_tune:
        # degree thresholds of the union store, tuned on PR iterations
        tune = M.autotune()
        log(f"autotuned thresholds {tune['thresholds']}: PR x2 "
            f"{tune['pr_start_ms']:.2f} -> {tune['pr_best_ms']:.2f} ms")
    # batched: one hipGraph replay per step
    for _ in range(args.warmup):
        M.run(sources, pr_iters=iters)
    S.synchronize()
    tb = [timed(lambda: M.run(sources, pr_iters=iters)) for _ in range(args.steps)]
    res64 = M.results()
    M.run(sources, pr_iters=iters, f32=True)
    res32 = M.results(f32=True)
    rel = np.abs(res32["rank"].astype(np.float64) - res64["rank"]) / res64["rank"]
    # plain: three calls per graph with device-resident outputs
    EMV = S.EdgeMapVertexMap
    outs = [(S.DeviceArray(pg.n_rows, np.int32), S.DeviceArray(pg.n_rows, np.uint32),
             S.DeviceArray(pg.n_rows, np.float64)) for pg in plain]

    def plain_pass():
        for pg, srcv, (dp, dl, dr) in zip(plain, sources, outs):
            EMV.BFS(pg, srcv, out=(None, dl))
            S._check(S.lib().sstg_cc(pg.handle, dp.ptr, S.SSTG_DEVICE_PTRS), "cc")
            EMV.PR_S(pg, iters, out=dr)
        S.synchronize()

    for _ in range(args.warmup):
        plain_pass()
    tp = [timed(plain_pass) for _ in range(max(1, args.steps // 2))]
    # same results alone and batched (spot check on a few graphs)
    for g in range(0, len(plain), max(1, len(plain) // 16)):
        _, dl, dr = outs[g]
        assert np.array_equal(M.part(res64["level"], g), dl.to_host())
        pr = dr.to_host()
        assert np.allclose(M.part(res64["rank"], g), pr, rtol=1e-9, atol=0)
    t_b, t_p = float(np.median(tb)), float(np.median(tp))
    # update batches of every graph: one insert_batches + one remove_batches
    # call over the union store, against one insert and one remove call per
    # graph on the same device data
    gids, counts, us, ud = c4_update_batches(S, specs)
    L = S.lib()
    m_union = M.edges()
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)

    def upd_batched():
        M.insert_batches_device(gids, counts, us, ud)
        S.synchronize()
        ti = time.perf_counter()
        M.remove_batches_device(gids, counts, us, ud)
        S.synchronize()
        return ti

    def upd_plain():
        for g in gids:
            ps = ctypes.c_void_p(us.ptr.value + 4 * int(offs[g]))
            pd = ctypes.c_void_p(ud.ptr.value + 4 * int(offs[g]))
            S._check(L.sstg_multi_insert_batch(M._h, g, ps, pd, counts[g],
                                               S.SSTG_DEVICE_PTRS), "insert")
        S.synchronize()
        ti = time.perf_counter()
        for g in gids:
            ps = ctypes.c_void_p(us.ptr.value + 4 * int(offs[g]))
            pd = ctypes.c_void_p(ud.ptr.value + 4 * int(offs[g]))
            gi = np.array([g], np.uint32)
            cnt = np.array([counts[g]], np.uint64)
            S._check(L.sstg_multi_remove_batches(M._h, 1, S._ptr(gi), S._ptr(cnt),
                                                 ps, pd, S.SSTG_DEVICE_PTRS), "remove")
        S.synchronize()
        return ti

    def timed_pair(fn):
        t0 = time.perf_counter()
        ti = fn()
        return ti - t0, time.perf_counter() - ti

    for fn in (upd_batched, upd_plain):
        for _ in range(args.warmup):
            fn()
    ub = [timed_pair(upd_batched) for _ in range(args.steps)]
    up = [timed_pair(upd_plain) for _ in range(max(1, args.steps // 2))]
    assert M.edges() <= m_union, "insert+remove of fresh batches keeps the union"
    us.free()
    ud.free()
    upd = dict(batch_edges=int(sum(counts)),
               batched_insert_ms=1e3 * float(np.median([x[0] for x in ub])),
               batched_remove_ms=1e3 * float(np.median([x[1] for x in ub])),
               plain_insert_ms=1e3 * float(np.median([x[0] for x in up])),
               plain_remove_ms=1e3 * float(np.median([x[1] for x in up])))
    upd["speedup_insert"] = upd["plain_insert_ms"] / upd["batched_insert_ms"]
    upd["speedup_remove"] = upd["plain_remove_ms"] / upd["batched_remove_ms"]
    log(f"update batches of all graphs ({upd['batch_edges']} edges): insert "
        f"{upd['batched_insert_ms']:.2f} ms in one call vs {upd['plain_insert_ms']:.2f} ms "
        f"per graph ({upd['speedup_insert']:.1f}x); remove {upd['batched_remove_ms']:.2f} "
        f"vs {upd['plain_remove_ms']:.2f} ms ({upd['speedup_remove']:.1f}x)")
    out = {
        "metric": "batched BFS+CC+PR(10) over 512 graphs: graph-algorithm passes/s",
        "value": 1.0 / t_b, "unit": "passes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * t_b, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 (PR), u32",
        "data": "synthetic (rMat + seeded Erdos-Renyi; " + RMAT_NOTE[RMAT_FLAGS != 0] + ")",
        "config": {"workload": f"{len(specs)} graphs, scale 8..18, half rMat (edge "
                               "factor 8..32), half ER (avg degree 4..64), seed 3",
                   "vertices": int(sum(sizes)), "directed_edges": int(m_total)},
        "extra": {"batched_ms": 1e3 * t_b, "plain_ms": 1e3 * t_p,
                  "speedup_vs_plain": t_p / t_b,
                  "edges_per_s_batched": m_total * (iters + 2) / t_b,
                  "bfs_levels": res64["levels_used"],
                  "autotune": tune,
                  "pr_f32_kahan_max_rel_err_vs_f64": float(rel.max()),
                  "pr_f32_kahan_mean_rel_err_vs_f64": float(rel.mean()),
                  "update_batches": upd},
    }
    log(f"batched {1e3 * t_b:.2f} ms vs plain {1e3 * t_p:.2f} ms per pass "
        f"({t_p / t_b:.1f}x); FP32+Kahan PR max rel err {rel.max():.2e}")
    print(json.dumps(out))


def serve_once(S, g, t_arr, reqs, batches, slots, coalesce, snapshots,
               one_heavy=False, maintain=False):
    """Submits the trace (arrival offsets t_arr, or all at once when None)
    and returns latency and throughput figures."""
    srv = S.Server(g, slots=slots, coalesce=coalesce, snapshots=snapshots,
                   one_heavy=one_heavy, maintain=maintain)
    tickets = []
    t0 = time.perf_counter()
    for k, (kind, arg) in enumerate(reqs):
        if t_arr is not None:
            delay = t0 + t_arr[k] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
        if kind in ("insert", "remove"):
            bs, bd = batches[arg]
            tickets.append(srv.insert(bs, bd) if kind == "insert" else srv.remove(bs, bd))
        elif kind == "bfs":
            tickets.append(srv.bfs(arg))
        elif kind == "pr":
            tickets.append(srv.pagerank(arg))
        else:
            tickets.append(srv.tc())
    infos = []
    for t in tickets:
        infos.append(srv.wait(t))
        srv.release(t)
    st = srv.stats()
    srv.close()
    assert all(i["status"] == 0 for i in infos), [i["status"] for i in infos]
    lat = np.array([i["t_done"] - i["t_arrive"] for i in infos])
    svc = np.array([i["t_done"] - i["t_start"] for i in infos])
    first = min(i["t_arrive"] for i in infos)
    span = max(i["t_done"] for i in infos) - first
    per_kind = {}
    for k in ("insert", "remove", "bfs", "pr", "tc"):
        sel = np.array([r[0] == k for r in reqs])
        if sel.any():
            per_kind[k] = dict(count=int(sel.sum()),
                               p50_ms=1e3 * float(np.percentile(lat[sel], 50)),
                               p99_ms=1e3 * float(np.percentile(lat[sel], 99)),
                               service_p50_ms=1e3 * float(np.percentile(svc[sel], 50)))
    return dict(p50_ms=1e3 * float(np.percentile(lat, 50)),
                p99_ms=1e3 * float(np.percentile(lat, 99)),
                mean_ms=1e3 * float(lat.mean()),
                throughput_rps=len(reqs) / span, makespan_s=span,
                device_update_batches=st["device_batches"],
                store_snapshots=st["snapshots"], per_kind=per_kind)


SERVE_MODES = (
    ("serialised", dict(slots=1, coalesce=False, snapshots=False)),
    ("concurrent", dict(slots=4, coalesce=True, snapshots=False)),
    ("snapshots", dict(slots=4, coalesce=True, snapshots=True)),
    ("served", dict(slots=4, coalesce=True, snapshots=True, one_heavy=True,
                    maintain=True)),
)


def serve_compare(S, scale, edge_factor, requests, load, modes):
    """The served-stream comparison on rMat scale `scale` (seed 4,
    symmetrised): each mode first takes the whole seeded trace at once
    (capacity, requests/s), then the trace again with Poisson arrivals at
    `load` times the first mode's capacity (latency from arrival to result,
    p50/p99).  Each mode starts from a freshly built graph.  Returns
    (results per mode, Poisson rate, directed edges, requests)."""
    n = 1 << scale
    count = edge_factor * n

    def fresh_graph():
        src, dst = device_symmetric_rmat(S, n, 4, count)
        g = S.SparseMatrixV(n)
        g.insert_batch(src, dst)
        src.free()
        dst.free()
        return g

    g = fresh_graph()
    m0 = g.M()
    deg = g.degrees()
    _, reqs = serve_trace(n, deg, requests, 1.0)
    # the served graph stays undirected: an update batch of k rMat edges is
    # applied as its 2k directed pairs
    batches = serve_batches(
        lambda nn, sd, c: S.rmat(nn, sd, c, normalized=RMAT_FLAGS != 0), n, reqs)
    log(f"serve: scale {scale}, m={m0}, {len(reqs)} requests")
    res, rate, t_arr = {}, None, None
    for name, mode in modes:
        if g is None:
            g = fresh_graph()
        burst = serve_once(S, g, None, reqs, batches, **mode)
        if rate is None:
            rate = load * burst["throughput_rps"]
            t_arr, _ = serve_trace(n, deg, requests, rate)
        g.close()
        g = fresh_graph()
        paced = serve_once(S, g, t_arr, reqs, batches, **mode)
        res[name] = dict(burst=burst, poisson=paced, **mode)
        log(f"{name}: capacity {burst['throughput_rps']:.1f} req/s; at {rate:.1f} req/s "
            f"p50 {paced['p50_ms']:.1f} ms p99 {paced['p99_ms']:.1f} ms; "
            f"{paced['device_update_batches']} update batches, "
            f"{paced['store_snapshots']} snapshots")
        g.close()
        g = None
    return res, rate, m0, reqs


def run_serve(args):
    """Served stream on a device-resident rMat graph (scale --serve-scale,
    seed 4, symmetrised).  Four servers see the same seeded trace: the
    serialised plain-launch loop (one request at a time, one device batch
    per update), concurrent slots with coalesced updates, the same with
    store snapshots (updates do not wait for queries), and the served mode
    that also runs one PageRank/TC at a time while lighter requests pass.
    Each first takes the whole trace at once (capacity, requests/s), then
    the trace again with Poisson arrivals at 70% of the serialised loop's
    capacity (latency from arrival to result, p50/p99)."""
    import sstgraph_amd as S
    scale = args.serve_scale
    res, rate, m0, reqs = serve_compare(S, scale, args.edge_factor, args.requests,
                                        args.load, SERVE_MODES)
    sv, se = res["served"], res["serialised"]
    out = {
        "metric": "served stream p99 latency (updates + BFS + PR + TC)",
        "value": sv["poisson"]["p99_ms"], "unit": "ms", "n_gpus": 1,
        "steps": len(reqs), "warmup": 0, "ms_per_step": sv["poisson"]["mean_ms"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32 / f64 (PR)", "data": "synthetic (rMat, seeded trace; " + RMAT_NOTE[RMAT_FLAGS != 0] + ")",
        "config": {"workload": f"served stream on rMat scale {scale} (seed 4, edge "
                               f"factor {args.edge_factor}, symmetrised), Poisson "
                               f"arrivals at {rate:.1f}/s ({args.load:.0%} of the "
                               f"serialised loop's capacity), {len(reqs)} requests",
                   "directed_edges": int(m0)},
        "extra": {"modes": res,
                  "p99_ratio_serialised_over_served":
                      se["poisson"]["p99_ms"] / sv["poisson"]["p99_ms"],
                  "capacity_ratio_served_over_serialised":
                      sv["burst"]["throughput_rps"] / se["burst"]["throughput_rps"]},
    }
    print(json.dumps(out))


def run_algo(args):
    """One algorithm alone on the scale-`--scale` rMat graph (seed 2), outputs
    kept in HBM: `pagerank` runs PR_S(g, --pr-iters) per step, `bfs` runs BFS
    from the largest hub per step.  Used for per-kernel profiles."""
    import sstgraph_amd as S
    EMV = S.EdgeMapVertexMap
    n = 1 << args.scale
    count = args.edge_factor * n
    src, dst = device_symmetric_rmat(S, n, 2, count)
    g = S.SparseMatrixV(n)
    g.insert_batch(src, dst)
    src.free()
    dst.free()
    m = g.M()
    deg = g.degrees().astype(np.int64)
    extra = {}
    if args.workload == "pagerank":
        out = S.DeviceArray(n, np.float64)
        if args.pr_segments is not None:
            g.set_pagerank_segments(args.pr_segments)
        step = lambda: EMV.PR_S(g, args.pr_iters, out=out)
        work = m * args.pr_iters
        name = f"PageRank FP64 x{args.pr_iters}"
        # the first call on a graph version builds the column-segment layout
        S.synchronize()
        extra["first_call_ms"] = 1e3 * timed(lambda: (step(), S.synchronize()))
        extra["segment_bits"] = g.pagerank_segments()
    elif args.workload == "bc":
        hub = int(np.argmax(deg))
        d_bc = S.DeviceArray(n, np.float64)
        L = S.lib()
        step = lambda: S._check(L.sstg_bc(g.handle, hub, d_bc.ptr, S.SSTG_DEVICE_PTRS), "bc")
        d_lev = S.DeviceArray(n, np.uint32)
        EMV.BFS(g, hub, out=(None, d_lev))
        lev = d_lev.to_host()
        work = int(deg[lev != S.UNREACHED].sum())
        name = f"BC from vertex {hub}"
    else:
        hub = int(np.argmax(deg))
        d_par = S.DeviceArray(n, np.int32)
        d_lev = S.DeviceArray(n, np.uint32)
        step = lambda: EMV.BFS(g, hub, out=(d_par, d_lev))
        step()
        lev = d_lev.to_host()
        work = int(deg[lev != S.UNREACHED].sum())
        name = f"BFS from vertex {hub}"
    for _ in range(args.warmup):
        step()
    S.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    S.synchronize()
    t = (time.perf_counter() - t0) / max(args.steps, 1)
    log(f"{name}: {1e3 * t:.3f} ms per step, {work / t / 1e9:.2f} GTEPS")
    print(json.dumps({
        "metric": f"{name} traversed edges/s", "value": work / t, "unit": "edges/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.workload in ("pagerank", "bc") else "u32",
        "data": "synthetic (rMat generator of the reference, seeded; " + RMAT_NOTE[RMAT_FLAGS != 0] + ")",
        "config": {"workload": f"{name} on rMat scale {args.scale} (seed 2, edge "
                               f"factor {args.edge_factor}, symmetrised)",
                   "directed_edges": int(m)},
        "extra": extra}))


def run_c2(args):
    """The plain-method GPU configuration: rMat scale --scale (22 by default
    here), seed 1, edge factor 16, symmetrised; BFS (from vertex 0 and from
    the largest hub), PageRank x --pr-iters in FP64, CC, BC from vertex 0 and
    from the hub, and TC, outputs kept in HBM.  The line's value is
    PageRank's traversed edges/s; every algorithm's time and rate is in
    extra, beside the oracle's on the same graph (bounded CPU sample)."""
    import sstgraph_amd as S
    EMV = S.EdgeMapVertexMap
    L = S.lib()
    scale = args.scale
    n = 1 << scale
    src, dst = device_symmetric_rmat(S, n, 1, args.edge_factor * n)
    g = S.SparseMatrixV(n)
    g.insert_batch(src, dst)
    src.free()
    dst.free()
    m = g.M()
    off, cols = g.export_csr()
    deg = np.diff(off).astype(np.int64)
    hub = int(np.argmax(deg))
    if not args.no_tune:
        tune = g.autotune(pr=True, bfs_source=hub)
        log(f"autotuned thresholds {tune['thresholds']}, BFS {tune['bfs_thresholds']}")
    else:
        tune = None
    d_i32 = S.DeviceArray(n, np.int32)
    d_u32 = S.DeviceArray(n, np.uint32)
    d_f64 = S.DeviceArray(n, np.float64)
    reps = max(3, min(args.steps, 10))

    def best(fn):
        for _ in range(args.warmup):
            fn()
        S.synchronize()
        return float(np.median([timed(lambda: (fn(), S.synchronize()))
                                for _ in range(reps)]))

    out = {}
    for name, sv in (("bfs_src0", 0), ("bfs_hub", hub)):
        t = best(lambda: EMV.BFS(g, sv, out=(d_i32, d_u32)))
        lev = d_u32.to_host()
        m_r = int(deg[lev != S.UNREACHED].sum())
        out[name] = dict(src=sv, ms=1e3 * t, reached=int((lev != S.UNREACHED).sum()),
                         gteps=m_r / t / 1e9, gbs_model=(4 * m_r + 12 * n) / t / 1e9)
    t = best(lambda: EMV.PR_S(g, args.pr_iters, out=d_f64))
    out["pagerank"] = dict(iters=args.pr_iters, ms=1e3 * t,
                           gteps=m * args.pr_iters / t / 1e9,
                           gbs_model=(4 * m + 24 * n) * args.pr_iters / t / 1e9)
    pr_value = m * args.pr_iters / t
    t = best(lambda: S._check(L.sstg_cc(g.handle, d_u32.ptr, S.SSTG_DEVICE_PTRS), "cc"))
    labels = d_u32.to_host()
    out["cc"] = dict(ms=1e3 * t, components=int(np.unique(labels).size),
                     gteps=m / t / 1e9)
    for name, sv in (("bc_src0", 0), ("bc_hub", hub)):
        t = best(lambda: S._check(L.sstg_bc(g.handle, sv, d_f64.ptr,
                                            S.SSTG_DEVICE_PTRS), "bc"))
        out[name] = dict(src=sv, ms=1e3 * t)
    t = best(lambda: EMV.TC(g))
    out["tc"] = dict(ms=1e3 * t, triangles=int(EMV.TC(g)))
    for k, v in out.items():
        log(f"{k}: " + ", ".join(f"{a}={b:.4g}" if isinstance(b, float) else f"{a}={b}"
                                 for a, b in v.items()))
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O  # CPU baseline leg only
        if not os.path.exists(O.LIB_PATH):
            O.build()
        t_pr = timed(lambda: O.pagerank(off, cols, 2)) / 2
        t_bfs = timed(lambda: O.bfs(off, cols, hub))
        t_cc = timed(lambda: O.cc(off, cols))
        t_bc = timed(lambda: O.bc(off, cols, hub))
        cpu = dict(value=m / t_pr, unit="edges/s", cores=O.threads(), kind="port",
                   sample=f"oracle PageRank (2 iterations, OpenMP) on the same scale-"
                          f"{scale} graph; BFS and BC from the hub (one thread) and CC "
                          f"(OpenMP) beside it",
                   pagerank_ms_per_iter=1e3 * t_pr, bfs_hub_ms=1e3 * t_bfs,
                   cc_ms=1e3 * t_cc, bc_hub_ms=1e3 * t_bc)
        log(f"cpu: PR {1e3 * t_pr:.1f} ms/iter, BFS {1e3 * t_bfs:.0f} ms, "
            f"CC {1e3 * t_cc:.0f} ms, BC {1e3 * t_bc:.0f} ms")
    print(json.dumps({
        "metric": f"PageRank FP64 traversed edges/s on rMat scale {scale} (with BFS, CC, "
                  "BC, TC in extra)",
        "value": pr_value, "unit": "edges/s", "n_gpus": 1, "steps": reps,
        "warmup": args.warmup, "ms_per_step": out["pagerank"]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (PR, BC), u32",
        "data": "synthetic (rMat generator of the reference, seeded; "
                + RMAT_NOTE[RMAT_FLAGS != 0] + ")",
        "config": {"workload": f"plain graph algorithms on rMat scale {scale} (seed 1, "
                               f"edge factor {args.edge_factor}, symmetrised): BFS, "
                               f"PageRank x{args.pr_iters}, CC, BC, TC",
                   "directed_edges": int(m)},
        "cpu_baseline": cpu,
        "extra": dict(out, autotune=tune)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=None,
                    help="rMat scale (24 by default; 22 for --workload algos)")
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--batch", type=float, default=1e7)
    ap.add_argument("--pr-iters", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the served-stream comparison of the default run")
    ap.add_argument("--grow-batches", type=int, default=24,
                    help="fresh headline-size inserts of the growing sequence")
    ap.add_argument("--serve-requests", type=int, default=60,
                    help="requests of the default run's served-stream comparison")
    ap.add_argument("--no-algos", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the degree-threshold autotuner")
    ap.add_argument("--rmat-normalized", action="store_true",
                    help="rMat draws in [0,1) (probabilities 0.5/0.1/0.1/0.3) "
                         "instead of the reference's [0,2)")
    ap.add_argument("--workload", choices=["updates", "batched", "serve", "pagerank", "bfs",
                                           "bc", "algos"],
                    default="updates")
    ap.add_argument("--pr-segments", type=int, default=None,
                    help="PageRank column-segment width in log2 vertices "
                         "(0: vertex pull; default: the library's)")
    ap.add_argument("--graphs", type=int, default=512)
    ap.add_argument("--serve-scale", type=int, default=25)
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--load", type=float, default=0.7,
                    help="Poisson rate as a fraction of the serialised capacity")
    args = ap.parse_args()
    if args.scale is None:
        args.scale = 22 if args.workload == "algos" else 24
    global RMAT_FLAGS
    if args.rmat_normalized:
        RMAT_FLAGS = 0x10  # SSTG_RMAT_NORMALIZED
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batched":
        return run_batched(args)
    if args.workload == "serve":
        return run_serve(args)
    if args.workload in ("pagerank", "bfs", "bc"):
        return run_algo(args)
    if args.workload == "algos":
        return run_c2(args)
    import sstgraph_amd as S
    L = S.lib()
    n = 1 << args.scale
    count = args.edge_factor * n
    B = int(args.batch)

    # ---------------------------------------------------------- base graph
    log(f"generating rMat scale {args.scale} seed 2: {2 * count} directed pairs")
    src, dst = device_symmetric_rmat(S, n, 2, count)
    g = S.SparseMatrixV(n)
    S.synchronize()
    t_build = timed(lambda: g.insert_batch(src, dst))
    m = g.M()
    log(f"built: m={m} in {t_build:.3f}s ({2 * count / t_build:.3e} pairs/s)")
    src.free()
    dst.free()
    off, cols = g.export_csr()
    extra = {"build_seconds": t_build, "build_pairs_per_s": 2 * count / t_build,
             "m_directed": m, "n": n, "rmat_probabilities": RMAT_NOTE[RMAT_FLAGS != 0]}

    # ------------------------------------------------ BFS / PR on the graph
    # Outputs stay in HBM (DeviceArray); the reference generator leaves vertex
    # 0 isolated at this scale, so BFS is also run from the largest hub.
    if not args.no_algos:
        EMV = S.EdgeMapVertexMap
        deg = np.diff(off).astype(np.int64)
        hub = int(np.argmax(deg))
        d_par = S.DeviceArray(n, np.int32)
        d_lev = S.DeviceArray(n, np.uint32)
        for name, src_v in (("bfs_src0", 0), ("bfs_hub", hub)):
            EMV.BFS(g, src_v, out=(d_par, d_lev))
            S.synchronize()
            tb = [timed(lambda: EMV.BFS(g, src_v, out=(d_par, d_lev)))
                  for _ in range(5)]
            level = d_lev.to_host()
            reached = level != S.UNREACHED
            m_r = int(deg[reached].sum())
            t = float(np.median(tb))
            extra[name] = dict(src=src_v, ms=1e3 * t, reached=int(reached.sum()),
                               m_reached=m_r, gteps=m_r / t / 1e9,
                               levels=int(level[reached].max()) + 1,
                               bytes_model=4 * m_r + 12 * n,
                               gbs_model=(4 * m_r + 12 * n) / t / 1e9)
            log(f"{name} (src {src_v}): {1e3 * t:.3f} ms, reached {int(reached.sum())}, "
                f"{m_r / t / 1e9:.2f} GTEPS")
        # the same BFS from the hub replayed from its recording (opt-in,
        # SSTG_GRAPH_CAPTURE=1): the first call records, the timed calls replay
        os.environ["SSTG_GRAPH_CAPTURE"] = "1"
        try:
            EMV.BFS(g, hub, out=(d_par, d_lev))
            S.synchronize()
            tb = [timed(lambda: EMV.BFS(g, hub, out=(d_par, d_lev))) for _ in range(5)]
        finally:
            del os.environ["SSTG_GRAPH_CAPTURE"]
        t = float(np.median(tb))
        ok = bool(np.array_equal(d_lev.to_host(), level))
        extra["bfs_hub_replayed"] = dict(src=hub, ms=1e3 * t, levels_match=ok,
                                         gteps=extra["bfs_hub"]["m_reached"] / t / 1e9)
        log(f"bfs_hub replayed: {1e3 * t:.3f} ms (levels {'match' if ok else 'DIFFER'})")
        # BC from the hub, launched plainly and replayed from its recording
        d_bc = S.DeviceArray(n, np.float64)
        Lb = S.lib()
        bc_step = lambda: S._check(Lb.sstg_bc(g.handle, hub, d_bc.ptr,
                                              S.SSTG_DEVICE_PTRS), "bc")
        bc_step()
        S.synchronize()
        t = float(np.median([timed(bc_step) for _ in range(5)]))
        bc_plain = d_bc.to_host()
        extra["bc_hub"] = dict(src=hub, ms=1e3 * t)
        os.environ["SSTG_GRAPH_CAPTURE"] = "1"
        try:
            bc_step()
            S.synchronize()
            t = float(np.median([timed(bc_step) for _ in range(5)]))
        finally:
            del os.environ["SSTG_GRAPH_CAPTURE"]
        bc_rep = d_bc.to_host()
        diff = float(np.max(np.abs(bc_rep - bc_plain) / np.maximum(np.abs(bc_plain), 1e-12)))
        extra["bc_hub_replayed"] = dict(src=hub, ms=1e3 * t, max_rel_diff_vs_plain=diff)
        log(f"bc_hub: {extra['bc_hub']['ms']:.3f} ms, replayed {1e3 * t:.3f} ms "
            f"(max rel diff {diff:.1e})")
        d_bc.free()
        d_pr = S.DeviceArray(n, np.float64)
        EMV.PR_S(g, args.pr_iters, out=d_pr)
        S.synchronize()
        tp = [timed(lambda: EMV.PR_S(g, args.pr_iters, out=d_pr)) for _ in range(3)]
        t = float(np.median(tp))
        per_it = t / args.pr_iters
        extra["pagerank"] = dict(iters=args.pr_iters, ms=1e3 * t,
                                 gteps=m * args.pr_iters / t / 1e9,
                                 ms_per_iter=1e3 * per_it,
                                 bytes_model_per_iter=4 * m + 24 * n,
                                 gbs_model=(4 * m + 24 * n) / per_it / 1e9)
        log(f"PR x{args.pr_iters}: {1e3 * t:.2f} ms ({m * args.pr_iters / t / 1e9:.2f} GTEPS, "
            f"{(4 * m + 24 * n) / per_it / 1e9:.0f} GB/s model)")
        if not args.no_tune:
            # degree thresholds picked by the autotuner on this graph (PR
            # iterations + BFS from the hub), then both measured again
            rep = g.autotune(pr=True, bfs_source=hub)
            EMV.BFS(g, hub, out=(d_par, d_lev))
            EMV.PR_S(g, args.pr_iters, out=d_pr)
            S.synchronize()
            tb = float(np.median([timed(lambda: EMV.BFS(g, hub, out=(d_par, d_lev)))
                                  for _ in range(5)]))
            tp = float(np.median([timed(lambda: EMV.PR_S(g, args.pr_iters, out=d_pr))
                                  for _ in range(3)]))
            m_r = extra["bfs_hub"]["m_reached"]
            rep.update(bfs_hub_ms=1e3 * tb, bfs_hub_gteps=m_r / tb / 1e9,
                       pagerank_ms=1e3 * tp,
                       pagerank_gteps=m * args.pr_iters / tp / 1e9,
                       pagerank_gbs_model=(4 * m + 24 * n) / (tp / args.pr_iters) / 1e9)
            extra["autotune"] = rep
            log(f"autotuned thresholds {rep['thresholds']}, BFS {rep['bfs_thresholds']} "
                f"alpha {rep['alpha']:g} beta {rep['beta']:g} ({rep['evaluations']} "
                f"settings): BFS hub {1e3 * tb:.3f} ms, PR x{args.pr_iters} "
                f"{1e3 * tp:.2f} ms")
            g.set_degree_thresholds([8, 64, 256])
            g.set_degree_thresholds([32, 128, 2048], bfs=True)
            g.set_bfs_switch(15, 18)
        for d in (d_par, d_lev, d_pr):
            d.free()

    # ------------------------------------------------------- update batches
    sizes = [1000, 10000, 100000, 1000000, B]
    per_size = {}
    host_batches = {}
    headline = None
    for k, size in enumerate(sizes):
        ds, dd, hs, hd = make_batch(S, n, off, cols, size, 100 + k)
        host_batches[size] = (hs, hd)
        steps = args.steps if size == B else max(3, min(args.steps, 20))
        for _ in range(args.warmup):
            g.insert_batch(ds, dd)
            g.remove_batch(ds, dd)
        S.synchronize()
        ti, tr = [], []
        for _ in range(steps):
            ti.append(timed(lambda: g.insert_batch(ds, dd)))
            tr.append(timed(lambda: g.remove_batch(ds, dd)))
        S.synchronize()
        # insert-then-remove of the same batch also deletes the batch edges
        # that were present before (test.hpp:668-677), so after the first
        # step the edge count is stable at m minus those edges
        m_after = g.M()
        # the row-merge model of this insert on the graph as each timed step
        # found it: 8 B per batch pair, 4 B per old and new element of every
        # touched row, 16 B per touched row
        uk = np.unique(hs.astype(np.int64) * n + hd.astype(np.int64))
        ur, uc = (uk // n), (uk % n)
        keep = ur != uc
        ur, uc = ur[keep], uc[keep]
        present = g.has_batch(ur.astype(np.uint32), uc.astype(np.uint32)).astype(bool)
        deg_now = g.degrees().astype(np.int64)
        trow, tcount = np.unique(ur, return_counts=True)
        _, tnew = np.unique(ur[~present], return_counts=True) if (~present).any() \
            else (None, np.zeros(0, np.int64))
        d_old = int(deg_now[trow].sum())
        d_new = d_old + int(tnew.sum())
        op_bytes = 8 * size + 4 * (d_old + d_new) + 16 * len(trow)
        g.insert_batch(ds, dd)
        new_edges = g.last_update_stats()["effective"]
        g.remove_batch(ds, dd)
        assert new_edges == d_new - d_old, "model and store disagree on new edges"
        assert g.M() == m_after, "update steps must leave the graph stable"
        t_ins = float(np.mean(ti))
        rec = dict(steps=steps, insert_ms=1e3 * t_ins,
                   remove_ms=1e3 * float(np.mean(tr)),
                   insert_p99_ms=1e3 * float(np.percentile(ti, 99)),
                   insert_max_ms=1e3 * float(np.max(ti)),
                   inserts_per_s=size / t_ins,
                   removes_per_s=size / float(np.mean(tr)),
                   # edges each timed insert actually added (distinct, not
                   # self-loops, not already present)
                   unique_new_edges=new_edges,
                   unique_new_per_s=new_edges / t_ins,
                   touched_rows=int(len(trow)),
                   model_bytes=op_bytes,
                   model_gbs=op_bytes / t_ins / 1e9,
                   model_frac=op_bytes / t_ins / 1e9 / HBM_PEAK_GBS)
        rec["m_after"] = m_after
        per_size[str(size)] = rec
        log(f"batch {size}: insert {rec['insert_ms']:.3f} ms "
            f"({rec['inserts_per_s']:.3e}/s, {rec['model_gbs']:.0f} GB/s model), "
            f"remove {rec['remove_ms']:.3f} ms")
        if size == B:
            headline = (ds, dd, hs, hd, ti, tr, op_bytes)
        else:
            ds.free()
            dd.free()
    extra["batches"] = per_size

    # ------------------------------- kernel timings of the headline batch
    # HIP events around each kernel launch on the graph's stream; the
    # insert launches are profiled on their own.
    ds, dd, hs, hd, ti, tr, op_bytes = headline
    S.Profile.reset()
    prof_steps = max(3, min(args.steps, 5))
    stats = None
    for _ in range(prof_steps):
        S.Profile.enable(True)
        g.insert_batch(ds, dd)
        S.synchronize()
        S.Profile.enable(False)
        stats = g.last_update_stats()
        g.remove_batch(ds, dd)
    S.synchronize()
    kernels = sorted(S.Profile.dump(), key=lambda r: -r[2])
    extra["insert_kernels_ms"] = {name: ms / prof_steps for name, _, ms in kernels[:12]}
    extra["insert_stats"] = stats
    dom_name, dom_n, dom_ms = kernels[0]
    ms_per = dom_ms / max(dom_n, 1)
    # algorithmic bytes: the row-merge model of the whole insert of the batch
    # (8B + 4 sum(d_old + d_new) + 16|T|), attributed to its dominant
    # kernel, which carries the merge of nearly every touched row; the
    # tile-input count (4 B in + 4 B out per element the LDS tiles merge)
    # beside it
    achieved = op_bytes / (ms_per * 1e-3) / 1e9 if ms_per > 0 else 0.0
    tile_bytes = 8 * stats["tile_inputs"]
    traffic, traffic_src = None, None
    tfile = os.path.join(ROOT, "profiles", "k_merge_tile_scale24_batch1e7.traffic.json")
    if os.path.exists(tfile) and args.scale == 24 and B == 10_000_000:
        traffic = json.load(open(tfile)).get("bytes_per_launch")
        traffic_src = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this command, "
                       "profiles/k_merge_tile_scale24_batch1e7.traffic.json")
    roofline = dict(kernel=dom_name, bound="hbm", achieved=achieved,
                    peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                    traffic=traffic, traffic_source=traffic_src,
                    algorithmic_bytes=op_bytes,
                    algorithmic_model="8B + 4 sum_T(d_old + d_new) + 16|T| of the "
                                      "whole insert, attributed to its dominant kernel",
                    avg_launch_ms=ms_per,
                    tile_input_bytes=tile_bytes,
                    tile_input_gbs=tile_bytes / (ms_per * 1e-3) / 1e9 if ms_per > 0 else 0.0,
                    whole_insert_ms=1e3 * float(np.mean(ti)),
                    whole_insert_frac=op_bytes / float(np.mean(ti)) / 1e9 / HBM_PEAK_GBS)
    extra["insert_op_model_bytes"] = op_bytes
    extra["insert_op_model_gbs"] = op_bytes / float(np.mean(ti)) / 1e9

    # ------------------------- a growing sequence that crosses rebuilds
    # fresh batches of the headline size inserted one after another (no
    # removes): rows outgrow their slots and move to the pool tail, and the
    # pool is compacted or grown inside whichever insert crosses the
    # threshold; the per-batch latency distribution shows what those
    # inserts cost
    ds.free()
    dd.free()
    seq_ms, seq_rebuild = [], []
    r0 = g.last_update_stats()["rebuilds"]
    for i in range(args.grow_batches):
        gs, gd = S.rmat(n, 5000 + i, B, device=True, normalized=RMAT_FLAGS != 0)
        S.synchronize()
        seq_ms.append(1e3 * timed(lambda: g.insert_batch(gs, gd)))
        seq_rebuild.append(g.last_update_stats()["rebuilds"])
        gs.free()
        gd.free()
    if seq_ms:
        sq = np.array(seq_ms)
        crossed = [i for i in range(len(seq_rebuild))
                   if seq_rebuild[i] != (seq_rebuild[i - 1] if i else r0)]
        extra["growing_sequence"] = dict(
            batches=len(sq), batch_edges=B,
            p50_ms=float(np.percentile(sq, 50)), p99_ms=float(np.percentile(sq, 99)),
            max_ms=float(sq.max()), rebuilds=int(seq_rebuild[-1] - r0),
            rebuild_batches_ms=[float(sq[i]) for i in crossed],
            m_end=g.M())
        log(f"growing sequence: {len(sq)} x {B} inserts, p50 {np.percentile(sq, 50):.3f} ms "
            f"p99 {np.percentile(sq, 99):.3f} ms max {sq.max():.3f} ms, "
            f"{seq_rebuild[-1] - r0} rebuilds")

    # ------------------------------------------------ served stream (C5)
    if not args.no_serve:
        g.close()
        res, rate, m5, reqs5 = serve_compare(
            S, args.serve_scale, args.edge_factor, args.serve_requests, args.load,
            [m for m in SERVE_MODES if m[0] in ("serialised", "served")])
        sv, se = res["served"], res["serialised"]
        extra["served"] = dict(
            workload=f"rMat scale {args.serve_scale} (seed 4, symmetrised, "
                     f"{m5} directed edges), {len(reqs5)} seeded requests "
                     f"(55% updates of 1e2..1e6 edges, 25% BFS, 12% PR x10, 8% TC), "
                     f"Poisson at {rate:.1f}/s ({args.load:.0%} of the serialised capacity)",
            served_p50_ms=sv["poisson"]["p50_ms"], served_p99_ms=sv["poisson"]["p99_ms"],
            serialised_p50_ms=se["poisson"]["p50_ms"],
            serialised_p99_ms=se["poisson"]["p99_ms"],
            served_capacity_rps=sv["burst"]["throughput_rps"],
            serialised_capacity_rps=se["burst"]["throughput_rps"],
            capacity_ratio=sv["burst"]["throughput_rps"] / se["burst"]["throughput_rps"],
            per_kind_served=sv["poisson"]["per_kind"],
            per_kind_serialised=se["poisson"]["per_kind"])

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(off, cols, n, host_batches, B)
        log(f"cpu baseline {cpu['value']:.3e} edges/s on {cpu['cores']} threads, "
            f"{cpu['single_core']['value']:.3e} on one")
        if not args.no_algos:
            hub = int(np.argmax(np.diff(off)))
            extra["cpu_algos"] = cpu_algos(off, cols, hub)
            ca = extra["cpu_algos"]
            log(f"cpu BFS from hub {ca['all_cores']['bfs_hub_ms']:.0f} ms "
                f"({ca['one_core']['bfs_hub_ms']:.0f} on one), BC "
                f"{ca['all_cores']['bc_hub_ms']:.0f} ms, PR "
                f"{ca['all_cores']['pagerank_ms_per_iter']:.0f} ms/iter")

    ins = float(np.mean(ti))
    out = {
        "metric": "edges inserted/s + BFS/PR GTEPS on scale-24 RMAT; p99 query latency served",
        "value": B / ins,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * (ins + float(np.mean(tr))),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (rMat generator of the reference, seeded; " + RMAT_NOTE[RMAT_FLAGS != 0] + ")",
        "config": {"workload": f"dynamic updates: rMat scale {args.scale} base graph "
                               f"(seed 2, edge factor {args.edge_factor}, symmetrised), "
                               f"insert+remove batches of {B} edges (seed 104, 20% existing)",
                   "scale": args.scale, "batch": B, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
