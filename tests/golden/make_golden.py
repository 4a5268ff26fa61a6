"""Writes the golden fixtures under tests/golden/ (run once, in a container
that has the reference tree; the fixtures are committed and the GPU box only
reads them).

Sources of truth, none of which is the oracle:
  * rmat_vectors.json / hash_kats.json -- printed by oracle/_ref/rmat_ref, the
    reference's own include/SSTGraph/internal/rmat_util.h compiled from the
    reference tree by oracle/Makefile.
  * graph_goldens.json -- BFS levels, the CC partition and the triangle count
    computed by networkx (3.4.2) on RMAT graphs whose edges come from
    rmat_ref, symmetrised, self-loops dropped, deduplicated.
  * johnson8-2-4.mtx, star.mtx -- copies of the reference's examples/ data
    files (MatrixMarket, pattern symmetric, 1-indexed).

Usage: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
RMAT_REF = os.path.join(ROOT, "oracle", "_ref", "rmat_ref")


def ref_rmat(n, seed, first, count, abc=None):
    args = [RMAT_REF, str(n), str(seed), str(first), str(count)]
    if abc is not None:
        args += [repr(x) for x in abc]
    out = subprocess.run(args, check=True, capture_output=True).stdout
    arr = np.array(out.split(), dtype=np.uint64).reshape(-1, 2)
    return arr[:, 0].astype(np.uint32), arr[:, 1].astype(np.uint32)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def rmat_vectors():
    cases = []
    specs = [
        # (n, seed, first, count, abc)
        (65536, 0, 0, 256, None),            # scale-16 base graph, seed 0
        (1 << 22, 1, 0, 128, None),          # scale-22 base graph, seed 1
        (1 << 22, 1, 100_000_000, 64, None), # far index
        (1 << 24, 2, 0, 128, None),          # scale-24 base graph, seed 2
        (1 << 24, 100, 0, 128, None),        # update batch seed
        (1 << 25, 4, 0, 128, None),          # scale-25 served graph, seed 4
        (1 << 25, 4, 4_000_000_000, 32, None),  # index near uint32 wrap
        (1024, 3, 0, 256, None),
        (2, 7, 0, 32, None),
        (1 << 20, 5, 0, 128, (0.57, 0.19, 0.19)),  # Graph500-style a/b/c
        (1 << 16, 0xFFFFFFFF, 0, 64, None),   # seed at uint32 max
    ]
    for n, seed, first, count, abc in specs:
        s, d = ref_rmat(n, seed, first, count, abc)
        a, b, c = abc if abc else (0.5, 0.1, 0.1)
        cases.append(dict(n=n, seed=seed, first=first, count=count, a=a, b=b,
                          c=c, src=s.tolist(), dst=d.tolist()))
    return cases


def hash_kats():
    out = subprocess.run([RMAT_REF, "hash"], check=True,
                         capture_output=True, text=True).stdout
    kats = {"h32": [], "h64": [], "aspen": [], "log2_up": []}
    for line in out.splitlines():
        k, x, v = line.split()
        kats[k].append([int(x), int(v)])
    return kats


def rmat_graph(scale, ef, seed):
    n = 1 << scale
    s, d = ref_rmat(n, seed, 0, ef * n)
    src = np.concatenate([s, d]).astype(np.int64)
    dst = np.concatenate([d, s]).astype(np.int64)
    keep = src != dst
    key = np.unique(src[keep] * n + dst[keep])
    return n, (key // n).astype(np.int64), (key % n).astype(np.int64)


def graph_goldens():
    import networkx as nx
    out = {}
    for scale, ef, seed, full in [(10, 16, 0, True), (12, 8, 7, True),
                                  (16, 16, 0, False)]:
        n, src, dst = rmat_graph(scale, ef, seed)
        G = nx.Graph()
        G.add_nodes_from(range(n))
        G.add_edges_from(zip(src.tolist(), dst.tolist()))
        lev = np.full(n, 0xFFFFFFFF, dtype=np.uint32)
        for v, l in nx.single_source_shortest_path_length(G, 0).items():
            lev[v] = l
        lab = np.empty(n, dtype=np.uint32)
        comps = list(nx.connected_components(G))
        for comp in comps:
            mn = min(comp)
            for v in comp:
                lab[v] = mn
        tri = sum(nx.triangles(G).values()) // 3
        deg = np.bincount(src, minlength=n).astype(np.uint32)
        rec = dict(scale=scale, edge_factor=ef, seed=seed, n=n,
                   m_directed=int(len(src)), triangles=int(tri),
                   components=len(comps),
                   largest_component=max(len(c) for c in comps),
                   bfs_reached=int((lev != 0xFFFFFFFF).sum()),
                   bfs_level_hist={str(k): int(v) for k, v in
                                   zip(*np.unique(lev, return_counts=True))},
                   sha_levels=sha(lev), sha_cc_labels=sha(lab),
                   sha_degrees=sha(deg),
                   edge_key_sum=int((src * n + dst).sum()))
        if full:
            rec["levels"] = lev.tolist()
            rec["cc_labels"] = lab.tolist()
        out[f"rmat_s{scale}_ef{ef}_seed{seed}"] = rec
        print(f"scale {scale}: m={len(src)} tri={tri} comps={len(comps)}",
              file=sys.stderr)
    return out


def main():
    if not os.path.exists(RMAT_REF):
        sys.exit("oracle/_ref/rmat_ref missing: run make -C oracle where the "
                 "reference tree is present")
    with open(os.path.join(HERE, "rmat_vectors.json"), "w") as f:
        json.dump(rmat_vectors(), f)
    with open(os.path.join(HERE, "hash_kats.json"), "w") as f:
        json.dump(hash_kats(), f, indent=1)
    with open(os.path.join(HERE, "graph_goldens.json"), "w") as f:
        json.dump(graph_goldens(), f)


if __name__ == "__main__":
    main()
