"""Writes the golden vectors under tests/golden/ (run from the repo root).

Independent libraries, not the oracle, compute the expected answers:
  * scipy.sparse.csgraph.shortest_path (unweighted) -> BFS distance per source,
    which fixes the Shortest layers {v : dist(s, v) = i};
  * networkx.single_source_shortest_path_length with a cutoff, as a second
    BFS implementation on the same graphs;
  * Boolean powers of the scipy adjacency matrix -> the Any layers
    {v : (A^i)[s, v] != 0} (walks of length exactly i).
Graphs: hand-made ones with closed forms (path, star, cycle, clique, two
components with isolated vertices, a directed chain) and two small seeded
random graphs.
"""
from __future__ import annotations

import json
from pathlib import Path

import networkx as nx
import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import shortest_path

OUT = Path(__file__).resolve().parent


def graphs():
    g = {}
    n = 70
    g["path70"] = (n, [(i, i + 1) for i in range(n - 1)], True)
    g["star33"] = (33, [(0, i) for i in range(1, 33)], True)
    g["cycle65"] = (65, [(i, (i + 1) % 65) for i in range(65)], True)
    g["clique9"] = (9, [(i, j) for i in range(9) for j in range(i + 1, 9)], True)
    g["two_comp"] = (130, [(i, i + 1) for i in range(0, 40)] +
                     [(i, i + 1) for i in range(60, 99)], True)
    g["dchain"] = (20, [(i, i + 1) for i in range(19)] + [(5, 0)], False)
    rng = np.random.default_rng(20261203)
    for name, V, m in (("rand200", 200, 600), ("rand1000", 1000, 2500)):
        e = rng.integers(0, V, size=(m, 2))
        e = [(int(a), int(b)) for a, b in e if a != b]
        g[name] = (V, e, True)
    return g


def csr_of(V, edges, symmetric):
    pairs = set()
    for a, b in edges:
        pairs.add((a, b))
        if symmetric:
            pairs.add((b, a))
    pairs = sorted(pairs)
    off = np.zeros(V + 1, dtype=np.uint64)
    for a, _ in pairs:
        off[a + 1] += 1
    off = np.cumsum(off, dtype=np.uint64)
    adj = np.array([b for _, b in pairs], dtype=np.uint32)
    return off, adj, pairs


def main():
    manifest = {}
    for name, (V, edges, sym) in graphs().items():
        off, adj, pairs = csr_of(V, edges, sym)
        rows = np.array([a for a, _ in pairs], dtype=np.int64)
        A = sp.csr_matrix((np.ones(len(pairs)), (rows, adj.astype(np.int64))), shape=(V, V))
        rng = np.random.default_rng(len(name) * 1000 + V)
        S = min(V, 70)
        sources = np.sort(rng.choice(V, size=S, replace=False)).astype(np.uint32)
        if name == "path70":
            sources[:3] = [0, 0, 69]  # duplicates and an end point
            sources = np.sort(sources)
        dist = shortest_path(A, method="D", unweighted=True, directed=True, indices=sources)
        dist = np.where(np.isinf(dist), -1, dist).astype(np.int32)
        # networkx cross-check (cutoff 6)
        G = nx.DiGraph()
        G.add_nodes_from(range(V))
        G.add_edges_from(pairs)
        for r, s in enumerate(sources[:10]):
            ln = nx.single_source_shortest_path_length(G, int(s), cutoff=6)
            for v in range(V):
                d = dist[r, v]
                exp = ln.get(v, -1)
                assert (d if 0 <= d <= 6 else -1) == exp, (name, s, v, d, exp)
        # Any layers for k = 0..5 via boolean matrix powers
        kmax = 5
        any_layers = np.zeros((kmax + 1, S, V), dtype=np.uint8)
        X = sp.csr_matrix((np.ones(S), (np.arange(S), sources.astype(np.int64))), shape=(S, V))
        for i in range(kmax + 1):
            any_layers[i] = (X.toarray() != 0)
            X = (X @ A)
            X.data[:] = 1.0
        np.savez_compressed(OUT / f"{name}.npz", offsets=off, adj=adj, sources=sources,
                            dist=dist, any_layers=np.packbits(any_layers, axis=2,
                                                              bitorder="little"),
                            V=np.array([V]), symmetric=np.array([int(sym)]))
        manifest[name] = {"V": V, "E": int(off[-1]), "S": int(S), "symmetric": sym}
    (OUT / "manifest.json").write_text(json.dumps(manifest, indent=1, sort_keys=True))
    print(json.dumps(manifest, indent=1))


if __name__ == "__main__":
    main()
