"""Writes tests/golden/scc_golden.npz: small directed graphs with their SCC labels
computed by two library routines independent of this project and of the
reference -- scipy.sparse.csgraph.connected_components(connection='strong')
and networkx.strongly_connected_components -- canonicalised to the minimum
vertex id per SCC.  Run once in a container that has scipy and networkx:

    python tests/golden/make_golden.py

The graphs are drawn with numpy's PCG64 from fixed seeds; each is stored as
its arc list (src, dst), 0-based, possibly with duplicates and self-loops.
"""
import os

import networkx as nx
import numpy as np
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scc_golden.npz")


def canon(labels):
    labels = np.asarray(labels)
    mn = {}
    for v, l in enumerate(labels):
        mn.setdefault(l, v)
    return np.array([mn[l] for l in labels], np.int32)


def cases():
    rng = np.random.default_rng(20241203)
    out = []
    for k in range(24):
        n = int(rng.integers(1, 400))
        m = int(rng.integers(0, 4 * n))
        out.append((f"er_{k}", n, rng.integers(0, n, m), rng.integers(0, n, m)))
    # sparse graphs near the SCC phase transition (mean degree ~1) and dense ones
    for k, deg in enumerate([0.8, 1.0, 1.2, 1.5, 3.0, 8.0]):
        n = 2000
        m = int(deg * n)
        out.append((f"deg{deg}", n, rng.integers(0, n, m), rng.integers(0, n, m)))
    # many 2-cycles (trim-2 territory) chained by one-way arcs
    n = 600
    src, dst = [], []
    for i in range(0, n, 2):
        src += [i, i + 1]
        dst += [i + 1, i]
        if i + 2 < n:
            src.append(i + 1)
            dst.append(i + 2)
    out.append(("pairs_chain", n, np.array(src), np.array(dst)))
    return out


def main():
    data = {}
    names = []
    for name, n, s, d in cases():
        s = np.asarray(s, np.int32)
        d = np.asarray(d, np.int32)
        A = csr_matrix((np.ones(len(s)), (s, d)), shape=(n, n))
        _, lab = connected_components(A, directed=True, connection="strong")
        a = canon(lab)
        G = nx.DiGraph()
        G.add_nodes_from(range(n))
        G.add_edges_from(zip(s.tolist(), d.tolist()))
        b = np.empty(n, np.int32)
        for comp in nx.strongly_connected_components(G):
            c = min(comp)
            for v in comp:
                b[v] = c
        assert np.array_equal(a, b), name
        data[f"{name}__n"] = np.array([n], np.int32)
        data[f"{name}__src"] = s
        data[f"{name}__dst"] = d
        data[f"{name}__labels"] = a
        names.append(name)
    data["names"] = np.array(names)
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT}: {len(names)} graphs")


if __name__ == "__main__":
    main()
