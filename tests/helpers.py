"""Shared helpers for the tests: small graph builders and label checks."""
import numpy as np

from gpuscc_code_amd import scc


def graph_from_arcs(n, arcs, drop_self=False):
    arcs = np.asarray(arcs, np.int64).reshape(-1, 2)
    return scc.csr_from_arcs(n, arcs[:, 0], arcs[:, 1], drop_self_loops=drop_self)


def cycle(n):
    return graph_from_arcs(n, [(i, (i + 1) % n) for i in range(n)])


def path_dag(n):
    return graph_from_arcs(n, [(i, i + 1) for i in range(n - 1)])


def complete(n):
    return graph_from_arcs(n, [(i, j) for i in range(n) for j in range(n) if i != j])


def bigrid(side):
    arcs = []
    for r in range(side):
        for c in range(side):
            v = r * side + c
            if c + 1 < side:
                arcs += [(v, v + 1), (v + 1, v)]
            if r + 1 < side:
                arcs += [(v, v + side), (v + side, v)]
    return graph_from_arcs(side * side, arcs)


def random_digraph(n, m, seed, self_loops=False):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, n, m)
    d = rng.integers(0, n, m)
    return scc.csr_from_arcs(n, s, d, drop_self_loops=not self_loops)


def planted(n_cycles, size, extra, seed):
    """Disjoint directed cycles of `size` plus `extra` arcs that only go from a
    lower cycle index to a higher one (a DAG between cycles): SCCs = the cycles."""
    rng = np.random.default_rng(seed)
    n = n_cycles * size
    perm = rng.permutation(n)
    arcs = []
    for k in range(n_cycles):
        vs = perm[k * size:(k + 1) * size]
        arcs += [(vs[i], vs[(i + 1) % size]) for i in range(size)]
    for _ in range(extra):
        a, b = sorted(rng.integers(0, n_cycles, 2))
        if a == b:
            continue
        arcs.append((perm[a * size + rng.integers(size)], perm[b * size + rng.integers(size)]))
    expect = np.empty(n, np.int64)
    for k in range(n_cycles):
        vs = perm[k * size:(k + 1) * size]
        expect[vs] = vs.min()
    return graph_from_arcs(n, arcs), expect


def is_canonical(labels):
    """labels[v] <= v and labels[labels[v]] == labels[v]"""
    lab = np.asarray(labels)
    idx = np.arange(len(lab))
    return bool(np.all(lab <= idx) and np.all(lab[lab] == lab))


def settled(words):
    """The vertex words with the direction and expanded bits (B 1, F 2, BX 32,
    FX 64) of eliminated vertices cleared.  Nothing reads those bits once a vertex
    is eliminated, and they depend on timing: in the first FB step the short-row
    expansions OR a direction bit into every target without reading it first
    (every arc joins one colour there), eliminated targets included, and which
    rows are expanded that way rather than by a path that reads the target first
    depends on which workgroup reaches a vertex first.  Colours and the E, T, P
    and pair bits stay, as in the comparisons with the oracle."""
    w = np.asarray(words, np.uint32)
    return np.where((w & 4) != 0, w & np.uint32(~0x63 & 0xFFFFFFFF), w)
