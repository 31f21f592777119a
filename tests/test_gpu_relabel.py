"""Renumbered uploads (alg_opt.relabel): inside the library the vertices are
numbered by decreasing degree, and every result is still that of the graph as
given.  Bar, as for the other parity tests: labels bit-exact against the
oracle's Tarjan, and the final colour, E/T/P state, root, step count, WCC step
and trim-2 pairs of every vertex bit-exact against the oracle's sequential FB
restatement, which never sees the renumbering (the elections, roots, trim-2 and
the WCC labels order vertices by the caller's ids).  The whole GPU suite can
also run renumbered: pytest --relabel on.
"""
import numpy as np
import pytest

import oracle
from gpuscc_code_amd import scc

from helpers import random_digraph, settled
from test_gpu_parity import DEMO_EXPECT, batch_oracle, check_against_oracle, default_slot_bits, demo_graph

pytestmark = pytest.mark.gpu

SCHEDS = ["q", "t", "d", "l", "h"]


@pytest.fixture
def rctx(ctx):
    """The session context with renumbering forced on for its uploads."""
    ctx.set_opts(scc.alg_opt(relabel=1))
    yield ctx
    ctx.set_batch(None)
    ctx.set_opts(scc.alg_opt())


def graphs():
    return {
        "demo": demo_graph(),
        "rmat12": scc.gen_rmat(12, 8, 0.57, 0.19, 0.19, seed=1),
        "mesh48": scc.gen_mesh(48, 0.6, seed=2),
        "random": random_digraph(3000, 4500, 11, self_loops=True),
    }


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("sched", SCHEDS)
def test_renumbered_full_state(rctx, sched, variant):
    for name, g in graphs().items():
        print(f"[relabel] {name} n={g.n} schedule {sched} variant {variant}", flush=True)
        labels, st = check_against_oracle(rctx, g, sched, variant)
        assert st["renumbered"] == 1
        if name == "demo":
            assert np.array_equal(labels, DEMO_EXPECT)


def test_renumbered_combine_best_and_reference_hybrid(rctx):
    g = scc.gen_rmat(12, 8, 0.57, 0.19, 0.19, seed=3)
    for sched in ("c", "b"):
        check_against_oracle(rctx, g, sched, 0)
    rctx.set_opts(scc.alg_opt(relabel=1, hybrid_policy=1))
    for variant in (0, 1, 2):
        check_against_oracle(rctx, g, "h", variant)


def test_renumbered_learned_rule_matches_plain_numbering(ctx):
    """The learned election orders by score, then by the caller's id: the same
    pivots with and without renumbering, so the same steps and labels."""
    from gpuscc_code_amd import learned
    g = scc.gen_rmat(13, 8, 0.57, 0.19, 0.19, seed=4)
    ref, _ = oracle.tarjan(g.row, g.col)
    out = {}
    try:
        ctx.mlp_set(learned.load_params())
        for relabel in (-1, 1):
            ctx.set_opts(scc.alg_opt(relabel=relabel))
            ctx.upload(g)
            lab, st = ctx.run("h", 0, pivot=scc.PIVOT_LEARNED)
            assert np.array_equal(lab, ref)
            assert st["renumbered"] == (1 if relabel == 1 else 0)
            words, roots = ctx.get_state()
            out[relabel] = (st["steps"], st["first_pivot"], words, roots)
    finally:
        ctx.set_opts(scc.alg_opt())
    a, b = out[-1], out[1]
    assert a[0] == b[0] and a[1] == b[1]
    assert np.array_equal(settled(a[2]), settled(b[2])) and np.array_equal(a[3], b[3])


def test_renumbered_state_and_phases(rctx):
    """set_state / get_state in the caller's order; trim-2 and the WCC pass by
    the caller's ids (the lower id of a pair, the least id of a class)."""
    rng = np.random.default_rng(12)
    g = random_digraph(5000, 6000, 13)
    rctx.upload(g)
    colour = rng.integers(1, 3, g.n).astype(np.uint32)
    state = np.where(rng.random(g.n) < 0.1, 4, 0).astype(np.uint8)
    roots = rng.integers(0, g.n, g.n).astype(np.int32)
    rctx.set_state(scc.make_word(colour, state), roots)
    w2, r2 = rctx.get_state()
    assert np.array_equal(w2, scc.make_word(colour, state)) and np.array_equal(r2, roots)
    roots = np.arange(g.n, dtype=np.int32)
    rctx.set_state(scc.make_word(colour, state), roots)
    pairs = rctx.trim2()
    ost, oroot = state.copy(), roots.copy()
    opairs = oracle.trim2(g.row, g.col, g.rowT, g.colT, colour, ost, oroot)
    words, groots = rctx.get_state()
    assert pairs == opairs
    assert np.array_equal(scc.split_word(words)[1], ost)
    assert np.array_equal(groots, oroot)
    rctx.set_state(scc.make_word(colour, ost), oroot)
    rctx.update_wcc()
    words, _ = rctx.get_state()
    lab = oracle.wcc_labels(g.rowT, g.colT, colour, ost)
    live = (ost & 4) == 0
    maxc = int(colour.max())
    assert np.array_equal(scc.split_word(words)[0][live], ((maxc + 1 + lab[live]) & 0xFFFFFF).astype(np.uint32))


@pytest.mark.parametrize("solver", [0, 1])
def test_renumbered_batch(rctx, solver):
    """A batch is renumbered within each graph (graphs keep their ranges): the
    level kernels (full state against the oracle on the whole batch) and the
    persistent solver (against the oracle on each graph alone)."""
    g = scc.gen_batch(120, 64, 6000, 1.5, 6.0, 0.05, 0.5, seed=6)
    off = g.graph_offsets
    ref, _ = oracle.tarjan(g.row, g.col)
    rctx.set_opts(scc.alg_opt(relabel=1, batch_solver=solver))
    rctx.upload(g)
    rctx.set_batch(off)
    lab, st = rctx.run("h", 0)
    assert np.array_equal(lab, ref) and st["renumbered"] == 1
    words, roots = rctx.get_state()
    colour, state = scc.split_word(words)
    if solver == 0:
        gid = np.repeat(np.arange(len(off) - 1), np.diff(off)).astype(np.uint32)
        o = oracle.fb(g.row, g.col, g.rowT, g.colT, variant=0, init_colour=gid + 1,
                      slot_bits=default_slot_bits(g.n), hash_kind=1)
        oc, ost, oroot, steps = o["colour"], o["state"], o["root"], o["steps"]
    else:
        oc, ost, oroot, steps = batch_oracle(g, default_slot_bits(g.n))
    assert st["steps"] == steps
    assert np.array_equal(colour, oc)
    assert np.array_equal(state & 0x1C, ost & 0x1C)
    assert np.array_equal(roots, oroot)


def test_relabel_auto_and_edges(ctx):
    """relabel 0 renumbers a skewed graph of 65,536 vertices or more by degree,
    numbers a mesh of 65,536 vertices tile by tile (high diameter; with tiles off it
    keeps its numbering) and keeps the numbering of a small graph; graphs of 0, 1
    and 2 vertices and graphs without arcs run renumbered too."""
    cases = [(scc.gen_rmat(16, 8, 0.57, 0.19, 0.19, seed=7), 1, 0), (scc.gen_mesh(256, 0.6, seed=8), 2, 0),
             (scc.gen_mesh(256, 0.6, seed=8), 0, -1), (scc.gen_rmat(10, 8, 0.57, 0.19, 0.19, seed=9), 0, 0)]
    for g, want, tile in cases:
        ctx.set_opts(scc.alg_opt(relabel=0, tile_reach=tile))
        ctx.upload(g)
        lab, st = ctx.run("h", 0)
        assert st["renumbered"] == want
        assert np.array_equal(lab, oracle.tarjan(g.row, g.col)[0])
    try:
        ctx.set_opts(scc.alg_opt(relabel=1))
        for n, arcs in ((0, []), (1, []), (1, [(0, 0)]), (2, [(0, 1), (1, 0)]), (5, [])):
            g = scc.csr_from_arcs(n, np.array([a for a, _ in arcs], np.int32),
                                  np.array([b for _, b in arcs], np.int32))
            ctx.upload(g)
            lab, st = ctx.run("h", 0)
            assert np.array_equal(lab[:n], oracle.tarjan(g.row, g.col)[0])
    finally:
        ctx.set_opts(scc.alg_opt())
