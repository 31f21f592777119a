"""GPU parity: the HIP path (through the C ABI) against the oracle.

Bar: bit-exact.  Whole runs are compared as canonical labels (minimum vertex id
of each SCC) with the oracle's Tarjan; because every election in the HIP path
is deterministic (lowest id per slot), the final colour, state and class
representative of every vertex must also equal the oracle's sequential FB
restatement run with the same election table size.
"""
import ctypes

import numpy as np
import pytest

import oracle
from gpuscc_code_amd import scc

from helpers import bigrid, complete, cycle, graph_from_arcs, is_canonical, path_dag, planted, \
    random_digraph, settled

pytestmark = pytest.mark.gpu

SCHEDS = ["q", "t", "d", "l", "h"]
DEMO_EXPECT = np.array([1, 2, 3, 4, 2, 2, 7, 8, 7, 7, 7, 7, 8]) - 1  # SCCs of demo.graph


def default_slot_bits(n):
    sb = 10
    while sb < 22 and (1 << sb) < n:
        sb += 1
    return sb


def demo_graph():
    # the reference's demo.graph (13 vertices): line "a b" is the arc b -> a
    lines = [(1, 1), (1, 12), (2, 5), (2, 11), (4, 3), (5, 4), (5, 6), (6, 2), (7, 8), (7, 11),
             (8, 13), (9, 12), (10, 7), (10, 13), (11, 7), (11, 9), (11, 10), (12, 11), (13, 8)]
    return graph_from_arcs(13, [(b - 1, a - 1) for a, b in lines])


def check_against_oracle(ctx, g, sched, variant, full_state=True):
    ref, _ = oracle.tarjan(g.row, g.col)
    ctx.upload(g)
    labels, st = ctx.run(sched, variant)
    assert np.array_equal(labels, ref), f"labels differ (sched {sched}, variant {variant})"
    assert st["num_scc"] == len(np.unique(ref))
    if full_state:
        o = oracle.fb(g.row, g.col, g.rowT, g.colT, variant=variant,
                      slot_bits=default_slot_bits(g.n), hash_kind=1)
        words, roots = ctx.get_state()
        colour, state = scc.split_word(words)
        assert st["steps"] == o["steps"]
        assert np.array_equal(colour, o["colour"])
        assert np.array_equal(state & 0x1C, o["state"] & 0x1C)  # E, T, P bits
        assert np.array_equal(roots, o["root"])
        assert st["wcc_step"] == o["wcc_step"]
        assert st["trim2_pairs"] == o["trim2_pairs"]
    return labels, st


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("sched", SCHEDS)
def test_demo_graph(ctx, sched, variant):
    g = demo_graph()
    labels, st = check_against_oracle(ctx, g, sched, variant)
    assert np.array_equal(labels, DEMO_EXPECT)
    assert st["num_scc"] == 6 and st["trivial"] == 3 and st["biggest"] == 5


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("sched", SCHEDS)
def test_rmat14_config(ctx, sched, variant):
    # R-MAT scale 14, edge factor 8, (0.57, 0.19, 0.19), seed 1, dedup, no self-loops
    g = scc.gen_rmat(14, 8, 0.57, 0.19, 0.19, seed=1)
    check_against_oracle(ctx, g, sched, variant)


@pytest.mark.parametrize("sched", ["c", "b"])
def test_combine_and_best(ctx, sched):
    # COMBINE (l, then q after n/100 FB-eliminated; scc.cu:565-638) and BEST
    # (direction-optimising thresholds): whole runs, every variant, full state
    check_against_oracle(ctx, demo_graph(), sched, 0)
    for g in (scc.gen_rmat(14, 8, 0.57, 0.19, 0.19, seed=1), scc.gen_mesh(128, 0.6, seed=3),
              scc.gen_batch(48, 64, 2048, seed=4)):
        for variant in (0, 1, 2):
            check_against_oracle(ctx, g, sched, variant)
    ctx.set_batch(None)


@pytest.mark.parametrize("sched", SCHEDS)
def test_mesh_small(ctx, sched):
    g = scc.gen_mesh(128, 0.6, seed=3)
    for variant in (0, 1, 2):
        check_against_oracle(ctx, g, sched, variant)


@pytest.mark.parametrize("sched", SCHEDS)
def test_batch_small(ctx, sched):
    g = scc.gen_batch(48, 64, 2048, seed=4)
    for variant in (0, 1, 2):
        check_against_oracle(ctx, g, sched, variant)


@pytest.mark.parametrize("sched", SCHEDS)
def test_closed_forms(ctx, sched):
    ctx.upload(cycle(1000))
    lab, st = ctx.run(sched)
    assert np.all(lab == 0) and st["num_scc"] == 1
    ctx.upload(path_dag(777))
    lab, st = ctx.run(sched)
    assert np.array_equal(lab, np.arange(777)) and st["num_scc"] == 777
    ctx.upload(complete(40))
    lab, st = ctx.run(sched)
    assert np.all(lab == 0) and st["num_scc"] == 1
    ctx.upload(bigrid(40))
    lab, st = ctx.run(sched)
    assert np.all(lab == 0) and st["num_scc"] == 1
    g, expect = planted(300, 7, 2000, seed=5)
    ctx.upload(g)
    lab, st = ctx.run(sched)
    assert np.array_equal(lab, expect) and st["num_scc"] == 300


@pytest.mark.parametrize("sched", SCHEDS)
def test_random_and_self_loops(ctx, sched):
    for seed in range(3):
        g = random_digraph(3000, 4500, seed, self_loops=True)
        check_against_oracle(ctx, g, sched, seed % 3)


@pytest.mark.parametrize("sched", SCHEDS)
def test_edge_cases(ctx, sched):
    # no vertices
    ctx.upload(scc.csr_from_arcs(0, [], []))
    lab, st = ctx.run(sched)
    assert len(lab) == 0 and st["num_scc"] == 0
    # one vertex, with and without a self-loop
    ctx.upload(scc.csr_from_arcs(1, [], []))
    lab, st = ctx.run(sched)
    assert list(lab) == [0] and st["num_scc"] == 1
    ctx.upload(graph_from_arcs(1, [(0, 0)]))
    lab, st = ctx.run(sched)
    assert list(lab) == [0] and st["num_scc"] == 1
    # vertices without arcs
    ctx.upload(scc.csr_from_arcs(500, [], []))
    lab, st = ctx.run(sched)
    assert np.array_equal(lab, np.arange(500))
    # a single 2-cycle plus isolated vertices; trim-2 pair in scc_2_w
    g = graph_from_arcs(6, [(4, 1), (1, 4)])
    for variant in (0, 1, 2):
        check_against_oracle(ctx, g, sched, variant)


def test_upload_errors(ctx):
    with pytest.raises(scc.SccError):
        ctx.upload(scc.Graph(n=3, row=np.array([0, 1, 2, 5], np.int32), col=np.array([1, 2, 0], np.int32),
                             rowT=np.array([0, 1, 2, 3], np.int32), colT=np.array([2, 0, 1], np.int32)))
    with pytest.raises(scc.SccError):  # column id out of range
        ctx.upload(scc.Graph(n=2, row=np.array([0, 1, 2], np.int32), col=np.array([1, 7], np.int32),
                             rowT=np.array([0, 1, 2], np.int32), colT=np.array([1, 0], np.int32)))
    # a good upload afterwards still works; transpose built by the library
    g = cycle(10)
    ctx.upload(g, transpose=False)
    lab, _ = ctx.run("q")
    assert np.all(lab == 0)


def test_upload_out_of_memory_leaves_context_usable(ctx):
    """An allocation that fails inside scc_upload (forced from the host: the nth
    device allocation reports out of memory) gives SCC_E_OOM, and the context
    takes the next upload and run as if nothing had happened; the reference
    exits instead (GPUerrchk.h:6-12)."""
    g = scc.gen_rmat(10, 8, seed=3)
    ref, _ = oracle.tarjan(g.row, g.col)
    for nth in (1, 2, 7, 15, 26):
        ctx.fail_alloc(nth)
        with pytest.raises(scc.SccError) as e:
            ctx.upload(g)
        assert e.value.status == 2, f"allocation {nth}"
        with pytest.raises(scc.SccError) as e:  # nothing uploaded: a run is out of order
            ctx.run("h")
        assert e.value.status == 5
        ctx.upload(g)
        lab, _ = ctx.run("h")
        assert np.array_equal(lab, ref)
    ctx.fail_alloc(1000)  # beyond the upload's allocations: never fires
    ctx.upload(g)
    ctx.fail_alloc(0)
    lab, _ = ctx.run("d")
    assert np.array_equal(lab, ref)


def test_upload_sizes_beyond_int32(ctx):
    """Sizes an int32 CSR cannot hold give SCC_E_OVERFLOW before any array is
    read (the reference's worklists drop what does not fit, worklistc.h:80-81)."""
    from gpuscc_code_amd import _lib
    row = np.zeros(8, np.int32)
    col = np.zeros(8, np.int32)
    L = _lib.lib()
    p = lambda a: a.ctypes.data_as(_lib._i32p)  # noqa: E731
    assert L.scc_upload(ctx._h, 4, 1 << 31, p(row), p(col), None, None) == 4
    assert L.scc_upload(ctx._h, 4, (1 << 31) - 1, p(row), p(col), None, None) == 4
    assert L.scc_upload(ctx._h, (1 << 31) - 1, 4, p(row), p(col), None, None) == 4
    assert L.scc_upload(ctx._h, -1, 4, p(row), p(col), None, None) == 3
    g = cycle(5)  # the context still works
    ctx.upload(g)
    lab, _ = ctx.run("t")
    assert np.all(lab == 0)


# ---------------------------------------------------------------------------
# phase-level parity on random states
# ---------------------------------------------------------------------------
def random_state(n, rng, ncol=3, p_elim=0.2):
    colour = rng.integers(1, ncol + 1, n).astype(np.uint32)
    state = np.where(rng.random(n) < p_elim, 4, 0).astype(np.uint8)
    return colour, state


@pytest.mark.parametrize("sched", SCHEDS)
@pytest.mark.parametrize("direction", [0, 1, 2])
def test_phase_reach(ctx, sched, direction):
    rng = np.random.default_rng(7 + direction)
    g = random_digraph(20000, 60000, 11)
    ctx.upload(g)
    colour, state = random_state(g.n, rng)
    bits = {0: [2], 1: [1], 2: [2, 1]}[direction]
    for bit in bits:  # separate seed sets per direction
        seeds = rng.choice(g.n, 30, replace=False)
        state[seeds] = np.where(state[seeds] & 4, state[seeds], state[seeds] | bit)
    ctx.set_state(scc.make_word(colour, state))
    depth = ctx.reach(direction, sched)
    assert depth >= 1
    words, _ = ctx.get_state()
    _, gst = scc.split_word(words)
    ost = state.copy()
    for bit in bits:
        if bit == 2:
            oracle.reach(g.row, g.col, colour, ost, bit)
        else:
            oracle.reach(g.rowT, g.colT, colour, ost, bit)
    assert np.array_equal(gst & 0x1F, ost & 0x1F)
    assert np.array_equal(scc.split_word(words)[0], colour)


@pytest.mark.parametrize("sched", SCHEDS)
def test_phase_reach_pull_levels(sched):
    """A tiny pull threshold forces the hybrid's bottom-up levels everywhere."""
    rng = np.random.default_rng(17)
    g = random_digraph(30000, 120000, 19)
    with scc.Context(0, scc.alg_opt(hybrid_threshold=1)) as c:
        c.upload(g)
        colour, state = random_state(g.n, rng, ncol=2)
        for bit in (2, 1):
            seeds = rng.choice(g.n, 5, replace=False)
            state[seeds] = np.where(state[seeds] & 4, state[seeds], state[seeds] | bit)
        c.set_state(scc.make_word(colour, state))
        c.reach(2, sched)
        words, _ = c.get_state()
        ost = state.copy()
        oracle.reach(g.row, g.col, colour, ost, 2)
        oracle.reach(g.rowT, g.colT, colour, ost, 1)
        assert np.array_equal(scc.split_word(words)[1] & 0x1F, ost & 0x1F)
        ref, _ = oracle.tarjan(g.row, g.col)
        lab, _ = c.run(sched)
        assert np.array_equal(lab, ref)


@pytest.mark.parametrize("sched", SCHEDS)
def test_sequential_reach_matches(sched):
    """Forward-then-backward chains (the reference's order) give the same run."""
    g = scc.gen_rmat(13, 8, seed=3)
    ref, _ = oracle.tarjan(g.row, g.col)
    o = oracle.fb(g.row, g.col, g.rowT, g.colT, slot_bits=default_slot_bits(g.n), hash_kind=1)
    with scc.Context(0, scc.alg_opt(sequential_reach=1)) as c:
        c.upload(g)
        for variant in (0, 1, 2):
            lab, st = c.run(sched, variant)
            assert np.array_equal(lab, ref)
        lab, st = c.run(sched, 0)
        assert st["steps"] == o["steps"]


def test_phase_trim(ctx):
    rng = np.random.default_rng(3)
    g = random_digraph(20000, 30000, 12, self_loops=True)
    ctx.upload(g)
    colour, state = random_state(g.n, rng)
    roots = np.arange(g.n, dtype=np.int32)
    ctx.set_state(scc.make_word(colour, state), roots)
    it = ctx.trim()
    assert it >= 1
    words, groots = ctx.get_state()
    ost, oroot = state.copy(), roots.copy()
    oracle.trim(g.row, g.col, g.rowT, g.colT, colour, ost, oroot)
    assert np.array_equal(scc.split_word(words)[1], ost)
    assert np.array_equal(groots, oroot)


def test_phase_update_elect():
    rng = np.random.default_rng(5)
    g = random_digraph(5000, 10000, 13)
    sb = 6  # a small table makes slot collisions certain
    with scc.Context(0, scc.alg_opt(slot_bits=sb)) as c:
        c.upload(g)
        colour, state = random_state(g.n, rng, ncol=200)
        roots = np.arange(g.n, dtype=np.int32)
        c.set_state(scc.make_word(colour, state), roots)
        npiv = c.update(recolour=False)
        ocol, ost, oroot = colour.copy(), state.copy(), roots.copy()
        onp, owin = oracle.update_elect(ocol, ost, oroot, oracle.COLOUR_MASK_PACKED, sb, 1, None, False)
        words, _ = c.get_state()
        assert npiv == onp
        assert np.array_equal(scc.split_word(words)[1], ost)
        # mark F/B on vertices of pivoted colours, then split colours
        pivc = set(ocol[ost == 19].tolist())
        fb = rng.integers(0, 4, g.n).astype(np.uint8)
        live = (ost & 4) == 0
        sel = live & np.isin(ocol, list(pivc)) & (ost != 19)
        ost2 = ost.copy()
        ost2[sel] = fb[sel]
        c.set_state(scc.make_word(ocol, ost2), oroot)
        npiv2 = c.update(recolour=True)
        words, groots = c.get_state()
        onp2, _ = oracle.update_elect(ocol, ost2, oroot, oracle.COLOUR_MASK_PACKED, sb, 1, owin, True)
        gc, gs = scc.split_word(words)
        assert npiv2 == onp2
        assert np.array_equal(gc, ocol)
        assert np.array_equal(gs, ost2)
        assert np.array_equal(groots, oroot)


def test_phase_count_elimmed_and_unset_pivots(ctx):
    """countElimmed (scc.cu:42-51) counts the FB-eliminated vertices, state & 12
    == 4 (E without T), against the oracle's restatement; unsetPivots
    (scc_routines.cu:10-27) takes exactly the vertices in state 19 back to 0
    and leaves every other word alone."""
    rng = np.random.default_rng(11)
    g = random_digraph(5000, 20000, seed=12)
    ctx.upload(g)
    L = oracle.lib()
    for trial in range(4):
        colour = rng.integers(1, 1 << 24, g.n).astype(np.uint32)
        state = rng.choice(np.array([0, 1, 2, 3, 4, 6, 12, 19, 20, 28, 35, 19, 19], np.uint8), g.n)
        ctx.set_state(scc.make_word(colour, state))
        cnt = ctx.count_elimmed()
        st8 = np.ascontiguousarray(state, np.uint8)
        want = L.orc_count_elimmed(g.n, st8.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        assert cnt == want == int(np.count_nonzero((state & 12) == 4))
        ctx.unset_pivots()
        words, _ = ctx.get_state()
        c2, s2 = scc.split_word(words)
        assert np.array_equal(c2, colour)
        assert np.array_equal(s2, np.where(state == 19, 0, state))


def test_hybrid_reference_rule(ctx):
    """hybrid_policy 1: the reference's HYBRID rule (scc.cu:493-563, switch at
    scc.cu:877-891: t until n/100 vertices are FB-eliminated, then q), every
    variant, full state against the oracle, on graphs that cross the switch in
    their first step (R-MAT, batch) and ones that take many steps (mesh)."""
    try:
        ctx.set_opts(scc.alg_opt(hybrid_policy=1))
        for g in (demo_graph(), scc.gen_rmat(14, 8, 0.57, 0.19, 0.19, seed=1), scc.gen_mesh(128, 0.6, seed=3),
                  planted(400, 5, 3000, seed=2)[0]):
            for variant in (0, 1, 2):
                check_against_oracle(ctx, g, "h", variant)
    finally:
        ctx.set_opts(scc.alg_opt())


def test_best_direction_optimising_first_step(ctx):
    """BEST (reach_b.cu:374-725): the first FB step of R-MAT switches to
    bottom-up once its frontier holds more than the unexplored arcs / 15 (its
    giant SCC makes the frontier explode), so at least one level pulls.  Labels
    and full state against the oracle; a long cycle (one vertex per frontier)
    as a closed form."""
    g = scc.gen_rmat(14, 8, 0.57, 0.19, 0.19, seed=1)
    try:
        ctx.profile(1)
        ctx.profile_read(reset=True)
        check_against_oracle(ctx, g, "b", 0)
        prof = ctx.profile_read(reset=True)
        assert prof["sweep_levels"] >= 1
        ctx.upload(cycle(3000))
        lab, st = ctx.run("b")
        assert np.all(lab == 0) and st["steps"] == 1 and st["num_scc"] == 1
    finally:
        ctx.profile(0)


def test_phase_trim2_and_wcc(ctx):
    rng = np.random.default_rng(9)
    # sparse random arcs among 20000 vertices plus 2-cycles on 4000 further vertices, some
    # of them with one extra arc in or out (still a pair by one of the two tests)
    n0, n1 = 20000, 4000
    base = random_digraph(n0, 22000, 14)
    src = list(np.repeat(np.arange(n0), np.diff(base.row)))
    dst = list(base.col)
    for k in range(0, n1, 2):
        a, b = n0 + k, n0 + k + 1
        src += [a, b]
        dst += [b, a]
        if k % 6 == 2:
            src.append(b)
            dst.append(int(rng.integers(0, n0)))
    g = scc.csr_from_arcs(n0 + n1, np.array(src), np.array(dst))
    ctx.upload(g)
    colour, state = random_state(g.n, rng, ncol=2, p_elim=0.1)
    colour[n0:] = 1
    roots = np.arange(g.n, dtype=np.int32)
    ctx.set_state(scc.make_word(colour, state), roots)
    pairs = ctx.trim2()
    ost, oroot = state.copy(), roots.copy()
    opairs = oracle.trim2(g.row, g.col, g.rowT, g.colT, colour, ost, oroot)
    words, groots = ctx.get_state()
    assert pairs == opairs and pairs > 0
    assert np.array_equal(scc.split_word(words)[1], ost)
    assert np.array_equal(groots, oroot)
    # WCC labels -> fresh colours: members of a class share (maxc + 1 + root)
    ctx.set_state(scc.make_word(colour, ost), oroot)
    ctx.update_wcc()
    words, _ = ctx.get_state()
    gc, gs = scc.split_word(words)
    lab = oracle.wcc_labels(g.rowT, g.colT, colour, ost)
    live = (ost & 4) == 0
    maxc = int(colour.max())
    assert np.array_equal(gc[live], ((maxc + 1 + lab[live]) & 0xFFFFFF).astype(np.uint32))


def test_canonical_property(ctx):
    g = random_digraph(50000, 150000, 21)
    ctx.upload(g)
    for sched in SCHEDS:
        lab, _ = ctx.run(sched)
        assert is_canonical(lab)


@pytest.mark.parametrize("sched", SCHEDS)
def test_huge_rows_split(sched):
    """A tiny huge-row threshold sends most rows through the grid-wide split."""
    for huge in (2, 5):
        with scc.Context(0, scc.alg_opt(huge_row=huge)) as c:
            for g in (scc.gen_rmat(13, 8, seed=5), scc.gen_mesh(64, 0.6, seed=2), demo_graph()):
                for variant in (0, 1, 2):
                    ref, _ = oracle.tarjan(g.row, g.col)
                    c.upload(g)
                    lab, st = c.run(sched, variant)
                    assert np.array_equal(lab, ref), (huge, variant)
                    o = oracle.fb(g.row, g.col, g.rowT, g.colT, variant=variant,
                                  slot_bits=default_slot_bits(g.n), hash_kind=1)
                    assert st["steps"] == o["steps"]


def batch_oracle(g, slot_bits):
    """Each graph of a batch solved alone by the oracle's FB restatement (colour
    g + 1 to start, the batch's election table size): the colours, states and
    roots of all graphs, and the most FB steps any graph takes."""
    off = g.graph_offsets
    colour = np.empty(g.n, np.uint32)
    state = np.empty(g.n, np.uint8)
    root = np.empty(g.n, np.int32)
    steps = 0
    for k in range(len(off) - 1):
        a, b = int(off[k]), int(off[k + 1])
        if b == a:
            continue
        r0, r1, t0, t1 = int(g.row[a]), int(g.row[b]), int(g.rowT[a]), int(g.rowT[b])
        o = oracle.fb(g.row[a:b + 1] - r0, g.col[r0:r1] - a, g.rowT[a:b + 1] - t0, g.colT[t0:t1] - a,
                      variant=0, init_colour=np.full(b - a, k + 1, np.uint32), slot_bits=slot_bits, hash_kind=1)
        colour[a:b], state[a:b], root[a:b] = o["colour"], o["state"], o["root"] + a
        steps = max(steps, o["steps"])
    return colour, state, root, steps


@pytest.mark.parametrize("sched", SCHEDS)
def test_batch_persistent_solver(ctx, sched):
    """With batch_solver 1 the plain variant of a batch runs as one persistent launch, every graph
    solved on its own by one workgroup (largest first): labels against Tarjan,
    and the colour, E/T/P state and root of every vertex against the oracle run
    on each graph alone; the step count is the largest of any graph.  Graphs of
    a few vertices, empty graphs (no vertex) and graphs too big for the LDS copy
    are in the batch."""
    ctx.set_opts(scc.alg_opt(batch_solver=1))
    for g in (scc.gen_batch(160, 64, 8192, 1.5, 6.0, 0.05, 0.5, seed=4),
              scc.gen_batch(40, 2, 40000, 0.5, 4.0, 0.2, 0.6, seed=7)):
        off = g.graph_offsets
        ref, _ = oracle.tarjan(g.row, g.col)
        ctx.upload(g)
        ctx.set_batch(off)
        lab, st = ctx.run(sched, 0)
        assert np.array_equal(lab, ref)
        assert st["num_scc"] == len(np.unique(ref)) and st["host_syncs"] == 1
        colour, state, root, steps = batch_oracle(g, default_slot_bits(g.n))
        words, roots = ctx.get_state()
        c2, s2 = scc.split_word(words)
        assert np.array_equal(c2, colour)
        assert np.array_equal(s2 & 0x1C, state & 0x1C)
        assert np.array_equal(roots, root)
        assert st["steps"] == steps
    # an empty graph inside a batch
    g = scc.gen_batch(30, 50, 400, seed=11)
    off = np.concatenate([g.graph_offsets[:5], g.graph_offsets[4:]])
    ref, _ = oracle.tarjan(g.row, g.col)
    ctx.upload(g)
    ctx.set_batch(off)
    lab, _ = ctx.run(sched, 0)
    assert np.array_equal(lab, ref)
    ctx.set_batch(None)
    ctx.set_opts(scc.alg_opt())


@pytest.mark.parametrize("sched", SCHEDS)
def test_batched_graphs(ctx, sched):
    """Batch of uneven graphs advancing together in the level kernels
    (the default): colour = graph id + 1, one election table for all;
    labels and final state vs the oracle."""
    g = scc.gen_batch(160, 64, 8192, 1.5, 6.0, 0.05, 0.5, seed=4)
    off = g.graph_offsets
    gid = np.repeat(np.arange(len(off) - 1), np.diff(off)).astype(np.uint32)
    ref, _ = oracle.tarjan(g.row, g.col)
    ctx.set_opts(scc.alg_opt())
    ctx.upload(g)
    ctx.set_batch(off)
    for variant in (0, 1, 2):
        lab, st = ctx.run(sched, variant)
        assert np.array_equal(lab, ref)
        o = oracle.fb(g.row, g.col, g.rowT, g.colT, variant=variant, init_colour=gid + 1,
                      slot_bits=default_slot_bits(g.n), hash_kind=1)
        words, roots = ctx.get_state()
        colour, state = scc.split_word(words)
        assert st["steps"] == o["steps"]
        assert np.array_equal(colour, o["colour"])
        assert np.array_equal(roots, o["root"])
    # labels never leave their graph
    assert np.all(np.searchsorted(off, lab, side="right") == np.searchsorted(off, np.arange(g.n), side="right"))
    ctx.set_batch(None)
    ctx.set_opts(scc.alg_opt())


def test_batch_pull_bitmaps(ctx):
    """The first FB step of a batch pulls against visited bitmaps (bits set by
    the first election and by every claim): forced pull levels (small
    thresholds) on batches with trimmed vertices, full state against the oracle
    for every variant, and the same with pulls off."""
    try:
        for g in (scc.gen_batch(48, 64, 2048, seed=4), scc.gen_batch(200, 16, 600, 1.0, 3.0, 0.1, 0.3, seed=9)):
            for opts in (scc.alg_opt(hybrid_threshold=16, hybrid_arc_div=1000),
                         scc.alg_opt(hybrid_threshold=16, hybrid_arc_div=1000, local_rounds=-1),
                         scc.alg_opt(hybrid_threshold=1 << 30, hybrid_arc_div=-1)):
                ctx.set_opts(opts)
                ctx.upload(g)
                ctx.set_batch(g.graph_offsets)
                for variant in (0, 1, 2):
                    ctx.profile(4)  # level log: the pulls happen where they are forced
                    lab, st = ctx.run("h", variant)
                    lv = ctx.profile_levels(reset=True)
                    ctx.profile(0)
                    pulled = lv[:, scc.LEVEL_FIELDS.index("pulled")][lv[:, 0] == 0]
                    assert (pulled.sum() > 0) == (opts.hybrid_threshold == 16)
                    ref, _ = oracle.tarjan(g.row, g.col)
                    assert np.array_equal(lab, ref)
                    off = g.graph_offsets
                    gid = np.repeat(np.arange(len(off) - 1), np.diff(off)).astype(np.uint32)
                    o = oracle.fb(g.row, g.col, g.rowT, g.colT, variant=variant, init_colour=gid + 1,
                                  slot_bits=default_slot_bits(g.n), hash_kind=1)
                    words, roots = ctx.get_state()
                    colour, state = scc.split_word(words)
                    assert st["steps"] == o["steps"]
                    assert np.array_equal(colour, o["colour"]) and np.array_equal(roots, o["root"])
                    assert np.array_equal(state & 0x1C, o["state"] & 0x1C)
                ctx.set_batch(None)
    finally:
        ctx.set_opts(scc.alg_opt())


def test_batch_rejects_crossing_arcs(ctx):
    g = graph_from_arcs(6, [(0, 1), (1, 0), (2, 4), (4, 3)])
    ctx.upload(g)
    with pytest.raises(scc.SccError):
        ctx.set_batch([0, 3, 6])   # arc 2 -> 4 crosses
    ctx.set_batch([0, 2, 6])
    lab, _ = ctx.run("t")
    assert list(lab) == [0, 0, 2, 3, 4, 5]
    with pytest.raises(scc.SccError):
        ctx.set_batch([0, 4, 3, 6])  # decreasing offsets


@pytest.mark.parametrize("sched", SCHEDS)
def test_speculative_steps_match_waiting_path(ctx, sched):
    """A step launched without waits (guards abort a chain that may need more
    levels; the host resumes it), in every variant (the WCC pass of scc_w and
    scc_2_w follows a speculative step), ends exactly where the waiting path ends: the
    first run after an upload guesses the chain lengths from other graphs (and
    so exercises the resume path), the second uses the lengths of the first."""
    # level counts are left out: with the workgroup-local rounds of d, l and h
    # they depend on timing (the reached set of every chain does not)
    keys = ("steps", "num_scc", "pivots_total", "first_pivot", "trivial", "biggest", "wcc_step", "trim2_pairs")
    for graphs in (0, 1):  # reach chains launched level by level / replayed as captured HIP graphs
        ctx.set_opts(scc.alg_opt(reach_graphs=graphs))
        for g in (demo_graph(), scc.gen_rmat(13, 8, seed=21), scc.gen_mesh(96, 0.6, seed=5), cycle(3000),
                  path_dag(500), cycle_with_tails()):
            ref, _ = oracle.tarjan(g.row, g.col)
            ctx.upload(g)
            for variant in (0, 1, 2):  # scc, scc_w, scc_2_w (the WCC pass inside speculative steps)
                ctx.profile(2)  # event timing: every chain waits
                lab0, st0 = ctx.run(sched, variant)
                w0, r0 = ctx.get_state()
                ctx.profile(0)
                for _ in range(3):  # replays with eager work and waits between them
                    lab, st = ctx.run(sched, variant)
                    w, r = ctx.get_state()
                    assert np.array_equal(lab, ref) and np.array_equal(lab0, ref)
                    assert np.array_equal(settled(w), settled(w0)) and np.array_equal(r, r0), f"variant {variant}"
                    assert {k: st[k] for k in keys} == {k: st0[k] for k in keys}, f"variant {variant}"
                ctx.profile_read(reset=True)
    ctx.set_opts(scc.alg_opt())


def cycle_with_tails(k=50, t=300):
    """A k-cycle with a t-vertex path into it and one out of it: trim-1 peels the
    paths one vertex per level (about t levels) and leaves the cycle live."""
    arcs = [(i, (i + 1) % k) for i in range(k)]
    arcs += [(k + i, k + i + 1) for i in range(t - 1)] + [(k + t - 1, 0)]
    arcs += [(1, k + t)] + [(k + t + i, k + t + i + 1) for i in range(t - 1)]
    return graph_from_arcs(k + 2 * t, arcs)


@pytest.mark.parametrize("variant", [0, 1])
def test_trim1_resumed_after_short_guess(ctx, variant):
    """With speculative steps the rest of trim-1 goes out with the first step, as
    long as the last run's trim-1 (a guess after an upload); a deeper trim-1
    stops everything behind its guard and the host resumes it with waits.  Full
    state against the oracle when stopped, when the depth is known, and when
    trim-1 leaves nothing live."""
    ctx.upload(cycle(100))
    ctx.run("h")  # a shallow trim-1: the next guess is short
    _, st1 = check_against_oracle(ctx, cycle_with_tails(), "h", variant)  # stopped and resumed
    _, st2 = check_against_oracle(ctx, cycle_with_tails(), "h", variant)  # depth known from the run before
    if variant == 0:  # the speculative path: the resumed trim-1 costs waits of its own
        assert st1["host_syncs"] > st2["host_syncs"]
    lab, st = check_against_oracle(ctx, path_dag(400), "h", variant)  # no live vertex after trim-1
    assert st["steps"] == 0 and st["first_pivot"] == -1


def test_pull_cost_policy(ctx):
    """hybrid_policy 2 (pull when the frontier's arcs reach twice the live
    unreached vertices) changes only which levels pull: full state parity."""
    ctx.set_opts(scc.alg_opt(hybrid_policy=2))
    try:
        for g in (scc.gen_rmat(14, 8, 0.57, 0.19, 0.19, seed=1), scc.gen_mesh(128, 0.6, seed=3),
                  scc.gen_batch(48, 64, 2048, seed=4)):
            for variant in (0, 1, 2):
                check_against_oracle(ctx, g, "h", variant)
        g = scc.gen_batch(64, 64, 4096, seed=9)
        ref, _ = oracle.tarjan(g.row, g.col)
        ctx.upload(g)
        ctx.set_batch(g.graph_offsets)
        for variant in (0, 1, 2):
            labels, _ = ctx.run("h", variant)
            assert np.array_equal(labels, ref)
        ctx.set_batch(None)
    finally:
        ctx.set_opts(scc.alg_opt())


@pytest.mark.parametrize("sched", SCHEDS)
def test_wide_level_workgroups(ctx, sched):
    """reach_block 512 (512-wide workgroups for the level kernels: item and
    staging capacities at their limits) changes only the launch shape: full
    state parity, and a forced huge-row split and forced pulls on top."""
    try:
        for opts in (scc.alg_opt(reach_block=512), scc.alg_opt(reach_block=512, huge_row=8, hybrid_threshold=64)):
            ctx.set_opts(opts)
            for g in (scc.gen_rmat(14, 8, 0.57, 0.19, 0.19, seed=1), scc.gen_mesh(128, 0.6, seed=3),
                      demo_graph(), random_digraph(3000, 12000, seed=12, self_loops=True)):
                for variant in (0, 1, 2):
                    check_against_oracle(ctx, g, sched, variant)
        with pytest.raises(scc.SccError):
            ctx.set_opts(scc.alg_opt(reach_block=384))
    finally:
        ctx.set_opts(scc.alg_opt())


@pytest.mark.parametrize("sched", SCHEDS)
def test_tail_settings(ctx, sched):
    """The one-workgroup tail (tail_live) ends where the multi-kernel steps end:
    the same words, roots, steps and pivots as with it off, and the oracle's
    full state, whether it starts at once, late or never.  The planted graph's
    live arcs first exceed the tail's LDS copy (steps over global memory until
    the live set has split), the larger mesh starts it with more live vertices
    than the copy holds."""
    keys = ("steps", "num_scc", "pivots_total", "first_pivot", "trivial", "biggest")
    try:
        for g in (demo_graph(), scc.gen_mesh(96, 0.6, seed=5), planted(1500, 2, 75000, seed=7)[0],
                  scc.gen_mesh(256, 0.6, seed=8), random_digraph(3000, 12000, seed=12, self_loops=True),
                  cycle(3000)):
            ref, _ = oracle.tarjan(g.row, g.col)
            ctx.upload(g)
            ctx.set_opts(scc.alg_opt(tail_live=-1))
            lab0, st0 = ctx.run(sched)
            w0, r0 = ctx.get_state()
            assert np.array_equal(lab0, ref) and st0["tail_steps"] == 0
            for tl in (0, 1, 64, 1 << 30):
                ctx.set_opts(scc.alg_opt(tail_live=tl))
                lab, st = ctx.run(sched)
                w, r = ctx.get_state()
                assert np.array_equal(lab, ref)
                dw, dr = np.nonzero(settled(w) != settled(w0))[0], np.nonzero(r != r0)[0]
                xbits = int(np.bitwise_or.reduce(w[dw] ^ w0[dw])) if len(dw) else 0
                assert len(dw) == 0 and len(dr) == 0, (
                    f"n={g.n} tail_live={tl}: {len(dw)} words differ (bits {xbits:#x}, "
                    f"first {dw[:4].tolist()}: {w[dw[:4]].tolist()} vs {w0[dw[:4]].tolist()}), {len(dr)} roots differ "
                    f"(first {dr[:4].tolist()}: {r[dr[:4]].tolist()} vs {r0[dr[:4]].tolist()}); "
                    f"steps {st['steps']} vs {st0['steps']}, tail steps {st['tail_steps']}")
                assert {k: st[k] for k in keys} == {k: st0[k] for k in keys}
                if tl == 1 << 30 and st0["steps"] > 1:
                    assert st["tail_steps"] == st["steps"] - 1 and st["tail_live"] > 0
        ctx.set_opts(scc.alg_opt(tail_live=1 << 30))
        check_against_oracle(ctx, planted(1500, 2, 75000, seed=7)[0], sched, 0)
        check_against_oracle(ctx, planted(2000, 2, 2500, seed=3)[0], sched, 0)
        with pytest.raises(scc.SccError):
            ctx.set_opts(scc.alg_opt(tail_live=-2))
    finally:
        ctx.set_opts(scc.alg_opt())


def test_tail_hands_back_at_tag_wrap(ctx):
    """Election tags are 16 bits and the elections of every run of a context
    share them: the tail stops before the election whose tag would wrap, and
    the host refills the table and goes on.  Runs until more than 2^16
    elections have passed; every run ends with the first run's state."""
    g = planted(2000, 2, 2500, seed=3)[0]  # 810 FB steps
    ref, _ = oracle.tarjan(g.row, g.col)
    ctx.upload(g)
    lab0, st0 = ctx.run("h")
    w0, r0 = ctx.get_state()
    assert np.array_equal(lab0, ref) and st0["steps"] >= 100 and st0["tail_steps"] > 0
    elections, runs = st0["steps"] + 1, 1
    while elections < (1 << 16) + 3 * st0["steps"]:
        lab, st = ctx.run("h")
        w, r = ctx.get_state()
        assert np.array_equal(lab, ref) and np.array_equal(settled(w), settled(w0)) and np.array_equal(r, r0), \
            f"run {runs}"
        assert st["steps"] == st0["steps"] and st["pivots_total"] == st0["pivots_total"]
        elections += st["steps"] + 1
        runs += 1


def test_level_log(ctx):
    """Profile flag 4 logs one row per level launched: each chain is a run of
    levels 0, 1, 2, ...; one reach chain per FB step; a level's longest serial
    row scan is bounded by its scanned arcs, is zero on push-only levels and
    positive on levels that pulled arcs in both directions.  Logging changes
    no label."""
    g = scc.gen_rmat(14, 8, seed=7)
    ref, _ = oracle.tarjan(g.row, g.col)
    ctx.upload(g)
    ctx.set_opts(scc.alg_opt(hybrid_threshold=64))  # frontiers of 64 or more pull
    ctx.profile(4)
    try:
        labels, st = ctx.run("h")
        rows = ctx.profile_levels()
    finally:
        ctx.profile(0)
        ctx.profile_read(reset=True)
        ctx.set_opts(scc.alg_opt())
    assert np.array_equal(labels, ref)
    f = {k: i for i, k in enumerate(scc.LEVEL_FIELDS)}
    assert rows.shape[1] == len(f) and len(rows) > 0
    kind, level = rows[:, f["kind"]], rows[:, f["level"]]
    assert set(np.unique(kind)) <= {0.0, 1.0}
    for i in range(len(rows)):
        if level[i] > 0:
            assert kind[i - 1] == kind[i] and level[i - 1] == level[i] - 1
    assert (rows[:, f["ms"]] > 0).all()
    assert int(np.sum((kind == 0) & (level == 0))) == st["steps"]
    reach = rows[kind == 0]
    scan, edges, pulled = reach[:, f["max_scan"]], reach[:, f["edges"]], reach[:, f["pulled"]]
    assert (scan <= edges).all()
    assert (scan[pulled == 0] == 0).all()
    both = (pulled == 3) & (edges > 0)
    assert both.any() and (scan[both] >= 1).all()
    assert rows[0, f["kind"]] == 1  # trim1 first (its degree-test level runs inside the initialisation)


def _device_i32(addr, n):
    """Copy n int32 from a device address (hipMemcpy through the HIP runtime)."""
    import ctypes
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    out = np.empty(n, np.int32)
    rc = hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(addr), ctypes.c_size_t(4 * n),
                       ctypes.c_int(2))  # hipMemcpyDeviceToHost
    assert rc == 0
    return out


def test_labels_device_across_phase_calls(ctx):
    """After a plain-rule run the labels live in the root array; calls that
    write roots move them to the label buffer first, so the labels at the
    address asked for after such a call are still the run's."""
    from gpuscc_code_amd import learned
    g = scc.gen_rmat(12, 8, seed=11)
    ref, _ = oracle.tarjan(g.row, g.col)
    ctx.upload(g)
    ctx.mlp_set(learned.init_params(7))
    for pivot in (scc.PIVOT_PLAIN, scc.PIVOT_LEARNED):
        labels, _ = ctx.run("h", 0, pivot)
        assert np.array_equal(labels, ref)
        assert np.array_equal(_device_i32(ctx.labels_device(), g.n), ref)
        w, r = ctx.get_state()
        ctx.set_state(w, np.zeros_like(r))  # writes every root
        assert np.array_equal(_device_i32(ctx.labels_device(), g.n), ref)
        ctx.trim()
        assert np.array_equal(_device_i32(ctx.labels_device(), g.n), ref)
