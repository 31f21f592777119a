"""The tile reach (alg_opt.tile_reach): the upload numbers the vertices tile by
tile, and the first FB step's reach (schedule h) runs inside tiles in LDS with
claims across tile boundaries between iterations.  Bar, as for the other
parity tests: labels bit-exact against the oracle's Tarjan, and the final
colour, E/T/P state, root, step count, WCC step and trim-2 pairs of every
vertex bit-exact against the oracle's sequential FB restatement -- the reach is
the same closure, so the state is the level kernels'.  Small tile sizes put
many tile boundaries on small graphs.
"""
import numpy as np
import pytest

import oracle
from gpuscc_code_amd import scc

from helpers import random_digraph
from test_gpu_parity import DEMO_EXPECT, check_against_oracle, demo_graph

pytestmark = pytest.mark.gpu


@pytest.fixture
def tctx(ctx):
    yield ctx
    ctx.set_batch(None)
    ctx.set_opts(scc.alg_opt())


def graphs():
    return {
        "demo": demo_graph(),
        "mesh64": scc.gen_mesh(64, 0.6, seed=5),
        "mesh150": scc.gen_mesh(150, 0.4, seed=6),
        "random": random_digraph(4000, 7000, 17, self_loops=True),
        "rmat12": scc.gen_rmat(12, 8, 0.57, 0.19, 0.19, seed=8),
    }


@pytest.mark.parametrize("tile", [1, 64, 300])
@pytest.mark.parametrize("variant", [0, 1, 2])
def test_tile_reach_full_state(tctx, tile, variant):
    """Every variant with the tile reach (h) and, on the same renumbered graph, a
    level schedule (t), full state against the oracle."""
    tctx.set_opts(scc.alg_opt(tile_reach=tile))
    for name, g in graphs().items():
        for sched in ("h", "t"):
            print(f"[tiles] {name} n={g.n} tile {tile} schedule {sched} variant {variant}", flush=True)
            labels, st = check_against_oracle(tctx, g, sched, variant)
            assert st["renumbered"] == 2
            if name == "demo":
                assert np.array_equal(labels, DEMO_EXPECT)


def test_tile_reach_matches_level_kernels(tctx):
    """The same graph with tiles and without: identical labels, steps, colours, E/T/P
    bits and roots, through the speculative path and with waits (profiling forces
    the waits).  The F and B bits of vertices trimmed before the reach may differ:
    the level kernels' local rounds claim without reading the target's word."""
    g = scc.gen_mesh(200, 0.6, seed=9)
    out = {}
    for tile in (-1, 128):
        tctx.set_opts(scc.alg_opt(tile_reach=tile, relabel=-1))
        tctx.upload(g)
        for prof in (0, 2):
            tctx.profile(prof)
            try:
                lab, st = tctx.run("h", 0)
            finally:
                tctx.profile(0)
            words, roots = tctx.get_state()
            out[(tile, prof)] = (lab, st["steps"], words & np.uint32(0xFFFFFF1C), roots)
    base = out[(-1, 0)]
    for k, v in out.items():
        assert np.array_equal(v[0], base[0]) and v[1] == base[1], k
        assert np.array_equal(v[2], base[2]) and np.array_equal(v[3], base[3]), k


def test_tile_reach_default_large_and_batch(tctx):
    """tile_reach 0 (the default) tiles a mesh of 2^20 vertices (high diameter), -1
    keeps it untiled, 1 tiles it with full-size tiles; R-MAT (skewed) is numbered by
    degree, not tiled; a batch drops the tiles."""
    g = scc.gen_mesh(1024, 0.6, seed=10)
    ref = oracle.tarjan(g.row, g.col)[0]
    r = scc.gen_rmat(17, 8, 0.57, 0.19, 0.19, seed=12)
    tctx.set_opts(scc.alg_opt())
    tctx.upload(r)
    lab, st = tctx.run("h", 0)
    assert st["renumbered"] == 1 and np.array_equal(lab, oracle.tarjan(r.row, r.col)[0])
    for tile, want in ((0, 2), (-1, 0), (1, 2)):
        tctx.set_opts(scc.alg_opt(tile_reach=tile))
        tctx.upload(g)
        lab, st = tctx.run("h", 0)
        assert st["renumbered"] == want
        assert np.array_equal(lab, ref)
    b = scc.gen_batch(60, 64, 3000, 1.5, 6.0, 0.05, 0.5, seed=11)
    tctx.set_opts(scc.alg_opt(tile_reach=256))
    tctx.upload(b)
    tctx.set_batch(b.graph_offsets)
    lab, st = tctx.run("h", 0)
    assert np.array_equal(lab, oracle.tarjan(b.row, b.col)[0]) and st["renumbered"] == 1
