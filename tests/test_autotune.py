"""Autotuner: graph classes, cached-table lookup, knob validity (CPU) and a
small tuning run on the GPU that must keep the labels bit-exact."""
import json

import numpy as np
import pytest

import oracle
from gpuscc_code_amd import autotune, scc

from helpers import bigrid, cycle


def test_graph_class_shapes():
    assert autotune.graph_class(scc.gen_rmat(12, 8, seed=1)).startswith("skewed-")
    assert autotune.graph_class(bigrid(32)).startswith("regular-")
    assert autotune.graph_class(cycle(100), batch=True).startswith("batch-")
    # size = m rounded to a power of two
    g = cycle(1000)
    assert autotune.graph_class(g) == "regular-m2^10"


def test_knob_candidates_are_valid_options():
    for n in (0, 100, 4194304):
        for cus in (1, 256):
            for k, f in autotune.KNOBS.items():
                vals = f(n, cus)
                assert vals[0] == 0, "first candidate is the library default"
                for v in vals:
                    if k == "block_size":
                        assert v == 0 or (64 <= v <= 1024 and v % 64 == 0)
                    if k == "grid_blocks":
                        assert v >= 0
    assert set(autotune.ORDER) == set(autotune.KNOBS)


def test_opts_for_lookup(tmp_path):
    g = scc.gen_rmat(12, 8, seed=1)
    cls = autotune.graph_class(g)
    shape, size = cls.split("-m2^")
    table = {"gfx950": {
        autotune.class_key(cls, "h", 0): {"opts": {"grid_blocks": 512, "local_rounds": 16}},
        autotune.class_key(f"{shape}-m2^{int(size) + 2}", "h", 0): {"opts": {"grid_blocks": 2048}},
        autotune.class_key(f"regular-m2^{size}", "h", 0): {"opts": {"huge_row": 1024}},
        autotune.class_key(cls, "d", 0): {"opts": {"block_size": 128}},
    }}
    path = tmp_path / "tuned.json"
    path.write_text(json.dumps(table))
    opt, key = autotune.opts_for(g, "gfx950", "h", 0, path=str(path))
    assert key == autotune.class_key(cls, "h", 0)
    assert opt.grid_blocks == 512 and opt.local_rounds == 16 and opt.huge_row == 0
    # other schedule, other target, no entry within a factor 4 in m: defaults
    opt, key = autotune.opts_for(g, "gfx950", "d", 0, path=str(path))
    assert opt.block_size == 128
    opt, key = autotune.opts_for(g, "gfx942", "h", 0, path=str(path))
    assert key is None and opt == scc.alg_opt()
    opt, key = autotune.opts_for(g, "gfx950", "h", 1, path=str(path))
    assert key is None
    # nearest size of the same shape
    del table["gfx950"][autotune.class_key(cls, "h", 0)]
    path.write_text(json.dumps(table))
    opt, key = autotune.opts_for(g, "gfx950", "h", 0, path=str(path))
    assert opt.grid_blocks == 2048


def test_committed_table_is_well_formed():
    table = autotune.load_table()
    fields = set(scc.alg_opt().__dict__)
    for arch, entries in table.items():
        assert arch.startswith("gfx")
        for key, ent in entries.items():
            cls, sched, var = key.split("/")
            assert sched in scc.SCHEDULES and int(var) in (0, 1, 2)
            assert set(ent["opts"]) <= fields
            assert ent["ms"] <= ent["default_ms"]


@pytest.mark.gpu
def test_set_opts_and_device_info(ctx):
    arch, cus = ctx.device_info()
    assert arch.startswith("gfx") and cus > 0
    g = scc.gen_rmat(12, 8, seed=3)
    ref, _ = oracle.tarjan(g.row, g.col)
    ctx.upload(g)
    for o in (scc.alg_opt(grid_blocks=64, block_size=128), scc.alg_opt(hybrid_arc_div=-1, local_rounds=-1),
              scc.alg_opt(huge_row=64, hybrid_threshold=1)):
        ctx.set_opts(o)
        for sched in ("t", "h"):
            labels, _ = ctx.run(sched)
            assert np.array_equal(labels, ref)
    with pytest.raises(scc.SccError):
        ctx.set_opts(scc.alg_opt(block_size=100))
    with pytest.raises(scc.SccError):
        ctx.set_opts(scc.alg_opt(slot_bits=3))  # fixed once uploaded
    ctx.set_opts(scc.alg_opt())


@pytest.mark.gpu
def test_tune_small_keeps_labels(ctx):
    g = scc.gen_rmat(13, 8, seed=4)
    ref, _ = oracle.tarjan(g.row, g.col)
    ctx.upload(g)
    res = autotune.tune(ctx, g, "h", 0, reps=1, rounds=1, knobs=["grid_blocks", "local_rounds"])
    assert res["ms"] <= res["default_ms"] and res["trials"] >= 1
    assert set(res["opts"]) <= {"grid_blocks", "local_rounds"}
    labels, _ = ctx.run("h")  # the context keeps the tuned options
    assert np.array_equal(labels, ref)
    ctx.set_opts(scc.alg_opt())
