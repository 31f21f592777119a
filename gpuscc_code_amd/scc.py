"""Host-side mirror of the reference's interface for the FB-SCC path.

Reference names kept where they exist: ``reach_strategy`` and ``alg_opt``
(scc.h:7-27), ``load`` (load.cpp:18-207), ``computeSCC`` (scc.cu:150-1022),
the phase functions ``trim``/``fwd_reach``/``bwd_reach``/``update``/``trim2``/
``update_wcc``/``countElimmed`` (scc_routines.h, reach_*.h, wcc.h, scc.cu:42).
Everything below calls libgpuscc.so; the FB run itself executes only on the
GPU.  Vertex ids are 0-based (the reference's files are 1-based).
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import SccError, SccHostGraph, SccOpts, SccStats, check, lib

# reach_strategy (scc.h:7-16) restricted to the schedules this library provides
QUADRATIC, TOPO, DATA, LINEAR, HYBRID, COMBINE, BEST = 0, 1, 2, 3, 4, 5, 6
SCHEDULES = {"q": QUADRATIC, "t": TOPO, "d": DATA, "l": LINEAR, "h": HYBRID, "c": COMBINE, "b": BEST}
VARIANTS = {"scc": 0, "scc_w": 1, "scc_2_w": 2}
PIVOT_PLAIN, PIVOT_LEARNED = 0, 1

ST_B, ST_F, ST_E, ST_T, ST_P, ST_BX, ST_FX = 1, 2, 4, 8, 16, 32, 64
MLP_PARAMS, MLP_FEATURES, MLP_HIDDEN = 1505, 12, 32
LEVEL_FIELDS = ("kind", "level", "ms", "queue_f", "queue_b", "arcs_f", "arcs_b", "pulled",
                "expanded", "edges", "visited", "max_scan")  # one row of scc_profile_levels
COLOUR_MASK = 0x00FFFFFF
# default of alg_opt.relabel (0: the library decides by the degree skew); a test
# session may set it to 1 so that every upload renumbers its vertices
DEFAULT_RELABEL = 0


@dataclass
class alg_opt:  # noqa: N801 -- the reference's name (scc.h:19-27)
    alg: int = TOPO
    block_size: int = 256
    hybrid_policy: int = 0
    hybrid_threshold: int = 0
    slot_bits: int = 0
    grid_blocks: int = 0
    sequential_reach: int = 0
    huge_row: int = 0
    local_rounds: int = 0
    hybrid_arc_div: int = 0
    reach_graphs: int = 0
    reach_block: int = 0
    tail_live: int = 0
    batch_solver: int = 0
    relabel: int = None  # None: DEFAULT_RELABEL
    tile_reach: int = 0

    def to_c(self):
        o = SccOpts()
        o.block_size = self.block_size
        o.hybrid_policy = self.hybrid_policy
        o.hybrid_threshold = self.hybrid_threshold
        o.slot_bits = self.slot_bits
        o.grid_blocks = self.grid_blocks
        o.sequential_reach = self.sequential_reach
        o.huge_row = self.huge_row
        o.local_rounds = self.local_rounds
        o.hybrid_arc_div = self.hybrid_arc_div
        o.reach_graphs = self.reach_graphs
        o.reach_block = self.reach_block
        o.tail_live = self.tail_live
        o.batch_solver = self.batch_solver
        o.relabel = DEFAULT_RELABEL if self.relabel is None else self.relabel
        o.tile_reach = self.tile_reach
        return o


@dataclass
class Graph:
    """Host CSR + CSR^T (0-based, int32, sorted duplicate-free rows)."""
    n: int
    row: np.ndarray
    col: np.ndarray
    rowT: np.ndarray
    colT: np.ndarray
    m_reported: int = 0
    info: dict = field(default_factory=dict)
    graph_offsets: np.ndarray = None

    @property
    def m(self):
        return int(self.row[-1]) if len(self.row) else 0


def _from_host(g: SccHostGraph) -> Graph:
    n, m = g.n, g.m

    def arr(p, k):
        if k == 0:
            return np.zeros(0, np.int32)
        return np.ctypeslib.as_array(p, shape=(k,)).copy()

    out = Graph(n=n, row=arr(g.row, n + 1), col=arr(g.col, m), rowT=arr(g.rowT, n + 1),
                colT=arr(g.colT, m), m_reported=g.m_reported,
                info=dict(pattern=g.pattern, symmetric=g.symmetric, mtx=g.mtx, max_in=g.max_in,
                          min_in=g.min_in, max_out=g.max_out, min_out=g.min_out,
                          average=g.average))
    if g.num_graphs > 0:
        out.graph_offsets = arr(g.graph_offsets, g.num_graphs + 1)
    lib().scc_graph_free(ctypes.byref(g))
    return out


def load(path) -> Graph:
    """load.cpp:18-207: edge list / MatrixMarket, line "a b" = arc b->a."""
    g = SccHostGraph()
    check(lib().scc_load_file(str(path).encode(), ctypes.byref(g)))
    return _from_host(g)


def save(g: Graph, path):
    """Writes g as a binary CSR file (scc_save_graph) that load() reads without parsing text."""
    row = np.ascontiguousarray(g.row, np.int32)
    col = np.ascontiguousarray(g.col, np.int32)
    info = g.info or {}
    h = SccHostGraph()
    h.n, h.m = g.n, len(col)
    h.row = row.ctypes.data_as(_lib._i32p)
    h.col = col.ctypes.data_as(_lib._i32p) if len(col) else None
    h.m_reported = g.m_reported or len(col)
    for k in ("pattern", "symmetric", "mtx", "max_in", "min_in", "max_out", "min_out"):
        setattr(h, k, int(info.get(k, 0)))
    h.average = float(info.get("average", 0.0))
    off = None
    if g.graph_offsets is not None and len(g.graph_offsets) > 1:
        off = np.ascontiguousarray(g.graph_offsets, np.int32)
        h.num_graphs = len(off) - 1
        h.graph_offsets = off.ctypes.data_as(_lib._i32p)
    check(lib().scc_save_graph(ctypes.byref(h), str(path).encode()))


def csr_from_arcs(n, src, dst, drop_self_loops=True) -> Graph:
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    g = SccHostGraph()
    check(lib().scc_csr_from_arcs(int(n), len(src), src.ctypes.data_as(_lib._i32p),
                                  dst.ctypes.data_as(_lib._i32p), 1 if drop_self_loops else 0,
                                  ctypes.byref(g)))
    return _from_host(g)


def gen_rmat(scale, edge_factor=8, a=0.57, b=0.19, c=0.19, seed=1, permute=True) -> Graph:
    g = SccHostGraph()
    check(lib().scc_gen_rmat(scale, edge_factor, a, b, c, seed, 1 if permute else 0, ctypes.byref(g)))
    return _from_host(g)


def gen_mesh(side, p_bidir=0.6, seed=3) -> Graph:
    g = SccHostGraph()
    check(lib().scc_gen_mesh(side, p_bidir, seed, ctypes.byref(g)))
    return _from_host(g)


def gen_batch(num_graphs=4096, n_min=64, n_max=65536, deg_lo=1.5, deg_hi=6.0, cycle_p=0.05,
              cycle_cover=0.5, seed=4) -> Graph:
    g = SccHostGraph()
    check(lib().scc_gen_batch(num_graphs, n_min, n_max, deg_lo, deg_hi, cycle_p, cycle_cover, seed,
                              ctypes.byref(g)))
    return _from_host(g)


def _p(a, t):
    return a.ctypes.data_as(t)


class Context:
    """One device context (scc_create): a graph resident in HBM, FB runs on it."""

    def __init__(self, device=0, opt: alg_opt = None):
        self._h = ctypes.c_void_p()
        o = (opt or alg_opt()).to_c()
        check(lib().scc_create(device, ctypes.byref(o), ctypes.byref(self._h)))
        self.n = 0

    def close(self):
        if self._h and self._h.value:
            lib().scc_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, st):
        check(st, self._h)

    def set_opts(self, opt: alg_opt):
        """New run-time options for the next runs (slot_bits stays fixed once uploaded)."""
        o = opt.to_c()
        self._chk(lib().scc_set_opts(self._h, ctypes.byref(o)))

    def device_info(self):
        """(gfx target, compute units) of the context's device."""
        buf = ctypes.create_string_buffer(64)
        cus = ctypes.c_int32()
        self._chk(lib().scc_device_info(self._h, buf, 64, ctypes.byref(cus)))
        return buf.value.decode(), cus.value

    def upload(self, g: Graph, transpose=True):
        row = np.ascontiguousarray(g.row, np.int32)
        col = np.ascontiguousarray(g.col, np.int32)
        if transpose:
            rowT = np.ascontiguousarray(g.rowT, np.int32)
            colT = np.ascontiguousarray(g.colT, np.int32)
            pT, cT = _p(rowT, _lib._i32p), _p(colT, _lib._i32p)
        else:
            pT = cT = None
        self._chk(lib().scc_upload(self._h, g.n, len(col), _p(row, _lib._i32p),
                                   _p(col, _lib._i32p), pT, cT))
        self.n = g.n

    def set_batch(self, graph_offsets):
        """Treat the uploaded CSR as a batch of graphs (None or [] clears)."""
        if graph_offsets is None or len(graph_offsets) == 0:
            self._chk(lib().scc_set_batch(self._h, None, 0))
            return
        off = np.ascontiguousarray(graph_offsets, np.int32)
        self._chk(lib().scc_set_batch(self._h, _p(off, _lib._i32p), len(off) - 1))

    def run(self, schedule="t", variant=0, pivot=PIVOT_PLAIN, labels=True):
        sched = SCHEDULES[schedule] if isinstance(schedule, str) else int(schedule)
        var = VARIANTS[variant] if isinstance(variant, str) else int(variant)
        st = SccStats()
        out = np.empty(max(self.n, 1), np.int32) if labels else None
        self._chk(lib().scc_run(self._h, sched, var, pivot,
                                None if out is None else _p(out, _lib._i32p), ctypes.byref(st)))
        return (None if out is None else out[:self.n]), st.as_dict()

    # phase-level entry points
    def set_state(self, words, roots=None):
        w = np.ascontiguousarray(words, np.uint32)
        r = None if roots is None else np.ascontiguousarray(roots, np.int32)
        self._chk(lib().scc_state_set(self._h, _p(w, _lib._u32p),
                                      None if r is None else _p(r, _lib._i32p)))

    def labels_device(self):
        """Device address (int) of the last run's labels (scc_labels_device)."""
        ptr = _lib._i32p()
        self._chk(lib().scc_labels_device(self._h, ctypes.byref(ptr)))
        return ctypes.cast(ptr, ctypes.c_void_p).value or 0

    def get_state(self):
        w = np.empty(max(self.n, 1), np.uint32)
        r = np.empty(max(self.n, 1), np.int32)
        self._chk(lib().scc_state_get(self._h, _p(w, _lib._u32p), _p(r, _lib._i32p)))
        return w[:self.n], r[:self.n]

    def _i64call(self, fn, *args):
        x = ctypes.c_int64()
        self._chk(fn(self._h, *args, ctypes.byref(x)))
        return x.value

    def trim(self):
        return self._i64call(lib().scc_phase_trim)

    def reach(self, direction, schedule="q"):
        sched = SCHEDULES[schedule] if isinstance(schedule, str) else int(schedule)
        return self._i64call(lib().scc_phase_reach, direction, sched)

    def update(self, recolour=True):
        return self._i64call(lib().scc_phase_update, 1 if recolour else 0)

    def trim2(self):
        return self._i64call(lib().scc_phase_trim2)

    def update_wcc(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self._chk(lib().scc_phase_wcc(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def count_elimmed(self):
        return self._i64call(lib().scc_phase_count_elimmed)

    def recolour(self):
        self._chk(lib().scc_phase_recolour(self._h))

    def unset_pivots(self):
        """unsetPivots (scc_routines.cu:10-27): state 19 -> 0."""
        self._chk(lib().scc_phase_unset_pivots(self._h))

    def fail_alloc(self, nth):
        """Fault injection: the nth device allocation of the next upload fails (0 disarms)."""
        self._chk(lib().scc_test_fail_alloc(self._h, int(nth)))

    def features(self, idx, step=0):
        idx = np.ascontiguousarray(idx, np.int32)
        X = np.empty((max(len(idx), 1), MLP_FEATURES), np.float32)
        self._chk(lib().scc_phase_features(self._h, int(step), _p(idx, _lib._i32p), len(idx),
                                           _p(X, _lib._fp)))
        return X[:len(idx)]

    # learned pivot scorer
    def mlp_set(self, params):
        p = np.ascontiguousarray(params, np.float32)
        self._chk(lib().scc_mlp_set_params(self._h, _p(p, _lib._fp), len(p)))

    def mlp_get(self):
        p = np.empty(MLP_PARAMS, np.float32)
        self._chk(lib().scc_mlp_get_params(self._h, _p(p, _lib._fp), MLP_PARAMS))
        return p

    def mlp_infer(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.empty(max(len(X), 1), np.float32)
        self._chk(lib().scc_mlp_infer(self._h, _p(X, _lib._fp), len(X), _p(out, _lib._fp)))
        return out[:len(X)]

    def mlp_gradient(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.empty(MLP_PARAMS, np.float32)
        loss = ctypes.c_float()
        self._chk(lib().scc_mlp_gradient(self._h, _p(X, _lib._fp), _p(y, _lib._fp), len(X),
                                         _p(g, _lib._fp), ctypes.byref(loss)))
        return g, loss.value

    def mlp_train(self, X, y, epochs=10, batch=8192, lr=3e-3, seed=0):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        losses = np.empty(max(epochs, 1), np.float32)
        self._chk(lib().scc_mlp_train(self._h, _p(X, _lib._fp), _p(y, _lib._fp), len(X), epochs,
                                      batch, lr, seed, _p(losses, _lib._fp)))
        return losses[:epochs]

    # measurement: flags 1 = device work counters, 2 = HIP-event kernel timing,
    # 4 = per-level log (levels launched one at a time)
    def profile(self, flags):
        self._chk(lib().scc_profile_set(self._h, int(flags)))

    def profile_copy(self, nbytes):
        """One run of the calibration copy kernel (k_stream_copy) over nbytes."""
        self._chk(lib().scc_profile_copy(self._h, int(nbytes)))

    def profile_read(self, reset=True):
        p = _lib.SccProfile()
        self._chk(lib().scc_profile_get(self._h, ctypes.byref(p), 1 if reset else 0))
        return p.as_dict()

    def profile_levels(self, reset=True):
        """The level log of profile flag 4: array (rows, len(LEVEL_FIELDS))."""
        rows = ctypes.c_int64()
        self._chk(lib().scc_profile_levels(self._h, None, 0, ctypes.byref(rows), 0))
        out = np.zeros((rows.value, len(LEVEL_FIELDS)), np.float64)
        self._chk(lib().scc_profile_levels(self._h, _p(out, _lib._dp), rows.value, ctypes.byref(rows),
                                           1 if reset else 0))
        return out


def computeSCC(g: Graph, opt: alg_opt = None, variant=0, ctx: Context = None):  # noqa: N802
    """computeSCC (scc.cu:150-1022): returns (#SCC, seconds, labels, stats)."""
    opt = opt or alg_opt()
    own = ctx is None
    ctx = ctx or Context(0, opt)
    try:
        ctx.upload(g)
        labels, st = ctx.run(opt.alg, variant)
    finally:
        if own:
            ctx.close()
    return st["num_scc"], st["ms_total"] / 1000.0, labels, st


def make_word(colour, state):
    return ((np.asarray(colour, np.uint32) & COLOUR_MASK) << 8) | np.asarray(state, np.uint32)


def split_word(words):
    w = np.asarray(words, np.uint32)
    return (w >> 8).astype(np.uint32), (w & 0xFF).astype(np.uint8)


__all__ = ["alg_opt", "Graph", "Context", "computeSCC", "load", "csr_from_arcs", "gen_rmat",
           "gen_mesh", "gen_batch", "SccError", "QUADRATIC", "TOPO", "DATA", "LINEAR", "HYBRID", "COMBINE", "BEST",
           "SCHEDULES", "VARIANTS", "make_word", "split_word"]
