"""ctypes binding of libgpuscc.so (the C ABI of include/gpuscc.h).

The shared library holds the HIP kernels for gfx950 and the host input path.
Loading fails loudly when it has not been built: there is no Python or CPU
fallback for any GPU entry point.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libgpuscc.so")

SCC_OK, SCC_E_HIP, SCC_E_OOM, SCC_E_INPUT, SCC_E_OVERFLOW, SCC_E_STATE = range(6)
STATUS_NAMES = {0: "SCC_OK", 1: "SCC_E_HIP", 2: "SCC_E_OOM", 3: "SCC_E_INPUT",
                4: "SCC_E_OVERFLOW", 5: "SCC_E_STATE"}


class SccError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")


class SccOpts(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_int32), ("hybrid_policy", ctypes.c_int32),
                ("hybrid_threshold", ctypes.c_int32), ("slot_bits", ctypes.c_int32),
                ("grid_blocks", ctypes.c_int32), ("sequential_reach", ctypes.c_int32),
                ("huge_row", ctypes.c_int32), ("local_rounds", ctypes.c_int32),
                ("hybrid_arc_div", ctypes.c_int32), ("reach_graphs", ctypes.c_int32),
                ("reach_block", ctypes.c_int32),
                ("tail_live", ctypes.c_int32), ("batch_solver", ctypes.c_int32),
                ("relabel", ctypes.c_int32), ("tile_reach", ctypes.c_int32)]


_STAT_INTS = ["steps", "fwd_depths", "bwd_depths", "trim_iters", "num_scc", "pivots_total",
              "wcc_step", "trim2_pairs", "trivial", "biggest", "host_syncs", "first_pivot",
              "tail_steps", "tail_live", "renumbered"]
_STAT_DBLS = ["ms_total", "ms_fwd", "ms_bwd", "ms_trim", "ms_update"]


class SccStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in _STAT_INTS] + [(k, ctypes.c_double) for k in _STAT_DBLS]

    def as_dict(self):
        return {k: getattr(self, k) for k in _STAT_INTS + _STAT_DBLS}


_i32p = ctypes.POINTER(ctypes.c_int32)

_PROF_INTS = ["reach_launches", "reach_levels", "sweep_levels", "expanded", "edges", "visited",
              "queue_levels", "trim_launches"]
_PROF_DBLS = ["reach_kernel_ms", "trim_kernel_ms"]
_PROF_TAIL = ["tile_runs", "tile_rounds"]


class SccProfile(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int64) for k in _PROF_INTS] + [(k, ctypes.c_double) for k in _PROF_DBLS]
                + [(k, ctypes.c_int64) for k in _PROF_TAIL] + [("tile_phase_ms", ctypes.c_double * 4),
                                                               ("tile_busy_ms", ctypes.c_double),
                                                               ("tile_idle_polls", ctypes.c_int64),
                                                               ("tail_phase_ms", ctypes.c_double * 4),
                                                               ("tail_local_live", ctypes.c_int64),
                                                               ("tail_local_step", ctypes.c_int64)])

    def as_dict(self):
        d = {k: getattr(self, k) for k in _PROF_INTS + _PROF_DBLS + _PROF_TAIL
             + ["tile_busy_ms", "tile_idle_polls", "tail_local_live", "tail_local_step"]}
        d["tile_phase_ms"] = list(self.tile_phase_ms)
        d["tail_phase_ms"] = list(self.tail_phase_ms)
        return d


class SccHostGraph(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("m", ctypes.c_int64),
                ("row", _i32p), ("col", _i32p), ("rowT", _i32p), ("colT", _i32p),
                ("m_reported", ctypes.c_int64),
                ("pattern", ctypes.c_int32), ("symmetric", ctypes.c_int32), ("mtx", ctypes.c_int32),
                ("max_in", ctypes.c_int32), ("min_in", ctypes.c_int32),
                ("max_out", ctypes.c_int32), ("min_out", ctypes.c_int32),
                ("average", ctypes.c_double),
                ("num_graphs", ctypes.c_int32), ("graph_offsets", _i32p)]


# every symbol declared in include/gpuscc.h with its ctypes signature
_vp = ctypes.c_void_p
_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_gp = ctypes.POINTER(SccHostGraph)
_st = ctypes.c_int
SIGNATURES = {
    "scc_create": (_st, [ctypes.c_int, ctypes.POINTER(SccOpts), ctypes.POINTER(_vp)]),
    "scc_destroy": (None, [_vp]),
    "scc_set_opts": (_st, [_vp, ctypes.POINTER(SccOpts)]),
    "scc_device_info": (_st, [_vp, ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "scc_status_string": (ctypes.c_char_p, [_st]),
    "scc_last_error": (ctypes.c_char_p, [_vp]),
    "scc_upload": (_st, [_vp, ctypes.c_int32, ctypes.c_int64, _i32p, _i32p, _i32p, _i32p]),
    "scc_run": (_st, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, ctypes.POINTER(SccStats)]),
    "scc_labels_device": (_st, [_vp, ctypes.POINTER(_i32p)]),
    "scc_set_batch": (_st, [_vp, _i32p, ctypes.c_int32]),
    "scc_state_set": (_st, [_vp, _u32p, _i32p]),
    "scc_state_get": (_st, [_vp, _u32p, _i32p]),
    "scc_phase_trim": (_st, [_vp, _i64p]),
    "scc_phase_reach": (_st, [_vp, ctypes.c_int, ctypes.c_int, _i64p]),
    "scc_phase_update": (_st, [_vp, ctypes.c_int, _i64p]),
    "scc_phase_trim2": (_st, [_vp, _i64p]),
    "scc_phase_wcc": (_st, [_vp, _i64p, _i64p]),
    "scc_phase_count_elimmed": (_st, [_vp, _i64p]),
    "scc_phase_recolour": (_st, [_vp]),
    "scc_phase_unset_pivots": (_st, [_vp]),
    "scc_test_fail_alloc": (_st, [_vp, ctypes.c_int32]),
    "scc_run_noalloc": (_st, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                              ctypes.c_int64, ctypes.c_void_p, _i64p, ctypes.c_void_p]),
    "scc_phase_features": (_st, [_vp, ctypes.c_int32, _i32p, ctypes.c_int64, _fp]),
    "scc_mlp_set_params": (_st, [_vp, _fp, ctypes.c_int32]),
    "scc_mlp_get_params": (_st, [_vp, _fp, ctypes.c_int32]),
    "scc_mlp_infer": (_st, [_vp, _fp, ctypes.c_int64, _fp]),
    "scc_mlp_gradient": (_st, [_vp, _fp, _fp, ctypes.c_int64, _fp, _fp]),
    "scc_mlp_train": (_st, [_vp, _fp, _fp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                            ctypes.c_float, ctypes.c_uint64, _fp]),
    "scc_profile_set": (_st, [_vp, ctypes.c_int]),
    "scc_profile_get": (_st, [_vp, ctypes.POINTER(SccProfile), ctypes.c_int]),
    "scc_profile_copy": (_st, [_vp, ctypes.c_int64]),
    "scc_profile_levels": (_st, [_vp, _dp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), ctypes.c_int]),
    "scc_load_file": (_st, [ctypes.c_char_p, _gp]),
    "scc_save_graph": (_st, [_gp, ctypes.c_char_p]),
    "scc_csr_from_arcs": (_st, [ctypes.c_int32, ctypes.c_int64, _i32p, _i32p, ctypes.c_int, _gp]),
    "scc_gen_rmat": (_st, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                           ctypes.c_double, ctypes.c_uint64, ctypes.c_int, _gp]),
    "scc_gen_mesh": (_st, [ctypes.c_int, ctypes.c_double, ctypes.c_uint64, _gp]),
    "scc_gen_batch": (_st, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _gp]),
    "scc_graph_free": (None, [_gp]),
}

_lib = None


def lib():
    """Load libgpuscc.so; raises if it is missing (build with __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, ctx=None):
    if status != SCC_OK:
        msg = ""
        if ctx is not None and ctx.value:
            msg = (lib().scc_last_error(ctx) or b"").decode(errors="replace")
        raise SccError(status, msg)
