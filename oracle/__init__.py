"""Oracle for the FB-SCC path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this package, and only as the checker / the timed CPU baseline.  The
product (``gpuscc_code_amd``: the HIP library and the ``scc`` CLI) never
imports, links or calls it.

``scc_oracle.c`` restates the reference's algorithm in plain C (each function
cites the reference file:line it follows); this module binds it with ctypes and
adds numpy helpers.  ``ref_load`` binds the reference's own ``load.cpp`` built
into ``oracle/_ref/`` (only where /root/reference exists).

Pinning: every function here is checked in tests/test_oracle.py against the
known answer of the reference's ``demo.graph``, golden labels written by scipy
and networkx (tests/golden/), brute-force reachability closure and closed forms
(cycles, DAGs, complete digraphs, bidirectional grids, planted cycles).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref_load.so")

ST_B, ST_F, ST_E, ST_T, ST_P, ST_BX, ST_FX = 1, 2, 4, 8, 16, 32, 64
COLOUR_MASK_REF = 0xFFFFFFFF
COLOUR_MASK_PACKED = 0x00FFFFFF

_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i64p = ctypes.POINTER(ctypes.c_int64)


class FBStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("num_scc", ctypes.c_int64),
                ("pivots_total", ctypes.c_int64), ("wcc_step", ctypes.c_int64),
                ("trimmed", ctypes.c_int64), ("trim2_pairs", ctypes.c_int64)]


def build():
    """Compile liboracle.so (and the reference loader when /root/reference exists)."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_tarjan.restype = ctypes.c_int64
        L.orc_tarjan.argtypes = [ctypes.c_int32, _i32p, _i32p, _i32p, _i64p, _i64p, _i64p,
                                 ctypes.POINTER(ctypes.c_double)]
        L.orc_canonicalize.argtypes = [ctypes.c_int32, _i32p, _i32p]
        L.orc_brute_scc.argtypes = [ctypes.c_int32, _i32p, _i32p, _i32p]
        L.orc_trim.restype = ctypes.c_int64
        L.orc_trim.argtypes = [ctypes.c_int32, _i32p, _i32p, _i32p, _i32p, _u32p, _u8p, _i32p,
                               ctypes.c_int]
        L.orc_reach.restype = ctypes.c_int64
        L.orc_reach.argtypes = [ctypes.c_int32, _i32p, _i32p, _u32p, _u8p, ctypes.c_int]
        L.orc_update_elect.restype = ctypes.c_int64
        L.orc_update_elect.argtypes = [ctypes.c_int32, _u32p, _u8p, _i32p, ctypes.c_uint32,
                                       ctypes.c_int, ctypes.c_int, _i32p, _i32p, ctypes.c_int]
        L.orc_count_elimmed.restype = ctypes.c_int64
        L.orc_count_elimmed.argtypes = [ctypes.c_int32, _u8p]
        L.orc_trim2.restype = ctypes.c_int64
        L.orc_trim2.argtypes = [ctypes.c_int32, _i32p, _i32p, _i32p, _i32p, _u32p, _u8p, _i32p]
        L.orc_wcc_labels.argtypes = [ctypes.c_int32, _i32p, _i32p, _u32p, _u8p, _i32p]
        L.orc_wcc_recolour.restype = ctypes.c_int64
        L.orc_wcc_recolour.argtypes = [ctypes.c_int32, _u32p, _u8p, _i32p, ctypes.c_uint32,
                                       ctypes.c_uint32]
        L.orc_fb.restype = ctypes.c_int64
        L.orc_fb.argtypes = [ctypes.c_int32, _i32p, _i32p, _i32p, _i32p, _u32p, ctypes.c_uint32,
                             ctypes.c_int, ctypes.c_int, ctypes.c_int, _u32p, _u8p, _i32p,
                             ctypes.POINTER(FBStats)]
        L.orc_hash_slot.restype = ctypes.c_uint32
        L.orc_hash_slot.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def tarjan(row, col):
    """(labels canonical = min id, stats dict).  Restates tarjan.cpp:72-240."""
    row, col = _i32(row), _i32(col)
    n = len(row) - 1
    root = np.empty(max(n, 1), np.int32)
    triv, nontriv, big = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    sec = ctypes.c_double()
    cnt = lib().orc_tarjan(n, _p(row, _i32p), _p(col, _i32p), _p(root, _i32p), ctypes.byref(triv),
                           ctypes.byref(nontriv), ctypes.byref(big), ctypes.byref(sec))
    root = root[:n]
    return canonicalize(root), dict(num_scc=int(cnt), trivial=triv.value,
                                    nontrivial=nontriv.value, biggest=big.value,
                                    seconds=sec.value)


def tarjan_time(row, col):
    """Seconds of the DFS only (tarjan.cpp:108->192) and the SCC count."""
    row, col = _i32(row), _i32(col)
    n = len(row) - 1
    root = np.empty(max(n, 1), np.int32)
    sec = ctypes.c_double()
    cnt = lib().orc_tarjan(n, _p(row, _i32p), _p(col, _i32p), _p(root, _i32p), None, None, None,
                           ctypes.byref(sec))
    return sec.value, int(cnt)


def canonicalize(root):
    root = _i32(root)
    n = len(root)
    out = np.empty(max(n, 1), np.int32)
    lib().orc_canonicalize(n, _p(root, _i32p), _p(out, _i32p))
    return out[:n]


def brute_scc(row, col):
    row, col = _i32(row), _i32(col)
    n = len(row) - 1
    out = np.empty(max(n, 1), np.int32)
    lib().orc_brute_scc(n, _p(row, _i32p), _p(col, _i32p), _p(out, _i32p))
    return out[:n]


def trim(row, col, rowT, colT, colour, state, root, colour_blind=False):
    row, col, rowT, colT = _i32(row), _i32(col), _i32(rowT), _i32(colT)
    n = len(row) - 1
    return lib().orc_trim(n, _p(row, _i32p), _p(col, _i32p), _p(rowT, _i32p), _p(colT, _i32p),
                          _p(colour, _u32p), _p(state, _u8p), _p(root, _i32p),
                          1 if colour_blind else 0)


def reach(row, col, colour, state, bit):
    row, col = _i32(row), _i32(col)
    n = len(row) - 1
    return lib().orc_reach(n, _p(row, _i32p), _p(col, _i32p), _p(colour, _u32p),
                           _p(state, _u8p), bit)


def update_elect(colour, state, root, colour_mask, slot_bits, hash_kind, prev_winner, recolour):
    n = len(colour)
    winner = np.empty(1 << slot_bits, np.int32)
    pw = None if prev_winner is None else _p(prev_winner, _i32p)
    np_ = lib().orc_update_elect(n, _p(colour, _u32p), _p(state, _u8p), _p(root, _i32p),
                                 colour_mask, slot_bits, hash_kind, pw, _p(winner, _i32p),
                                 1 if recolour else 0)
    return np_, winner


def trim2(row, col, rowT, colT, colour, state, root):
    row, col, rowT, colT = _i32(row), _i32(col), _i32(rowT), _i32(colT)
    n = len(row) - 1
    return lib().orc_trim2(n, _p(row, _i32p), _p(col, _i32p), _p(rowT, _i32p), _p(colT, _i32p),
                           _p(colour, _u32p), _p(state, _u8p), _p(root, _i32p))


def wcc_labels(rowT, colT, colour, state):
    rowT, colT = _i32(rowT), _i32(colT)
    n = len(rowT) - 1
    lab = np.empty(max(n, 1), np.int32)
    lib().orc_wcc_labels(n, _p(rowT, _i32p), _p(colT, _i32p), _p(colour, _u32p),
                         _p(state, _u8p), _p(lab, _i32p))
    return lab[:n]


def hash_slot(c, slot_bits, kind):
    return int(lib().orc_hash_slot(c, slot_bits, kind))


def fb(row, col, rowT, colT, variant=0, init_colour=None, colour_mask=COLOUR_MASK_PACKED,
       slot_bits=20, hash_kind=1):
    """Whole FB run (scc.cu:150-1022 restated).  Returns dict with canonical labels,
    final colour/state arrays, root representatives and stats."""
    row, col, rowT, colT = _i32(row), _i32(col), _i32(rowT), _i32(colT)
    n = len(row) - 1
    colour = np.empty(max(n, 1), np.uint32)
    state = np.empty(max(n, 1), np.uint8)
    root = np.empty(max(n, 1), np.int32)
    ic = None
    if init_colour is not None:
        ic = np.ascontiguousarray(init_colour, dtype=np.uint32)
    st = FBStats()
    lib().orc_fb(n, _p(row, _i32p), _p(col, _i32p), _p(rowT, _i32p), _p(colT, _i32p),
                 None if ic is None else _p(ic, _u32p), colour_mask, slot_bits, hash_kind,
                 variant, _p(colour, _u32p), _p(state, _u8p), _p(root, _i32p), ctypes.byref(st))
    root = root[:n]
    return dict(labels=canonicalize(root), colour=colour[:n], state=state[:n], root=root,
                steps=st.steps, num_scc=st.num_scc, pivots_total=st.pivots_total,
                wcc_step=st.wcc_step, trimmed=st.trimmed, trim2_pairs=st.trim2_pairs)


# ---------------------------------------------------------------------------
# the reference's own loader (oracle/_ref), present only where it was built
# ---------------------------------------------------------------------------
_ref = None


def ref_available():
    return os.path.exists(_REF_PATH)


def ref_load(path):
    """Run the reference's load() on `path`; returns 1-based arrays exactly as it
    builds them: (n, m_counted, nodes[1..n+1], edges, nodesT[1..n+1], edgesT)."""
    global _ref
    if _ref is None:
        R = ctypes.CDLL(_REF_PATH)
        pp = ctypes.POINTER(ctypes.POINTER(ctypes.c_uint))
        R.ref_load.restype = ctypes.c_int
        R.ref_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint),
                               ctypes.POINTER(ctypes.c_uint), pp, pp, pp, pp]
        R.ref_free.argtypes = [ctypes.c_void_p]
        _ref = R
    n, m = ctypes.c_uint(), ctypes.c_uint()
    a, b, c, d = (ctypes.POINTER(ctypes.c_uint)() for _ in range(4))
    ok = _ref.ref_load(path.encode(), ctypes.byref(n), ctypes.byref(m), ctypes.byref(a),
                       ctypes.byref(b), ctypes.byref(c), ctypes.byref(d))
    if not ok:
        raise RuntimeError("reference load() failed")
    N = n.value
    nodes = np.ctypeslib.as_array(a, shape=(N + 2,)).astype(np.int64).copy()
    E = int(nodes[N + 1])
    edges = np.ctypeslib.as_array(b, shape=(max(E, 1),))[:E].astype(np.int64).copy()
    nodesT = np.ctypeslib.as_array(c, shape=(N + 2,)).astype(np.int64).copy()
    ET = int(nodesT[N + 1])
    edgesT = np.ctypeslib.as_array(d, shape=(max(ET, 1),))[:ET].astype(np.int64).copy()
    same = ctypes.addressof(a.contents) == ctypes.addressof(c.contents)
    _ref.ref_free(ctypes.cast(a, ctypes.c_void_p))
    _ref.ref_free(ctypes.cast(b, ctypes.c_void_p))
    if not same:
        _ref.ref_free(ctypes.cast(c, ctypes.c_void_p))
        _ref.ref_free(ctypes.cast(d, ctypes.c_void_p))
    return N, m.value, nodes, edges, nodesT, edgesT
