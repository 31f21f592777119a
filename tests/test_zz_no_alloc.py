"""scc_run_noalloc: the decomposition that allocates nothing, through the C
ABI, against scc_run (the context entry point) on the same graphs: labels
bit-exact and the same SCC count, on calls in a row whose inputs shrink (the
shared device array keeps the last call's contents), on a graph whose
elections cross a 16-bit election-tag wrap over successive calls, and the
size and input checks.  The test puts the CSR pair in device memory with the
HIP runtime the library itself links (/opt/rocm/lib/libamdhip64.so.7)."""
import ctypes

import numpy as np
import pytest

import oracle
from gpuscc_code_amd import _lib, scc

from helpers import planted

pytestmark = pytest.mark.gpu

MAX_N = 262144  # SCC_NOALLOC_MAX_VERTICES
_hip = None


def hip():
    global _hip
    if _hip is None:
        _lib.lib()  # the library first: the runtime below is the one it loaded
        _hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    return _hip


class DevBuf:
    """int32 device buffer: hipMalloc + hipMemcpy (kind 1 host->device, 2 device->host)."""

    def __init__(self, a=None, count=None):
        a = None if a is None else np.ascontiguousarray(a, np.int32)
        self.count = len(a) if a is not None else count
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(4 * max(self.count, 1))) == 0
        if a is not None and len(a):
            assert hip().hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(4 * len(a)), 1) == 0

    def get(self):
        out = np.empty(max(self.count, 1), np.int32)
        assert hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(4 * self.count), 2) == 0
        return out[:self.count]

    def __del__(self):
        if self.p:
            hip().hipFree(self.p)


def noalloc(g, n=None, row=None, labels_len=None):
    """Calls scc_run_noalloc on g's CSR pair in device memory; (status, labels, count)."""
    n = g.n if n is None else n
    bufs = [DevBuf(g.row if row is None else row), DevBuf(g.col), DevBuf(g.rowT), DevBuf(g.colT)]
    lab = DevBuf(np.full(max(labels_len or g.n, 1), -7, np.int32))
    cnt = ctypes.c_int64(-1)
    print(f"[noalloc] n={n} m={len(g.col)}", flush=True)
    st = _lib.lib().scc_run_noalloc(*[b.p for b in bufs], n, len(g.col), lab.p, ctypes.byref(cnt), None)
    return st, lab.get()[:g.n], cnt.value


def check_same_as_run(ctx, g):
    st, lab, cnt = noalloc(g)
    assert st == 0, f"status {st}"
    ctx.upload(g)
    want, stats = ctx.run("h")
    assert np.array_equal(lab, want)
    assert cnt == stats["num_scc"] == len(np.unique(want))
    return lab


def test_noalloc_matches_scc_run_on_shrinking_inputs(ctx):
    big = scc.gen_rmat(14, 8, 0.57, 0.19, 0.19, seed=1)   # 16,384 vertices (config 1)
    small = scc.gen_mesh(60, 0.6, seed=3)                # 3,600 vertices, many FB steps
    tiny = planted(50, 3, 80, seed=4)[0]                 # 150 vertices
    for g in (big, small, tiny, big):
        lab = check_same_as_run(ctx, g)
        ref, _ = oracle.tarjan(g.row, g.col)
        assert np.array_equal(lab, ref)


def test_noalloc_largest_input_and_limits(ctx):
    g = scc.gen_rmat(18, 4, 0.57, 0.19, 0.19, seed=6)    # 262,144 vertices: the maximum
    assert g.n == MAX_N
    check_same_as_run(ctx, g)
    # beyond the maximum: refused before anything runs
    st, _, _ = noalloc(g, n=MAX_N + 1, labels_len=MAX_N + 1, row=np.concatenate([g.row, [g.row[-1]]]))
    assert st == 4
    # offsets that do not end at m, ids out of range: the set-up kernel's checks
    h = planted(20, 3, 10, seed=5)[0]
    bad_row = h.row.copy()
    bad_row[-1] -= 1
    st, _, _ = noalloc(h, row=bad_row)
    assert st == 3
    st, lab, cnt = noalloc(h)  # a good call afterwards still works
    assert st == 0 and np.array_equal(lab, oracle.tarjan(h.row, h.col)[0])


def test_noalloc_tag_wrap_across_calls(ctx):
    """Election tags are 16 bits and go on from call to call: enough calls on a
    graph of 810 FB steps cross a wrap, which the kernel handles by clearing
    the next election table; every call's labels stay exact."""
    g = planted(2000, 2, 2500, seed=3)[0]
    ref, _ = oracle.tarjan(g.row, g.col)
    ctx.upload(g)
    want, stats = ctx.run("h")
    assert stats["steps"] >= 100
    calls = (1 << 16) // stats["steps"] + 3
    for k in range(calls):
        st, lab, cnt = noalloc(g)
        assert st == 0 and np.array_equal(lab, ref) and cnt == stats["num_scc"], f"call {k}"
