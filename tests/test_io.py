"""Host input path: the load.cpp-compatible parser and the CSR builder.

The parser is compared with the reference's own load() (load.cpp:18-207),
compiled from the reference sources into oracle/_ref/ where the reference is
present, and with the format rules restated by hand on small files.
"""
import os

import numpy as np
import pytest

import oracle
from gpuscc_code_amd import scc

HERE = os.path.dirname(os.path.abspath(__file__))
DEMO = os.path.join(HERE, "golden", "demo.graph")


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def to_ref_layout(g):
    """0-based CSR -> the reference's 1-based arrays (nodes[1..n+1], edges 1-based)."""
    return g.row.astype(np.int64), g.col.astype(np.int64) + 1, g.rowT.astype(np.int64), \
        g.colT.astype(np.int64) + 1


def assert_same_as_reference(path):
    if not oracle.ref_available():
        pytest.skip("reference loader not built (needs /root/reference)")
    N, m, nodes, edges, nodesT, edgesT = oracle.ref_load(path)
    g = scc.load(path)
    row, col, rowT, colT = to_ref_layout(g)
    assert g.n == N
    assert g.m_reported == m
    assert np.array_equal(nodes[1:], row)
    assert np.array_equal(edges, col)
    assert np.array_equal(nodesT[1:], rowT)
    assert np.array_equal(edgesT, colT)


def test_demo_graph_matches_reference_loader():
    assert_same_as_reference(DEMO)


def test_demo_graph_known_csr():
    g = scc.load(DEMO)
    assert g.n == 13 and g.m == 19 and g.m_reported == 19
    # "1 1" is a self-loop on vertex 1 (kept), "1 12" is the arc 12 -> 1
    assert 0 in g.col[g.row[0]:g.row[1]]
    assert 0 in g.col[g.row[11]:g.row[12]]
    assert g.info["max_out"] == 3 and g.info["min_out"] == 1
    assert g.info["max_in"] == 3 and g.info["min_in"] == 0


@pytest.mark.parametrize("seed", range(6))
def test_random_edge_lists_match_reference_loader(tmp_path, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 200))
    nnz = int(rng.integers(0, 600))
    a = rng.integers(1, n + 1, nnz)
    b = rng.integers(1, n + 1, nnz)
    lines = [f"{n} {nnz}"] + [f"{x} {y}" for x, y in zip(a, b)]
    p = write(tmp_path, "g.txt", "\n".join(lines) + "\n")
    assert_same_as_reference(p)
    g = scc.load(p)
    # arcs are b -> a, duplicates dropped, self-loops kept
    expect = sorted(set(zip((b - 1).tolist(), (a - 1).tolist())))
    got = [(v, int(u)) for v in range(n) for u in g.col[g.row[v]:g.row[v + 1]]]
    assert got == expect


def test_matrix_market_general_and_symmetric(tmp_path):
    gen = ("%%MatrixMarket matrix coordinate pattern general\n% comment\n4 4 5\n"
           "1 2\n2 3\n3 1\n4 4\n4 1\n")
    p = write(tmp_path, "a.mtx", gen)
    assert_same_as_reference(p)
    g = scc.load(p)
    assert g.info["pattern"] == 1 and g.info["symmetric"] == 0 and g.info["mtx"] == 1
    sym = ("%%MatrixMarket matrix coordinate real symmetric\n4 4 4\n1 2 0.5\n3 2 1.0\n4 4 2.0\n"
           "4 1 3.0\n")
    p = write(tmp_path, "b.mtx", sym)
    assert_same_as_reference(p)
    g = scc.load(p)
    assert g.info["symmetric"] == 1
    assert g.m_reported == 4 + 3  # load.cpp:83-86: each off-diagonal entry adds its mirror
    assert np.array_equal(g.row, g.rowT) and np.array_equal(g.col, g.colT)


@pytest.mark.parametrize("text,name", [
    ("3 3 2\n1 2\n2 3\n", "bad.mtx"),          # m == n required... here 3 3 ok but nnz 2 lines: fine
])
def test_mtx_ok(tmp_path, text, name):
    g = scc.load(write(tmp_path, name, text))
    assert g.n == 3 and g.m == 2


@pytest.mark.parametrize("text,name", [
    ("3 4 2\n1 2\n2 3\n", "rect.mtx"),          # not square: the reference exits (load.cpp:49-52)
    ("3 5\n1 2\n2 3\n", "short.txt"),           # file ends before nnz entries
    ("3 2\n1 2\n2 9\n", "range.txt"),           # vertex id above n
    ("3 2\n1 2\n0 1\n", "zero.txt"),            # vertex id 0 (ids are 1-based)
    ("3 2\n1 2\nx y\n", "junk.txt"),            # unparsable entry
    ("", "empty.txt"),
])
def test_malformed_inputs_are_errors(tmp_path, text, name):
    with pytest.raises(scc.SccError) as e:
        scc.load(write(tmp_path, name, text))
    assert e.value.status == 3  # SCC_E_INPUT


def test_missing_file():
    with pytest.raises(scc.SccError):
        scc.load("/nonexistent/graph.txt")


def test_csr_from_arcs_dedup_and_transpose():
    src = np.array([0, 0, 1, 2, 2, 2, 3, 3], np.int32)
    dst = np.array([1, 1, 2, 0, 2, 3, 3, 0], np.int32)
    g = scc.csr_from_arcs(4, src, dst, drop_self_loops=True)
    assert list(g.row) == [0, 1, 2, 4, 5]
    assert list(g.col) == [1, 2, 0, 3, 0]
    assert list(g.rowT) == [0, 2, 3, 4, 5]
    assert list(g.colT) == [2, 3, 0, 1, 2]
    g2 = scc.csr_from_arcs(4, src, dst, drop_self_loops=False)
    assert g2.m == 7
    with pytest.raises(scc.SccError):
        scc.csr_from_arcs(3, np.array([0, 5], np.int32), np.array([1, 1], np.int32))


def test_csr_from_arcs_large_random_matches_numpy():
    rng = np.random.default_rng(0)
    n, m = 200000, 1500000
    s = rng.integers(0, n, m).astype(np.int32)
    d = rng.integers(0, n, m).astype(np.int32)
    g = scc.csr_from_arcs(n, s, d, drop_self_loops=True)
    keep = s != d
    key = np.unique(s[keep].astype(np.int64) * n + d[keep])
    assert g.m == len(key)
    assert np.array_equal(g.col, (key % n).astype(np.int32))
    assert np.array_equal(g.row, np.searchsorted(key // n, np.arange(n + 1)).astype(np.int32))
    keyT = np.sort((key % n) * n + key // n)
    assert np.array_equal(g.colT, (keyT % n).astype(np.int32))


# ---------------------------------------------------------------------------
# parallel entry parsing: the text after the header is cut into chunks at line
# ends (one per host thread, at least 4 KiB each); results and error positions
# must not depend on the cut
# ---------------------------------------------------------------------------
def load_with_threads(path, threads, monkeypatch):
    monkeypatch.setenv("GPUSCC_HOST_THREADS", str(threads))
    return scc.load(path)


def big_edge_list(tmp_path, rng, n=5000, nnz=60000, name="big.txt", crlf=False, trailing_nl=True,
                  extra=""):
    a = rng.integers(1, n + 1, nnz)
    b = rng.integers(1, n + 1, nnz)
    sep = "\r\n" if crlf else "\n"
    # vary the white space the way sscanf accepts it
    lines = [f"{n} {nnz}"] + [(f"{x}\t{y}" if i % 7 == 0 else f"  {x}  {y} 1.5") for i, (x, y) in enumerate(zip(a, b))]
    text = sep.join(lines) + (sep if trailing_nl else "") + extra
    return write(tmp_path, name, text), a, b


@pytest.mark.parametrize("crlf,trailing_nl", [(False, True), (True, True), (False, False)])
def test_parallel_parse_matches_single_thread(tmp_path, monkeypatch, crlf, trailing_nl):
    rng = np.random.default_rng(7)
    p, a, b = big_edge_list(tmp_path, rng, crlf=crlf, trailing_nl=trailing_nl)
    g1 = load_with_threads(p, 1, monkeypatch)
    g8 = load_with_threads(p, 8, monkeypatch)
    for k in ("row", "col", "rowT", "colT"):
        assert np.array_equal(getattr(g1, k), getattr(g8, k))
    assert g1.m_reported == g8.m_reported == len(a)
    expect = sorted(set(zip((b - 1).tolist(), (a - 1).tolist())))
    got = [(v, int(u)) for v in range(g8.n) for u in g8.col[g8.row[v]:g8.row[v + 1]]]
    assert got == expect
    if oracle.ref_available() and not crlf:
        assert_same_as_reference(p)


def test_parallel_parse_ignores_lines_after_nnz(tmp_path, monkeypatch):
    rng = np.random.default_rng(8)
    p, a, b = big_edge_list(tmp_path, rng, extra="garbage line\n1 2\n")
    g = load_with_threads(p, 8, monkeypatch)
    assert g.m_reported == len(a)


def test_parallel_parse_symmetric_mtx(tmp_path, monkeypatch):
    rng = np.random.default_rng(9)
    n, nnz = 3000, 40000
    a = rng.integers(1, n + 1, nnz)
    b = rng.integers(1, n + 1, nnz)
    text = f"%%MatrixMarket matrix coordinate pattern symmetric\n{n} {n} {nnz}\n" + \
        "".join(f"{x} {y}\n" for x, y in zip(a, b))
    p = write(tmp_path, "s.mtx", text)
    g1 = load_with_threads(p, 1, monkeypatch)
    g8 = load_with_threads(p, 8, monkeypatch)
    assert np.array_equal(g1.col, g8.col) and np.array_equal(g1.row, g8.row)
    assert g8.m_reported == nnz + int(np.sum(a != b))
    if oracle.ref_available():
        assert_same_as_reference(p)


BIN = os.path.join(os.path.dirname(HERE), "gpuscc_code_amd", "lib")


def cli_load_error(path, threads):
    import subprocess
    env = dict(os.environ, GPUSCC_HOST_THREADS=str(threads))
    r = subprocess.run([os.path.join(BIN, "scc"), "-f", path], capture_output=True, text=True, timeout=120,
                       env=env)
    assert r.returncode != 0
    return r.stderr.strip()


@pytest.mark.parametrize("where", [10, 30000, 59990])
def test_parallel_parse_reports_the_first_bad_line(tmp_path, where):
    """The first failing line (as a sequential reader meets it) is reported, in
    whichever chunk it falls; a later bad line does not mask it."""
    rng = np.random.default_rng(10)
    n, nnz = 5000, 60000
    a = rng.integers(1, n + 1, nnz).tolist()
    b = rng.integers(1, n + 1, nnz).tolist()
    lines = [f"{x} {y}" for x, y in zip(a, b)]
    lines[where] = "12 x"
    lines[-1] = "1 99999"  # out of range, after the bad line
    p = write(tmp_path, "bad.txt", f"{n} {nnz}\n" + "\n".join(lines) + "\n")
    for threads in (1, 8):
        assert cli_load_error(p, threads) == f"error: bad entry line {where + 1}"


def test_parallel_parse_short_file(tmp_path):
    rng = np.random.default_rng(11)
    n, nnz = 5000, 60000
    lines = [f"{x} {y}" for x, y in zip(rng.integers(1, n + 1, nnz - 5), rng.integers(1, n + 1, nnz - 5))]
    p = write(tmp_path, "short.txt", f"{n} {nnz}\n" + "\n".join(lines) + "\n")
    for threads in (1, 8):
        assert cli_load_error(p, threads) == "error: file ends before nnz entries"
    lines[20000] = "0 5"  # an error before the end of the file comes first
    p = write(tmp_path, "short2.txt", f"{n} {nnz}\n" + "\n".join(lines) + "\n")
    for threads in (1, 8):
        assert cli_load_error(p, threads) == "error: vertex id out of range on entry line 20001"


# ---------------------------------------------------------------------------
# binary graph files (scc_save_graph / scc_load_file)
# ---------------------------------------------------------------------------
def same_graph(g, h):
    for k in ("row", "col", "rowT", "colT"):
        assert np.array_equal(getattr(g, k), getattr(h, k)), k
    assert g.n == h.n and g.m_reported == h.m_reported
    assert g.info == h.info


def test_binary_round_trip_of_loaded_graphs(tmp_path):
    for src in (DEMO, write(tmp_path, "m.mtx", "%%MatrixMarket matrix coordinate pattern symmetric\n"
                                                 "4 4 4\n1 2\n3 2\n4 4\n4 1\n")):
        g = scc.load(src)
        out = str(tmp_path / "g.bin")
        scc.save(g, out)
        same_graph(g, scc.load(out))


def test_binary_round_trip_generated_and_batch(tmp_path):
    g = scc.gen_rmat(12, 8, seed=3)
    scc.save(g, tmp_path / "r.bin")
    same_graph(g, scc.load(tmp_path / "r.bin"))
    b = scc.gen_batch(20, 16, 256, seed=2)
    scc.save(b, tmp_path / "b.bin")
    h = scc.load(tmp_path / "b.bin")
    same_graph(b, h)
    assert np.array_equal(b.graph_offsets, h.graph_offsets)
    e = scc.csr_from_arcs(0, np.zeros(0, np.int32), np.zeros(0, np.int32))
    scc.save(e, tmp_path / "e.bin")
    assert scc.load(tmp_path / "e.bin").n == 0


def test_binary_files_are_checked(tmp_path):
    g = scc.gen_rmat(10, 4, seed=1)
    p = tmp_path / "ok.bin"
    scc.save(g, p)
    raw = bytearray(p.read_bytes())

    def corrupt(name, data):
        q = tmp_path / name
        q.write_bytes(bytes(data))
        with pytest.raises(scc.SccError) as e:
            scc.load(q)
        assert e.value.status == 3

    corrupt("trunc.bin", raw[:-4])                       # size does not match the header
    corrupt("head.bin", raw[:40])                        # truncated header
    col0 = 64 + 4 * (g.n + 1)
    bad = bytearray(raw)
    bad[col0:col0 + 4] = (g.n + 5).to_bytes(4, "little")  # column id out of range
    corrupt("range.bin", bad)
    r0 = int(np.argmax(np.diff(g.row) >= 2))              # a row with two entries: swap them
    e = col0 + 4 * int(g.row[r0])
    bad = bytearray(raw)
    bad[e:e + 4], bad[e + 4:e + 8] = raw[e + 4:e + 8], raw[e:e + 4]
    corrupt("order.bin", bad)
    bad = bytearray(raw)
    bad[64:68] = (1).to_bytes(4, "little")                 # row[0] != 0
    corrupt("row0.bin", bad)
    # interior offsets outside [0, m]: far past the arcs, and negative at the
    # start of a later thread's share of the rows
    for name, v, val in (("far.bin", 1, 10 ** 6), ("neg.bin", g.n // 2, -5), ("past.bin", g.n - 1, g.m + 1)):
        bad = bytearray(raw)
        bad[64 + 4 * v:68 + 4 * v] = val.to_bytes(4, "little", signed=True)
        corrupt(name, bad)


def test_cli_writes_a_binary_graph(tmp_path):
    import subprocess
    out = tmp_path / "demo.bin"
    r = subprocess.run([os.path.join(BIN, "scc"), "-t", "-o", str(out), "-f", DEMO], capture_output=True,
                       text=True, timeout=120)
    assert f"saved binary graph to {out}" in r.stdout
    assert "    ->count of SCC: 6" in r.stdout  # Tarjan's line (the GPU part needs a device)
    same_graph(scc.load(DEMO), scc.load(out))
    r2 = subprocess.run([os.path.join(BIN, "scc"), "-t", "-f", str(out)], capture_output=True, text=True,
                        timeout=120)
    assert "nodes=13\tedges=19" in r2.stdout and "    ->count of SCC: 6" in r2.stdout


def test_sizes_beyond_int32_are_overflow(tmp_path):
    """Sizes an int32 CSR cannot hold are SCC_E_OVERFLOW (4), not a silent drop
    (the reference's worklists drop pushes past their capacity,
    worklistc.h:80-81) and not a crash; negative sizes are invalid input (3).
    Every case is refused before an array is read or allocated."""
    import ctypes
    from gpuscc_code_amd import _lib
    L = _lib.lib()
    src = np.zeros(4, np.int32)
    g = _lib.SccHostGraph()
    p = lambda a: a.ctypes.data_as(_lib._i32p)  # noqa: E731
    assert L.scc_csr_from_arcs(8, 1 << 31, p(src), p(src), 1, ctypes.byref(g)) == 4
    assert L.scc_csr_from_arcs(8, -1, p(src), p(src), 1, ctypes.byref(g)) == 3
    assert L.scc_gen_mesh(50000, 0.6, 3, ctypes.byref(g)) == 4   # 2.5e9 vertices
    assert L.scc_gen_rmat(31, 1, 0.57, 0.19, 0.19, 1, 1, ctypes.byref(g)) == 3  # scale above 30
    assert L.scc_gen_rmat(28, 8, 0.57, 0.19, 0.19, 1, 1, ctypes.byref(g)) == 4  # 2^31 arcs drawn
    for name, text, status in (("big.graph", "3000000000 5\n1 2\n", 4),
                               ("bigm.graph", "5 3000000000\n1 2\n", 4),
                               ("neg.graph", "-5 3\n1 2\n", 3),
                               ("big.mtx", "%%MatrixMarket matrix coordinate pattern general\n"
                                           "3000000000 3000000000 2\n1 2\n2 1\n", 4)):
        with pytest.raises(scc.SccError) as e:
            scc.load(write(tmp_path, name, text))
        assert e.value.status == status, name
    # a binary graph whose header claims 2^31 vertices
    gg = scc.gen_rmat(6, 4, seed=1)
    q = tmp_path / "ok.bin"
    scc.save(gg, q)
    raw = bytearray(q.read_bytes())
    raw[8:16] = (1 << 31).to_bytes(8, "little")
    q.write_bytes(bytes(raw))
    with pytest.raises(scc.SccError) as e:
        scc.load(q)
    assert e.value.status == 4
