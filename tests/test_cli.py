"""The drop-in binaries keep the reference's flags and stdout lines
(scc.cu:57-148, load.cpp prints, tarjan.cpp:193,237)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "gpuscc_code_amd", "lib")
DEMO = os.path.join(ROOT, "tests", "golden", "demo.graph")


def run(binary, *args):
    return subprocess.run([os.path.join(BIN, binary), *args], capture_output=True, text=True,
                          timeout=120)


@pytest.mark.parametrize("binary", ["scc", "scc_w", "scc_2_w"])
def test_load_and_tarjan_lines(binary):
    r = run(binary, "-t", "-a", "q", "-f", DEMO)
    out = r.stdout
    assert "loading graph..." in out
    assert "nodes=13\tedges=19" in out
    assert "    ->nodes: 13" in out and "    ->edges: 19" in out
    assert "tarjan decomposition..." in out
    assert "trivial = 3\tnon_trivial=3\tnSCC = 6\tbiggestSCC = 5" in out
    assert "    ->count of SCC: 6" in out


def test_usage_and_bad_strategy():
    r = run("scc")
    assert r.returncode != 0 and "usage" in r.stdout
    r = run("scc", "-a", "x", "-f", DEMO)
    assert r.returncode != 0 and "not provided" in r.stderr


def test_missing_file_is_an_error():
    r = run("scc", "-f", "/nonexistent.graph")
    assert r.returncode != 0 and "error" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("binary", ["scc", "scc_w", "scc_2_w"])
@pytest.mark.parametrize("alg", ["q", "t", "d", "l", "h", "c", "b"])
def test_gpu_run_lines(binary, alg):
    r = run(binary, "-t", "-a", alg, "-j", "-f", DEMO)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "FB done!" in out
    assert out.count("    ->count of SCC: 6") == 2  # Tarjan's and the GPU's
    assert "    ->steps:" in out and "    ->time:" in out
    assert '"num_scc": 6' in out


@pytest.mark.gpu
def test_gpu_all_strategies():
    # -a a runs QUADRATIC..BEST in turn (scc.cu:128-139)
    r = run("scc", "-a", "a", "-j", "-f", DEMO)
    assert r.returncode == 0, r.stderr
    for k in range(7):
        assert f"computing SCC in parallel - alg: {k}..." in r.stdout
    assert r.stdout.count("    ->count of SCC: 6") == 7
    for name in ("QUADRATIC", "TOPO", "DATA", "LINEAR", "HYBRID", "COMBINE", "BEST"):
        assert f"running {name} strategy" in r.stdout


@pytest.mark.gpu
def test_gpu_verify_files(tmp_path):
    # VERIFY outputs (scc.cu:920-995): the demo graph's SCCs, 1-based
    r = subprocess.run([os.path.join(BIN, "scc"), "-V", "-a", "h", "-f", DEMO], capture_output=True,
                       text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    sccs = sorted(tuple(int(x) for x in line.split()) for line in (tmp_path / "parallel_scc.txt").read_text().splitlines())
    assert sccs == sorted([(1,), (2, 5, 6), (3,), (4,), (7, 9, 10, 11, 12), (8, 13)])
    assert sorted(int(x) for x in (tmp_path / "trivial.txt").read_text().split()) == [1, 3, 4]
    dist = dict(tuple(map(int, line.split())) for line in (tmp_path / "scc_dist.txt").read_text().splitlines())
    assert dist == {1: 3, 2: 1, 3: 1, 5: 1}
    detail = (tmp_path / "state_range_root.txt").read_text().splitlines()
    assert detail[0].startswith("ID") and len(detail) == 14
    assert "number=6" in r.stdout


# ---------------------------------------------------------------------------
# the log tabulator (process_rst.cpp:7-41) and its batch script (runall)
# ---------------------------------------------------------------------------
REF_TAB = os.path.join(ROOT, "oracle", "_ref", "process_rst")


def tabulate(tool, log_text, tmp_path, name):
    log = tmp_path / f"{name}.log"
    log.write_text(log_text)
    out = tmp_path / f"{name}.tab"
    r = subprocess.run([tool, str(log), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return out.read_text()


def run_log(path, times, binary="./scc"):
    lines = [f"{binary} -t -a a -f {path}", "loading graph...", "    ->nodes: 13", "    ->edges: 19",
             "tarjan decomposition..."]
    for t in times:
        lines += ["    ->count of SCC: 6", f"    ->time: {t}"]
    return "\n".join(lines) + "\n"


def test_process_rst_rows():
    """Seven "->time:" values per row after the graph's file name (the sixth field
    of the command line), written as the reference's ostream prints floats."""
    import tempfile
    from pathlib import Path
    with tempfile.TemporaryDirectory() as d:
        log = run_log("../scc_dataset/cage13.mtx", [0.5, 0.25, 1.0, 13.633, 0.000012, 2.5, 100.0]) + \
            run_log("../scc_dataset/web.txt", [1, 2, 3, 4, 5, 6, 7])
        got = tabulate(os.path.join(BIN, "process_rst"), log, Path(d), "a")
    assert got == ("../scc_dataset/cage13.mtx 0.5 0.25 1 13.633 1.2e-05 2.5 100 \n"
                   "../scc_dataset/web.txt 1 2 3 4 5 6 7 \n")


def test_process_rst_matches_reference_tabulator(tmp_path):
    if not os.path.exists(REF_TAB):
        pytest.skip("reference tabulator not built (needs /root/reference)")
    rng = __import__("numpy").random.default_rng(3)
    logs = {
        "aligned": "".join(run_log(f"d/g{k}.txt", rng.random(7).round(6).tolist()) for k in range(4)),
        # -t -a a prints eight times per graph (Tarjan and seven strategies): rows drift, as in the reference
        "eight": "".join(run_log(f"d/g{k}.mtx", rng.random(8).round(5).tolist()) for k in range(5)),
        "odd": "./scc -f short\n    ->time: 1.5\n    ->time: oops\ntime: 2\n" + run_log("x y", [3] * 6, "./scc_w") +
               "no match here\n",
        "empty": "",
    }
    for name, text in logs.items():
        assert tabulate(os.path.join(BIN, "process_rst"), text, tmp_path, "b_" + name) == \
            tabulate(REF_TAB, text, tmp_path, "r_" + name), name


def test_process_rst_errors(tmp_path):
    r = subprocess.run([os.path.join(BIN, "process_rst")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "usage" in r.stderr
    r = subprocess.run([os.path.join(BIN, "process_rst"), "/nonexistent.log", str(tmp_path / "o")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "cannot open" in r.stderr


def test_runall_prints_each_command_line(tmp_path):
    """runall DIR runs scc and scc_w on every file with -t -a a; the Tarjan part
    runs on the CPU, so its lines appear even where the GPU part cannot run."""
    d = tmp_path / "ds"
    d.mkdir()
    (d / "demo.graph").write_text(open(DEMO).read())
    r = subprocess.run([os.path.join(BIN, "runall"), str(d)], capture_output=True, text=True, timeout=120)
    out = r.stdout
    assert f"./scc -t -a a -f {d}/demo.graph" in out and f"./scc_w -t -a a -f {d}/demo.graph" in out
    assert out.count("    ->count of SCC: 6") >= 2
