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
    r = run("scc", "-a", "c", "-f", DEMO)
    assert r.returncode != 0


def test_missing_file_is_an_error():
    r = run("scc", "-f", "/nonexistent.graph")
    assert r.returncode != 0 and "error" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("binary", ["scc", "scc_w", "scc_2_w"])
@pytest.mark.parametrize("alg", ["q", "t", "d", "l", "h"])
def test_gpu_run_lines(binary, alg):
    r = run(binary, "-t", "-a", alg, "-j", "-f", DEMO)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "FB done!" in out
    assert out.count("    ->count of SCC: 6") == 2  # Tarjan's and the GPU's
    assert "    ->steps:" in out and "    ->time:" in out
    assert '"num_scc": 6' in out
