import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_addoption(parser):
    parser.addoption("--relabel", choices=("auto", "on", "off"), default="auto",
                     help="renumbering of uploaded graphs for every context the tests create "
                          "(alg_opt.relabel 0 / 1 / -1)")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgpuscc.so on cuda:0)")
    config.addinivalue_line("markers", "slow: large inputs (full benchmark sizes)")
    mode = config.getoption("--relabel")
    if mode != "auto":
        from gpuscc_code_amd import scc
        scc.DEFAULT_RELABEL = 1 if mode == "on" else -1


@pytest.fixture(scope="session")
def ctx():
    """One device context for the GPU tests (scc_create on device 0)."""
    from gpuscc_code_amd import scc
    c = scc.Context(0)
    yield c
    c.close()
