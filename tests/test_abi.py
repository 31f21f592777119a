"""The C ABI library loads (no GPU needed) and exports every entry point that
include/gpuscc.h declares; error paths that need no device return codes."""
import ctypes
import os
import re
import subprocess

import pytest

from gpuscc_code_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gpuscc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(scc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("scc_create", "scc_upload", "scc_run", "scc_destroy", "scc_phase_reach",
              "scc_phase_trim", "scc_phase_update", "scc_phase_trim2", "scc_phase_wcc",
              "scc_load_file", "scc_gen_rmat"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding covers every one of them
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_exported_symbols_are_c_names():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    for s in declared_symbols():
        assert s in exported


def test_status_strings_and_null_args():
    L = _lib.lib()
    assert L.scc_status_string(0) == b"ok"
    assert L.scc_status_string(3) == b"invalid input"
    assert L.scc_create(0, None, None) == _lib.SCC_E_INPUT
    assert L.scc_upload(None, 0, 0, None, None, None, None) == _lib.SCC_E_INPUT
    assert L.scc_run(None, 0, 0, 0, None, None) == _lib.SCC_E_INPUT
    assert L.scc_load_file(b"/nonexistent", ctypes.byref(_lib.SccHostGraph())) == _lib.SCC_E_INPUT
    L.scc_destroy(None)


def test_create_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    st = _lib.lib().scc_create(0, None, ctypes.byref(h))
    assert st == _lib.SCC_E_HIP and not h.value


def header_struct_fields(name):
    """Field names, in order, of `typedef struct { ... } name;` in gpuscc.h."""
    text = open(os.path.join(ROOT, "include", "gpuscc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    body = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + ";", text).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if decl:
            fields += [re.sub(r"[\s*]|\[.*\]", "", f) for f in re.sub(r"^[A-Za-z0-9_ ]+?\s+(?=[*A-Za-z_])", "", decl,
                                                                   count=1).split(",")]
    return fields


@pytest.mark.parametrize("name,cls", [("scc_opts", _lib.SccOpts), ("scc_stats", _lib.SccStats),
                                      ("scc_profile_data", _lib.SccProfile), ("scc_host_graph", _lib.SccHostGraph)])
def test_binding_structs_match_the_header(name, cls):
    assert header_struct_fields(name) == [f[0] for f in cls._fields_]
