"""gpuscc_code_amd -- Forward-Backward strongly connected components on MI355X (gfx950).

The product is libgpuscc.so (HIP kernels + C ABI, include/gpuscc.h) and the
drop-in binaries lib/scc, lib/scc_w, lib/scc_2_w.  This package binds the
library with ctypes (``gpuscc_code_amd.scc``) and builds it (``build``).
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")


def build(jobs=8):
    """Compile every HIP source for gfx950 and link libgpuscc.so + the CLI binaries."""
    subprocess.run(["make", "-s", "-C", CSRC, f"-j{jobs}", "all"], check=True)


from .scc import (Context, Graph, alg_opt, computeSCC, csr_from_arcs, gen_batch,  # noqa: E402
                  gen_mesh, gen_rmat, load)

__all__ = ["build", "Context", "Graph", "alg_opt", "computeSCC", "csr_from_arcs", "gen_batch",
           "gen_mesh", "gen_rmat", "load"]
