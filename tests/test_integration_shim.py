"""The reference-side C++ shim shown in INTEGRATION.md compiles against the
reference's own headers (smatrix.h, slvparam.h, svector.h) and this library's
include/bcmhip.h: the callback casts, the CSlvParam fields the shim copies and
the CSVector calls the operator adapters make all exist with those types.

CPU only; skipped where the reference sources are not present (they are
read here as headers, never built into the product).
"""

import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
BSLIB = Path("/root/reference/Bslib")

pytestmark = pytest.mark.skipif(not (BSLIB / "smatrix.h").exists() or shutil.which("g++") is None,
                                reason="needs the reference's Bslib headers and g++")


def _cpp_blocks():
    text = (ROOT / "INTEGRATION.md").read_text()
    return re.findall(r"```cpp\n(.*?)```", text, re.S)


def test_shim_blocks_are_present():
    blocks = _cpp_blocks()
    assert len(blocks) >= 4
    assert "AbstractSymmSolverGpu" in blocks[0] and "bcm_abstract_par_solver" in blocks[0]
    assert "CBLOCKROWC" in blocks[1] and "bcm_abstract_solver_c" in blocks[1]
    assert "GpuMvm" in blocks[2] and "GpuSolveLU" in blocks[2]
    assert "bcm_solve_block" in blocks[3]


def test_shim_compiles_against_reference_headers(tmp_path):
    blocks = _cpp_blocks()
    # the BlockGMRES fragment runs inside a function with the variables it names
    fragment = (
        "void BlockGmresGpu(bcm_context c, bcm_matrix a, bcm_factor f, CSVector& b, CSVector& x,\n"
        "                   CSlvParam& param, int n, int nrhs, double* db, double* dx) {\n"
        + blocks[3] + "}\n")
    src = tmp_path / "bslib_gpu.cpp"
    src.write_text(blocks[0] + "\n" + blocks[1] + "\n" + blocks[2] + "\n" + fragment)
    cmd = ["g++", "-std=c++11", "-fsyntax-only", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
           f"-I{BSLIB}", f"-I{ROOT / 'include'}", str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-4000:]
