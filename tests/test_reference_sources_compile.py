"""The reference's own tutorials on this path compile unchanged against the drop-in headers
(viennacl-dev_amd/viennacl, include/): examples/tutorial/iterative-custom.cpp (compressed_matrix,
MatrixMarket reader, prod, CG/BiCGStab/GMRES with and without Jacobi, solver objects with an
initial guess) and examples/tutorial/matrix-free.cpp (a user operator with apply() in solve() with
the three tags; its host worker reaches MAIN_MEMORY arrays through extract_raw_pointer).  Compiled
with g++ -fsyntax-only from where they lie in the reference tree; skipped where that tree is
absent.  Running them needs their data file (iterative-custom) or a HIP branch in the operator
(matrix-free: tests/cpp/reference_call_sites.cpp runs that case)."""
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
REFERENCE = pathlib.Path("/root/reference")
SOURCES = ["examples/tutorial/iterative-custom.cpp", "examples/tutorial/matrix-free.cpp"]


@pytest.mark.skipif(not (REFERENCE / "examples").is_dir() or shutil.which("g++") is None,
                    reason="reference tree or g++ not present")
@pytest.mark.parametrize("src", SOURCES)
def test_reference_tutorial_compiles_against_the_drop_in_headers(src):
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-I", str(ROOT / "viennacl-dev_amd"), "-I", str(ROOT / "include"),
           str(REFERENCE / src)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
