"""The drop-in C++ facade (sparse::SMatrix, block::BlockMatrix) driven by
reference-style programs (cxx/examples), each of which checks residuals,
the reference's thresholds and its exception contract and exits non-zero on
any failure."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cpp-fem-solver_amd", "cxx", "bin")


@pytest.mark.parametrize("prog", ["smatrix_solvers", "blockmatrix_checks"])
def test_driver(prog):
    exe = os.path.join(BIN, prog)
    assert os.path.exists(exe), "build the facade first (__graft_entry__.build())"
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(p.stdout[-4000:])
    print(p.stderr[-2000:])
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert "PASSED" in p.stdout
