"""The header-only C++ layer (include/cabana_amd_grid.hpp) on the GPU: runs
tests/cpp/test_grid_shim (built by __graft_entry__.build()), which repeats
Cabana's own tstInterpolation3d known answers through Cabana-shaped calls and
checks slices vs views, binned vs unbinned, fused vs single functors and the
adjoint transpose identity."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_grid_shim")


def test_cpp_grid_shim():
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/test_grid_shim is not built (run __graft_entry__.build())")
    print("running", BIN, flush=True)
    proc = subprocess.run([BIN], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                          timeout=120)
    print(proc.stdout)
    assert proc.returncode == 0, proc.stdout
    assert "ALL PASSED" in proc.stdout
