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


SAN_BIN = BIN + "_san"


def test_cpp_grid_shim_under_asan_and_ubsan():
    """The same program over libcabana_amd_san.so, whose C-ABI host layer is
    built with host AddressSanitizer and UndefinedBehaviorSanitizer (device
    code unchanged): every descriptor check, plan choice, scratch carve and
    part-table cache path the program reaches runs instrumented.  Leak
    reports are off because the HIP runtime keeps allocations until exit."""
    if not os.path.exists(SAN_BIN):
        pytest.fail("tests/cpp/test_grid_shim_san is not built (run __graft_entry__.build())")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    print("running", SAN_BIN, flush=True)
    proc = subprocess.run([SAN_BIN], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                          timeout=300, env=env)
    print(proc.stdout)
    assert proc.returncode == 0, proc.stdout
    assert "ALL PASSED" in proc.stdout
    assert "AddressSanitizer" not in proc.stdout and "runtime error" not in proc.stdout


CUSTOM_BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_custom_functor")


def test_cpp_custom_functors():
    """User-written functors through include/cabana_amd_custom.hpp (the
    functor escape hatch of Cabana's p2g / g2p): every built-in operator
    written again with the P2G:: / G2P:: helpers matches the library's own
    kernels for Spline<1> to Spline<3>."""
    if not os.path.exists(CUSTOM_BIN):
        pytest.fail("tests/cpp/test_custom_functor is not built (run __graft_entry__.build())")
    print("running", CUSTOM_BIN, flush=True)
    proc = subprocess.run([CUSTOM_BIN], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True, timeout=120)
    print(proc.stdout)
    assert proc.returncode == 0, proc.stdout
    assert "ALL PASSED" in proc.stdout
