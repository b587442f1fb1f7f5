"""The C-ABI library loads without a GPU and exports every entry point that
include/cabana_amd.h declares (no compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cabana_amd.h")
LIB = os.path.join(ROOT, "cabana_amd", "libcabana_amd.so")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(cabana_amd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for name in ("cabana_amd_p2g", "cabana_amd_g2p", "cabana_amd_p2g_adjoint",
                 "cabana_amd_g2p_adjoint", "cabana_amd_halo_gather",
                 "cabana_amd_halo_scatter_sum", "cabana_amd_linked_cell_build",
                 "cabana_amd_permute"):
        assert name in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libcabana_amd.so is not built (python -m cabana_amd.build)")
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_bindings_cover_the_header():
    from cabana_amd import _native
    assert sorted(_native.SIGNATURES) == declared_symbols()


def test_abi_version_and_host_helpers():
    lib = ctypes.CDLL(LIB)
    lib.cabana_amd_abi_version.restype = ctypes.c_int32
    assert lib.cabana_amd_abi_version() == 3
    d3 = ctypes.c_double * 3
    nb = (ctypes.c_int32 * 3)()
    lib.cabana_amd_linked_cell_num_bins.restype = ctypes.c_int32
    rc = lib.cabana_amd_linked_cell_num_bins(d3(0.5, 0.125, 0.25), d3(-1.0, -0.5, -0.6),
                                             d3(2.5, 1.5, 1.9), nb)
    assert rc == 0 and tuple(nb) == (7, 16, 10)   # tstCartesianGrid.cpp:22-40
    rc = lib.cabana_amd_linked_cell_num_bins(d3(0.0, 1.0, 1.0), d3(0, 0, 0), d3(1, 1, 1), nb)
    assert rc == 1
    lib.cabana_amd_last_error.restype = ctypes.c_char_p
    assert b"Cartesian" in lib.cabana_amd_last_error(None)


STRUCTS = {
    "cabana_amd_mesh": "Mesh",
    "cabana_amd_particle_field": "ParticleField",
    "cabana_amd_grid_field": "GridField",
    "cabana_amd_binning": "Binning",
    "cabana_amd_p2g_term": "P2GTerm",
    "cabana_amd_g2p_term": "G2PTerm",
    "cabana_amd_p2g_adjoint_term": "P2GAdjointTerm",
    "cabana_amd_g2p_adjoint_term": "G2PAdjointTerm",
}


def test_ctypes_structs_match_the_header(tmp_path):
    """Size and field offsets of every struct of the header, compiled by the
    C compiler, equal those of the ctypes mirrors."""
    import shutil
    import subprocess
    from cabana_amd import _native
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "cabana_amd.h"',
             "int main(void) {"]
    for cname, pname in STRUCTS.items():
        cls = getattr(_native, pname)
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for field, _ in cls._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));'
                         % (cname, field, cname, field))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True,
                               text=True).stdout.splitlines():
        a, b, v = line.split()
        got[(a, b)] = int(v)
    for cname, pname in STRUCTS.items():
        cls = getattr(_native, pname)
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert got[(cname, field)] == getattr(cls, field).offset, (cname, field)


def test_binning_serials_are_issued_by_the_library():
    """Serials come from one library-wide counter: never 0, never repeated,
    and independent of the caller (Python here, the C++ layer elsewhere)."""
    from cabana_amd import _native
    lib = _native.load()
    got = [lib.cabana_amd_new_binning_serial() for _ in range(100)]
    assert all(s > 0 for s in got)
    assert len(set(got)) == len(got)
    assert got == sorted(got)
