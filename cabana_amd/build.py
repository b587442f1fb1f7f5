"""Builds the cabana_amd HIP library (and the CPU oracle used by the tests).

The library is compiled for gfx950 with hipcc directly: one object per spline
order for the interpolation kernels (compiled in parallel), one for the grid
and binning kernels and one for the C-ABI layer, linked into
``cabana_amd/libcabana_amd.so``.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcabana_amd.so")
ARCH = os.environ.get("CABANA_AMD_ARCH", "gfx950")

COMMON = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    # Packed FP32 instructions stay disabled for all device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall",
]


# roctx ranges around the entry points (rocprofiler-sdk's marker library).
ROCTX = ["-L" + os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib"),
         "-lrocprofiler-sdk-roctx"]


def _hipcc():
    cand = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")
    if os.path.exists(cand):
        return cand
    found = shutil.which("hipcc")
    if not found:
        raise RuntimeError("hipcc not found; cannot build the HIP library")
    return found


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line
    )
    if proc.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    # Per-particle kernels for Spline<0> .. Spline<6>; cell-tiled kernels for
    # orders 1 to 3.
    for order in range(7):
        for src in ("ca_kernels", "ca_tiled") if 1 <= order <= 3 else ("ca_kernels",):
            obj = os.path.join(BUILD, "%s_o%d.o" % (src, order))
            jobs.append([hipcc] + COMMON + ["-DCA_ORDER=%d" % order, "-c",
                                            os.path.join(CSRC, src + ".hip"), "-o", obj])
    for src in ("ca_grid", "ca_probe"):
        jobs.append([hipcc] + COMMON + ["-c", os.path.join(CSRC, src + ".hip"), "-o",
                                        os.path.join(BUILD, src + ".o")])
    for src in ("ca_api", "ca_no_alloc"):
        jobs.append([hipcc] + COMMON + ["-x", "hip", "-c", os.path.join(CSRC, src + ".cpp"),
                                        "-o", os.path.join(BUILD, src + ".o")])
    workers = max(1, min(len(jobs), os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ROCTX)
    os.replace(tmp, LIB)
    return LIB


def build_cpp_tests():
    """Host-only C++ programs over include/cabana_amd_grid.hpp (g++, linked
    against the library and the HIP runtime); run by the GPU tests."""
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    src = os.path.join(ROOT, "tests", "cpp", "test_grid_shim.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "test_grid_shim")
    _run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
          "-I", os.path.join(ROOT, "include"), "-I", os.path.join(rocm, "include"),
          src, "-o", out, "-L", HERE, "-lcabana_amd", "-L", os.path.join(rocm, "lib"),
          "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../cabana_amd",
          "-Wl,-rpath," + os.path.join(rocm, "lib")])
    # User functors through include/cabana_amd_custom.hpp: a HIP program.
    src = os.path.join(ROOT, "tests", "cpp", "test_custom_functor.hip")
    custom = os.path.join(ROOT, "tests", "cpp", "test_custom_functor")
    _run([_hipcc()] + COMMON + ["-I", os.path.join(ROOT, "include"),
                                "-I", os.path.join(ROOT, "tests", "cpp"), src, "-o", custom,
                                "-L", HERE, "-lcabana_amd", "-Wl,-rpath,$ORIGIN/../../cabana_amd"])
    return out


# Host-side AddressSanitizer + UndefinedBehaviorSanitizer (device code is
# untouched: every flag applies to the host compilation only).
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined",
            "-Xarch_host", "-fno-omit-frame-pointer"]
SAN_LIB = os.path.join(HERE, "libcabana_amd_san.so")


def _clang_runtime_dir():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    base = os.path.join(rocm, "lib", "llvm", "lib", "clang")
    for ver in sorted(os.listdir(base)) if os.path.isdir(base) else []:
        d = os.path.join(base, ver, "lib", "linux")
        if os.path.exists(os.path.join(d, "libclang_rt.asan-x86_64.so")):
            return d
    raise RuntimeError("clang's shared AddressSanitizer runtime was not found")


def build_sanitized():
    """libcabana_amd_san.so: the C-ABI host layer (ca_api.cpp, ca_no_alloc.cpp)
    built with host AddressSanitizer and UndefinedBehaviorSanitizer, linked
    with the kernel objects of build_library() and clang's shared sanitizer
    runtime; and tests/cpp/test_grid_shim_san, the C++ layer's test program
    over it.  Needs build_library() first."""
    hipcc = _hipcc()
    rt = _clang_runtime_dir()
    sdir = os.path.join(BUILD, "san")
    os.makedirs(sdir, exist_ok=True)
    objs = []
    for src in ("ca_api", "ca_no_alloc"):
        obj = os.path.join(sdir, src + ".o")
        _run([hipcc] + COMMON + SAN_HOST + ["-x", "hip", "-c", os.path.join(CSRC, src + ".cpp"),
                                            "-o", obj])
        objs.append(obj)
    for name in sorted(os.listdir(BUILD)):
        if name.endswith(".o") and name not in ("ca_api.o", "ca_no_alloc.o"):
            objs.append(os.path.join(BUILD, name))
    tmp = SAN_LIB + ".tmp"
    _run([hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ROCTX +
         ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
          "-Xarch_host", "-shared-libasan", "-Wl,-rpath," + rt])
    os.replace(tmp, SAN_LIB)
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    src = os.path.join(ROOT, "tests", "cpp", "test_grid_shim.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "test_grid_shim_san")
    _run([os.path.join(rocm, "llvm", "bin", "clang++"), "-std=c++17", "-O1", "-g",
          "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
          "-I", os.path.join(rocm, "include"), src, "-o", out, "-L", HERE, "-lcabana_amd_san",
          "-L", os.path.join(rocm, "lib"), "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../cabana_amd",
          "-Wl,-rpath," + os.path.join(rocm, "lib"), "-Wl,-rpath," + rt] + SAN_HOST +
         ["-Xarch_host", "-shared-libasan"])
    return SAN_LIB, out


def build_oracle():
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "-s"])
    return os.path.join(odir, "liboracle.so")


def main():
    lib = build_library(verbose="-v" in sys.argv)
    print("built", lib)
    print("built", build_oracle())
    print("built", build_cpp_tests())
    print("built", *build_sanitized())


if __name__ == "__main__":
    main()
