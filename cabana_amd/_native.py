"""ctypes bindings of the cabana_amd C ABI (include/cabana_amd.h).

The shared library is built in-tree by ``cabana_amd/build.py``.  Loading it
fails loudly when it is missing: there is no fallback implementation.
"""
import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime this library binds to)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcabana_amd.so")

c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p

FLOAT64, FLOAT32, INT32, INT64 = 0, 1, 2, 3
NODE, CELL = 0, 1

OK = 0
ERROR_INVALID_ARGUMENT = 1
ERROR_UNSUPPORTED = 2
ERROR_OUT_OF_MEMORY = 3
ERROR_DEVICE = 4
ERROR_OUT_OF_DOMAIN = 5

SCALAR_VALUE_P2G = 0
VECTOR_VALUE_P2G = 1
SCALAR_GRADIENT_P2G = 2
VECTOR_DIVERGENCE_P2G = 3
TENSOR_DIVERGENCE_P2G = 4

SCALAR_VALUE_G2P = 0
VECTOR_VALUE_G2P = 1
SCALAR_GRADIENT_G2P = 2
VECTOR_GRADIENT_G2P = 3
VECTOR_DIVERGENCE_G2P = 4


class Mesh(ctypes.Structure):
    _fields_ = [
        ("low_corner", c_double * 3),
        ("high_corner", c_double * 3),
        ("cell_size", c_double * 3),
        ("halo_cell_width", c_int32),
        ("periodic", c_int32 * 3),
        ("entity", c_int32),
    ]


class ParticleField(ctypes.Structure):
    _fields_ = [
        ("data", c_void_p),
        ("num_particle", c_int64),
        ("num_comp", c_int32),
        ("dtype", c_int32),
        ("vector_length", c_int64),
        ("soa_stride", c_int64),
    ]


class GridField(ctypes.Structure):
    _fields_ = [
        ("data", c_void_p),
        ("extent", c_int32 * 4),
        ("stride", c_int64 * 4),
        ("dtype", c_int32),
    ]


class Binning(ctypes.Structure):
    _fields_ = [
        ("counts", c_void_p),
        ("offsets", c_void_p),
        ("begin", c_int64),
        ("end", c_int64),
        ("num_bin", c_int32 * 3),
        ("serial", ctypes.c_uint64),
    ]


class P2GTerm(ctypes.Structure):
    _fields_ = [
        ("op", c_int32),
        ("multiplier", c_double),
        ("values", ParticleField),
        ("grid", GridField),
    ]


class G2PTerm(ctypes.Structure):
    _fields_ = [
        ("op", c_int32),
        ("multiplier", c_double),
        ("grid", GridField),
        ("values", ParticleField),
    ]


class P2GAdjointTerm(ctypes.Structure):
    _fields_ = [
        ("op", c_int32),
        ("multiplier", c_double),
        ("values", ParticleField),
        ("grid_bar", GridField),
        ("values_bar", ParticleField),
    ]


class G2PAdjointTerm(ctypes.Structure):
    _fields_ = [
        ("op", c_int32),
        ("multiplier", c_double),
        ("grid", GridField),
        ("values_bar", ParticleField),
        ("grid_bar", GridField),
    ]


P = ctypes.POINTER

# name -> (restype, argtypes); every symbol declared in include/cabana_amd.h.
SIGNATURES = {
    "cabana_amd_abi_version": (c_int32, []),
    "cabana_amd_new_binning_serial": (ctypes.c_uint64, []),
    "cabana_amd_context_create": (c_int32, [c_int32, P(c_void_p)]),
    "cabana_amd_context_destroy": (c_int32, [c_void_p]),
    "cabana_amd_last_error": (ctypes.c_char_p, [c_void_p]),
    "cabana_amd_check_domain": (c_int32, [c_void_p, c_void_p, P(c_int64)]),
    "cabana_amd_p2g": (c_int32, [c_void_p, P(Mesh), c_int32, P(ParticleField), c_int32,
                                 P(P2GTerm), c_int32, P(Binning), c_void_p]),
    "cabana_amd_p2g_no_alloc": (c_int32, [c_void_p, P(Mesh), c_int32, P(ParticleField),
                                          c_int32, P(P2GTerm), c_int32, P(Binning),
                                          c_void_p]),
    "cabana_amd_g2p": (c_int32, [c_void_p, P(Mesh), c_int32, P(ParticleField), c_int32,
                                 P(G2PTerm), c_int32, P(Binning), c_void_p]),
    "cabana_amd_g2p2g": (c_int32, [c_void_p, P(Mesh), c_int32, P(ParticleField), c_int32,
                                   P(G2PTerm), c_int32, P(P2GTerm), c_int32, P(Binning),
                                   c_void_p]),
    "cabana_amd_p2g_adjoint": (c_int32, [c_void_p, P(Mesh), c_int32, P(ParticleField),
                                         c_int32, P(P2GAdjointTerm), c_int32,
                                         P(ParticleField), P(Binning), c_void_p]),
    "cabana_amd_g2p_adjoint": (c_int32, [c_void_p, P(Mesh), c_int32, P(ParticleField),
                                         c_int32, P(G2PAdjointTerm), c_int32,
                                         P(ParticleField), P(Binning), c_void_p]),
    "cabana_amd_probe_lds_add": (c_int32, [c_void_p, c_int32, c_int32, P(c_double)]),
    "cabana_amd_adjoint_gather": (c_int32, [c_void_p, P(Mesh), c_int32, P(ParticleField),
                                            c_int32, P(P2GAdjointTerm), c_int32,
                                            P(G2PAdjointTerm), c_int32, P(ParticleField),
                                            P(Binning), c_void_p]),
    "cabana_amd_halo_gather": (c_int32, [c_void_p, P(Mesh), P(GridField), c_int32, c_void_p]),
    "cabana_amd_halo_scatter_sum": (c_int32, [c_void_p, P(Mesh), P(GridField), c_int32,
                                              c_void_p]),
    "cabana_amd_linked_cell_num_bins": (c_int32, [P(c_double), P(c_double), P(c_double),
                                                  P(c_int32)]),
    "cabana_amd_linked_cell_build": (c_int32, [c_void_p, P(ParticleField), c_int64, c_int64,
                                               P(c_double), P(c_double), P(c_double),
                                               c_void_p, c_void_p, c_void_p, c_void_p]),
    "cabana_amd_permute": (c_int32, [c_void_p, c_void_p, c_int64, c_int64,
                                     P(ParticleField), c_void_p]),
    "cabana_amd_permute_members": (c_int32, [c_void_p, c_void_p, c_int64, c_int64,
                                             P(ParticleField), c_int32, c_void_p]),
    "cabana_amd_bov_write_time_step": (c_int32, [c_void_p, P(Mesh), P(GridField), c_int32,
                                                 ctypes.c_char_p, ctypes.c_char_p, c_int32,
                                                 ctypes.c_double, c_int32, c_void_p]),
    "cabana_amd_particle_migrate": (c_int32, [c_void_p, P(Mesh), P(ParticleField),
                                              P(ParticleField), c_int32, c_int32, c_int32,
                                              P(c_int64), P(c_int32), c_void_p]),
    "cabana_amd_set_auto_binning": (c_int32, [c_void_p, c_int32]),
    "cabana_amd_set_deterministic": (c_int32, [c_void_p, c_int32]),
    "cabana_amd_profile_enable": (c_int32, [c_void_p, c_int32]),
    "cabana_amd_profile_reset": (c_int32, [c_void_p]),
    "cabana_amd_profile_count": (c_int32, [c_void_p, P(c_int32)]),
    "cabana_amd_profile_entry": (c_int32, [c_void_p, c_int32, ctypes.c_char_p, c_int32,
                                           P(c_int64), P(c_double)]),
}

_lib = None


def library_path():
    return LIB_PATH


# CABANA_AMD_ABI_VERSION of include/cabana_amd.h that the structs below mirror.
ABI_VERSION = 3


def load():
    """Loads libcabana_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    # CABANA_AMD_LIBRARY selects another build of the same library.
    path = os.environ.get("CABANA_AMD_LIBRARY", LIB_PATH)
    if not os.path.exists(path):
        raise RuntimeError(
            "cabana_amd: %s is missing; build it with `python -m cabana_amd.build` "
            "(there is no fallback path)" % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.cabana_amd_abi_version() != ABI_VERSION:
        raise RuntimeError("cabana_amd: %s has ABI version %d, these bindings need %d; rebuild it"
                           % (path, lib.cabana_amd_abi_version(), ABI_VERSION))
    _lib = lib
    return lib


class CabanaAmdError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("cabana_amd error %d: %s" % (status, message))
        self.status = status


def check(status, ctx=None):
    if status != OK:
        msg = load().cabana_amd_last_error(ctx)
        raise CabanaAmdError(status, msg.decode() if msg else "")


_contexts = {}


def context(device=None):
    """Returns the process-wide context of a HIP device (created lazily)."""
    if device is None:
        device = torch.cuda.current_device()
    ctx = _contexts.get(device)
    if ctx is None:
        lib = load()
        handle = c_void_p()
        check(lib.cabana_amd_context_create(device, ctypes.byref(handle)))
        ctx = handle
        _contexts[device] = ctx
    return ctx


def current_stream():
    return c_void_p(torch.cuda.current_stream().cuda_stream)
