"""ctypes binding of the C oracle (oracle/oracle.c) over numpy arrays.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check of the
package entry point and by bench.py's cpu_baseline leg, always as the
checker.  Build it with ``make -C oracle``.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

c_int = ctypes.c_int
c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
P = ctypes.POINTER

P2G_OPS = {"ScalarValue": 0, "VectorValue": 1, "ScalarGradient": 2, "VectorDivergence": 3,
           "TensorDivergence": 4}
G2P_OPS = {"ScalarValue": 0, "VectorValue": 1, "ScalarGradient": 2, "VectorGradient": 3,
           "VectorDivergence": 4}
# (particle components, grid dofs)
P2G_SHAPES = {0: (1, 1), 1: (3, 3), 2: (1, 3), 3: (3, 1), 4: (9, 3)}
G2P_SHAPES = {0: (1, 1), 1: (3, 3), 2: (3, 1), 3: (9, 3), 4: (1, 3)}
# Knots of the highest spline order (ORC_MAX_KNOT, Spline<6>).
MAX_KNOT = 7


class Mesh(ctypes.Structure):
    _fields_ = [("global_low", c_double * 3), ("global_high", c_double * 3),
                ("cell_size", c_double * 3), ("num_cell", c_int32 * 3), ("halo", c_int32),
                ("periodic", c_int32 * 3), ("ghost_low", c_double * 3),
                ("num_node", c_int32 * 3), ("entity", c_int32), ("stag", c_int32 * 3)]


# Entity codes (0 Node, 1 Cell, 2-4 Face<I,J,K>, 5-7 Edge<I,J,K>).
ENTITIES = {"node": 0, "cell": 1, "face_i": 2, "face_j": 3, "face_k": 4, "edge_i": 5,
            "edge_j": 6, "edge_k": 7}


class PField(ctypes.Structure):
    _fields_ = [("data", P(c_double)), ("n", c_int64), ("ncomp", c_int32), ("vl", c_int64),
                ("soa_stride", c_int64)]


class Grid(ctypes.Structure):
    _fields_ = [("data", P(c_double)), ("ext", c_int32 * 4), ("stride", c_int64 * 4)]


class SplineData(ctypes.Structure):
    _fields_ = [("order", c_int), ("num_knot", c_int), ("dx", c_double * 3),
                ("rdx", c_double * 3), ("x", c_double * 3), ("s", (c_int * MAX_KNOT) * 3),
                ("w", (c_double * MAX_KNOT) * 3), ("g", (c_double * MAX_KNOT) * 3),
                ("d", (c_double * MAX_KNOT) * 3)]


class CartGrid(ctypes.Structure):
    _fields_ = [("min", c_double * 3), ("max", c_double * 3), ("dx", c_double * 3),
                ("rdx", c_double * 3), ("nx", c_int32 * 3)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        lib = ctypes.CDLL(LIB)
        lib.orc_mesh_init.argtypes = [P(Mesh), P(c_double), P(c_double), P(c_double), c_int,
                                      P(c_int)]
        lib.orc_mesh_set_entity.argtypes = [P(Mesh), c_int]
        lib.orc_spline_map.restype = c_double
        lib.orc_spline_map.argtypes = [c_int, c_double, c_double, c_double]
        lib.orc_spline_stencil.argtypes = [c_int, c_double, P(c_int)]
        lib.orc_spline_value.argtypes = [c_int, c_double, P(c_double)]
        lib.orc_spline_gradient.argtypes = [c_int, c_double, c_double, P(c_double)]
        lib.orc_spline_hessian.argtypes = [c_int, c_double, c_double, P(c_double)]
        lib.orc_evaluate_spline.argtypes = [P(Mesh), c_int, P(c_double), P(SplineData)]
        lib.orc_node_coordinates.argtypes = [P(Mesh), P(c_int), P(c_double)]
        for name in ("orc_p2g", "orc_g2p"):
            pass
        lib.orc_p2g.argtypes = [P(Mesh), c_int, c_int, c_double, P(PField), P(PField), P(Grid),
                                c_int]
        lib.orc_g2p.argtypes = [P(Mesh), c_int, c_int, c_double, P(Grid), c_int, P(PField),
                                P(PField)]
        lib.orc_p2g_omp.argtypes = [P(Mesh), c_int, c_int, c_double, P(PField), P(PField),
                                    P(Grid), c_int, c_int]
        lib.orc_g2p_omp.argtypes = [P(Mesh), c_int, c_int, c_double, P(Grid), c_int,
                                    P(PField), P(PField), c_int]
        lib.orc_halo_gather.argtypes = [P(Mesh), P(Grid), c_int]
        lib.orc_halo_scatter_sum.argtypes = [P(Mesh), P(Grid), c_int]
        lib.orc_p2g_adjoint.argtypes = [P(Mesh), c_int, c_int, c_double, P(PField), P(PField),
                                        P(Grid), c_int, P(PField), P(PField), c_int]
        lib.orc_g2p_adjoint.argtypes = [P(Mesh), c_int, c_int, c_double, P(Grid), c_int,
                                        P(PField), P(PField), P(Grid), P(PField), c_int]
        lib.orc_cart_grid_init.argtypes = [P(CartGrid), P(c_double), P(c_double), P(c_double)]
        lib.orc_cart_locate.argtypes = [P(CartGrid), P(c_double), P(c_int)]
        lib.orc_cart_cardinal.restype = c_int64
        lib.orc_cart_cardinal.argtypes = [P(CartGrid), P(c_int)]
        lib.orc_lcl_build.argtypes = [P(PField), c_int64, c_int64, P(c_double), P(c_double),
                                      P(c_double), P(c_int32), P(c_int64), P(c_int64)]
        lib.orc_permute.argtypes = [P(c_int64), c_int64, c_int64, P(PField)]
        lib.orc_bov_write.argtypes = [P(Mesh), P(Grid), c_int, ctypes.c_char_p, ctypes.c_char_p,
                                      c_int, c_double, c_int]
        lib.orc_particle_migrate.restype = c_int64
        lib.orc_particle_migrate.argtypes = [P(Mesh), P(c_double), c_int64, c_int, c_int,
                                             P(c_int32)]
        _lib = lib
    return _lib


def _d3(v):
    return (c_double * 3)(*[float(x) for x in v])


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("oracle %s failed with code %d" % (what, rc))


class OMesh:
    """Single-rank uniform mesh: createUniformGlobalMesh + createGlobalGrid +
    createLocalGrid."""

    def __init__(self, low, high, cell_size, halo=1, periodic=(True, True, True),
                 entity="node"):
        if isinstance(cell_size, (int, float)):
            cell_size = [cell_size] * 3
        self.m = Mesh()
        per = (c_int * 3)(*[1 if p else 0 for p in periodic])
        _check(load().orc_mesh_init(ctypes.byref(self.m), _d3(low), _d3(high), _d3(cell_size),
                                    int(halo), per), "mesh init")
        _check(load().orc_mesh_set_entity(ctypes.byref(self.m), ENTITIES[entity]), "entity")

    @property
    def num_node(self):
        return tuple(self.m.num_node)

    @property
    def num_cell(self):
        return tuple(self.m.num_cell)

    @property
    def halo(self):
        return self.m.halo

    def grid(self, dofs):
        return np.zeros(self.num_node + (dofs,), dtype=np.float64)

    def owned_slices(self):
        sl = []
        for d in range(3):
            h = self.m.halo
            closing = 0 if (self.m.periodic[d] or self.m.stag[d]) else 1
            sl.append(slice(h, h + self.m.num_cell[d] + closing))
        return tuple(sl)


def pfield(arr, vl=None, soa_stride=None, ncomp=None, n=None):
    """PField over a float64 numpy array.  By default arr is (n, ...)
    row-major (vl = 1); pass vl/soa_stride/ncomp/n for AoSoA storage."""
    if arr.dtype != np.float64 or not arr.flags.c_contiguous:
        raise ValueError("oracle fields must be C-contiguous float64")
    f = PField()
    f.data = arr.ctypes.data_as(P(c_double))
    if vl is None:
        f.n = arr.shape[0]
        f.ncomp = int(np.prod(arr.shape[1:])) if arr.ndim > 1 else 1
        f.vl = 1
        f.soa_stride = f.ncomp
    else:
        f.n = n
        f.ncomp = ncomp
        f.vl = vl
        f.soa_stride = soa_stride
    f._keep = arr
    return f


def gridfield(arr):
    if arr.dtype != np.float64 or arr.ndim != 4:
        raise ValueError("oracle grids are float64 (ni, nj, nk, dofs)")
    g = Grid()
    g.data = arr.ctypes.data_as(P(c_double))
    for d in range(4):
        g.ext[d] = arr.shape[d]
        g.stride[d] = arr.strides[d] // 8
    g._keep = arr
    return g


def p2g(mesh, order, op, multiplier, points, values, grid, halo_width=-1, threads=0):
    """p2g (Serial order when threads == 0, OpenMP duplicated grids otherwise)."""
    pts = pfield(np.ascontiguousarray(points, dtype=np.float64))
    vals = values if isinstance(values, PField) else pfield(
        np.ascontiguousarray(values, dtype=np.float64))
    g = gridfield(grid)
    lib = load()
    if threads:
        rc = lib.orc_p2g_omp(ctypes.byref(mesh.m), order, op, multiplier, ctypes.byref(pts),
                             ctypes.byref(vals), ctypes.byref(g), halo_width, threads)
    else:
        rc = lib.orc_p2g(ctypes.byref(mesh.m), order, op, multiplier, ctypes.byref(pts),
                         ctypes.byref(vals), ctypes.byref(g), halo_width)
    _check(rc, "p2g")


def g2p(mesh, order, op, multiplier, grid, points, values, halo_width=-1, threads=0):
    """g2p; values (float64, C-contiguous) are accumulated in place."""
    pts = pfield(np.ascontiguousarray(points, dtype=np.float64))
    vals = values if isinstance(values, PField) else pfield(values)
    g = gridfield(grid)
    lib = load()
    if threads:
        rc = lib.orc_g2p_omp(ctypes.byref(mesh.m), order, op, multiplier, ctypes.byref(g),
                             halo_width, ctypes.byref(pts), ctypes.byref(vals), threads)
    else:
        rc = lib.orc_g2p(ctypes.byref(mesh.m), order, op, multiplier, ctypes.byref(g),
                         halo_width, ctypes.byref(pts), ctypes.byref(vals))
    _check(rc, "g2p")


def p2g_adjoint(mesh, order, op, multiplier, points, values, grid_bar, values_bar=None,
                points_bar=None, halo_width=-1, threads=0):
    pts = pfield(np.ascontiguousarray(points, dtype=np.float64))
    vals = pfield(np.ascontiguousarray(values, dtype=np.float64))
    g = gridfield(grid_bar)
    vb = pfield(values_bar) if values_bar is not None else None
    xb = pfield(points_bar) if points_bar is not None else None
    _check(load().orc_p2g_adjoint(ctypes.byref(mesh.m), order, op, multiplier,
                                  ctypes.byref(pts), ctypes.byref(vals), ctypes.byref(g),
                                  halo_width, ctypes.byref(vb) if vb else None,
                                  ctypes.byref(xb) if xb else None, threads), "p2g adjoint")


def g2p_adjoint(mesh, order, op, multiplier, grid, points, values_bar, grid_bar=None,
                points_bar=None, halo_width=-1, threads=0):
    pts = pfield(np.ascontiguousarray(points, dtype=np.float64))
    g = gridfield(grid)
    vb = pfield(np.ascontiguousarray(values_bar, dtype=np.float64))
    gb = gridfield(grid_bar) if grid_bar is not None else None
    xb = pfield(points_bar) if points_bar is not None else None
    _check(load().orc_g2p_adjoint(ctypes.byref(mesh.m), order, op, multiplier, ctypes.byref(g),
                                  halo_width, ctypes.byref(pts), ctypes.byref(vb),
                                  ctypes.byref(gb) if gb else None,
                                  ctypes.byref(xb) if xb else None, threads), "g2p adjoint")


def halo_gather(mesh, grid, halo_width=-1):
    _check(load().orc_halo_gather(ctypes.byref(mesh.m), ctypes.byref(gridfield(grid)),
                                  halo_width), "halo gather")


def halo_scatter_sum(mesh, grid, halo_width=-1):
    _check(load().orc_halo_scatter_sum(ctypes.byref(mesh.m), ctypes.byref(gridfield(grid)),
                                       halo_width), "halo scatter")


def spline_1d(order, xp, rdx, low_x):
    """(x0, stencil, values, gradients, hessians) of one coordinate."""
    lib = load()
    x0 = lib.orc_spline_map(order, xp, rdx, low_x)
    k = order + 1
    s = (c_int * k)()
    w = (c_double * k)()
    g = (c_double * k)()
    h = (c_double * k)()
    lib.orc_spline_stencil(order, x0, s)
    lib.orc_spline_value(order, x0, w)
    lib.orc_spline_gradient(order, x0, rdx, g)
    lib.orc_spline_hessian(order, x0, rdx, h)
    return x0, list(s), list(w), list(g), list(h)


def evaluate_spline(mesh, order, p):
    sd = SplineData()
    load().orc_evaluate_spline(ctypes.byref(mesh.m), order, _d3(p), ctypes.byref(sd))
    k = order + 1
    return {
        "dx": list(sd.dx), "rdx": list(sd.rdx), "x": list(sd.x),
        "s": [list(sd.s[d])[:k] for d in range(3)],
        "w": [list(sd.w[d])[:k] for d in range(3)],
        "g": [list(sd.g[d])[:k] for d in range(3)],
        "d": [list(sd.d[d])[:k] for d in range(3)],
    }


def cart_grid(grid_min, grid_max, grid_delta):
    cg = CartGrid()
    load().orc_cart_grid_init(ctypes.byref(cg), _d3(grid_min), _d3(grid_max), _d3(grid_delta))
    return cg


def cart_locate(cg, p):
    ijk = (c_int * 3)()
    load().orc_cart_locate(ctypes.byref(cg), _d3(p), ijk)
    return tuple(ijk)


def lcl_build(positions, grid_delta, grid_min, grid_max, begin=0, end=None, vl=None,
              soa_stride=None, n=None):
    """LinkedCellList::build in Serial order -> (counts, offsets, permutes, nbin)."""
    if vl is None:
        pos = pfield(np.ascontiguousarray(positions, dtype=np.float64))
    else:
        pos = pfield(positions, vl=vl, soa_stride=soa_stride, ncomp=3, n=n)
    if end is None:
        end = pos.n
    cg = cart_grid(grid_min, grid_max, grid_delta)
    nbin = tuple(cg.nx)
    ncell = nbin[0] * nbin[1] * nbin[2]
    counts = np.zeros(ncell, dtype=np.int32)
    offsets = np.zeros(ncell, dtype=np.int64)
    permutes = np.zeros(max(end - begin, 1), dtype=np.int64)
    _check(load().orc_lcl_build(ctypes.byref(pos), begin, end, _d3(grid_delta), _d3(grid_min),
                                _d3(grid_max), counts.ctypes.data_as(P(c_int32)),
                                offsets.ctypes.data_as(P(c_int64)),
                                permutes.ctypes.data_as(P(c_int64))), "lcl build")
    return counts, offsets, permutes[:end - begin], nbin


def permute(permutes, begin, end, field):
    """In-place permute of a PField (see pfield)."""
    perm = np.ascontiguousarray(permutes, dtype=np.int64)
    _check(load().orc_permute(perm.ctypes.data_as(P(c_int64)), begin, end,
                              ctypes.byref(field)), "permute")


def particle_migrate(mesh, positions, min_halo_width, force=False):
    """particleMigrate on one rank -> (migrated, shifted positions, keep
    flags).  The positions array is not modified; keep is None when nothing
    migrates."""
    x = np.ascontiguousarray(positions, dtype=np.float64).copy()
    n = x.shape[0]
    keep = np.zeros(max(n, 1), dtype=np.int32)
    kept = load().orc_particle_migrate(ctypes.byref(mesh.m), x.ctypes.data_as(P(c_double)), n,
                                       int(min_halo_width), 1 if force else 0,
                                       keep.ctypes.data_as(P(c_int32)))
    if kept < 0:
        return False, x, None
    return True, x, keep[:n].astype(bool)


def bov_write(mesh, grid, entity, label, prefix, time_step_index, time, gather_array=True):
    """BovWriter::writeTimeStep on one rank over an FP64 (i, j, k, dofs)
    numpy grid (node or cell extents; entity "node" or "cell")."""
    rc = load().orc_bov_write(ctypes.byref(mesh.m), ctypes.byref(gridfield(grid)),
                              0 if entity == "node" else 1, label.encode(), prefix.encode(),
                              int(time_step_index), float(time), 1 if gather_array else 0)
    _check(rc, "bov write")
