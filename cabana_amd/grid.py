"""Host-side mirror of Cabana::Grid's interpolation interface over the C ABI.

Names, argument meaning and error behaviour follow the reference
(grid/src/Cabana_Grid_Interpolation.hpp, Cabana_Grid_GlobalMesh.hpp,
Cabana_Grid_LocalGrid.hpp, Cabana_Grid_Array.hpp, Cabana_Grid_Halo*.hpp,
core/src/Cabana_LinkedCellList.hpp, core/src/Cabana_Sort.hpp).  Device memory
is held in torch tensors; every computation runs in the HIP library.
"""
import ctypes
import math

import torch

from . import _native as N

# ---------------------------------------------------------------------------
# Tags (grid/src/Cabana_Grid_Types.hpp)
# ---------------------------------------------------------------------------


class Own:
    pass


class Ghost:
    pass


class Local:
    pass


class Node:
    code = 0


class Cell:
    code = 1


class Dim:
    I, J, K = 0, 1, 2


class Face:
    """Face<Dir> tag: faces normal to dimension Dir."""

    def __init__(self, dim):
        if dim not in (0, 1, 2):
            raise ValueError("face dimension must be 0, 1 or 2")
        self.dim = dim
        self.code = 2 + dim


class Edge:
    """Edge<Dir> tag: edges along dimension Dir."""

    def __init__(self, dim):
        if dim not in (0, 1, 2):
            raise ValueError("edge dimension must be 0, 1 or 2")
        self.dim = dim
        self.code = 5 + dim


def _stag(entity, d):
    """1 where the entity sits at cell centres in dimension d
    (Cabana_Grid_LocalMesh.hpp:176-237)."""
    if isinstance(entity, Node):
        return 0
    if isinstance(entity, Cell):
        return 1
    if isinstance(entity, Face):
        return 0 if d == entity.dim else 1
    return 1 if d == entity.dim else 0


class NodeHaloPattern:
    """26-neighbour node halo pattern (Cabana_Grid_HaloBase.hpp:83-99)."""


class ScatterReduce:
    class Sum:
        pass


class Spline:
    """Spline<Order> tag (Cabana_Grid_Splines.hpp)."""

    def __init__(self, order):
        if order not in range(7):
            raise ValueError("cabana_amd supports Spline<0> to Spline<6>")
        self.order = order
        self.num_knot = order + 1


# ---------------------------------------------------------------------------
# Mesh / grid geometry
# ---------------------------------------------------------------------------


class GlobalMesh:
    """Uniform global mesh (Cabana_Grid_GlobalMesh.hpp:59-165)."""

    def __init__(self, low, high, cell_size):
        self._low = [float(v) for v in low]
        self._high = [float(v) for v in high]
        if isinstance(cell_size, (int, float)):
            cell_size = [float(cell_size)] * 3
        self._cell = [float(v) for v in cell_size]
        for d in range(3):
            ext = self.globalNumCell(d) * self._cell[d]
            if abs(ext - self.extent(d)) > 100.0 * 2.220446049250313e-16:
                raise RuntimeError(
                    "Cabana::Grid::GlobalMesh: Extent not evenly divisible by uniform cell size")

    def lowCorner(self, d):
        return self._low[d]

    def highCorner(self, d):
        return self._high[d]

    def extent(self, d):
        return self._high[d] - self._low[d]

    def cellSize(self, d):
        return self._cell[d]

    def globalNumCell(self, d):
        # std::rint rounds half to even, as does Python's round().
        return int(round(self.extent(d) / self._cell[d]))


def createUniformGlobalMesh(low, high, cell_size):
    """createUniformGlobalMesh(low, high, cell_size) with one cell size or one
    per dimension, or with the number of cells per dimension (a sequence of
    ints: cell size (high - low) / num_cell, Cabana_Grid_GlobalMesh.hpp:105-130)."""
    if not isinstance(cell_size, (int, float)) and all(
            isinstance(v, int) and not isinstance(v, bool) for v in cell_size):
        cell_size = [(float(high[d]) - float(low[d])) / cell_size[d] for d in range(3)]
    return GlobalMesh(low, high, cell_size)


class GlobalGrid:
    """Global grid on one rank (no MPI): every block is block 0 of 1."""

    def __init__(self, global_mesh, is_dim_periodic):
        self._mesh = global_mesh
        self._periodic = [bool(p) for p in is_dim_periodic]

    def globalMesh(self):
        return self._mesh

    def isPeriodic(self, d):
        return self._periodic[d]

    def ownedNumCell(self, d):
        return self._mesh.globalNumCell(d)

    def globalNumEntity(self, entity, d):
        n = self._mesh.globalNumCell(d)
        return n if isinstance(entity, Cell) or self._periodic[d] else n + 1


def createGlobalGrid(global_mesh, is_dim_periodic, partitioner=None):
    return GlobalGrid(global_mesh, is_dim_periodic)


class IndexSpace:
    def __init__(self, mins, maxs):
        self._min = list(mins)
        self._max = list(maxs)

    def min(self, d):
        return self._min[d]

    def max(self, d):
        return self._max[d]

    def extent(self, d):
        return self._max[d] - self._min[d]

    def size(self):
        s = 1
        for d in range(len(self._min)):
            s *= self.extent(d)
        return s


class LocalGrid:
    """Local grid with ghost layers (Cabana_Grid_LocalGrid_impl.hpp)."""

    def __init__(self, global_grid, halo_cell_width):
        self._gg = global_grid
        self._h = int(halo_cell_width)

    def globalGrid(self):
        return self._gg

    def haloCellWidth(self):
        return self._h

    def totalNumCell(self, d):
        return self._gg.ownedNumCell(d) + 2 * self._h

    def indexSpace(self, decomposition, entity, index_type=None):
        """Own / Ghost local index space of any entity
        (Cabana_Grid_LocalGrid_impl.hpp:239-280, 430-470, 534-560): per
        dimension N entities plus, where the entity is node-centred, the
        ghosted extra node and, in a non-periodic dimension, the owned
        closing node."""
        h = self._h
        stag = [_stag(entity, d) for d in range(3)]
        if isinstance(decomposition, Own):
            return IndexSpace([h] * 3, [
                h + self._gg.ownedNumCell(d) + (0 if (stag[d] or self._gg.isPeriodic(d)) else 1)
                for d in range(3)])
        return IndexSpace([0] * 3, [self.totalNumCell(d) + 1 - stag[d] for d in range(3)])

    def mesh_descriptor(self, entity=None):
        gm = self._gg.globalMesh()
        m = N.Mesh()
        for d in range(3):
            m.low_corner[d] = gm.lowCorner(d)
            m.high_corner[d] = gm.highCorner(d)
            m.cell_size[d] = gm.cellSize(d)
            m.periodic[d] = 1 if self._gg.isPeriodic(d) else 0
        m.halo_cell_width = self._h
        m.entity = entity.code if entity is not None else 0
        return m


def createLocalGrid(global_grid, halo_cell_width):
    return LocalGrid(global_grid, halo_cell_width)


class LocalMesh:
    """Uniform local mesh geometry (Cabana_Grid_LocalMesh.hpp:61-200)."""

    def __init__(self, local_grid):
        gm = local_grid.globalGrid().globalMesh()
        self._cell = [gm.cellSize(d) for d in range(3)]
        self._own_low = [gm.lowCorner(d) + self._cell[d] * 0 for d in range(3)]
        self._ghost_low = [self._own_low[d] - local_grid.haloCellWidth() * self._cell[d]
                           for d in range(3)]

    def lowCorner(self, decomposition, d):
        return self._own_low[d] if isinstance(decomposition, Own) else self._ghost_low[d]

    def coordinates(self, entity, index):
        if isinstance(entity, Cell):
            return [self._ghost_low[d] + (float(index[d]) + 0.5) * self._cell[d]
                    for d in range(3)]
        return [self._ghost_low[d] + float(index[d]) * self._cell[d] for d in range(3)]


def createLocalMesh(local_grid):
    return LocalMesh(local_grid)


# ---------------------------------------------------------------------------
# Arrays
# ---------------------------------------------------------------------------


class ArrayLayout:
    """Array layout of any entity (Node, Cell, Face(d), Edge(d))."""

    def __init__(self, local_grid, dofs, entity):
        if not isinstance(entity, (Node, Cell, Face, Edge)):
            raise TypeError("entity must be Node(), Cell(), Face(d) or Edge(d)")
        self._lg = local_grid
        self._dofs = int(dofs)
        self.entity = entity

    def localGrid(self):
        return self._lg

    def dofsPerEntity(self):
        return self._dofs

    def indexSpace(self, decomposition, index_type=None):
        s = self._lg.indexSpace(decomposition, self.entity)
        return IndexSpace(list(s._min) + [0], list(s._max) + [self._dofs])


def createArrayLayout(local_grid, dofs_per_entity, entity):
    return ArrayLayout(local_grid, dofs_per_entity, entity)


class Array:
    """Node array over the ghosted index space; view() is a device tensor
    (n_i, n_j, n_k, dofs).  layout="right" gives a row-major tensor
    (Kokkos LayoutRight), layout="left" a column-major one (LayoutLeft)."""

    def __init__(self, label, layout, dtype=torch.float64, device=None, memory_layout="right"):
        self.label = label
        self._layout = layout
        lg = layout.localGrid()
        shape = [lg.totalNumCell(d) + 1 - _stag(layout.entity, d) for d in range(3)] + \
            [layout.dofsPerEntity()]
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        if memory_layout == "right":
            self._view = torch.zeros(shape, dtype=dtype, device=device)
        elif memory_layout == "left":
            self._view = torch.zeros(shape[::-1], dtype=dtype, device=device).permute(3, 2, 1, 0)
        else:
            raise ValueError("memory_layout must be 'right' or 'left'")

    def layout(self):
        return self._layout

    def view(self):
        return self._view

    def descriptor(self):
        return grid_descriptor(self._view)

    def mesh_descriptor(self):
        """The local grid's mesh with this array's entity."""
        return self._layout.localGrid().mesh_descriptor(self._layout.entity)


def _call_mesh(arrays):
    """Mesh descriptor of a call: every array of one call has one entity."""
    codes = {a.layout().entity.code for a in arrays}
    if len(codes) != 1:
        raise ValueError("the arrays of one call must share one entity")
    return arrays[0].mesh_descriptor()


def createArray(label, layout, dtype=torch.float64, device=None, memory_layout="right"):
    return Array(label, layout, dtype, device, memory_layout)


def _dtype_code(t):
    if t.dtype == torch.float64:
        return N.FLOAT64
    if t.dtype == torch.float32:
        return N.FLOAT32
    if t.dtype == torch.int32:
        return N.INT32
    if t.dtype == torch.int64:
        return N.INT64
    raise TypeError("unsupported dtype %s" % t.dtype)


def _require_device(t, what):
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError("%s must be a HIP device tensor" % what)


def grid_descriptor(view):
    if view is None:
        return N.GridField()
    _require_device(view, "grid array")
    if view.dim() != 4:
        raise ValueError("grid arrays are (n_i, n_j, n_k, dofs)")
    g = N.GridField()
    g.data = view.data_ptr()
    for d in range(4):
        g.extent[d] = view.shape[d]
        g.stride[d] = view.stride(d)
    g.dtype = _dtype_code(view)
    return g


class ArrayOp:
    @staticmethod
    def assign(array, value, decomposition):
        """ArrayOp::assign over the owned or the ghosted index space
        (Cabana_Grid_Array.hpp:431-441)."""
        v = array.view()
        if isinstance(decomposition, Own):
            s = array.layout().localGrid().indexSpace(Own(), array.layout().entity)
            v[s.min(0):s.max(0), s.min(1):s.max(1), s.min(2):s.max(2), :] = value
        else:
            v.fill_(value)


# ---------------------------------------------------------------------------
# Particle fields (slices)
# ---------------------------------------------------------------------------


class AoSoASlice:
    """A Cabana slice over raw AoSoA storage: element (p, c) at
    data[(p // VL) * soa_stride + p % VL + VL * c] (Cabana_Slice.hpp:119-190)."""

    def __init__(self, data, num_particle, shape, vector_length, soa_stride):
        _require_device(data, "slice data")
        self.data = data
        self.num_particle = int(num_particle)
        self.shape = tuple(shape)
        self.num_comp = 1
        for s in self.shape:
            self.num_comp *= s
        self.vector_length = int(vector_length)
        self.soa_stride = int(soa_stride)

    def offsets(self):
        p = torch.arange(self.num_particle, device=self.data.device)
        base = (p // self.vector_length) * self.soa_stride + p % self.vector_length
        c = torch.arange(self.num_comp, device=self.data.device)
        return base[:, None] + self.vector_length * c[None, :]

    def to_tensor(self):
        """(n, *shape) copy of the slice values."""
        flat = self.data.reshape(-1)[self.offsets().reshape(-1)]
        return flat.reshape((self.num_particle,) + self.shape)

    def from_tensor(self, t):
        flat = self.data.reshape(-1)
        flat[self.offsets().reshape(-1)] = t.reshape(-1).to(self.data.dtype)


def particle_descriptor(x, num_comp=None):
    """ParticleField descriptor of a torch tensor (n,), (n, c) or (n, a, b)
    with row-major or column-major particle data, or of an AoSoASlice."""
    f = N.ParticleField()
    if x is None:
        return f
    if isinstance(x, AoSoASlice):
        f.data = x.data.data_ptr()
        f.num_particle = x.num_particle
        f.num_comp = x.num_comp
        f.dtype = _dtype_code(x.data)
        f.vector_length = x.vector_length
        f.soa_stride = x.soa_stride
        return f
    _require_device(x, "particle field")
    n = x.shape[0] if x.dim() > 0 else 0
    comp_shape = tuple(x.shape[1:])
    nc = 1
    for s in comp_shape:
        nc *= s
    f.data = x.data_ptr()
    f.num_particle = n
    f.num_comp = nc
    f.dtype = _dtype_code(x)
    if n == 0:
        f.vector_length = 1
        f.soa_stride = nc
        return f
    if nc == 1:
        if x.dim() > 1:
            x = x.reshape(n)
        if x.stride(0) != 1:
            raise ValueError("scalar particle fields must be contiguous")
        f.vector_length = max(n, 1)
        f.soa_stride = max(n, 1)
        return f
    # Components must be laid out as one flattened, row-major index c.
    inner = x.reshape(n, nc) if x.is_contiguous() else None
    if inner is not None:
        f.vector_length = 1
        f.soa_stride = nc
        return f
    strides = x.stride()
    # Column-major (component-major) storage: (p, c) at p + n * c.
    comp_ok = all(strides[1 + i] == n * int(math.prod(comp_shape[i + 1:]))
                  for i in range(len(comp_shape)))
    if strides[0] == 1 and comp_ok:
        f.vector_length = n
        f.soa_stride = n * nc
        return f
    # Row-major components inside a wider row: (p, c) at p * s0 + c.
    row_ok = all(strides[1 + i] == int(math.prod(comp_shape[i + 1:]))
                 for i in range(len(comp_shape)))
    if row_ok and strides[0] >= nc:
        f.vector_length = 1
        f.soa_stride = strides[0]
        return f
    raise ValueError("unsupported particle field layout %s" % (strides,))


# ---------------------------------------------------------------------------
# Halo
# ---------------------------------------------------------------------------


class Halo:
    """Node halo of one rank whose periodic neighbours are itself
    (Cabana_Grid_HaloBase.hpp, impl/Cabana_Grid_Halo_Mpi.hpp)."""

    def __init__(self, pattern, width, arrays):
        self._width = int(width)
        self._arrays = list(arrays)
        lg = self._arrays[0].layout().localGrid()
        hw = lg.haloCellWidth() if self._width == -1 else self._width
        if hw > lg.haloCellWidth():
            raise RuntimeError("Requested halo width larger than local grid halo")
        self._local_grid = lg

    def width(self):
        return self._width

    def gather(self, *arrays):
        for a in arrays or self._arrays:
            m = a.mesh_descriptor()
            ctx = N.context()
            g = a.descriptor()
            N.check(N.load().cabana_amd_halo_gather(ctx, ctypes.byref(m), ctypes.byref(g),
                                                    self._width, N.current_stream()), ctx)

    def scatter(self, reduce_op, *arrays):
        if reduce_op is not ScatterReduce.Sum and not isinstance(reduce_op, ScatterReduce.Sum):
            raise NotImplementedError("cabana_amd implements ScatterReduce::Sum")
        for a in arrays or self._arrays:
            m = a.mesh_descriptor()
            ctx = N.context()
            g = a.descriptor()
            N.check(N.load().cabana_amd_halo_scatter_sum(ctx, ctypes.byref(m), ctypes.byref(g),
                                                         self._width, N.current_stream()), ctx)


def createHalo(pattern, width, *arrays):
    return Halo(pattern, width, arrays)


# ---------------------------------------------------------------------------
# Interpolation functors (Cabana_Grid_Interpolation.hpp:938-1566)
# ---------------------------------------------------------------------------


class _Functor:
    op = None
    rank = None
    direction = None

    def __init__(self, x, multiplier):
        shape = x.shape if isinstance(x, AoSoASlice) else tuple(x.shape[1:])
        rank = 1 + len(x.shape) if isinstance(x, AoSoASlice) else x.dim()
        if rank != self.rank:
            raise ValueError("%s: View must be of rank %d" % (type(self).__name__, self.rank))
        self._x = x
        self._multiplier = float(multiplier)
        self._shape = shape


class ScalarValueP2G(_Functor):
    op, rank, direction = N.SCALAR_VALUE_P2G, 1, "p2g"


class VectorValueP2G(_Functor):
    op, rank, direction = N.VECTOR_VALUE_P2G, 2, "p2g"


class ScalarGradientP2G(_Functor):
    op, rank, direction = N.SCALAR_GRADIENT_P2G, 1, "p2g"


class VectorDivergenceP2G(_Functor):
    op, rank, direction = N.VECTOR_DIVERGENCE_P2G, 2, "p2g"


class TensorDivergenceP2G(_Functor):
    op, rank, direction = N.TENSOR_DIVERGENCE_P2G, 3, "p2g"


class ScalarValueG2P(_Functor):
    op, rank, direction = N.SCALAR_VALUE_G2P, 1, "g2p"


class VectorValueG2P(_Functor):
    op, rank, direction = N.VECTOR_VALUE_G2P, 2, "g2p"


class ScalarGradientG2P(_Functor):
    op, rank, direction = N.SCALAR_GRADIENT_G2P, 2, "g2p"


class VectorGradientG2P(_Functor):
    op, rank, direction = N.VECTOR_GRADIENT_G2P, 3, "g2p"


class VectorDivergenceG2P(_Functor):
    op, rank, direction = N.VECTOR_DIVERGENCE_G2P, 1, "g2p"


def createScalarValueP2G(x, multiplier):
    return ScalarValueP2G(x, multiplier)


def createVectorValueP2G(x, multiplier):
    return VectorValueP2G(x, multiplier)


def createScalarGradientP2G(x, multiplier):
    return ScalarGradientP2G(x, multiplier)


def createVectorDivergenceP2G(x, multiplier):
    return VectorDivergenceP2G(x, multiplier)


def createTensorDivergenceP2G(x, multiplier):
    return TensorDivergenceP2G(x, multiplier)


def createScalarValueG2P(x, multiplier):
    return ScalarValueG2P(x, multiplier)


def createVectorValueG2P(x, multiplier):
    return VectorValueG2P(x, multiplier)


def createScalarGradientG2P(x, multiplier):
    return ScalarGradientG2P(x, multiplier)


def createVectorGradientG2P(x, multiplier):
    return VectorGradientG2P(x, multiplier)


def createVectorDivergenceG2P(x, multiplier):
    return VectorDivergenceG2P(x, multiplier)


def _points_descriptor(points, num_point):
    f = particle_descriptor(points)
    if f.num_comp != 3:
        raise ValueError("points must be indexed (point, dim) with 3 dims")
    if num_point > f.num_particle:
        raise ValueError("num_point exceeds the number of points")
    f.num_particle = int(num_point)
    return f


def _values_descriptor(x, num_point):
    f = particle_descriptor(x)
    if f.num_particle < num_point:
        raise ValueError("particle field shorter than num_point")
    return f


def _bins_descriptor(bins):
    if bins is None:
        return None
    return bins.binning_descriptor()


def _halo_width(halo):
    return -1 if halo is None else halo.width()


def p2g(functor, points, num_point, spline, halo, array, bins=None):
    """p2g(functor, points, num_point, Spline<N>, halo, array)
    (Cabana_Grid_Interpolation.hpp:1217-1313).  ``functor``/``array`` may be
    lists to interpolate several functors in one pass."""
    functors = functor if isinstance(functor, (list, tuple)) else [functor]
    arrays = array if isinstance(array, (list, tuple)) else [array] * len(functors)
    if len(arrays) != len(functors):
        raise ValueError("one array per functor")
    lg = arrays[0].layout().localGrid()
    m = _call_mesh(arrays)
    pts = _points_descriptor(points, num_point)
    terms = (N.P2GTerm * len(functors))()
    for i, (f, a) in enumerate(zip(functors, arrays)):
        if f.direction != "p2g":
            raise TypeError("p2g needs P2G functors")
        terms[i].op = f.op
        terms[i].multiplier = f._multiplier
        terms[i].values = _values_descriptor(f._x, num_point)
        terms[i].grid = a.descriptor()
    b = _bins_descriptor(bins)
    ctx = N.context()
    N.check(N.load().cabana_amd_p2g(ctx, ctypes.byref(m), spline.order, ctypes.byref(pts),
                                    len(functors), terms, _halo_width(halo),
                                    ctypes.byref(b) if b is not None else None,
                                    N.current_stream()), ctx)


def g2p(array, halo, points, num_point, spline, functor, bins=None):
    """g2p(array, halo, points, num_point, Spline<N>, functor)
    (Cabana_Grid_Interpolation.hpp:831-927).  ``array``/``functor`` may be
    lists."""
    functors = functor if isinstance(functor, (list, tuple)) else [functor]
    arrays = array if isinstance(array, (list, tuple)) else [array] * len(functors)
    if len(arrays) != len(functors):
        raise ValueError("one array per functor")
    lg = arrays[0].layout().localGrid()
    m = _call_mesh(arrays)
    pts = _points_descriptor(points, num_point)
    terms = (N.G2PTerm * len(functors))()
    for i, (f, a) in enumerate(zip(functors, arrays)):
        if f.direction != "g2p":
            raise TypeError("g2p needs G2P functors")
        terms[i].op = f.op
        terms[i].multiplier = f._multiplier
        terms[i].values = _values_descriptor(f._x, num_point)
        terms[i].grid = a.descriptor()
    b = _bins_descriptor(bins)
    ctx = N.context()
    N.check(N.load().cabana_amd_g2p(ctx, ctypes.byref(m), spline.order, ctypes.byref(pts),
                                    len(functors), terms, _halo_width(halo),
                                    ctypes.byref(b) if b is not None else None,
                                    N.current_stream()), ctx)


def _terms(kind, functors, arrays, num_point):
    terms = (N.P2GTerm if kind == "p2g" else N.G2PTerm) * len(functors)
    out = terms()
    for i, (f, a) in enumerate(zip(functors, arrays)):
        if f.direction != kind:
            raise TypeError("expected %s functors" % kind.upper())
        out[i].op = f.op
        out[i].multiplier = f._multiplier
        out[i].values = _values_descriptor(f._x, num_point)
        out[i].grid = a.descriptor()
    return out


def g2p2g(p2g_functor, points, num_point, spline, halo, array, g2p_functor, g2p_array=None,
          bins=None):
    """g2p2g(scalar_p2g, points, num_point, halo, scalar_array, scalar_g2p)
    (Cabana_Grid_InterpolationPerformance.cpp:130-174) for any functors and
    spline order: gather the G2P arrays (``g2p_array``, else ``array``),
    G2P of ``g2p_functor`` (each adds into its particle field), then P2G of
    ``p2g_functor``, reading the particle fields as the G2P left them, into
    ``array``, then fold.  Every G2P read sees the arrays as gathered at the
    start of the call, as the reference's Serial and OpenMP scatter views
    do.  Functors and arrays may be lists; without a cell binning it is one
    pass over the particles."""
    pf = p2g_functor if isinstance(p2g_functor, (list, tuple)) else [p2g_functor]
    pa = array if isinstance(array, (list, tuple)) else [array] * len(pf)
    gf = g2p_functor if isinstance(g2p_functor, (list, tuple)) else [g2p_functor]
    ga = g2p_array if g2p_array is not None else array
    ga = ga if isinstance(ga, (list, tuple)) else [ga] * len(gf)
    if len(pa) != len(pf) or len(ga) != len(gf):
        raise ValueError("one array per functor")
    m = _call_mesh(list(pa) + list(ga))
    pts = _points_descriptor(points, num_point)
    gt = _terms("g2p", gf, ga, num_point)
    pt = _terms("p2g", pf, pa, num_point)
    b = _bins_descriptor(bins)
    ctx = N.context()
    N.check(N.load().cabana_amd_g2p2g(ctx, ctypes.byref(m), spline.order, ctypes.byref(pts),
                                      len(gf), gt, len(pf), pt, _halo_width(halo),
                                      ctypes.byref(b) if b is not None else None,
                                      N.current_stream()), ctx)


def p2g_adjoint(functor, points, num_point, spline, halo, grid_bar, values_bar,
                points_bar=None, bins=None):
    """Adjoint of p2g: values_bar (+=) and points_bar (+=) from the cotangent
    of the folded grid.  Lists give a fused pass."""
    functors = functor if isinstance(functor, (list, tuple)) else [functor]
    gbars = grid_bar if isinstance(grid_bar, (list, tuple)) else [grid_bar]
    vbars = values_bar if isinstance(values_bar, (list, tuple)) else [values_bar]
    lg = gbars[0].layout().localGrid()
    m = _call_mesh(gbars)
    pts = _points_descriptor(points, num_point)
    terms = (N.P2GAdjointTerm * len(functors))()
    for i, f in enumerate(functors):
        terms[i].op = f.op
        terms[i].multiplier = f._multiplier
        terms[i].values = _values_descriptor(f._x, num_point)
        terms[i].grid_bar = gbars[i].descriptor()
        terms[i].values_bar = particle_descriptor(vbars[i]) if vbars[i] is not None \
            else N.ParticleField()
    xb = particle_descriptor(points_bar) if points_bar is not None else None
    b = _bins_descriptor(bins)
    ctx = N.context()
    N.check(N.load().cabana_amd_p2g_adjoint(
        ctx, ctypes.byref(m), spline.order, ctypes.byref(pts), len(functors), terms,
        _halo_width(halo), ctypes.byref(xb) if xb is not None else None,
        ctypes.byref(b) if b is not None else None, N.current_stream()), ctx)


def g2p_adjoint(array, halo, points, num_point, spline, functor, values_bar, grid_bar,
                points_bar=None, bins=None):
    """Adjoint of g2p: grid_bar (+= on owned nodes) and points_bar (+=) from the
    cotangent of the particle output.  Lists give a fused pass."""
    functors = functor if isinstance(functor, (list, tuple)) else [functor]
    arrays = array if isinstance(array, (list, tuple)) else [array] * len(functors)
    vbars = values_bar if isinstance(values_bar, (list, tuple)) else [values_bar]
    gbars = grid_bar if isinstance(grid_bar, (list, tuple)) else [grid_bar] * len(functors)
    lg = arrays[0].layout().localGrid()
    m = _call_mesh(list(arrays) + [g for g in gbars if g is not None])
    pts = _points_descriptor(points, num_point)
    terms = (N.G2PAdjointTerm * len(functors))()
    for i, f in enumerate(functors):
        terms[i].op = f.op
        terms[i].multiplier = f._multiplier
        terms[i].grid = arrays[i].descriptor()
        terms[i].values_bar = _values_descriptor(vbars[i], num_point)
        terms[i].grid_bar = gbars[i].descriptor() if gbars[i] is not None else N.GridField()
    xb = particle_descriptor(points_bar) if points_bar is not None else None
    b = _bins_descriptor(bins)
    ctx = N.context()
    N.check(N.load().cabana_amd_g2p_adjoint(
        ctx, ctypes.byref(m), spline.order, ctypes.byref(pts), len(functors), terms,
        _halo_width(halo), ctypes.byref(xb) if xb is not None else None,
        ctypes.byref(b) if b is not None else None, N.current_stream()), ctx)


def adjoint_gather(p2g_functor, array, halo, points, num_point, spline, grid_bar,
                   p2g_values_bar, g2p_functor, g2p_values_bar, points_bar, bins=None):
    """The gather-shaped parts of both adjoints of a transfer pair in one
    call: p2g_adjoint(p2g_functor, ..., grid_bar, p2g_values_bar, points_bar)
    and the points_bar part of g2p_adjoint(array, ..., g2p_functor,
    g2p_values_bar, grid_bar=None, points_bar).  ``array`` is the G2P
    forward grid (or one per G2P functor).  For the MPM set (P2G ScalarValue,
    VectorValue, TensorDivergence; G2P VectorValue, VectorGradient of one
    grid) it is one pass over the particles."""
    pf = p2g_functor if isinstance(p2g_functor, (list, tuple)) else [p2g_functor]
    gbars = grid_bar if isinstance(grid_bar, (list, tuple)) else [grid_bar]
    pvb = p2g_values_bar if isinstance(p2g_values_bar, (list, tuple)) else [p2g_values_bar]
    gf = g2p_functor if isinstance(g2p_functor, (list, tuple)) else [g2p_functor]
    arrays = array if isinstance(array, (list, tuple)) else [array] * len(gf)
    gvb = g2p_values_bar if isinstance(g2p_values_bar, (list, tuple)) else [g2p_values_bar]
    if len(gbars) != len(pf) or len(pvb) != len(pf) or len(arrays) != len(gf) or \
            len(gvb) != len(gf):
        raise ValueError("one grid_bar and values_bar per P2G functor, one array and "
                         "values_bar per G2P functor")
    m = _call_mesh(list(gbars) + list(arrays))
    pts = _points_descriptor(points, num_point)
    pt = (N.P2GAdjointTerm * len(pf))()
    for i, f in enumerate(pf):
        pt[i].op = f.op
        pt[i].multiplier = f._multiplier
        pt[i].values = _values_descriptor(f._x, num_point)
        pt[i].grid_bar = gbars[i].descriptor()
        pt[i].values_bar = particle_descriptor(pvb[i]) if pvb[i] is not None \
            else N.ParticleField()
    gt = (N.G2PAdjointTerm * len(gf))()
    for i, f in enumerate(gf):
        gt[i].op = f.op
        gt[i].multiplier = f._multiplier
        gt[i].grid = arrays[i].descriptor()
        gt[i].values_bar = _values_descriptor(gvb[i], num_point)
        gt[i].grid_bar = N.GridField()
    xb = particle_descriptor(points_bar)
    b = _bins_descriptor(bins)
    ctx = N.context()
    N.check(N.load().cabana_amd_adjoint_gather(
        ctx, ctypes.byref(m), spline.order, ctypes.byref(pts), len(pf), pt, len(gf), gt,
        _halo_width(halo), ctypes.byref(xb), ctypes.byref(b) if b is not None else None,
        N.current_stream()), ctx)


def check_domain():
    """Waits for the stream and raises if particles left the ghosted grid."""
    ctx = N.context()
    n = ctypes.c_int64(0)
    N.check(N.load().cabana_amd_check_domain(ctx, N.current_stream(), ctypes.byref(n)), ctx)
    return n.value


# ---------------------------------------------------------------------------
# Particle migration (one rank)
# ---------------------------------------------------------------------------


def particleMigrate(local_grid, positions, particles, min_halo_width, force_migrate=False):
    """particleMigrate(local_grid, positions, particles, min_halo_width,
    force_migrate) (Cabana_Grid_ParticleDistributor.hpp:330-352) on one rank.

    Unless force_migrate is set, nothing changes while every particle lies
    within the ghosted domain moved in by min_halo_width cells.  Otherwise
    periodic coordinates are shifted back into the global box (in place in
    ``positions``) and particles beyond a non-periodic boundary are removed
    from every member of ``particles`` (a sequence of particle fields,
    compacted in place in their order; include ``positions`` when it belongs
    to the particle set).  Returns ``(migrated, num_particle)``: whether the
    migration ran (the reference's return value) and how many particles
    remain, which are the first ``num_particle`` rows of every member (a
    torch tensor cannot shrink in place as an AoSoA resizes)."""
    pos = particle_descriptor(positions, 3)
    descs = [particle_descriptor(f) for f in particles]
    if len({d.data for d in descs}) != len(descs):
        raise ValueError("particleMigrate: the same field was passed twice")
    arr = (type(descs[0]) * len(descs))(*descs) if descs else None
    m = local_grid.mesh_descriptor()
    ctx = N.context()
    num = ctypes.c_int64(0)
    migrated = ctypes.c_int32(0)
    N.check(N.load().cabana_amd_particle_migrate(
        ctx, ctypes.byref(m), ctypes.byref(pos), arr, len(descs), int(min_halo_width),
        1 if force_migrate else 0, ctypes.byref(num), ctypes.byref(migrated),
        N.current_stream()), ctx)
    return bool(migrated.value), int(num.value)


# ---------------------------------------------------------------------------
# Grid output
# ---------------------------------------------------------------------------


class BovWriter:
    """Experimental::BovWriter (Cabana_Grid_BovWriter.hpp) on one rank."""

    @staticmethod
    def writeTimeStep(*args, gather_array=True):
        """writeTimeStep(prefix, time_step_index, time, array[, gather_array])
        or writeTimeStep(time_step_index, time, array[, gather_array]) with
        the prefix "grid_" + label (:180-338, :389-395).  Writes
        <prefix>_<index:06d>.dat and .bov."""
        if isinstance(args[0], str):
            prefix, step, time, array = args[:4]
            rest = args[4:]
        else:
            step, time, array = args[:3]
            prefix = "grid_" + array.label
            rest = args[3:]
        if rest:
            gather_array = rest[0]
        layout = array.layout()
        if not isinstance(layout.entity, (Node, Cell)):
            raise NotImplementedError("BOV centering is nodal or zonal (Node or Cell arrays)")
        lg = layout.localGrid()
        m = lg.mesh_descriptor()
        g = grid_descriptor(array.view())
        entity = N.NODE if isinstance(layout.entity, Node) else N.CELL
        ctx = N.context()
        N.check(N.load().cabana_amd_bov_write_time_step(
            ctx, ctypes.byref(m), ctypes.byref(g), entity, array.label.encode(),
            str(prefix).encode(), int(step), float(time), 1 if gather_array else 0,
            N.current_stream()), ctx)


class Experimental:
    BovWriter = BovWriter


# ---------------------------------------------------------------------------
# Kernel timing
# ---------------------------------------------------------------------------


def profile_enable(enable=True):
    """Turns the library's per-launch HIP-event timing on or off (the
    counters are kept; profile_reset clears them)."""
    ctx = N.context()
    N.check(N.load().cabana_amd_profile_enable(ctx, 1 if enable else 0), ctx)


def profile_reset():
    ctx = N.context()
    N.check(N.load().cabana_amd_profile_reset(ctx), ctx)


def profile_report():
    """{kernel family: {"launches", "total_ms", "avg_ms"}} of the launches
    timed so far (waits for their events)."""
    lib = N.load()
    ctx = N.context()
    cnt = ctypes.c_int32()
    N.check(lib.cabana_amd_profile_count(ctx, ctypes.byref(cnt)), ctx)
    out = {}
    for i in range(cnt.value):
        name = ctypes.create_string_buffer(64)
        launches = ctypes.c_int64()
        ms = ctypes.c_double()
        N.check(lib.cabana_amd_profile_entry(ctx, i, name, 64, ctypes.byref(launches),
                                             ctypes.byref(ms)), ctx)
        out[name.value.decode()] = {"launches": launches.value, "total_ms": ms.value,
                                    "avg_ms": ms.value / max(launches.value, 1)}
    return out


def set_auto_binning(enable=True):
    """Whether calls without a binning bin the particles by cell themselves
    and run the cell-tiled kernels (on by default) or run the per-particle
    kernels with global atomics."""
    ctx = N.context()
    N.check(N.load().cabana_amd_set_auto_binning(ctx, 1 if enable else 0), ctx)


def set_deterministic(enable=True):
    """Deterministic P2G: every P2G call bins the particles and sums each
    node's contributions in a fixed order, so repeated calls give bitwise
    equal grids (Node arrays; slower; off by default)."""
    ctx = N.context()
    N.check(N.load().cabana_amd_set_deterministic(ctx, 1 if enable else 0), ctx)
