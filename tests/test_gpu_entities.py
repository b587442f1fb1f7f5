"""Arrays of every mesh entity (Cell, Face<D>, Edge<D> besides Node) on the
HIP path against the oracle: the halo gather/scatter patterns of the
reference's tstHalo3d.hpp for all eight entities, random-value halos with
mixed periodicity, and P2G / G2P onto Cell, Face and Edge arrays (the
splines weigh the entity positions, Cabana_Grid_Interpolation.hpp:874 and
Cabana_Grid_LocalMesh.hpp:176-237), with and without a cell binning."""
import numpy as np
import pytest
import torch

from test_gpu_interpolation import (G2P_FACTORY, G2P_SHAPE, P2G_FACTORY, P2G_SHAPE, TOL, rel_l2,
                                    shaped)

pytestmark = pytest.mark.gpu

ENTITIES = ["node", "cell", "face_i", "face_j", "face_k", "edge_i", "edge_j", "edge_k"]


def tag(ca, name):
    if name == "node":
        return ca.Node()
    if name == "cell":
        return ca.Cell()
    kind, d = name.split("_")
    dim = "ijk".index(d)
    return ca.Face(dim) if kind == "face" else ca.Edge(dim)


def local_grid(ca, low, high, cell, halo, periodic):
    gm = ca.createUniformGlobalMesh(low, high, cell)
    return ca.createLocalGrid(ca.createGlobalGrid(gm, list(periodic)), halo)


@pytest.mark.parametrize("entity", ENTITIES)
@pytest.mark.parametrize("periodic", [(True, True, True), (True, False, True)])
@pytest.mark.parametrize("width", [1, 3])
def test_halo_pattern(ca, oracle, entity, periodic, width):
    """tstHalo3d.hpp:181-230 geometry (17 x 20 x 21 cells of 0.23, halo 3,
    4 dofs): owned 1, gather, then scatter; bit for bit with the oracle,
    whose patterns are pinned by the reference's checkGather/checkScatter."""
    cell = 0.23
    low = np.array([1.2, 3.3, -2.8])
    high = low + cell * np.array([17, 20, 21])
    lg = local_grid(ca, low, high, cell, 3, periodic)
    arr = ca.createArray("a", ca.createArrayLayout(lg, 4, tag(ca, entity)))
    ca.ArrayOp.assign(arr, 0.0, ca.Ghost())
    ca.ArrayOp.assign(arr, 1.0, ca.Own())
    mesh = oracle.OMesh(low, high, cell, halo=3, periodic=periodic, entity=entity)
    assert arr.view().shape == mesh.num_node + (4,)
    ref = mesh.grid(4)
    ref[mesh.owned_slices()] = 1.0
    np.testing.assert_array_equal(arr.view().cpu().numpy(), ref)
    halo = ca.createHalo(ca.NodeHaloPattern(), width, arr)
    halo.gather(arr)
    oracle.halo_gather(mesh, ref, width)
    np.testing.assert_array_equal(arr.view().cpu().numpy(), ref)
    halo.scatter(ca.ScatterReduce.Sum, arr)
    oracle.halo_scatter_sum(mesh, ref, width)
    np.testing.assert_array_equal(arr.view().cpu().numpy(), ref)


@pytest.mark.parametrize("entity", ENTITIES[1:])
@pytest.mark.parametrize("periodic", [(False, True, False), (False, False, False)])
def test_halo_random_values(ca, oracle, entity, periodic):
    rng = np.random.default_rng(ENTITIES.index(entity))
    cell = 0.5
    low = np.array([0.0, 1.0, -2.0])
    high = low + cell * np.array([6, 7, 5])
    lg = local_grid(ca, low, high, cell, 2, periodic)
    arr = ca.createArray("a", ca.createArrayLayout(lg, 3, tag(ca, entity)), memory_layout="left")
    vals = rng.normal(size=arr.view().shape)
    arr.view().copy_(torch.tensor(vals))
    mesh = oracle.OMesh(low, high, cell, halo=2, periodic=periodic, entity=entity)
    ref = np.ascontiguousarray(vals)
    halo = ca.createHalo(ca.NodeHaloPattern(), -1, arr)
    halo.scatter(ca.ScatterReduce.Sum, arr)
    oracle.halo_scatter_sum(mesh, ref)
    np.testing.assert_allclose(arr.view().cpu().numpy(), ref, rtol=1e-14, atol=1e-14)
    halo.gather(arr)
    oracle.halo_gather(mesh, ref)
    np.testing.assert_allclose(arr.view().cpu().numpy(), ref, rtol=1e-14, atol=1e-14)


class EntityCase:
    def __init__(self, ca, oracle, entity, order, n=3000, seed=0, num_cell=(10, 9, 11),
                 cell=0.25):
        self.ca = ca
        low = np.array([0.3, -1.1, 2.0])
        high = low + cell * np.array(num_cell)
        halo = max(1, order // 2) + 1
        self.lg = local_grid(ca, low, high, cell, halo, (True, True, True))
        self.mesh = oracle.OMesh(low, high, cell, halo=halo, entity=entity)
        self.entity = entity
        self.rng = np.random.default_rng(seed)
        self.pts_np = low + self.rng.uniform(0, 1, size=(n, 3)) * (high - low)
        self.pts = torch.tensor(self.pts_np, device="cuda")
        self.n = n

    def array(self, dofs, dtype):
        return self.ca.createArray("a", self.ca.createArrayLayout(self.lg, dofs,
                                                                  tag(self.ca, self.entity)),
                                   dtype=dtype)

    def halo(self, array):
        return self.ca.createHalo(self.ca.NodeHaloPattern(), -1, array)

    def bins(self):
        gm = self.lg.globalGrid().globalMesh()
        lcl = self.ca.createLinkedCellList(self.pts, [gm.cellSize(d) for d in range(3)],
                                           [gm.lowCorner(d) for d in range(3)],
                                           [gm.highCorner(d) for d in range(3)])
        perm = lcl.permutes.cpu().numpy()
        self.pts = self.pts[lcl.permutes].contiguous()
        self.pts_np = self.pts_np[perm]
        return lcl


@pytest.mark.parametrize("binned", [False, True])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("entity", ["cell", "face_i", "edge_k"])
@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("op", range(5))
def test_p2g_matches_oracle(ca, oracle, op, order, entity, dtype, binned):
    c = EntityCase(ca, oracle, entity, order, seed=op * 7 + order)
    bins = c.bins() if binned else None
    nc, nd = P2G_SHAPE[op]
    vals = c.rng.normal(size=(c.n, nc))
    grid = c.array(nd, dtype)
    f = getattr(ca, P2G_FACTORY[op])(shaped(torch.tensor(vals, dtype=dtype, device="cuda"),
                                            nc), -0.6)
    ca.p2g(f, c.pts, c.n, ca.Spline(order), c.halo(grid), grid, bins=bins)
    ref = c.mesh.grid(nd)
    oracle.p2g(c.mesh, order, op, -0.6, c.pts_np, vals, ref)
    own = c.mesh.owned_slices()
    assert rel_l2(grid.view().cpu().numpy()[own], ref[own]) <= TOL[dtype]
    assert ca.check_domain() == 0


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("entity", ["cell", "face_j", "edge_i"])
@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("op", range(5))
def test_g2p_matches_oracle(ca, oracle, op, order, entity, dtype):
    c = EntityCase(ca, oracle, entity, order, seed=100 + op * 7 + order)
    nc, nd = G2P_SHAPE[op]
    grid = c.array(nd, dtype)
    g = c.rng.normal(size=grid.view().shape)
    grid.view().copy_(torch.tensor(g, dtype=dtype))
    out = shaped(torch.zeros(c.n, nc, dtype=dtype, device="cuda"), nc)
    ca.g2p(grid, c.halo(grid), c.pts, c.n, ca.Spline(order), getattr(ca, G2P_FACTORY[op])(out, 1.7))
    ref = np.zeros((c.n, nc))
    gref = g.astype(np.float32).astype(np.float64) if dtype == torch.float32 else g.copy()
    oracle.g2p(c.mesh, order, op, 1.7, gref, c.pts_np, ref)
    assert rel_l2(out.cpu().numpy().reshape(c.n, nc), ref) <= TOL[dtype]
    np.testing.assert_allclose(grid.view().cpu().numpy().astype(np.float64), gref, rtol=0,
                               atol=1e-6 if dtype == torch.float32 else 0)


def test_arrays_of_one_call_share_an_entity(ca, oracle):
    c = EntityCase(ca, oracle, "cell", 2)
    node = ca.createArray("n", ca.createArrayLayout(c.lg, 1, ca.Node()))
    cell = c.array(1, torch.float64)
    v = torch.ones(c.n, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        ca.p2g([ca.createScalarValueP2G(v, 1.0), ca.createScalarValueP2G(v, 1.0)], c.pts, c.n,
               ca.Spline(2), c.halo(node), [node, cell])
