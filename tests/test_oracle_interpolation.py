"""Oracle p2g / g2p / halo / adjoints against the reference's known answers
(grid/unit_test/tstInterpolation3d.hpp, grid/unit_test/tstHalo3d.hpp) and
against what the mathematics fixes: dense brute-force weights built from
textbook B-splines, mass conservation, transpose identities and central
finite differences."""
import numpy as np
import pytest

from bspline_ref import dense_weights

P2G = {"ScalarValue": 0, "VectorValue": 1, "ScalarGradient": 2, "VectorDivergence": 3,
       "TensorDivergence": 4}
G2P = {"ScalarValue": 0, "VectorValue": 1, "ScalarGradient": 2, "VectorGradient": 3,
       "VectorDivergence": 4}
P2G_SHAPE = {0: (1, 1), 1: (3, 3), 2: (1, 3), 3: (3, 1), 4: (9, 3)}
G2P_SHAPE = {0: (1, 1), 1: (3, 3), 2: (3, 1), 3: (9, 3), 4: (1, 3)}


def cell_centre_points(mesh, cell):
    """One point per owned cell centre, ordered as tstInterpolation3d.hpp:62-77."""
    nc = mesh.num_cell
    h = mesh.halo
    pts = np.zeros((nc[0] * nc[1] * nc[2], 3))
    gl = list(mesh.m.ghost_low)
    for i in range(nc[0]):
        for j in range(nc[1]):
            for k in range(nc[2]):
                pid = i + nc[0] * (j + nc[1] * k)
                pts[pid] = [gl[0] + (i + h + 0.5) * cell, gl[1] + (j + h + 0.5) * cell,
                            gl[2] + (k + h + 0.5) * cell]
    return pts


@pytest.fixture(scope="module")
def interp_case(oracle):
    cell = 0.1
    mesh = oracle.OMesh([-1.2, 0.1, 1.1], [-0.3, 9.5, 2.3], cell, halo=1)
    return mesh, cell_centre_points(mesh, cell)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_interpolation3d_known_answers(oracle, interp_case, order):
    """tstInterpolation3d.hpp:118-250: constant 3.5 with multiplier -0.5 gives
    -1.75 for the value operators and 0 for gradients and divergences (the
    reference runs Spline<1>; uniform cell-centre points make the same
    answers exact for Spline<2> and Spline<3> by the partition of unity)."""
    mesh, pts = interp_case
    n = len(pts)
    own = mesh.owned_slices()
    tol = 1e-12
    for name, op in P2G.items():
        nc, nd = P2G_SHAPE[op]
        vals = np.full((n, nc), 3.5)
        grid = mesh.grid(nd)
        oracle.p2g(mesh, order, op, -0.5, pts, vals, grid)
        expect = -1.75 if name in ("ScalarValue", "VectorValue") else 0.0
        np.testing.assert_allclose(grid[own], expect, atol=tol, rtol=0)
    for name, op in G2P.items():
        nc, nd = G2P_SHAPE[op]
        grid = mesh.grid(nd)
        grid[own] = 3.5
        vals = np.zeros((n, nc))
        oracle.g2p(mesh, order, op, -0.5, grid, pts, vals)
        expect = -1.75 if name in ("ScalarValue", "VectorValue") else 0.0
        np.testing.assert_allclose(vals, expect, atol=1e-10, rtol=0)


def min_halo(order):
    """Smallest halo cell width that holds every stencil of owned points:
    Spline<N>::stencil reaches N / 2 nodes below the point's node."""
    return max(1, order // 2)


ORDERS = list(range(7))


def random_case(oracle, order, n=9, seed=0, num_cell=(5, 4, 6), cell=0.37, halo=None,
                entity="node"):
    rng = np.random.default_rng(seed)
    halo = min_halo(order) + (1 if entity != "node" else 0) if halo is None else halo
    low = np.array([0.3, -1.1, 2.0])
    high = low + cell * np.array(num_cell)
    mesh = oracle.OMesh(low, high, cell, halo=halo, entity=entity)
    pts = low + rng.uniform(0, 1, size=(n, 3)) * (high - low)
    return mesh, pts, rng


def dense(oracle, mesh, order, pts, derivs):
    rdx = [1.0 / ((mesh.m.ghost_low[d] + mesh.m.cell_size[d]) - mesh.m.ghost_low[d])
           for d in range(3)]
    # Entity index units: entity i of a cell-centred dimension sits at
    # ghost_low + (i + 1/2) dx.
    xnode = np.stack([(pts[:, d] - mesh.m.ghost_low[d]) * rdx[d] - 0.5 * mesh.m.stag[d]
                      for d in range(3)], axis=1)
    return dense_weights(order, xnode, mesh.num_cell, mesh.halo, rdx, derivs)


def dense_p2g(oracle, mesh, order, op, mult, pts, vals):
    """Owned-node result of p2g from the dense folded weight matrices."""
    W = dense(oracle, mesh, order, pts, (0, 0, 0))
    G = [dense(oracle, mesh, order, pts, tuple(1 if e == d else 0 for e in range(3)))
         for d in range(3)]
    v = vals * mult
    nn = W.shape[1]
    if op == 0:
        out = W.T @ v[:, 0]
        return out.reshape(-1, 1)
    if op == 1:
        return np.stack([W.T @ v[:, d] for d in range(3)], axis=1)
    if op == 2:
        return np.stack([G[d].T @ v[:, 0] for d in range(3)], axis=1)
    if op == 3:
        return (sum(G[d].T @ v[:, d] for d in range(3))).reshape(nn, 1)
    # Tensor divergence: grid[a] = sum_b dW_b sigma[a][b]
    return np.stack([sum(G[b].T @ v[:, 3 * a + b] for b in range(3)) for a in range(3)],
                    axis=1)


def dense_g2p(oracle, mesh, order, op, mult, pts, g_owned):
    W = dense(oracle, mesh, order, pts, (0, 0, 0))
    G = [dense(oracle, mesh, order, pts, tuple(1 if e == d else 0 for e in range(3)))
         for d in range(3)]
    if op == 0:
        return mult * (W @ g_owned[:, 0])[:, None]
    if op == 1:
        return mult * np.stack([W @ g_owned[:, d] for d in range(3)], axis=1)
    if op == 2:
        return mult * np.stack([G[d] @ g_owned[:, 0] for d in range(3)], axis=1)
    if op == 3:
        return mult * np.stack([G[a] @ g_owned[:, b] for a in range(3) for b in range(3)],
                               axis=1)
    return mult * (sum(G[d] @ g_owned[:, d] for d in range(3)))[:, None]


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("op", list(range(5)))
def test_p2g_matches_dense_bruteforce(oracle, order, op):
    mesh, pts, rng = random_case(oracle, order, seed=order * 10 + op)
    _p2g_dense_check(oracle, mesh, pts, rng, order, op)


@pytest.mark.parametrize("entity", ["cell", "face_j", "edge_k"])
@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("op", list(range(5)))
def test_p2g_on_other_entities_matches_dense_bruteforce(oracle, order, op, entity):
    """p2g onto Cell / Face / Edge arrays (SplineData of the array's entity,
    Cabana_Grid_Interpolation.hpp:874; entity coordinates,
    Cabana_Grid_LocalMesh.hpp:176-237): weights of the entity positions."""
    mesh, pts, rng = random_case(oracle, order, seed=500 + order * 10 + op, entity=entity)
    _p2g_dense_check(oracle, mesh, pts, rng, order, op)


def _p2g_dense_check(oracle, mesh, pts, rng, order, op):
    nc, nd = P2G_SHAPE[op]
    vals = rng.normal(size=(len(pts), nc))
    grid = mesh.grid(nd)
    oracle.p2g(mesh, order, op, -0.7, pts, vals, grid)
    got = grid[mesh.owned_slices()].reshape(-1, nd)
    ref = dense_p2g(oracle, mesh, order, op, -0.7, pts, vals)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-11 * max(1.0, np.abs(ref).max()))


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("op", list(range(5)))
def test_g2p_matches_dense_bruteforce(oracle, order, op):
    mesh, pts, rng = random_case(oracle, order, seed=100 + order * 10 + op)
    _g2p_dense_check(oracle, mesh, pts, rng, order, op)


@pytest.mark.parametrize("entity", ["cell", "face_i", "edge_j"])
@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("op", list(range(5)))
def test_g2p_on_other_entities_matches_dense_bruteforce(oracle, order, op, entity):
    mesh, pts, rng = random_case(oracle, order, seed=600 + order * 10 + op, entity=entity)
    _g2p_dense_check(oracle, mesh, pts, rng, order, op)


def _g2p_dense_check(oracle, mesh, pts, rng, order, op):
    nc, nd = G2P_SHAPE[op]
    grid = mesh.grid(nd)
    own = mesh.owned_slices()
    grid[own] = rng.normal(size=grid[own].shape)
    g_owned = grid[own].reshape(-1, nd).copy()
    base = rng.normal(size=(len(pts), nc))
    vals = base.copy()
    oracle.g2p(mesh, order, op, 1.3, grid, pts, vals)
    ref = base + dense_g2p(oracle, mesh, order, op, 1.3, pts, g_owned)
    np.testing.assert_allclose(vals, ref, rtol=0, atol=1e-11 * max(1.0, np.abs(ref).max()))


@pytest.mark.parametrize("order", ORDERS)
def test_p2g_conserves_mass(oracle, order):
    mesh, pts, rng = random_case(oracle, order, n=200, seed=5, num_cell=(7, 6, 5))
    m = rng.uniform(0.5, 1.5, size=(len(pts), 1))
    grid = mesh.grid(1)
    oracle.p2g(mesh, order, 0, 2.0, pts, m, grid)
    assert grid[mesh.owned_slices()].sum() == pytest.approx(2.0 * m.sum(), rel=1e-13)


@pytest.mark.parametrize("order", ORDERS[1:])
def test_g2p_gradient_of_linear_field_is_exact(oracle, order):
    """A linear node field is reproduced with its exact gradient away from the
    periodic seam (non-periodic mesh, interior points)."""
    rng = np.random.default_rng(9)
    cell = 0.5
    mesh = oracle.OMesh([0, 0, 0], [6, 6, 6], cell, halo=max(2, min_halo(order)),
                        periodic=(False,) * 3)
    pts = rng.uniform(1.5, 4.5, size=(50, 3))
    slope = np.array([0.7, -1.3, 2.1])
    grid = mesh.grid(1)
    gl = np.array(mesh.m.ghost_low)
    idx = np.indices(mesh.num_node).reshape(3, -1).T
    coords = gl + idx * cell
    grid[..., 0] = (coords @ slope + 0.4).reshape(mesh.num_node)
    grad = np.zeros((len(pts), 3))
    oracle.g2p(mesh, order, 2, 1.0, grid, pts, grad)
    np.testing.assert_allclose(grad, np.tile(slope, (len(pts), 1)), rtol=1e-11, atol=1e-11)
    val = np.zeros((len(pts), 1))
    oracle.g2p(mesh, order, 0, 1.0, grid, pts, val)
    np.testing.assert_allclose(val[:, 0], pts @ slope + 0.4, rtol=1e-12, atol=1e-12)


# (P2G op, G2P op, tensor transpose): <P2G(v), g>_owned = <v, G2P(g)>
TRANSPOSE_PAIRS = [(0, 0, False), (1, 1, False), (3, 2, False), (2, 4, False),
                   (4, 3, True)]


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("pair", TRANSPOSE_PAIRS)
def test_transpose_identities(oracle, order, pair):
    p_op, g_op, transpose = pair
    mesh, pts, rng = random_case(oracle, order, n=40, seed=21 + p_op)
    nc, nd = P2G_SHAPE[p_op]
    v = rng.normal(size=(len(pts), nc))
    grid = mesh.grid(nd)
    oracle.p2g(mesh, order, p_op, 0.9, pts, v, grid)
    own = mesh.owned_slices()
    g = mesh.grid(nd)
    g[own] = rng.normal(size=g[own].shape)
    lhs = np.sum(grid[own] * g[own])
    out = np.zeros((len(pts), nc))
    oracle.g2p(mesh, order, g_op, 0.9, g, pts, out)
    if transpose:
        out = out.reshape(-1, 3, 3).transpose(0, 2, 1).reshape(-1, 9)
    rhs = np.sum(v * out)
    assert lhs == pytest.approx(rhs, rel=1e-12, abs=1e-12)


def _loss_p2g(oracle, mesh, order, op, mult, pts, vals, gbar):
    grid = mesh.grid(gbar.shape[-1])
    oracle.p2g(mesh, order, op, mult, pts, vals, grid)
    own = mesh.owned_slices()
    return np.sum(grid[own] * gbar[own])


def _loss_g2p(oracle, mesh, order, op, mult, pts, grid, ubar):
    out = np.zeros(ubar.shape)
    oracle.g2p(mesh, order, op, mult, grid.copy(), pts, out)
    return np.sum(out * ubar)


@pytest.mark.parametrize("order", ORDERS[2:])
@pytest.mark.parametrize("op", list(range(5)))
def test_p2g_adjoint_matches_finite_differences(oracle, order, op):
    mesh, pts, rng = random_case(oracle, order, n=6, seed=300 + op)
    nc, nd = P2G_SHAPE[op]
    vals = rng.normal(size=(len(pts), nc))
    gbar = mesh.grid(nd)
    gbar[...] = rng.normal(size=gbar.shape)
    vbar = np.zeros_like(vals)
    xbar = np.zeros_like(pts)
    oracle.p2g_adjoint(mesh, order, op, 0.8, pts, vals, gbar.copy(), vbar, xbar)
    h = 1e-6
    for p in range(len(pts)):
        for e in range(3):
            pp = pts.copy()
            pm = pts.copy()
            pp[p, e] += h
            pm[p, e] -= h
            fd = (_loss_p2g(oracle, mesh, order, op, 0.8, pp, vals, gbar) -
                  _loss_p2g(oracle, mesh, order, op, 0.8, pm, vals, gbar)) / (2 * h)
            assert xbar[p, e] == pytest.approx(fd, rel=1e-6, abs=1e-6)
        for c in range(nc):
            vp = vals.copy()
            vm = vals.copy()
            vp[p, c] += h
            vm[p, c] -= h
            fd = (_loss_p2g(oracle, mesh, order, op, 0.8, pts, vp, gbar) -
                  _loss_p2g(oracle, mesh, order, op, 0.8, pts, vm, gbar)) / (2 * h)
            assert vbar[p, c] == pytest.approx(fd, rel=1e-6, abs=1e-6)


@pytest.mark.parametrize("order", ORDERS[2:])
@pytest.mark.parametrize("op", list(range(5)))
def test_g2p_adjoint_matches_finite_differences(oracle, order, op):
    mesh, pts, rng = random_case(oracle, order, n=6, seed=400 + op)
    nc, nd = G2P_SHAPE[op]
    grid = mesh.grid(nd)
    own = mesh.owned_slices()
    grid[own] = rng.normal(size=grid[own].shape)
    ubar = rng.normal(size=(len(pts), nc))
    gbar = mesh.grid(nd)
    xbar = np.zeros_like(pts)
    oracle.g2p_adjoint(mesh, order, op, 1.1, grid.copy(), pts, ubar, gbar, xbar)
    h = 1e-6
    for p in range(len(pts)):
        for e in range(3):
            pp = pts.copy()
            pm = pts.copy()
            pp[p, e] += h
            pm[p, e] -= h
            fd = (_loss_g2p(oracle, mesh, order, op, 1.1, pp, grid, ubar) -
                  _loss_g2p(oracle, mesh, order, op, 1.1, pm, grid, ubar)) / (2 * h)
            assert xbar[p, e] == pytest.approx(fd, rel=1e-6, abs=1e-6)
    # Grid cotangent: the loss is linear in the owned grid values.
    idx = [tuple(rng.integers(mesh.halo, mesh.halo + mesh.num_cell[d]) for d in range(3))
           for _ in range(8)]
    for (i, j, k) in idx:
        for d in range(nd):
            gp = grid.copy()
            gm = grid.copy()
            gp[i, j, k, d] += h
            gm[i, j, k, d] -= h
            fd = (_loss_g2p(oracle, mesh, order, op, 1.1, pts, gp, ubar) -
                  _loss_g2p(oracle, mesh, order, op, 1.1, pts, gm, ubar)) / (2 * h)
            assert gbar[i, j, k, d] == pytest.approx(fd, rel=1e-6, abs=1e-6)


def check_scatter_pattern(mesh, grid, width):
    """tstHalo3d.hpp:123-177: owned value = 2^(dims in which the node is in a
    halo-shared band)."""
    h = mesh.halo
    own = mesh.owned_slices()
    for i in range(own[0].start, own[0].stop):
        for j in range(own[1].start, own[1].stop):
            for k in range(own[2].start, own[2].stop):
                num = 0
                for d, x in enumerate((i, j, k)):
                    lo = x < h + width + 1
                    hi = x >= h + mesh.num_cell[d] - width
                    if lo or hi:
                        num += 1
                assert np.all(grid[i, j, k] == 2.0 ** num)


@pytest.mark.parametrize("width", [1, 2, 3])
def test_halo_gather_scatter_pattern(oracle, width):
    # tstHalo3d.hpp:181-230 geometry (cell 0.23, 17 x 20 x 21 cells, halo 3).
    cell = 0.23
    low = np.array([1.2, 3.3, -2.8])
    high = low + cell * np.array([17, 20, 21])
    mesh = oracle.OMesh(low, high, cell, halo=3)
    grid = mesh.grid(4)
    grid[mesh.owned_slices()] = 1.0
    oracle.halo_gather(mesh, grid, width)
    # Ghosts within the band now hold 1, the rest 0.
    oracle.halo_scatter_sum(mesh, grid, width)
    check_scatter_pattern(mesh, grid, width)


def halo_pad(mesh, d):
    """tstHalo3d.hpp:32-46 haloPad: 1 in a node-centred dimension, 0 in a
    cell-centred one (Cell 0, Node 1, Face<D> 1 only in D, Edge<D> 0 only
    in D)."""
    return 1 - mesh.m.stag[d]


def check_gather_pattern(mesh, grid, width, periodic):
    """tstHalo3d.hpp:52-117 (checkGather) on one rank: after the gather an
    entity holds 1 if it is owned or lies within `width` (+ the halo pad on
    the high side) of the owned entities in every dimension, 0 otherwise;
    the boundary halo of a non-periodic dimension stays 0."""
    own = mesh.owned_slices()
    for i in range(mesh.num_node[0]):
        for j in range(mesh.num_node[1]):
            for k in range(mesh.num_node[2]):
                idx = (i, j, k)
                boundary = any(not periodic[d] and not (own[d].start <= idx[d] < own[d].stop)
                               for d in range(3))
                outside = any(idx[d] < own[d].start - width or
                              idx[d] >= own[d].stop + width + halo_pad(mesh, d)
                              for d in range(3))
                want = 0.0 if (boundary or outside) else 1.0
                assert np.all(grid[i, j, k] == want), (idx, grid[i, j, k], want)


def check_scatter_pattern_periodicity(mesh, grid, width, periodic):
    """tstHalo3d.hpp:123-177 (checkScatter) on one rank: an owned node holds
    2^n, n = the periodic dimensions in which it lies in a halo-shared band
    (a non-periodic dimension of a single rank has no neighbour)."""
    own = mesh.owned_slices()
    for i in range(own[0].start, own[0].stop):
        for j in range(own[1].start, own[1].stop):
            for k in range(own[2].start, own[2].stop):
                num = 0
                for d, x in enumerate((i, j, k)):
                    if not periodic[d]:
                        continue
                    if x < own[d].start + width + halo_pad(mesh, d) or \
                            x >= own[d].stop - width:
                        num += 1
                assert np.all(grid[i, j, k] == 2.0 ** num), ((i, j, k), grid[i, j, k], num)


@pytest.mark.parametrize("entity", ["node", "cell", "face_i", "face_j", "face_k", "edge_i",
                                    "edge_j", "edge_k"])
@pytest.mark.parametrize("periodic", [(True, True, True), (False, True, False),
                                      (True, False, True), (False, False, False)])
@pytest.mark.parametrize("width", [1, 2, 3])
def test_halo_gather_scatter_pattern_periodicity(oracle, width, periodic, entity):
    """tstHalo3d.hpp:181-230 geometry with periodic and non-periodic
    dimensions: the non-periodic branch of the oracle's halo exchange
    (no neighbour across a non-periodic boundary) against the reference's
    checkGather / checkScatter expectations."""
    cell = 0.23
    low = np.array([1.2, 3.3, -2.8])
    high = low + cell * np.array([17, 20, 21])
    mesh = oracle.OMesh(low, high, cell, halo=3, periodic=periodic, entity=entity)
    grid = mesh.grid(4)
    grid[mesh.owned_slices()] = 1.0
    oracle.halo_gather(mesh, grid, width)
    check_gather_pattern(mesh, grid, width, periodic)
    oracle.halo_scatter_sum(mesh, grid, width)
    check_scatter_pattern_periodicity(mesh, grid, width, periodic)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_openmp_g2p_and_adjoints_equal_serial(oracle, order):
    """The threaded G2P and adjoints (the CPU baseline's code) against the
    Serial ones, every operator."""
    mesh, pts, rng = random_case(oracle, order, n=3000, seed=78, num_cell=(8, 8, 8))
    n = len(pts)
    for op in range(5):
        nc, nd = G2P_SHAPE[op]
        g = rng.normal(size=mesh.num_node + (nd,))
        a = np.zeros((n, nc))
        b = np.zeros((n, nc))
        oracle.g2p(mesh, order, op, 0.7, g.copy(), pts, a)
        oracle.g2p(mesh, order, op, 0.7, g.copy(), pts, b, threads=4)
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)
        vb = rng.normal(size=(n, nc))
        ga, gb = mesh.grid(nd), mesh.grid(nd)
        xa, xb = np.zeros((n, 3)), np.zeros((n, 3))
        oracle.g2p_adjoint(mesh, order, op, 0.7, g.copy(), pts, vb, ga, xa)
        oracle.g2p_adjoint(mesh, order, op, 0.7, g.copy(), pts, vb, gb, xb, threads=4)
        np.testing.assert_allclose(ga, gb, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(xa, xb, rtol=1e-12, atol=1e-12)
    for op in range(5):
        nc, nd = P2G_SHAPE[op]
        vals = rng.normal(size=(n, nc))
        gbar = rng.normal(size=mesh.num_node + (nd,))
        va, vb2 = np.zeros((n, nc)), np.zeros((n, nc))
        xa, xb = np.zeros((n, 3)), np.zeros((n, 3))
        oracle.p2g_adjoint(mesh, order, op, -0.4, pts, vals, gbar.copy(), va, xa)
        oracle.p2g_adjoint(mesh, order, op, -0.4, pts, vals, gbar.copy(), vb2, xb, threads=4)
        np.testing.assert_allclose(va, vb2, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(xa, xb, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_openmp_equals_serial(oracle, order):
    mesh, pts, rng = random_case(oracle, order, n=2000, seed=77, num_cell=(8, 8, 8))
    for op in range(5):
        nc, nd = P2G_SHAPE[op]
        vals = rng.normal(size=(len(pts), nc))
        a = mesh.grid(nd)
        b = mesh.grid(nd)
        oracle.p2g(mesh, order, op, 1.0, pts, vals, a)
        oracle.p2g(mesh, order, op, 1.0, pts, vals, b, threads=4)
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)
