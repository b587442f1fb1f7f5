"""Oracle splines against the reference's known answers
(grid/unit_test/tstSplines.cpp, grid/unit_test/tstSplineEvaluation3d.hpp) and
against textbook B-splines (tests/bspline_ref.py)."""
import numpy as np
import pytest

from bspline_ref import bspline, dbspline, ddbspline

LOW_X = -3.43
DX = 0.27
RDX = 1.0 / DX


def grid_func(x):
    return 4.32 * x - 0.31


@pytest.mark.parametrize("order", range(7))
def test_partition_of_unity(oracle, order):
    # tstSplines.cpp:27-55 (zero), :86-116 (linear), :148-178 (quadratic),
    # :216-246 (cubic), :286-314 (quartic), :361-389 (quintic), :439-467
    # (sextic)
    for xp in (-1.4, 2.1789, LOW_X + 5 * DX):
        _, _, w, _, _ = oracle.spline_1d(order, xp, RDX, LOW_X)
        assert sum(w) == pytest.approx(1.0, rel=1e-14)


def test_linear_stencil_and_reproduction(oracle):
    # tstSplines.cpp:118-143: point in the centre of primal cell 4.
    cell_id = 4
    xp = LOW_X + (cell_id + 0.5) * DX
    x0, s, w, g, _ = oracle.spline_1d(1, xp, RDX, LOW_X)
    assert int(x0) + 0 == cell_id and int(x0) + 1 == cell_id + 1
    assert s == [cell_id, cell_id + 1]
    field = [grid_func(LOW_X + cell_id * DX), grid_func(LOW_X + (cell_id + 1) * DX)]
    assert np.dot(field, w) == pytest.approx(grid_func(xp), rel=1e-13)
    assert np.dot(field, g) == pytest.approx(4.32, rel=1e-12)


def test_quadratic_stencil_and_reproduction(oracle):
    # tstSplines.cpp:180-211: point at node 4 + 0.25 dx (dual-cell centre).
    node_id = 4
    xp = LOW_X + (node_id + 0.25) * DX
    x0, s, w, g, _ = oracle.spline_1d(2, xp, RDX, LOW_X)
    assert [int(x0) + o for o in (-1, 0, 1)] == [node_id - 1, node_id, node_id + 1]
    assert s == [node_id - 1, node_id, node_id + 1]
    field = [grid_func(LOW_X + (node_id + o) * DX) for o in (-1, 0, 1)]
    assert np.dot(field, w) == pytest.approx(grid_func(xp), rel=1e-13)
    assert np.dot(field, g) == pytest.approx(4.32, rel=1e-12)


def test_cubic_stencil_and_reproduction(oracle):
    # tstSplines.cpp:248-281: point at cell 4 + 0.75 dx.
    cell_id = 4
    xp = LOW_X + (cell_id + 0.75) * DX
    x0, s, w, g, _ = oracle.spline_1d(3, xp, RDX, LOW_X)
    assert [int(x0) + o for o in (-1, 0, 1, 2)] == [cell_id - 1, cell_id, cell_id + 1,
                                                      cell_id + 2]
    assert s == [cell_id - 1, cell_id, cell_id + 1, cell_id + 2]
    field = [grid_func(LOW_X + (cell_id + o) * DX) for o in (-1, 0, 1, 2)]
    assert np.dot(field, w) == pytest.approx(grid_func(xp), rel=1e-13)
    assert np.dot(field, g) == pytest.approx(4.32, rel=1e-12)


def test_zero_stencil_and_reproduction(oracle):
    # tstSplines.cpp:57-81: point at node 4 + 0.25 dx; the one knot takes the
    # node's value and the gradient vanishes.
    node_id = 4
    xp = LOW_X + (node_id + 0.25) * DX
    x0, s, w, g, h = oracle.spline_1d(0, xp, RDX, LOW_X)
    assert int(x0) == node_id and s == [node_id]
    field = grid_func(LOW_X + node_id * DX)
    assert field * w[0] == field
    assert field * g[0] == 0.0 and h == [0.0]


@pytest.mark.parametrize("order,offset,first", [(4, 0.25, -2), (5, 0.75, -2), (6, 0.25, -3)])
def test_high_order_stencil_and_reproduction(oracle, order, offset, first):
    # tstSplines.cpp:316-356 (quartic: node 4 + 0.25 dx, nodes 2..6),
    # :391-434 (quintic: cell 4 + 0.75 dx, nodes 2..7), :469-517 (sextic:
    # node 4 + 0.25 dx, nodes 1..7); linear functions are reproduced.
    cid = 4
    xp = LOW_X + (cid + offset) * DX
    x0, s, w, g, _ = oracle.spline_1d(order, xp, RDX, LOW_X)
    nodes = [cid + first + n for n in range(order + 1)]
    assert [int(x0) + first + n for n in range(order + 1)] == nodes
    assert s == nodes
    field = [grid_func(LOW_X + n * DX) for n in nodes]
    assert np.dot(field, w) == pytest.approx(grid_func(xp), rel=1e-13)
    assert np.dot(field, g) == pytest.approx(4.32, rel=1e-12)


@pytest.mark.parametrize("order", range(7))
def test_weights_match_textbook_bsplines(oracle, order):
    """Weights, gradients and second derivatives equal B_p(x - n), B_p' rdx
    and B_p'' rdx^2 at the stencil nodes (truncated-power formula, whose
    alternating sum loses about 1e-13 to cancellation at degree 6)."""
    rng = np.random.default_rng(7 + order)
    atol = 1e-13 if order <= 3 else 1e-12
    for xp in rng.uniform(LOW_X + 3 * DX, LOW_X + 40 * DX, size=200):
        _, s, w, g, h = oracle.spline_1d(order, xp, RDX, LOW_X)
        xnode = (xp - LOW_X) * RDX
        t = xnode - np.array(s, dtype=np.float64)
        np.testing.assert_allclose(w, bspline(order, t), rtol=0, atol=atol)
        if order == 0:
            assert g == [0.0] and h == [0.0]
            continue
        np.testing.assert_allclose(g, dbspline(order, t) * RDX, rtol=0, atol=1e-11)
        if order >= 2:
            np.testing.assert_allclose(h, ddbspline(order, t) * RDX ** 2, rtol=0, atol=1e-9)
        else:
            assert h == [0.0] * (order + 1)


@pytest.mark.parametrize("order", [2, 3, 4, 5, 6])
def test_hessian_is_derivative_of_gradient(oracle, order):
    """Second derivatives (not in the reference) against central differences
    of Spline<N>::gradient."""
    rng = np.random.default_rng(11)
    eps = 1e-6
    for xp in rng.uniform(LOW_X + 3 * DX, LOW_X + 40 * DX, size=50):
        _, s0, _, _, h = oracle.spline_1d(order, xp, RDX, LOW_X)
        _, sp, _, gp, _ = oracle.spline_1d(order, xp + eps, RDX, LOW_X)
        _, sm, _, gm, _ = oracle.spline_1d(order, xp - eps, RDX, LOW_X)
        if sp != s0 or sm != s0:
            continue  # stencil switch inside the difference interval
        fd = (np.array(gp) - np.array(gm)) / (2 * eps)
        np.testing.assert_allclose(h, fd, rtol=1e-5, atol=1e-4)


def test_spline_evaluation_linear_cell_centres(oracle):
    """tstSplineEvaluation3d.hpp:263-497 (Spline<1>, Node, cell size 0.05,
    points at owned cell centres)."""
    cell = 0.05
    mesh = oracle.OMesh([-1.2, 0.1, 1.1], [-0.3, 9.5, 2.3], cell, halo=1)
    rng = np.random.default_rng(3)
    nc = mesh.num_cell
    for _ in range(100):
        idx = [int(rng.integers(1, 1 + nc[d])) for d in range(3)]
        gl = [mesh.m.ghost_low[d] for d in range(3)]
        p = [gl[d] + (idx[d] + 0.5) * cell for d in range(3)]
        sd = oracle.evaluate_spline(mesh, 1, p)
        for d in range(3):
            assert sd["dx"][d] == pytest.approx(cell, rel=1e-6)
            assert sd["rdx"][d] == pytest.approx(1.0 / cell, rel=1e-6)
            assert sd["x"][d] == pytest.approx(idx[d] + 0.5, rel=1e-6)
            assert sd["s"][d] == [idx[d], idx[d] + 1]
            assert sd["w"][d] == pytest.approx([0.5, 0.5], rel=1e-6)
            assert sd["g"][d] == pytest.approx([-1.0 / cell, 1.0 / cell], rel=1e-6)
        # Weight-gradient products of the 8 corners: +-0.25/dx.
        for i in range(2):
            for j in range(2):
                for k in range(2):
                    gx = sd["g"][0][i] * sd["w"][1][j] * sd["w"][2][k]
                    assert abs(gx) == pytest.approx(0.25 / cell, rel=1e-6)
        # Physical distances: sum_n w_n d_n = 0 (first moment).
        for d in range(3):
            assert np.dot(sd["w"][d], sd["d"][d]) == pytest.approx(0.0, abs=1e-12)
