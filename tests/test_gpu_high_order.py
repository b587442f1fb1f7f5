"""Spline<0> and Spline<4> .. Spline<6> (grid/src/Cabana_Grid_Splines.hpp
:38-121, :456-939) on the HIP path against the oracle, whose weights are pinned
against tstSplines.cpp and textbook B-splines in test_oracle_splines.py: every
P2G and G2P operator, their adjoints and the fused MPM term sets, FP64
(rel-L2 <= 1e-12) and FP32-mixed (rel-L2 <= 1e-5).  These orders run the
per-particle kernels; a cell binning passed with them must give the same
result.  The halo is the narrowest that holds every stencil (ORDER / 2
cells, at least 1)."""
import numpy as np
import pytest
import torch

from test_gpu_interpolation import (G2P_FACTORY, G2P_SHAPE, P2G_FACTORY, P2G_SHAPE, TOL, Case,
                                    mpm_fields, rel_l2, shaped, sort_case)

pytestmark = pytest.mark.gpu

ORDERS = [0, 4, 5, 6]
XTOL = {torch.float64: 1e-11, torch.float32: 1e-5}


def halo_of(order):
    return max(1, order // 2)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", ORDERS)
def test_known_answers(ca, oracle, order, dtype):
    """tstInterpolation3d.hpp:118-250: a constant 3.5 with multiplier -0.5
    gives -1.75 from the value operators and 0 from gradients and
    divergences (partition of unity), with points at owned cell centres."""
    c = Case(ca, oracle, num_cell=(8, 7, 9), halo=halo_of(order), n=8, seed=1)
    nc_ = c.mesh.num_cell
    h = c.mesh.halo
    gl = np.array(c.mesh.m.ghost_low)
    idx = np.stack(np.meshgrid(*[np.arange(n) for n in nc_], indexing="ij"), -1).reshape(-1, 3)
    pts = torch.tensor(gl + (idx + h + 0.5) * 0.25, device="cuda")
    n = len(pts)
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    for op in range(5):
        pc, nd = P2G_SHAPE[op]
        grid = c.array(nd, dtype)
        vals = shaped(torch.full((n, pc), 3.5, dtype=dtype, device="cuda"), pc)
        ca.p2g(getattr(ca, P2G_FACTORY[op])(vals, -0.5), pts, n, ca.Spline(order),
               c.halo(grid), grid)
        expect = -1.75 if op in (0, 1) else 0.0
        got = grid.view().cpu().numpy()[c.mesh.owned_slices()]
        np.testing.assert_allclose(got, expect, atol=tol * 8, rtol=0)
    for op in range(5):
        pc, nd = G2P_SHAPE[op]
        grid = c.array(nd, dtype)
        grid.view().fill_(3.5)
        out = shaped(torch.zeros(n, pc, dtype=dtype, device="cuda"), pc)
        ca.g2p(grid, c.halo(grid), pts, n, ca.Spline(order),
               getattr(ca, G2P_FACTORY[op])(out, -0.5))
        expect = -1.75 if op in (0, 1) else 0.0
        np.testing.assert_allclose(out.cpu().numpy(), expect, atol=tol * 64, rtol=0)


@pytest.mark.parametrize("sorted_bins", [False, True])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("op", range(5))
def test_p2g_matches_oracle(ca, oracle, op, order, dtype, sorted_bins):
    c = Case(ca, oracle, n=2000, halo=halo_of(order), seed=800 + op * 7 + order)
    bins = sort_case(ca, c)[0] if sorted_bins else None
    nc, nd = P2G_SHAPE[op]
    vals = c.rng.normal(size=(c.n, nc))
    grid = c.array(nd, dtype)
    init = c.rng.normal(size=grid.view().shape)
    grid.view().copy_(torch.tensor(init, dtype=dtype))
    f = getattr(ca, P2G_FACTORY[op])(shaped(torch.tensor(vals, dtype=dtype, device="cuda"),
                                            nc), -0.6)
    ca.p2g(f, c.pts, c.n, ca.Spline(order), c.halo(grid), grid, bins=bins)
    ref = init.copy()
    oracle.p2g(c.mesh, order, op, -0.6, c.pts_np, vals, ref)
    own = c.mesh.owned_slices()
    err = rel_l2(grid.view().cpu().numpy()[own], ref[own])
    assert err <= TOL[dtype], err
    assert ca.check_domain() == 0


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("op", range(5))
def test_g2p_matches_oracle(ca, oracle, op, order, dtype):
    c = Case(ca, oracle, n=2000, halo=halo_of(order), seed=900 + op * 7 + order)
    nc, nd = G2P_SHAPE[op]
    grid = c.array(nd, dtype)
    g = c.rng.normal(size=grid.view().shape)
    grid.view().copy_(torch.tensor(g, dtype=dtype))
    base = c.rng.normal(size=(c.n, nc))
    out = shaped(torch.tensor(base, dtype=dtype, device="cuda"), nc)
    # G2P adds into the values: compare the increment over the values as
    # stored (FP32 rounds them).
    stored = out.cpu().numpy().reshape(c.n, nc).astype(np.float64)
    ca.g2p(grid, c.halo(grid), c.pts, c.n, ca.Spline(order),
           getattr(ca, G2P_FACTORY[op])(out, 1.7))
    ref = base.copy()
    oracle.g2p(c.mesh, order, op, 1.7, g.copy(), c.pts_np, ref)
    got = out.cpu().numpy().reshape(c.n, nc) - stored
    if order == 0 and op in (2, 3, 4):
        # Spline<0>::gradient is zero (Cabana_Grid_Splines.hpp:111-117).
        assert np.abs(got).max() == 0.0 and np.abs(ref - base).max() == 0.0
    else:
        assert rel_l2(got, ref - base) <= TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("op", range(5))
def test_adjoints_match_oracle(ca, oracle, op, order, dtype):
    c = Case(ca, oracle, n=1500, halo=halo_of(order), seed=1000 + op * 7 + order)
    # P2G adjoint: value and position cotangents.
    nc, nd = P2G_SHAPE[op]
    vals = c.rng.normal(size=(c.n, nc))
    gbar_np = c.rng.normal(size=c.mesh.num_node + (nd,))
    gbar = c.array(nd, dtype)
    gbar.view().copy_(torch.tensor(gbar_np, dtype=dtype))
    vbar = shaped(torch.zeros(c.n, nc, dtype=dtype, device="cuda"), nc)
    xbar = torch.zeros(c.n, 3, dtype=torch.float64, device="cuda")
    f = getattr(ca, P2G_FACTORY[op])(shaped(torch.tensor(vals, dtype=dtype, device="cuda"),
                                            nc), -0.6)
    ca.p2g_adjoint(f, c.pts, c.n, ca.Spline(order), c.halo(gbar), gbar, vbar, xbar)
    vref = np.zeros((c.n, nc))
    xref = np.zeros((c.n, 3))
    oracle.p2g_adjoint(c.mesh, order, op, -0.6, c.pts_np, vals, gbar_np.copy(), vref, xref)
    if order == 0 and op in (2, 3, 4):
        assert np.abs(vbar.cpu().numpy()).max() == 0.0 and np.abs(vref).max() == 0.0
    else:
        assert rel_l2(vbar.cpu().numpy().reshape(c.n, nc), vref) <= TOL[dtype]
    if order == 0:
        assert np.abs(xbar.cpu().numpy()).max() == 0.0 and np.abs(xref).max() == 0.0
    else:
        assert rel_l2(xbar.cpu().numpy(), xref) <= XTOL[dtype]
    # G2P adjoint: grid and position cotangents.
    nc, nd = G2P_SHAPE[op]
    g_np = c.rng.normal(size=c.mesh.num_node + (nd,))
    grid = c.array(nd, dtype)
    grid.view().copy_(torch.tensor(g_np, dtype=dtype))
    ubar_np = c.rng.normal(size=(c.n, nc))
    ubar = shaped(torch.tensor(ubar_np, dtype=dtype, device="cuda"), nc)
    gbar = c.array(nd, dtype)
    gbar.view().zero_()
    xbar.zero_()
    ca.g2p_adjoint(grid, c.halo(grid), c.pts, c.n, ca.Spline(order),
                   getattr(ca, G2P_FACTORY[op])(ubar, 1.3), ubar, gbar, xbar)
    gref = np.zeros(c.mesh.num_node + (nd,))
    xref = np.zeros((c.n, 3))
    oracle.g2p_adjoint(c.mesh, order, op, 1.3, g_np.copy(), c.pts_np, ubar_np, gref, xref)
    own = c.mesh.owned_slices()
    got = gbar.view().cpu().numpy()[own]
    if order == 0 and op in (2, 3, 4):
        assert np.abs(got).max() == 0.0 and np.abs(gref[own]).max() == 0.0
    else:
        assert rel_l2(got, gref[own]) <= TOL[dtype]
    if order == 0:
        assert np.abs(xbar.cpu().numpy()).max() == 0.0
    else:
        assert rel_l2(xbar.cpu().numpy(), xref) <= XTOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("sorted_bins", [False, True])
def test_fused_mpm_transfer(ca, oracle, order, dtype, sorted_bins):
    """The fused MPM sets: P2G of (mass, momentum, stress divergence) and G2P
    of (velocity, velocity gradient) from one grid."""
    c = Case(ca, oracle, num_cell=(12, 10, 9), n=4000, halo=halo_of(order), seed=40 + order)
    m, q, sig = mpm_fields(c, dtype)
    bins = None
    if sorted_bins:
        bins, perm = sort_case(ca, c)
        m, q, sig = m[perm], q[perm], sig[perm]
    t = lambda a: torch.tensor(a, dtype=dtype, device="cuda")  # noqa: E731
    grids = [c.array(nd, dtype) for nd in (1, 3, 3)]
    fs = [ca.createScalarValueP2G(t(m[:, 0]), 1.0), ca.createVectorValueP2G(t(q), 1.0),
          ca.createTensorDivergenceP2G(t(sig.reshape(-1, 3, 3)), -0.25)]
    ca.p2g(fs, c.pts, c.n, ca.Spline(order), c.halo(grids[0]), grids, bins=bins)
    own = c.mesh.owned_slices()
    for arr, op, vals, mult, nd in ((grids[0], 0, m, 1.0, 1), (grids[1], 1, q, 1.0, 3),
                                    (grids[2], 4, sig, -0.25, 3)):
        ref = c.mesh.grid(nd)
        oracle.p2g(c.mesh, order, op, mult, c.pts_np, vals, ref)
        got = arr.view().cpu().numpy()[own]
        if order == 0 and op == 4:
            assert np.abs(got).max() == 0.0
        else:
            assert rel_l2(got, ref[own]) <= TOL[dtype], (op, rel_l2(got, ref[own]))
    u = c.array(3, dtype)
    un = c.rng.normal(size=u.view().shape)
    u.view().copy_(torch.tensor(un, dtype=dtype))
    vel = torch.zeros(c.n, 3, dtype=dtype, device="cuda")
    grad = torch.zeros(c.n, 3, 3, dtype=dtype, device="cuda")
    ca.g2p([u, u], c.halo(u), c.pts, c.n, ca.Spline(order),
           [ca.createVectorValueG2P(vel, 1.0), ca.createVectorGradientG2P(grad, 1.0)],
           bins=bins)
    vref = np.zeros((c.n, 3))
    gref = np.zeros((c.n, 9))
    oracle.g2p(c.mesh, order, 1, 1.0, un.copy(), c.pts_np, vref)
    oracle.g2p(c.mesh, order, 3, 1.0, un.copy(), c.pts_np, gref)
    assert rel_l2(vel.cpu().numpy(), vref) <= TOL[dtype]
    if order == 0:
        assert np.abs(grad.cpu().numpy()).max() == 0.0
    else:
        assert rel_l2(grad.cpu().numpy().reshape(-1, 9), gref) <= TOL[dtype]
    assert ca.check_domain() == 0


def test_halo_narrower_than_the_stencil_is_reported(ca, oracle):
    """A Spline<6> stencil reaches 3 nodes below the point's node: with a
    one-cell halo, points near the low edge leave the local grid, and the
    library reports them instead of writing outside the array."""
    from cabana_amd._native import CabanaAmdError
    c = Case(ca, oracle, n=500, halo=1, seed=3)
    grid = c.array(1, torch.float64)
    vals = torch.ones(c.n, dtype=torch.float64, device="cuda")
    assert ca.check_domain() == 0
    ca.p2g(ca.createScalarValueP2G(vals, 1.0), c.pts, c.n, ca.Spline(6), c.halo(grid), grid)
    with pytest.raises(CabanaAmdError):
        ca.check_domain()
    assert ca.check_domain() == 0
