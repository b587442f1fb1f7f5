"""HIP adjoint kernels against the oracle adjoints (which are pinned by
central finite differences in test_oracle_interpolation.py): value, grid
and position cotangents, single operators and the fused MPM term sets."""
import numpy as np
import pytest
import torch

from test_gpu_interpolation import (G2P_FACTORY, G2P_SHAPE, P2G_FACTORY, P2G_SHAPE, Case,
                                    mpm_fields, rel_l2, shaped)

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-12, torch.float32: 1e-5}
XTOL = {torch.float64: 1e-11, torch.float32: 1e-5}


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("op", range(5))
def test_p2g_adjoint_matches_oracle(ca, oracle, op, order, dtype):
    c = Case(ca, oracle, n=2000, seed=500 + op * 5 + order)
    nc, nd = P2G_SHAPE[op]
    vals = c.rng.normal(size=(c.n, nc))
    gbar_np = c.rng.normal(size=c.mesh.num_node + (nd,))
    gbar = c.array(nd, dtype)
    gbar.view().copy_(torch.tensor(gbar_np, dtype=dtype))
    tv = shaped(torch.tensor(vals, dtype=dtype, device="cuda"), nc)
    vbar0 = c.rng.normal(size=(c.n, nc))
    vbar = shaped(torch.tensor(vbar0, dtype=dtype, device="cuda"), nc)
    xbar0 = c.rng.normal(size=(c.n, 3))
    xbar = torch.tensor(xbar0, device="cuda")
    f = getattr(ca, P2G_FACTORY[op])(tv, -0.6)
    ca.p2g_adjoint(f, c.pts, c.n, ca.Spline(order), c.halo(gbar), gbar, vbar, xbar)
    vref = np.zeros((c.n, nc))
    xref = np.zeros((c.n, 3))
    oracle.p2g_adjoint(c.mesh, order, op, -0.6, c.pts_np, vals, gbar_np.copy(), vref, xref)
    assert rel_l2(vbar.cpu().numpy().reshape(c.n, nc) - vbar0, vref) <= TOL[dtype]
    assert rel_l2(xbar.cpu().numpy() - xbar0, xref) <= XTOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("op", range(5))
@pytest.mark.parametrize("sorted_bins", [False, True])
def test_g2p_adjoint_matches_oracle(ca, oracle, op, order, dtype, sorted_bins):
    c = Case(ca, oracle, n=2000, seed=600 + op * 5 + order)
    nc, nd = G2P_SHAPE[op]
    pts, pts_np = c.pts, c.pts_np
    bins = None
    if sorted_bins:
        gm = c.lg.globalGrid().globalMesh()
        lcl = ca.createLinkedCellList(pts, [gm.cellSize(d) for d in range(3)],
                                      [gm.lowCorner(d) for d in range(3)],
                                      [gm.highCorner(d) for d in range(3)])
        pts = pts[lcl.permutes].contiguous()
        pts_np = pts_np[lcl.permutes.cpu().numpy()]
        bins = lcl
    g_np = c.rng.normal(size=c.mesh.num_node + (nd,))
    grid = c.array(nd, dtype)
    grid.view().copy_(torch.tensor(g_np, dtype=dtype))
    ubar_np = c.rng.normal(size=(c.n, nc))
    ubar = shaped(torch.tensor(ubar_np, dtype=dtype, device="cuda"), nc)
    gbar0 = c.rng.normal(size=c.mesh.num_node + (nd,))
    gbar = c.array(nd, dtype)
    gbar.view().copy_(torch.tensor(gbar0, dtype=dtype))
    xbar = torch.zeros(c.n, 3, dtype=torch.float64, device="cuda")
    f = getattr(ca, G2P_FACTORY[op])(ubar, 1.3)
    ca.g2p_adjoint(grid, c.halo(grid), pts, c.n, ca.Spline(order), f, ubar, gbar, xbar,
                   bins=bins)
    gref = np.zeros_like(gbar0)
    xref = np.zeros((c.n, 3))
    oracle.g2p_adjoint(c.mesh, order, op, 1.3, g_np.copy(), pts_np, ubar_np, gref, xref)
    own = c.mesh.owned_slices()
    got = gbar.view().cpu().numpy()[own] - gbar0[own]
    assert rel_l2(got, gref[own]) <= TOL[dtype]
    assert rel_l2(xbar.cpu().numpy(), xref) <= XTOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("sorted_bins", [False, True])
def test_fused_mpm_adjoints(ca, oracle, order, dtype, sorted_bins):
    c = Case(ca, oracle, num_cell=(12, 10, 9), n=4000, seed=70 + order)
    m, q, sig = mpm_fields(c, dtype)
    pts, pts_np = c.pts, c.pts_np
    bins = None
    if sorted_bins:
        gm = c.lg.globalGrid().globalMesh()
        lcl = ca.createLinkedCellList(pts, [gm.cellSize(d) for d in range(3)],
                                      [gm.lowCorner(d) for d in range(3)],
                                      [gm.highCorner(d) for d in range(3)])
        perm = lcl.permutes.cpu().numpy()
        pts = pts[lcl.permutes].contiguous()
        pts_np = pts_np[perm]
        m, q, sig = m[perm], q[perm], sig[perm]
        bins = lcl
    t = lambda a: torch.tensor(a, dtype=dtype, device="cuda")  # noqa: E731
    tm, tq, ts = t(m[:, 0]), t(q), t(sig.reshape(-1, 3, 3))
    fs = [ca.createScalarValueP2G(tm, 1.0), ca.createVectorValueP2G(tq, 1.0),
          ca.createTensorDivergenceP2G(ts, -0.25)]
    gb = [c.rng.normal(size=c.mesh.num_node + (nd,)) for nd in (1, 3, 3)]
    gbars = []
    for a, nd in zip(gb, (1, 3, 3)):
        arr = c.array(nd, dtype)
        arr.view().copy_(torch.tensor(a, dtype=dtype))
        gbars.append(arr)
    vb = [torch.zeros(c.n, dtype=dtype, device="cuda"),
          torch.zeros(c.n, 3, dtype=dtype, device="cuda"),
          torch.zeros(c.n, 3, 3, dtype=dtype, device="cuda")]
    xbar = torch.zeros(c.n, 3, dtype=torch.float64, device="cuda")
    ca.p2g_adjoint(fs, pts, c.n, ca.Spline(order), c.halo(gbars[0]), gbars, vb, xbar,
                   bins=bins)
    xref = np.zeros((c.n, 3))
    for k, (op, vals, mult) in enumerate(((0, m, 1.0), (1, q, 1.0), (4, sig, -0.25))):
        vref = np.zeros_like(vals)
        oracle.p2g_adjoint(c.mesh, order, op, mult, pts_np, vals, gb[k].copy(), vref, xref)
        assert rel_l2(vb[k].cpu().numpy().reshape(vals.shape), vref) <= TOL[dtype]
    assert rel_l2(xbar.cpu().numpy(), xref) <= XTOL[dtype]

    # G2P adjoint of (velocity, velocity gradient) on one grid.
    u_np = c.rng.normal(size=c.mesh.num_node + (3,))
    u = c.array(3, dtype)
    u.view().copy_(torch.tensor(u_np, dtype=dtype))
    ubar_v = c.rng.normal(size=(c.n, 3))
    ubar_l = c.rng.normal(size=(c.n, 9))
    gbar = c.array(3, dtype)
    xbar.zero_()
    fv = ca.createVectorValueG2P(t(ubar_v), 1.0)
    fl = ca.createVectorGradientG2P(t(ubar_l.reshape(-1, 3, 3)), 1.0)
    ca.g2p_adjoint([u, u], c.halo(u), pts, c.n, ca.Spline(order), [fv, fl],
                   [t(ubar_v), t(ubar_l.reshape(-1, 3, 3))], [gbar, gbar], xbar, bins=bins)
    gref = np.zeros(c.mesh.num_node + (3,))
    xref = np.zeros((c.n, 3))
    oracle.g2p_adjoint(c.mesh, order, 1, 1.0, u_np.copy(), pts_np, ubar_v, gref, xref)
    oracle.g2p_adjoint(c.mesh, order, 3, 1.0, u_np.copy(), pts_np, ubar_l, gref, xref)
    own = c.mesh.owned_slices()
    assert rel_l2(gbar.view().cpu().numpy()[own], gref[own]) <= TOL[dtype]
    assert rel_l2(xbar.cpu().numpy(), xref) <= XTOL[dtype]
