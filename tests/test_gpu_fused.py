"""Fused passes against the oracle: the reference benchmark's five-functor
fused_p2g and fused_g2p (benchmark/grid/Cabana_Grid_InterpolationPerformance.cpp
:25-127) and its g2p2g (:130-174), with and without a cell binning, plus the
MPM term sets given in another order (still one fused tiled pass)."""
import numpy as np
import pytest
import torch

from test_gpu_interpolation import TOL, Case, rel_l2, sort_case

pytestmark = pytest.mark.gpu


def _fields(c, dtype):
    s = c.rng.normal(size=(c.n, 1))
    v = c.rng.normal(size=(c.n, 3))
    t = c.rng.normal(size=(c.n, 9))
    return s, v, t


def _cuda(a, dtype, shape=None):
    x = torch.tensor(a, dtype=dtype, device="cuda")
    return x.reshape(shape) if shape is not None else x


@pytest.mark.parametrize("sorted_bins", [False, True])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [1, 2, 4])
def test_five_functor_fused_p2g(ca, oracle, order, dtype, sorted_bins):
    """fused_p2g :25-77: ScalarValue and VectorDivergence into the scalar
    grid, VectorValue, ScalarGradient and TensorDivergence into the vector
    grid, multiplier -0.5."""
    c = Case(ca, oracle, n=4000, halo=max(1, order // 2), seed=70 + order)
    bins = None
    s, v, t = _fields(c, dtype)
    if sorted_bins:
        bins, perm = sort_case(ca, c)
        s, v, t = s[perm], v[perm], t[perm]
    sg = c.array(1, dtype)
    vg = c.array(3, dtype)
    ts, tv, tt = _cuda(s[:, 0], dtype), _cuda(v, dtype), _cuda(t, dtype, (-1, 3, 3))
    fs = [ca.createScalarValueP2G(ts, -0.5), ca.createVectorValueP2G(tv, -0.5),
          ca.createScalarGradientP2G(ts, -0.5), ca.createVectorDivergenceP2G(tv, -0.5),
          ca.createTensorDivergenceP2G(tt, -0.5)]
    ca.p2g(fs, c.pts, c.n, ca.Spline(order), c.halo(sg), [sg, vg, vg, sg, vg], bins=bins)
    # One fold per grid after all the terms: each term's oracle P2G goes
    # into a fresh grid (a second oracle call on one grid would fold the
    # first call's ghost partial sums again) and the owned nodes are summed.
    own = c.mesh.owned_slices()
    rs = c.mesh.grid(1)
    rv = c.mesh.grid(3)
    for op, vals, grid in ((0, s, rs), (1, v, rv), (2, s, rv), (3, v, rs), (4, t, rv)):
        one = c.mesh.grid(grid.shape[-1])
        oracle.p2g(c.mesh, order, op, -0.5, c.pts_np, vals, one)
        grid[own] += one[own]
    assert rel_l2(sg.view().cpu().numpy()[own], rs[own]) <= TOL[dtype]
    assert rel_l2(vg.view().cpu().numpy()[own], rv[own]) <= TOL[dtype]
    assert ca.check_domain() == 0


@pytest.mark.parametrize("sorted_bins", [False, True])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [1, 2, 4])
def test_five_functor_fused_g2p(ca, oracle, order, dtype, sorted_bins):
    """fused_g2p :81-127: ScalarValue (scalar grid) and VectorDivergence
    (vector grid) add into the scalar field, VectorValue (vector grid) and
    ScalarGradient (scalar grid) into the vector field, VectorGradient
    (vector grid) into the tensor field."""
    c = Case(ca, oracle, n=4000, halo=max(1, order // 2), seed=80 + order)
    bins = sort_case(ca, c)[0] if sorted_bins else None
    s0, v0, t0 = _fields(c, dtype)
    sg = c.array(1, dtype)
    vg = c.array(3, dtype)
    gs = c.rng.normal(size=sg.view().shape)
    gv = c.rng.normal(size=vg.view().shape)
    sg.view().copy_(torch.tensor(gs, dtype=dtype))
    vg.view().copy_(torch.tensor(gv, dtype=dtype))
    ts, tv, tt = _cuda(s0[:, 0], dtype), _cuda(v0, dtype), _cuda(t0, dtype, (-1, 3, 3))
    stored = [x.cpu().numpy().reshape(c.n, -1).astype(np.float64) for x in (ts, tv, tt)]
    fs = [ca.createScalarValueG2P(ts, -0.5), ca.createVectorValueG2P(tv, -0.5),
          ca.createScalarGradientG2P(tv, -0.5), ca.createVectorGradientG2P(tt, -0.5),
          ca.createVectorDivergenceG2P(ts, -0.5)]
    ca.g2p([sg, vg, sg, vg, vg], c.halo(sg), c.pts, c.n, ca.Spline(order), fs, bins=bins)
    rs, rv, rt = (np.zeros((c.n, 1)), np.zeros((c.n, 3)), np.zeros((c.n, 9)))
    for op, grid, out in ((0, gs, rs), (1, gv, rv), (2, gs, rv), (3, gv, rt), (4, gv, rs)):
        oracle.g2p(c.mesh, order, op, -0.5, grid.copy(), c.pts_np, out)
    for got, st, ref in ((ts, stored[0], rs), (tv, stored[1], rv), (tt, stored[2], rt)):
        assert rel_l2(got.cpu().numpy().reshape(c.n, -1) - st, ref) <= TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [2, 3])
def test_mpm_sets_in_another_order(ca, oracle, order, dtype):
    """(TensorDivergence, VectorValue, ScalarValue) P2G and (VectorGradient,
    VectorValue) G2P from one grid are the fused tiled passes too."""
    c = Case(ca, oracle, num_cell=(12, 10, 9), n=6000, seed=90 + order)
    bins, perm = sort_case(ca, c)
    s, v, t = [a[perm] for a in _fields(c, dtype)]
    grids = [c.array(nd, dtype) for nd in (3, 3, 1)]
    fs = [ca.createTensorDivergenceP2G(_cuda(t, dtype, (-1, 3, 3)), -0.25),
          ca.createVectorValueP2G(_cuda(v, dtype), 1.0),
          ca.createScalarValueP2G(_cuda(s[:, 0], dtype), 1.0)]
    ca.profile_reset()
    ca.profile_enable(True)
    ca.p2g(fs, c.pts, c.n, ca.Spline(order), c.halo(grids[0]), grids, bins=bins)
    own = c.mesh.owned_slices()
    for arr, op, vals, mult, nd in ((grids[0], 4, t, -0.25, 3), (grids[1], 1, v, 1.0, 3),
                                    (grids[2], 0, s, 1.0, 1)):
        ref = c.mesh.grid(nd)
        oracle.p2g(c.mesh, order, op, mult, c.pts_np, vals, ref)
        assert rel_l2(arr.view().cpu().numpy()[own], ref[own]) <= TOL[dtype]
    u = c.array(3, dtype)
    un = c.rng.normal(size=u.view().shape)
    u.view().copy_(torch.tensor(un, dtype=dtype))
    grad = torch.zeros(c.n, 3, 3, dtype=dtype, device="cuda")
    vel = torch.zeros(c.n, 3, dtype=dtype, device="cuda")
    ca.g2p([u, u], c.halo(u), c.pts, c.n, ca.Spline(order),
           [ca.createVectorGradientG2P(grad, 1.0), ca.createVectorValueG2P(vel, 1.0)], bins=bins)
    prof = ca.profile_report()
    ca.profile_enable(False)
    assert "p2g_mpm_cells" in prof and "g2p_mpm_tiled" in prof, sorted(prof)
    vref = np.zeros((c.n, 3))
    gref = np.zeros((c.n, 9))
    oracle.g2p(c.mesh, order, 1, 1.0, un.copy(), c.pts_np, vref)
    oracle.g2p(c.mesh, order, 3, 1.0, un.copy(), c.pts_np, gref)
    assert rel_l2(vel.cpu().numpy(), vref) <= TOL[dtype]
    assert rel_l2(grad.cpu().numpy().reshape(-1, 9), gref) <= TOL[dtype]


@pytest.mark.parametrize("sorted_bins", [False, True])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [1, 2, 5])
def test_scalar_g2p2g(ca, oracle, order, dtype, sorted_bins):
    """g2p2g :130-174 as the benchmark calls it: ScalarValueG2P and
    ScalarValueP2G of one particle field, multiplier -0.5, one scalar array
    that is gathered, read, added into and folded."""
    c = Case(ca, oracle, n=4000, halo=max(1, order // 2), seed=100 + order)
    bins = sort_case(ca, c)[0] if sorted_bins else None
    sv = c.rng.normal(size=(c.n, 1))
    grid = c.array(1, dtype)
    g0 = c.rng.normal(size=grid.view().shape)
    grid.view().copy_(torch.tensor(g0, dtype=dtype))
    ts = _cuda(sv[:, 0], dtype)
    s_stored = ts.cpu().numpy().astype(np.float64)[:, None]
    ca.g2p2g(ca.createScalarValueP2G(ts, -0.5), c.pts, c.n, ca.Spline(order), c.halo(grid),
             grid, ca.createScalarValueG2P(ts, -0.5), bins=bins)
    g = g0.astype(np.float32).astype(np.float64) if dtype == torch.float32 else g0.copy()
    s = s_stored.copy()
    oracle.halo_gather(c.mesh, g)
    oracle.g2p(c.mesh, order, 0, -0.5, g, c.pts_np, s)
    oracle.p2g(c.mesh, order, 0, -0.5, c.pts_np, s, g)
    assert rel_l2(ts.cpu().numpy()[:, None] - s_stored, s - s_stored) <= TOL[dtype]
    own = c.mesh.owned_slices()
    assert rel_l2(grid.view().cpu().numpy()[own], g[own]) <= TOL[dtype]
    assert ca.check_domain() == 0


@pytest.mark.parametrize("sorted_bins", [False, True])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_mpm_style_g2p2g(ca, oracle, dtype, sorted_bins):
    """G2P of (velocity, velocity gradient) from a velocity grid, then P2G of
    (mass, momentum, stress divergence) into three other grids, where the
    momentum P2G reads the velocity field the G2P just added into."""
    order = 2
    c = Case(ca, oracle, num_cell=(12, 10, 9), n=6000, seed=111)
    m, v, t = _fields(c, dtype)
    m = np.abs(m) + 0.5
    bins = None
    if sorted_bins:
        bins, perm = sort_case(ca, c)
        m, v, t = m[perm], v[perm], t[perm]
    u = c.array(3, dtype)
    un = c.rng.normal(size=u.view().shape)
    u.view().copy_(torch.tensor(un, dtype=dtype))
    tm, tv, tt = _cuda(m[:, 0], dtype), _cuda(v, dtype), _cuda(t, dtype, (-1, 3, 3))
    v_stored = tv.cpu().numpy().astype(np.float64)
    t_stored = tt.cpu().numpy().reshape(-1, 9).astype(np.float64)
    grids = [c.array(nd, dtype) for nd in (1, 3, 3)]
    ca.g2p2g([ca.createScalarValueP2G(tm, 1.0), ca.createVectorValueP2G(tv, 2.0),
              ca.createTensorDivergenceP2G(tt, -0.25)], c.pts, c.n, ca.Spline(order),
             c.halo(u), grids,
             [ca.createVectorValueG2P(tv, 1.0), ca.createVectorGradientG2P(tt, 1.0)],
             g2p_array=[u, u], bins=bins)
    vref = v_stored.copy()
    tref = t_stored.copy()
    ug = un.copy()
    oracle.g2p(c.mesh, order, 1, 1.0, ug, c.pts_np, vref)
    oracle.g2p(c.mesh, order, 3, 1.0, ug, c.pts_np, tref)
    assert rel_l2(tv.cpu().numpy(), vref) <= TOL[dtype]
    assert rel_l2(tt.cpu().numpy().reshape(-1, 9), tref) <= TOL[dtype]
    own = c.mesh.owned_slices()
    mm = tm.cpu().numpy().astype(np.float64)[:, None]
    for arr, op, vals, mult, nd in ((grids[0], 0, mm, 1.0, 1), (grids[1], 1, vref, 2.0, 3),
                                    (grids[2], 4, tref, -0.25, 3)):
        ref = c.mesh.grid(nd)
        oracle.p2g(c.mesh, order, op, mult, c.pts_np, vals, ref)
        assert rel_l2(arr.view().cpu().numpy()[own], ref[own]) <= TOL[dtype], op
