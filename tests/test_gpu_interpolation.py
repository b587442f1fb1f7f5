"""HIP path against the oracle on identical seeded inputs: every operator,
spline orders 1-3, FP64 (rel-L2 <= 1e-12) and FP32-mixed (rel-L2 <= 1e-5),
particle layouts (AoS, SoA, AoSoA), grid layouts (LayoutRight/Left), the
fused MPM term sets and the cell-sorted tiled scatter."""
import contextlib

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

P2G_SHAPE = {0: (1, 1), 1: (3, 3), 2: (1, 3), 3: (3, 1), 4: (9, 3)}
G2P_SHAPE = {0: (1, 1), 1: (3, 3), 2: (3, 1), 3: (9, 3), 4: (1, 3)}
P2G_FACTORY = ["createScalarValueP2G", "createVectorValueP2G", "createScalarGradientP2G",
               "createVectorDivergenceP2G", "createTensorDivergenceP2G"]
G2P_FACTORY = ["createScalarValueG2P", "createVectorValueG2P", "createScalarGradientG2P",
               "createVectorGradientG2P", "createVectorDivergenceG2P"]
TOL = {torch.float64: 1e-12, torch.float32: 1e-5}


def rel_l2(got, ref):
    got = np.asarray(got, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    den = np.linalg.norm(ref)
    return np.linalg.norm(got - ref) / (den if den > 0 else 1.0)


def shaped(arr, nc):
    n = arr.shape[0]
    if nc == 1:
        return arr.reshape(n)
    if nc == 3:
        return arr.reshape(n, 3)
    return arr.reshape(n, 3, 3)


class Case:
    def __init__(self, ca, oracle, num_cell=(10, 9, 11), cell=0.25, n=3000, halo=1, seed=0,
                 low=(0.3, -1.1, 2.0)):
        self.ca = ca
        self.oracle = oracle
        low = np.array(low, dtype=np.float64)
        high = low + cell * np.array(num_cell)
        gm = ca.createUniformGlobalMesh(low, high, cell)
        gg = ca.createGlobalGrid(gm, [True, True, True])
        self.lg = ca.createLocalGrid(gg, halo)
        self.mesh = oracle.OMesh(low, high, cell, halo=halo)
        self.rng = np.random.default_rng(seed)
        self.pts_np = low + self.rng.uniform(0, 1, size=(n, 3)) * (high - low)
        self.pts = torch.tensor(self.pts_np, device="cuda")
        self.n = n

    def array(self, dofs, dtype, memory_layout="right"):
        layout = self.ca.createArrayLayout(self.lg, dofs, self.ca.Node())
        return self.ca.createArray("a", layout, dtype=dtype, memory_layout=memory_layout)

    def halo(self, array):
        return self.ca.createHalo(self.ca.NodeHaloPattern(), -1, array)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_interpolation3d_known_answers(ca, oracle, order, dtype):
    """tstInterpolation3d.hpp:118-250 on the GPU: points at every owned cell
    centre, constant 3.5, multiplier -0.5."""
    cell = 0.1
    gm = ca.createUniformGlobalMesh([-1.2, 0.1, 1.1], [-0.3, 9.5, 2.3], cell)
    gg = ca.createGlobalGrid(gm, [True, True, True])
    lg = ca.createLocalGrid(gg, 1)
    lm = ca.createLocalMesh(lg)
    cs = lg.indexSpace(ca.Own(), ca.Cell(), ca.Local())
    pts = np.zeros((cs.size(), 3))
    for i in range(cs.min(0), cs.max(0)):
        for j in range(cs.min(1), cs.max(1)):
            for k in range(cs.min(2), cs.max(2)):
                pid = (i - 1) + cs.extent(0) * ((j - 1) + cs.extent(1) * (k - 1))
                pts[pid] = lm.coordinates(ca.Cell(), [i, j, k])
    points = torch.tensor(pts, device="cuda")
    n = len(pts)
    node_space = lg.indexSpace(ca.Own(), ca.Node(), ca.Local())
    own = tuple(slice(node_space.min(d), node_space.max(d)) for d in range(3))
    sl = ca.createArrayLayout(lg, 1, ca.Node())
    vl = ca.createArrayLayout(lg, 3, ca.Node())
    sgrid = ca.createArray("s", sl, dtype=dtype)
    vgrid = ca.createArray("v", vl, dtype=dtype)
    shalo = ca.createHalo(ca.NodeHaloPattern(), 1, sgrid)
    vhalo = ca.createHalo(ca.NodeHaloPattern(), 1, vgrid)
    spf = torch.full((n,), 3.5, dtype=dtype, device="cuda")
    vpf = torch.full((n, 3), 3.5, dtype=dtype, device="cuda")
    tpf = torch.full((n, 3, 3), 3.5, dtype=dtype, device="cuda")
    sp = ca.Spline(order)

    def run_p2g(functor, grid, halo, expect):
        ca.ArrayOp.assign(grid, 0.0, ca.Ghost())
        ca.p2g(functor, points, n, sp, halo, grid)
        got = grid.view()[own].cpu().numpy()
        if dtype == torch.float64:
            tol = 1e-10
        else:
            tol = 1e-4 if expect != 0.0 else 1e-3
        np.testing.assert_allclose(got, expect, atol=tol, rtol=0)

    run_p2g(ca.createScalarValueP2G(spf, -0.5), sgrid, shalo, -1.75)
    run_p2g(ca.createVectorValueP2G(vpf, -0.5), vgrid, vhalo, -1.75)
    run_p2g(ca.createScalarGradientP2G(spf, -0.5), vgrid, vhalo, 0.0)
    run_p2g(ca.createVectorDivergenceP2G(vpf, -0.5), sgrid, shalo, 0.0)
    run_p2g(ca.createTensorDivergenceP2G(tpf, -0.5), vgrid, vhalo, 0.0)

    ca.ArrayOp.assign(sgrid, 3.5, ca.Own())
    ca.ArrayOp.assign(vgrid, 3.5, ca.Own())
    tol_g = 1e-9 if dtype == torch.float64 else 1e-3

    def run_g2p(grid, halo, out, functor, expect):
        out.zero_()
        ca.g2p(grid, halo, points, n, sp, functor)
        np.testing.assert_allclose(out.cpu().numpy(), expect, atol=tol_g, rtol=0)

    run_g2p(sgrid, shalo, spf, ca.createScalarValueG2P(spf, -0.5), -1.75)
    run_g2p(vgrid, vhalo, vpf, ca.createVectorValueG2P(vpf, -0.5), -1.75)
    run_g2p(sgrid, shalo, vpf, ca.createScalarGradientG2P(vpf, -0.5), 0.0)
    run_g2p(vgrid, vhalo, tpf, ca.createVectorGradientG2P(tpf, -0.5), 0.0)
    run_g2p(vgrid, vhalo, spf, ca.createVectorDivergenceG2P(spf, -0.5), 0.0)
    assert ca.check_domain() == 0


def sort_case(ca, c):
    """Cell-sort the case's points with the library's LinkedCellList;
    returns the binning and the permutation applied."""
    gm = c.lg.globalGrid().globalMesh()
    lcl = ca.createLinkedCellList(c.pts, [gm.cellSize(d) for d in range(3)],
                                  [gm.lowCorner(d) for d in range(3)],
                                  [gm.highCorner(d) for d in range(3)])
    perm = lcl.permutes.cpu().numpy()
    c.pts = c.pts[lcl.permutes].contiguous()
    c.pts_np = c.pts_np[perm]
    return lcl, perm


@contextlib.contextmanager
def per_particle_path(ca, on):
    """Unbinned calls without the library's own binning (the per-particle
    kernels with global atomics) while `on`."""
    if on:
        ca.set_auto_binning(False)
    try:
        yield
    finally:
        if on:
            ca.set_auto_binning(True)


# sorted_bins: True = the caller's LinkedCellList; False = none (the
# library bins the particles itself); "atomic" = none, per-particle path.
@pytest.mark.parametrize("sorted_bins", [False, True, "atomic"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("op", range(5))
def test_p2g_matches_oracle(ca, oracle, op, order, dtype, sorted_bins):
    c = Case(ca, oracle, seed=op * 7 + order)
    bins = sort_case(ca, c)[0] if sorted_bins is True else None
    nc, nd = P2G_SHAPE[op]
    vals = c.rng.normal(size=(c.n, nc))
    grid = c.array(nd, dtype)
    # Pre-existing content is accumulated into (p2g adds).
    init = c.rng.normal(size=grid.view().shape)
    grid.view().copy_(torch.tensor(init, dtype=dtype))
    f = getattr(ca, P2G_FACTORY[op])(shaped(torch.tensor(vals, dtype=dtype, device="cuda"),
                                            nc), -0.6)
    with per_particle_path(ca, sorted_bins == "atomic"):
        ca.p2g(f, c.pts, c.n, ca.Spline(order), c.halo(grid), grid, bins=bins)
    ref = init.copy()
    oracle.p2g(c.mesh, order, op, -0.6, c.pts_np, vals, ref)
    own = c.mesh.owned_slices()
    err = rel_l2(grid.view().cpu().numpy()[own], ref[own])
    assert err <= TOL[dtype], err
    assert ca.check_domain() == 0


@pytest.mark.parametrize("sorted_bins", [False, True, "atomic"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("op", range(5))
def test_g2p_matches_oracle(ca, oracle, op, order, dtype, sorted_bins):
    c = Case(ca, oracle, seed=100 + op * 7 + order)
    bins = sort_case(ca, c)[0] if sorted_bins is True else None
    nc, nd = G2P_SHAPE[op]
    grid = c.array(nd, dtype)
    g = c.rng.normal(size=grid.view().shape)
    grid.view().copy_(torch.tensor(g, dtype=dtype))
    base = c.rng.normal(size=(c.n, nc))
    out = shaped(torch.tensor(base, dtype=dtype, device="cuda"), nc)
    f = getattr(ca, G2P_FACTORY[op])(out, 1.7)
    with per_particle_path(ca, sorted_bins == "atomic"):
        ca.g2p(grid, c.halo(grid), c.pts, c.n, ca.Spline(order), f, bins=bins)
    ref = base.copy()
    gref = g.copy()
    oracle.g2p(c.mesh, order, op, 1.7, gref, c.pts_np, ref)
    err = rel_l2(out.cpu().numpy().reshape(c.n, nc) - base, ref - base)
    assert err <= TOL[dtype], err
    # The ghost fill is a side effect on the array, as in the reference.
    np.testing.assert_allclose(grid.view().cpu().numpy(), gref,
                               rtol=0, atol=1e-6 if dtype == torch.float32 else 0)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_layouts(ca, oracle, order):
    """AoSoA slices (VL = 64, another member in the SoA), column-major SoA
    views and a LayoutLeft grid give the same FP64 result."""
    c = Case(ca, oracle, n=1000, seed=5)
    vl = 64
    nsoa = (c.n + vl - 1) // vl
    stride = vl * (3 + 3 + 9)        # position, velocity, stress members
    raw = torch.zeros(nsoa * stride, dtype=torch.float64, device="cuda")
    pos_slice = ca.AoSoASlice(raw[0:], c.n, (3,), vl, stride)
    vel_slice = ca.AoSoASlice(raw[3 * vl:], c.n, (3,), vl, stride)
    pos_slice.from_tensor(c.pts)
    vel = c.rng.normal(size=(c.n, 3))
    vel_slice.from_tensor(torch.tensor(vel, device="cuda"))
    grid = c.array(3, torch.float64, memory_layout="left")
    f = ca.createVectorValueP2G(vel_slice, 0.5)
    ca.p2g(f, pos_slice, c.n, ca.Spline(order), c.halo(grid), grid)
    ref = c.mesh.grid(3)
    oracle.p2g(c.mesh, order, 1, 0.5, c.pts_np, vel, ref)
    own = c.mesh.owned_slices()
    assert rel_l2(grid.view().cpu().numpy()[own], ref[own]) <= 1e-12
    # G2P into a column-major (SoA) view.
    out = torch.zeros(3, c.n, dtype=torch.float64, device="cuda").t()
    g2 = ca.createVectorValueG2P(out, 2.0)
    ca.g2p(grid, c.halo(grid), pos_slice, c.n, ca.Spline(order), g2)
    oref = np.zeros((c.n, 3))
    oracle.g2p(c.mesh, order, 1, 2.0, ref, c.pts_np, oref)
    assert rel_l2(out.cpu().numpy(), oref) <= 1e-12


def mpm_fields(c, dtype):
    m = c.rng.uniform(0.5, 1.5, size=(c.n, 1))
    v = c.rng.normal(size=(c.n, 3))
    q = m * v
    s6 = c.rng.normal(size=(c.n, 6))
    sig = np.stack([s6[:, 0], s6[:, 3], s6[:, 4], s6[:, 3], s6[:, 1], s6[:, 5],
                    s6[:, 4], s6[:, 5], s6[:, 2]], axis=1)
    return m, q, sig


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("sorted_bins", [False, True, "stale"])
def test_fused_mpm_transfer(ca, oracle, order, dtype, sorted_bins):
    """P2G (mass, momentum, stress divergence) and G2P (velocity, velocity
    gradient) as fused passes, with and without cell-sorted tiles, and with a
    stale binning (particles moved by up to 1.5 cells after the sort)."""
    c = Case(ca, oracle, num_cell=(12, 10, 9), n=6000, seed=9 + order)
    m, q, sig = mpm_fields(c, dtype)
    pts = c.pts
    pts_np = c.pts_np
    bins = None
    if sorted_bins:
        gm = c.lg.globalGrid().globalMesh()
        lo = [gm.lowCorner(d) for d in range(3)]
        hi = [gm.highCorner(d) for d in range(3)]
        lcl = ca.createLinkedCellList(pts, [gm.cellSize(d) for d in range(3)], lo, hi)
        perm = lcl.permutes.cpu().numpy()
        pts = pts[lcl.permutes].contiguous()
        pts_np = pts_np[perm]
        m, q, sig = m[perm], q[perm], sig[perm]
        bins = lcl
        if sorted_bins == "stale":
            lo_a, hi_a = np.array(lo), np.array(hi)
            step = c.rng.uniform(-1.5, 1.5, size=pts_np.shape) * gm.cellSize(0)
            pts_np = lo_a + np.mod(pts_np + step - lo_a, hi_a - lo_a)
            pts = torch.tensor(pts_np, device="cuda")
    tm = torch.tensor(m[:, 0], dtype=dtype, device="cuda")
    tq = torch.tensor(q, dtype=dtype, device="cuda")
    ts = torch.tensor(sig.reshape(-1, 3, 3), dtype=dtype, device="cuda")
    gmass = c.array(1, dtype)
    gmom = c.array(3, dtype)
    gstress = c.array(3, dtype)
    fs = [ca.createScalarValueP2G(tm, 1.0), ca.createVectorValueP2G(tq, 1.0),
          ca.createTensorDivergenceP2G(ts, -0.25)]
    ca.p2g(fs, pts, c.n, ca.Spline(order), c.halo(gmass), [gmass, gmom, gstress], bins=bins)
    own = c.mesh.owned_slices()
    for arr, op, vals, mult, nd in ((gmass, 0, m, 1.0, 1), (gmom, 1, q, 1.0, 3),
                                    (gstress, 4, sig, -0.25, 3)):
        ref = c.mesh.grid(nd)
        oracle.p2g(c.mesh, order, op, mult, pts_np, vals, ref)
        err = rel_l2(arr.view().cpu().numpy()[own], ref[own])
        assert err <= TOL[dtype], (op, err)
    # G2P of a grid velocity (fused VectorValue + VectorGradient on one grid).
    u = c.array(3, dtype)
    un = c.rng.normal(size=u.view().shape)
    u.view().copy_(torch.tensor(un, dtype=dtype))
    vel = torch.zeros(c.n, 3, dtype=dtype, device="cuda")
    grad = torch.zeros(c.n, 3, 3, dtype=dtype, device="cuda")
    ca.g2p([u, u], c.halo(u), pts, c.n, ca.Spline(order),
           [ca.createVectorValueG2P(vel, 1.0), ca.createVectorGradientG2P(grad, 1.0)],
           bins=bins)
    uref = un.copy()
    vref = np.zeros((c.n, 3))
    gref = np.zeros((c.n, 9))
    oracle.g2p(c.mesh, order, 1, 1.0, uref, pts_np, vref)
    oracle.g2p(c.mesh, order, 3, 1.0, uref, pts_np, gref)
    assert rel_l2(vel.cpu().numpy(), vref) <= TOL[dtype]
    assert rel_l2(grad.cpu().numpy().reshape(-1, 9), gref) <= TOL[dtype]
    assert ca.check_domain() == 0


@pytest.mark.parametrize("order", [1, 2, 3])
def test_fused_mpm_many_parts_per_workgroup(ca, oracle, order, monkeypatch):
    """Eight workgroups for ~1600 parts (16 particles per cell): every
    workgroup walks ~200 parts, so the scatter's part-table refills and
    run-table double buffers cycle many times.  P2G (mass, momentum, stress
    divergence) against the oracle, FP32-mixed."""
    monkeypatch.setenv("CABANA_AMD_MAX_TILE_GROUPS", "8")
    dtype = torch.float32
    c = Case(ca, oracle, num_cell=(36, 36, 34), n=700000, seed=77 + order)
    m, q, sig = mpm_fields(c, dtype)
    gm = c.lg.globalGrid().globalMesh()
    lo = [gm.lowCorner(d) for d in range(3)]
    hi = [gm.highCorner(d) for d in range(3)]
    lcl = ca.createLinkedCellList(c.pts, [gm.cellSize(d) for d in range(3)], lo, hi)
    perm = lcl.permutes.cpu().numpy()
    pts = c.pts[lcl.permutes].contiguous()
    pts_np = c.pts_np[perm]
    m, q, sig = m[perm], q[perm], sig[perm]
    tm = torch.tensor(m[:, 0], dtype=dtype, device="cuda")
    tq = torch.tensor(q, dtype=dtype, device="cuda")
    ts = torch.tensor(sig.reshape(-1, 3, 3), dtype=dtype, device="cuda")
    gmass, gmom, gstress = c.array(1, dtype), c.array(3, dtype), c.array(3, dtype)
    fs = [ca.createScalarValueP2G(tm, 1.0), ca.createVectorValueP2G(tq, 1.0),
          ca.createTensorDivergenceP2G(ts, -0.25)]
    ca.p2g(fs, pts, c.n, ca.Spline(order), c.halo(gmass), [gmass, gmom, gstress], bins=lcl)
    own = c.mesh.owned_slices()
    for arr, op, vals, mult, nd in ((gmass, 0, m, 1.0, 1), (gmom, 1, q, 1.0, 3),
                                    (gstress, 4, sig, -0.25, 3)):
        ref = c.mesh.grid(nd)
        oracle.p2g(c.mesh, order, op, mult, pts_np, vals, ref)
        err = rel_l2(arr.view().cpu().numpy()[own], ref[own])
        assert err <= TOL[dtype], (op, err)
    assert ca.check_domain() == 0


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_tiled_layout_left_grids_and_aosoa_attributes(ca, oracle, order, dtype):
    """Cell-sorted (tiled) MPM P2G and G2P with LayoutLeft grids and the
    particle attributes in one AoSoA (VL = 16: the interleaved path)."""
    c = Case(ca, oracle, num_cell=(12, 10, 9), n=20000, seed=41 + order)
    m, q, sig = mpm_fields(c, dtype)
    gm = c.lg.globalGrid().globalMesh()
    lo = [gm.lowCorner(d) for d in range(3)]
    hi = [gm.highCorner(d) for d in range(3)]
    lcl = ca.createLinkedCellList(c.pts, [gm.cellSize(d) for d in range(3)], lo, hi)
    perm = lcl.permutes.cpu().numpy()
    pts = c.pts[lcl.permutes].contiguous()
    pts_np = c.pts_np[perm]
    m, q, sig = m[perm], q[perm], sig[perm]
    vl = 16
    nsoa = (c.n + vl - 1) // vl
    stride = vl * (1 + 3 + 9)
    raw = torch.zeros(nsoa * stride, dtype=dtype, device="cuda")
    sm = ca.AoSoASlice(raw[0:], c.n, (), vl, stride)
    sq = ca.AoSoASlice(raw[vl:], c.n, (3,), vl, stride)
    ss = ca.AoSoASlice(raw[4 * vl:], c.n, (3, 3), vl, stride)
    sm.from_tensor(torch.tensor(m[:, 0], device="cuda"))
    sq.from_tensor(torch.tensor(q, device="cuda"))
    ss.from_tensor(torch.tensor(sig.reshape(-1, 3, 3), device="cuda"))
    gmass = c.array(1, dtype, memory_layout="left")
    gmom = c.array(3, dtype, memory_layout="left")
    gstress = c.array(3, dtype, memory_layout="left")
    fs = [ca.createScalarValueP2G(sm, 1.0), ca.createVectorValueP2G(sq, 1.0),
          ca.createTensorDivergenceP2G(ss, -0.25)]
    ca.p2g(fs, pts, c.n, ca.Spline(order), c.halo(gmass), [gmass, gmom, gstress], bins=lcl)
    own = c.mesh.owned_slices()
    for arr, op, vals, mult, nd in ((gmass, 0, m, 1.0, 1), (gmom, 1, q, 1.0, 3),
                                    (gstress, 4, sig, -0.25, 3)):
        ref = c.mesh.grid(nd)
        oracle.p2g(c.mesh, order, op, mult, pts_np, vals, ref)
        assert rel_l2(arr.view().cpu().numpy()[own], ref[own]) <= TOL[dtype], op
    u = c.array(3, dtype, memory_layout="left")
    un = c.rng.normal(size=u.view().shape)
    u.view().copy_(torch.tensor(un, dtype=dtype))
    vel = torch.zeros(c.n, 3, dtype=dtype, device="cuda")
    grad = torch.zeros(c.n, 3, 3, dtype=dtype, device="cuda")
    ca.g2p([u, u], c.halo(u), pts, c.n, ca.Spline(order),
           [ca.createVectorValueG2P(vel, 1.0), ca.createVectorGradientG2P(grad, 1.0)],
           bins=lcl)
    vref = np.zeros((c.n, 3))
    gref = np.zeros((c.n, 9))
    oracle.g2p(c.mesh, order, 1, 1.0, un.copy(), pts_np, vref)
    oracle.g2p(c.mesh, order, 3, 1.0, un.copy(), pts_np, gref)
    assert rel_l2(vel.cpu().numpy(), vref) <= TOL[dtype]
    assert rel_l2(grad.cpu().numpy().reshape(-1, 9), gref) <= TOL[dtype]
    assert ca.check_domain() == 0


@pytest.mark.parametrize("order", [1, 2, 3])
def test_tiled_scatter_handles_unsorted_and_partial_bins(ca, oracle, order):
    """With a binning that covers only [begin, end) and particles that are
    NOT permuted (so many leave their tile windows), the result still equals
    the oracle: out-of-window particles go straight to global memory."""
    c = Case(ca, oracle, num_cell=(16, 12, 10), n=5000, seed=31)
    gm = c.lg.globalGrid().globalMesh()
    lo = [gm.lowCorner(d) for d in range(3)]
    hi = [gm.highCorner(d) for d in range(3)]
    lcl = ca.createLinkedCellList(c.pts, [gm.cellSize(d) for d in range(3)], lo, hi,
                                  begin=700, end=4100)
    vals = c.rng.normal(size=(c.n, 3))
    grid = c.array(1, torch.float64)
    f = ca.createVectorDivergenceP2G(torch.tensor(vals, device="cuda"), 1.0)
    ca.p2g(f, c.pts, c.n, ca.Spline(order), c.halo(grid), grid, bins=lcl)
    ref = c.mesh.grid(1)
    oracle.p2g(c.mesh, order, 3, 1.0, c.pts_np, vals, ref)
    own = c.mesh.owned_slices()
    assert rel_l2(grid.view().cpu().numpy()[own], ref[own]) <= 1e-12


def test_out_of_domain_particles_are_reported(ca, oracle):
    c = Case(ca, oracle, n=100, seed=3)
    pts = c.pts.clone()
    pts[5, 0] = 1.0e6
    pts[7, 1] = float("nan")
    grid = c.array(1, torch.float64)
    vals = torch.ones(c.n, dtype=torch.float64, device="cuda")
    ca.p2g(ca.createScalarValueP2G(vals, 1.0), pts, c.n, ca.Spline(2), c.halo(grid), grid)
    from cabana_amd._native import CabanaAmdError
    with pytest.raises(CabanaAmdError):
        ca.check_domain()
    assert ca.check_domain() == 0   # the counter was cleared


def test_invalid_arguments_raise(ca, oracle):
    c = Case(ca, oracle, n=10)
    grid = c.array(3, torch.float64)
    vals = torch.ones(c.n, dtype=torch.float64, device="cuda")
    from cabana_amd._native import CabanaAmdError
    with pytest.raises(CabanaAmdError):   # scalar P2G into a 3-dof grid
        ca.p2g(ca.createScalarValueP2G(vals, 1.0), c.pts, c.n, ca.Spline(2), c.halo(grid),
               grid)
    f32 = torch.ones(c.n, 3, dtype=torch.float32, device="cuda")
    with pytest.raises(CabanaAmdError):   # FP32 values into an FP64 grid
        ca.p2g(ca.createVectorValueP2G(f32, 1.0), c.pts, c.n, ca.Spline(2), c.halo(grid),
               grid)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_dense_cells_multiple_batches(ca, oracle, order, dtype):
    """~200 particles per cell in a cluster: bricks exceed the per-batch sort
    capacity and base cells are flushed in several batches."""
    c = Case(ca, oracle, num_cell=(6, 6, 6), cell=0.5, n=1, seed=17)
    rng = np.random.default_rng(17)
    gm = c.lg.globalGrid().globalMesh()
    lo = np.array([gm.lowCorner(d) for d in range(3)])
    n = 60000
    centre = lo + 1.4
    pts = np.clip(centre + rng.normal(scale=0.35, size=(n, 3)), lo, lo + 2.999)
    c.pts_np = pts
    c.pts = torch.tensor(pts, device="cuda")
    c.n = n
    bins, perm = sort_case(ca, c)
    vals = rng.normal(size=(n, 9))
    grid = c.array(3, dtype)
    f = ca.createTensorDivergenceP2G(torch.tensor(vals.reshape(-1, 3, 3), dtype=dtype,
                                                  device="cuda"), 0.5)
    ca.p2g(f, c.pts, n, ca.Spline(order), c.halo(grid), grid, bins=bins)
    ref = c.mesh.grid(3)
    oracle.p2g(c.mesh, order, 4, 0.5, c.pts_np, vals, ref)
    own = c.mesh.owned_slices()
    assert rel_l2(grid.view().cpu().numpy()[own], ref[own]) <= TOL[dtype]
    assert ca.check_domain() == 0


class _NoSerial:
    """The same binning without a build serial (the part tables are rebuilt in
    every call)."""

    def __init__(self, lcl):
        self.lcl = lcl

    def binning_descriptor(self):
        b = self.lcl.binning_descriptor()
        b.serial = 0
        return b


@pytest.mark.parametrize("order", [1, 2])
def test_part_table_reuse_and_rebuild(ca, oracle, order):
    """A binning with a build serial reuses its brick part tables across
    calls: repeated tiled P2G calls equal those of the same binning without a
    serial to FP64 rounding (the grid flush adds through atomics, so the
    order of the additions varies), tiled G2P calls bitwise.  After the same
    LinkedCellList is rebuilt in
    place (same counts/offsets buffers) over a clustered cloud, whose bricks
    need other part counts, the tiled P2G still matches the oracle."""
    dtype = torch.float64
    c = Case(ca, oracle, num_cell=(16, 14, 12), n=60000, seed=91 + order)
    gm = c.lg.globalGrid().globalMesh()
    lo = np.array([gm.lowCorner(d) for d in range(3)])
    hi = np.array([gm.highCorner(d) for d in range(3)])
    delta = [gm.cellSize(d) for d in range(3)]
    lcl = ca.createLinkedCellList(c.pts, delta, list(lo), list(hi))
    pts = c.pts[lcl.permutes].contiguous()
    mass_np = c.rng.uniform(0.5, 1.5, size=c.n)
    mass = torch.tensor(mass_np, device="cuda")

    def p2g(bins, points):
        g = c.array(1, dtype)
        ca.p2g(ca.createScalarValueP2G(mass, 1.0), points, c.n, ca.Spline(order), c.halo(g), g,
               bins=bins)
        return g.view().cpu().numpy()

    def g2p(bins, points, grid):
        out = torch.zeros(c.n, 3, dtype=dtype, device="cuda")
        ca.g2p(grid, c.halo(grid), points, c.n, ca.Spline(order),
               ca.createVectorValueG2P(out, 1.0), bins=bins)
        return out.cpu().numpy()

    ref = p2g(_NoSerial(lcl), pts)
    own = c.mesh.owned_slices()
    for _ in range(3):
        assert rel_l2(p2g(lcl, pts)[own], ref[own]) <= TOL[dtype]
    gv = c.array(3, dtype)
    gv.view().copy_(torch.tensor(c.rng.normal(size=gv.view().shape), dtype=dtype))
    uref = g2p(_NoSerial(lcl), pts, gv)
    for _ in range(2):
        assert np.array_equal(g2p(lcl, pts, gv), uref)

    # In-place rebuild over a cloud clustered around the domain's centre.
    x2_np = lo + (hi - lo) * np.clip(0.5 + 0.1 * c.rng.normal(size=(c.n, 3)), 0.0, 0.999)
    lcl.build(torch.tensor(x2_np, device="cuda"), 0, c.n)
    perm2 = lcl.permutes.cpu().numpy()
    pts2_np = x2_np[perm2]
    got = p2g(lcl, torch.tensor(pts2_np, device="cuda"))
    refg = c.mesh.grid(1)
    oracle.p2g(c.mesh, order, 0, 1.0, pts2_np, mass_np[:, None], refg)
    assert rel_l2(got[own], refg[own]) <= TOL[dtype]
    assert ca.check_domain() == 0


@pytest.mark.parametrize("vl", [None, 16, 32, 64])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_mass_momentum_pair(ca, oracle, order, dtype, vl):
    """P2G of (ScalarValue, VectorValue) in one call: one fused pass over the
    cell-sorted particles of a binning of [700, n - 900), the particles
    outside it through the atomic pass; each grid against the oracle.  The
    attributes are contiguous tensors (vl None) or members of one AoSoA of
    vector length vl (the interleaved instantiation of the fused pass)."""
    c = Case(ca, oracle, num_cell=(14, 12, 11), n=40000, seed=131 + order)
    gm = c.lg.globalGrid().globalMesh()
    lo = [gm.lowCorner(d) for d in range(3)]
    hi = [gm.highCorner(d) for d in range(3)]
    begin, end = 700, c.n - 900
    lcl = ca.createLinkedCellList(c.pts, [gm.cellSize(d) for d in range(3)], lo, hi,
                                  begin, end)
    perm = np.arange(c.n)
    perm[begin:end] = lcl.permutes.cpu().numpy()
    pts_np = c.pts_np[perm]
    m = c.rng.uniform(0.5, 1.5, size=(c.n, 1))
    q = m * c.rng.normal(size=(c.n, 3))
    tm = torch.tensor(m[:, 0], dtype=dtype, device="cuda")
    tq = torch.tensor(q, dtype=dtype, device="cuda")
    if vl is not None:
        nsoa = (c.n + vl - 1) // vl
        stride = vl * (1 + 3)
        raw = torch.zeros(nsoa * stride, dtype=dtype, device="cuda")
        sm = ca.AoSoASlice(raw[0:], c.n, (), vl, stride)
        sq = ca.AoSoASlice(raw[vl:], c.n, (3,), vl, stride)
        sm.from_tensor(tm)
        sq.from_tensor(tq)
        tm, tq = sm, sq
    gmass, gmom = c.array(1, dtype), c.array(3, dtype)
    ca.p2g([ca.createScalarValueP2G(tm, 0.5), ca.createVectorValueP2G(tq, -2.0)],
           torch.tensor(pts_np, device="cuda"), c.n, ca.Spline(order), c.halo(gmass),
           [gmass, gmom], bins=lcl)
    own = c.mesh.owned_slices()
    for arr, op, vals, mult, nd in ((gmass, 0, m, 0.5, 1), (gmom, 1, q, -2.0, 3)):
        ref = c.mesh.grid(nd)
        oracle.p2g(c.mesh, order, op, mult, pts_np, vals, ref)
        err = rel_l2(arr.view().cpu().numpy()[own], ref[own])
        assert err <= TOL[dtype], (op, err)
    assert ca.check_domain() == 0


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_library_binning_with_particles_in_the_halo(ca, oracle, order, dtype):
    """Unbinned calls bin by owned cell themselves; particles in the ghost
    layer (their keys clamp into a boundary cell) and clustered particles
    still match the oracle, for the MPM P2G set and the G2P pair, and the
    library's own binning is what runs (its kernels appear in the profile)."""
    c = Case(ca, oracle, num_cell=(12, 10, 9), n=6000, halo=2, seed=31 + order)
    lo = np.array(c.mesh.m.global_low)
    hi = np.array(c.mesh.m.global_high)
    cell = 0.25
    pts = c.pts_np.copy()
    k = c.n // 10
    pts[:k] = lo + c.rng.uniform(-0.9, 0.0, size=(k, 3)) * cell
    pts[k:2 * k] = hi + c.rng.uniform(0.0, 0.9, size=(k, 3)) * cell
    pts[2 * k:4 * k] = lo + (hi - lo) * 0.5 + c.rng.normal(size=(2 * k, 3)) * cell
    c.pts_np = pts
    c.pts = torch.tensor(pts, device="cuda")
    m, q, sig = mpm_fields(c, dtype)
    t = lambda a: torch.tensor(a, dtype=dtype, device="cuda")  # noqa: E731
    grids = [c.array(nd, dtype) for nd in (1, 3, 3)]
    ca.profile_reset()
    ca.profile_enable(True)
    ca.p2g([ca.createScalarValueP2G(t(m[:, 0]), 1.0), ca.createVectorValueP2G(t(q), 1.0),
            ca.createTensorDivergenceP2G(t(sig.reshape(-1, 3, 3)), -0.25)], c.pts, c.n,
           ca.Spline(order), c.halo(grids[0]), grids)
    u = c.array(3, dtype)
    un = c.rng.normal(size=u.view().shape)
    u.view().copy_(torch.tensor(un, dtype=dtype))
    vel = torch.zeros(c.n, 3, dtype=dtype, device="cuda")
    grad = torch.zeros(c.n, 3, 3, dtype=dtype, device="cuda")
    ca.g2p([u, u], c.halo(u), c.pts, c.n, ca.Spline(order),
           [ca.createVectorValueG2P(vel, 1.0), ca.createVectorGradientG2P(grad, 1.0)])
    prof = ca.profile_report()
    ca.profile_enable(False)
    assert "auto_bin_keys" in prof and "p2g_mpm_cells" in prof and "g2p_mpm_tiled" in prof
    own = c.mesh.owned_slices()
    for arr, op, vals, mult, nd in ((grids[0], 0, m, 1.0, 1), (grids[1], 1, q, 1.0, 3),
                                    (grids[2], 4, sig, -0.25, 3)):
        ref = c.mesh.grid(nd)
        oracle.p2g(c.mesh, order, op, mult, c.pts_np, vals, ref)
        assert rel_l2(arr.view().cpu().numpy()[own], ref[own]) <= TOL[dtype], op
    vref = np.zeros((c.n, 3))
    gref = np.zeros((c.n, 9))
    oracle.g2p(c.mesh, order, 1, 1.0, un.copy(), c.pts_np, vref)
    oracle.g2p(c.mesh, order, 3, 1.0, un.copy(), c.pts_np, gref)
    assert rel_l2(vel.cpu().numpy(), vref) <= TOL[dtype]
    assert rel_l2(grad.cpu().numpy().reshape(-1, 9), gref) <= TOL[dtype]
    assert ca.check_domain() == 0
