"""cabana_amd_adjoint_gather: the P2G adjoint and the position part of the
G2P adjoint in one call.  For the MPM set (one fused pass) it must match
the two separate calls and the oracle adjoints, with and without a cell
binning, for the library's own binning and for the per-particle kernels;
other term sets run the two calls' passes."""
import numpy as np
import pytest
import torch

from test_gpu_interpolation import Case, mpm_fields, rel_l2

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-12, torch.float32: 1e-5}


class Problem:
    """MPM adjoint inputs: P2G (mass, momentum, stress divergence) values and
    grid cotangents, a forward velocity grid u and the cotangents of the G2P
    velocity and velocity gradient."""

    def __init__(self, ca, oracle, order, dtype, binned, seed, n=4000):
        c = Case(ca, oracle, num_cell=(12, 10, 9), n=n, seed=seed,
                 halo=max(1, order // 2) + 1)
        self.c = c
        self.ca = ca
        self.dtype = dtype
        self.order = order
        m, q, sig = mpm_fields(c, dtype)
        self.bins = None
        if binned:
            gm = c.lg.globalGrid().globalMesh()
            lcl = ca.createLinkedCellList(c.pts, [gm.cellSize(d) for d in range(3)],
                                          [gm.lowCorner(d) for d in range(3)],
                                          [gm.highCorner(d) for d in range(3)])
            perm = lcl.permutes.cpu().numpy()
            c.pts = c.pts[lcl.permutes].contiguous()
            c.pts_np = c.pts_np[perm]
            m, q, sig = m[perm], q[perm], sig[perm]
            self.bins = lcl
        self.m, self.q, self.sig = m, q, sig
        self.gb = [c.rng.normal(size=c.mesh.num_node + (nd,)) for nd in (1, 3, 3)]
        self.u_np = c.rng.normal(size=c.mesh.num_node + (3,))
        self.ubar_v = c.rng.normal(size=(c.n, 3))
        self.ubar_l = c.rng.normal(size=(c.n, 9))
        self.xbar0 = c.rng.normal(size=(c.n, 3))

    def t(self, a):
        return torch.tensor(a, dtype=self.dtype, device="cuda")

    def arrays(self):
        c = self.c
        gbars = []
        for a, nd in zip(self.gb, (1, 3, 3)):
            arr = c.array(nd, self.dtype)
            arr.view().copy_(torch.tensor(a, dtype=self.dtype))
            gbars.append(arr)
        u = c.array(3, self.dtype)
        u.view().copy_(torch.tensor(self.u_np, dtype=self.dtype))
        return gbars, u

    def terms(self):
        ca = self.ca
        fs = [ca.createScalarValueP2G(self.t(self.m[:, 0]), 1.0),
              ca.createVectorValueP2G(self.t(self.q), 1.0),
              ca.createTensorDivergenceP2G(self.t(self.sig.reshape(-1, 3, 3)), -0.25)]
        uv = self.t(self.ubar_v)
        ul = self.t(self.ubar_l.reshape(-1, 3, 3))
        gs = [ca.createVectorValueG2P(uv, 0.7), ca.createVectorGradientG2P(ul, -1.1)]
        vb = [torch.zeros(self.c.n, dtype=self.dtype, device="cuda"),
              torch.zeros(self.c.n, 3, dtype=self.dtype, device="cuda"),
              torch.zeros(self.c.n, 3, 3, dtype=self.dtype, device="cuda")]
        return fs, gs, [uv, ul], vb

    def fused(self):
        c, ca = self.c, self.ca
        gbars, u = self.arrays()
        fs, gs, ub, vb = self.terms()
        xbar = torch.tensor(self.xbar0, device="cuda")
        ca.adjoint_gather(fs, u, c.halo(u), c.pts, c.n, ca.Spline(self.order), gbars, vb, gs, ub,
                          xbar, bins=self.bins)
        return [v.cpu().numpy().reshape(c.n, -1) for v in vb], xbar.cpu().numpy()

    def separate(self):
        c, ca = self.c, self.ca
        gbars, u = self.arrays()
        fs, gs, ub, vb = self.terms()
        xbar = torch.tensor(self.xbar0, device="cuda")
        ca.p2g_adjoint(fs, c.pts, c.n, ca.Spline(self.order), c.halo(gbars[0]), gbars, vb, xbar,
                       bins=self.bins)
        ca.g2p_adjoint([u, u], c.halo(u), c.pts, c.n, ca.Spline(self.order), gs, ub,
                       [None, None], xbar, bins=self.bins)
        return [v.cpu().numpy().reshape(c.n, -1) for v in vb], xbar.cpu().numpy()

    def oracle(self, oracle):
        c = self.c
        xref = np.zeros((c.n, 3))
        vrefs = []
        for k, (op, vals, mult) in enumerate(((0, self.m, 1.0), (1, self.q, 1.0),
                                              (4, self.sig, -0.25))):
            vref = np.zeros_like(vals)
            oracle.p2g_adjoint(c.mesh, self.order, op, mult, c.pts_np, vals, self.gb[k].copy(),
                               vref, xref)
            vrefs.append(vref.reshape(c.n, -1))
        u = self.u_np.astype(np.float32).astype(np.float64) if self.dtype == torch.float32 \
            else self.u_np
        oracle.g2p_adjoint(c.mesh, self.order, 1, 0.7, u.copy(), c.pts_np, self.ubar_v, None,
                           xref)
        oracle.g2p_adjoint(c.mesh, self.order, 3, -1.1, u.copy(), c.pts_np, self.ubar_l, None,
                           xref)
        return vrefs, xref


def check(pr, oracle, got):
    vb, xb = got
    vrefs, xref = pr.oracle(oracle)
    for k in range(3):
        assert rel_l2(vb[k], vrefs[k]) <= TOL[pr.dtype], k
    assert rel_l2(xb - pr.xbar0, xref) <= TOL[pr.dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("binned", [False, True])
def test_mpm_set_matches_separate_calls_and_oracle(ca, oracle, order, dtype, binned):
    pr = Problem(ca, oracle, order, dtype, binned, seed=40 + order)
    got = pr.fused()
    ref = pr.separate()
    for a, b in zip(got[0], ref[0]):
        np.testing.assert_array_equal(a, b)   # same jets, same arithmetic
    np.testing.assert_allclose(got[1], ref[1], rtol=1e-13, atol=1e-13)
    check(pr, oracle, got)


@pytest.mark.parametrize("order", [2, 4])
def test_per_particle_kernels(ca, oracle, order):
    """The fused pass of the per-particle gather: without the library's own
    binning, and for an order without cell tiles."""
    pr = Problem(ca, oracle, order, torch.float64, False, seed=90 + order, n=2500)
    ca.set_auto_binning(False)
    try:
        got = pr.fused()
    finally:
        ca.set_auto_binning(True)
    check(pr, oracle, got)


def test_other_sets_run_both_calls(ca, oracle):
    """A set that is not the MPM one (two P2G terms, one G2P term) gives the
    two calls' results."""
    pr = Problem(ca, oracle, 2, torch.float64, True, seed=5)
    c = pr.c
    gbars, u = pr.arrays()
    fs, gs, ub, vb = pr.terms()
    xbar = torch.tensor(pr.xbar0, device="cuda")
    ca.adjoint_gather(fs[:2], u, c.halo(u), c.pts, c.n, ca.Spline(2), gbars[:2], vb[:2], gs[1:],
                      ub[1:], xbar, bins=pr.bins)
    xref = np.zeros((c.n, 3))
    for k, (op, vals) in enumerate(((0, pr.m), (1, pr.q))):
        vref = np.zeros_like(vals)
        oracle.p2g_adjoint(c.mesh, 2, op, 1.0, c.pts_np, vals, pr.gb[k].copy(), vref, xref)
        assert rel_l2(vb[k].cpu().numpy().reshape(vals.shape), vref) <= 1e-12
    oracle.g2p_adjoint(c.mesh, 2, 3, -1.1, pr.u_np.copy(), c.pts_np, pr.ubar_l, None, xref)
    assert rel_l2(xbar.cpu().numpy() - pr.xbar0, xref) <= 1e-12
    # grid_bar of the G2P terms is not touched: the velocity grid keeps its values.
    own = c.mesh.owned_slices()
    np.testing.assert_array_equal(u.view().cpu().numpy()[own], pr.u_np[own])


def test_points_bar_is_required(ca, oracle):
    from cabana_amd._native import CabanaAmdError
    pr = Problem(ca, oracle, 2, torch.float64, False, seed=6, n=100)
    c = pr.c
    gbars, u = pr.arrays()
    fs, gs, ub, vb = pr.terms()
    with pytest.raises(CabanaAmdError):
        ca.adjoint_gather(fs, u, c.halo(u), c.pts, c.n, ca.Spline(2), gbars, vb, gs, ub, None)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_mpm_set_with_aosoa_fields(ca, oracle, order, dtype):
    """Every particle field of the fused pass (values, value cotangents and
    G2P output cotangents) as members of one AoSoA with VL = 16: the
    interleaved-field form of the pass, which reads its inputs per slot."""
    pr = Problem(ca, oracle, order, dtype, True, seed=70 + order, n=6000)
    c = pr.c
    vl = 16
    nsoa = (c.n + vl - 1) // vl
    layout = {"m": (0, ()), "q": (1, (3,)), "s": (4, (3, 3)), "mb": (13, ()),
              "qb": (14, (3,)), "sb": (17, (3, 3)), "uv": (26, (3,)), "ul": (29, (3, 3))}
    stride = vl * 38
    raw = torch.zeros(nsoa * stride, dtype=dtype, device="cuda")
    sl = {k: ca.AoSoASlice(raw[vl * o:], c.n, shape, vl, stride) for k, (o, shape) in layout.items()}
    t = lambda a: torch.tensor(a, dtype=dtype, device="cuda")  # noqa: E731
    sl["m"].from_tensor(t(pr.m[:, 0]))
    sl["q"].from_tensor(t(pr.q))
    sl["s"].from_tensor(t(pr.sig.reshape(-1, 3, 3)))
    sl["uv"].from_tensor(t(pr.ubar_v))
    sl["ul"].from_tensor(t(pr.ubar_l.reshape(-1, 3, 3)))
    gbars, u = pr.arrays()
    fs = [ca.createScalarValueP2G(sl["m"], 1.0), ca.createVectorValueP2G(sl["q"], 1.0),
          ca.createTensorDivergenceP2G(sl["s"], -0.25)]
    gs = [ca.createVectorValueG2P(sl["uv"], 0.7), ca.createVectorGradientG2P(sl["ul"], -1.1)]
    xbar = torch.tensor(pr.xbar0, device="cuda")
    ca.adjoint_gather(fs, u, c.halo(u), c.pts, c.n, ca.Spline(order), gbars,
                      [sl["mb"], sl["qb"], sl["sb"]], gs, [sl["uv"], sl["ul"]], xbar, bins=pr.bins)
    vb = [sl[k].to_tensor().reshape(c.n, -1).double().cpu().numpy() for k in ("mb", "qb", "sb")]
    check(pr, oracle, (vb, xbar.cpu().numpy()))
    # The inputs are left as they were.
    np.testing.assert_array_equal(sl["uv"].to_tensor().cpu().numpy(),
                                  t(pr.ubar_v).cpu().numpy())
