"""torch.autograd wrappers over the HIP adjoint kernels.

* gradcheck: the full Jacobians of fused P2G and fused G2P (FP64, all
  operators, spline orders 1-3) against torch's own central differences.
* Differentiation workload: 64^3 periodic cells x 8 particles per cell,
  quadratic spline, L = sum_owned P2G(v)^2 + sum_p |G2P(g)(x_p)|^2; the
  position and value gradients from backward() are checked against FP64
  central differences (h = 1e-6) on 1000 seeded particles.  The differences
  are formed from each particle's own stencil contributions (the rest of the
  grid enters as a fixed residual), with spline weights from the oracle, so
  they do not cancel against the O(1e5) total loss.
* The FP32-mixed gradients against the FP64 ones on the same inputs.
* torch.utils.checkpoint over a multi-step loop gives the same gradients as
  the plain graph.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _grid(ca, cells, halo=1):
    gm = ca.createUniformGlobalMesh([0.0] * 3, [float(cells)] * 3, 1.0)
    return ca.createLocalGrid(ca.createGlobalGrid(gm, [True] * 3), halo)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_gradcheck_p2g(ca, order):
    from cabana_amd import autograd as AG
    torch.manual_seed(10 + order)
    lg = _grid(ca, 4)
    n = 6
    x = (torch.rand(n, 3, dtype=torch.float64, device="cuda") * 4.0).requires_grad_()
    vals = [torch.randn(n, dtype=torch.float64, device="cuda", requires_grad=True),
            torch.randn(n, 3, dtype=torch.float64, device="cuda", requires_grad=True),
            torch.randn(n, 3, 3, dtype=torch.float64, device="cuda", requires_grad=True),
            torch.randn(n, dtype=torch.float64, device="cuda", requires_grad=True),
            torch.randn(n, 3, dtype=torch.float64, device="cuda", requires_grad=True)]
    ops = ["ScalarValue", "VectorValue", "TensorDivergence", "ScalarGradient",
           "VectorDivergence"]
    mults = [1.0, -0.5, 0.75, 1.25, -2.0]

    def f(x, *v):
        return AG.p2g(x, list(v), lg, ops, mults, order=order)

    assert torch.autograd.gradcheck(f, (x, *vals), eps=1e-6, atol=1e-7, rtol=1e-5)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_gradcheck_g2p_shared_grid(ca, order):
    from cabana_amd import autograd as AG
    torch.manual_seed(20 + order)
    lg = _grid(ca, 4)
    n = 6
    x = (torch.rand(n, 3, dtype=torch.float64, device="cuda") * 4.0).requires_grad_()
    u = torch.randn(4, 4, 4, 3, dtype=torch.float64, device="cuda", requires_grad=True)
    phi = torch.randn(4, 4, 4, 1, dtype=torch.float64, device="cuda", requires_grad=True)
    ops = ["VectorValue", "VectorGradient", "VectorDivergence", "ScalarValue",
           "ScalarGradient"]
    mults = [1.0, 0.5, -1.5, 2.0, 0.25]

    def f(x, u, phi):
        return AG.g2p(x, [u, u, u, phi, phi], lg, ops, mults, order=order)

    assert torch.autograd.gradcheck(f, (x, u, phi), eps=1e-6, atol=1e-7, rtol=1e-5)


def _loss(ca, AG, lg, x, v, g, bins):
    grid = AG.p2g(x, v, lg, "ScalarValue", order=2, bins=bins)
    u = AG.g2p(x, g, lg, "VectorValue", order=2, bins=bins)
    return (grid * grid).sum() + (u * u).sum(), grid


def _differentiation_case(ca, seed=4242):
    cells, ppc = 64, 8
    n = cells ** 3 * ppc
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.rand(n, 3, dtype=torch.float64, device="cuda", generator=g) * cells
    lg = _grid(ca, cells)
    lcl = ca.createLinkedCellList(x, [1.0] * 3, [0.0] * 3, [float(cells)] * 3)
    x = x[lcl.permutes].contiguous()
    v = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    gf = torch.randn(cells, cells, cells, 3, dtype=torch.float64, device="cuda", generator=g)
    return cells, n, lg, lcl, x, v, gf


def test_differentiation_fd(ca, oracle):
    from cabana_amd import autograd as AG
    cells, n, lg, lcl, x, v, gf = _differentiation_case(ca)
    xr = x.clone().requires_grad_()
    vr = v.clone().requires_grad_()
    loss, grid = _loss(ca, AG, lg, xr, vr, gf, lcl)
    loss.backward()
    gx = xr.grad.cpu().numpy()
    gv = vr.grad.cpu().numpy()
    G = grid.detach().cpu().numpy()[..., 0]
    g_own = gf.cpu().numpy()
    x_np = x.cpu().numpy()
    v_np = v.cpu().numpy()
    mesh = oracle.OMesh([0.0] * 3, [float(cells)] * 3, 1.0, halo=1)
    h = 1e-6
    halo = 1

    def stencil(p):
        sd = oracle.evaluate_spline(mesh, 2, p)
        idx = [np.array([(s - halo) % cells for s in sd["s"][d]]) for d in range(3)]
        w = np.einsum("i,j,k->ijk", *[np.array(sd["w"][d]) for d in range(3)])
        return idx, w

    def contrib(p, val):
        idx, w = stencil(p)
        out = {}
        for a in range(3):
            for b in range(3):
                for c in range(3):
                    key = (idx[0][a], idx[1][b], idx[2][c])
                    out[key] = out.get(key, 0.0) + val * w[a, b, c]
        return out

    def u_at(p):
        idx, w = stencil(p)
        sub = g_own[np.ix_(idx[0], idx[1], idx[2])]
        return np.einsum("ijk,ijkd->d", w, sub)

    rng = np.random.default_rng(7)
    sample = rng.choice(n, size=1000, replace=False)
    fd_x = np.zeros((len(sample), 3))
    fd_v = np.zeros(len(sample))
    for r, p in enumerate(sample):
        w0 = contrib(x_np[p], 1.0)
        c0 = {k: v_np[p] * w for k, w in w0.items()}
        # d/dv_p: sum_n 2 G_n w_n(x_p) (L is quadratic in v_p, so the central
        # difference equals this for any h).
        fd_v[r] = sum(2.0 * G[k] * w for k, w in w0.items())
        for d in range(3):
            xp = x_np[p].copy()
            xm = x_np[p].copy()
            xp[d] += h
            xm[d] -= h
            cp = contrib(xp, v_np[p])
            cm = contrib(xm, v_np[p])
            keys = set(cp) | set(cm) | set(c0)
            t1 = 0.0
            for k in keys:
                R = G[k] - c0.get(k, 0.0)
                a, b = cp.get(k, 0.0), cm.get(k, 0.0)
                t1 += (a - b) * (2.0 * R + a + b)
            up, um = u_at(xp), u_at(xm)
            t2 = (up - um) @ (up + um)
            fd_x[r, d] = (t1 + t2) / (2.0 * h)
    err_x = np.abs(gx[sample] - fd_x)
    err_v = np.abs(gv[sample] - fd_v)
    scale_x = np.abs(fd_x).max()
    print("position grad: max |ad - fd| = %.3e (max |fd| = %.3e); value grad: %.3e"
          % (err_x.max(), scale_x, err_v.max()))
    assert np.all(err_x <= 1e-6 * np.maximum(1.0, np.abs(fd_x)))
    assert np.all(err_v <= 1e-6 * np.maximum(1.0, np.abs(fd_v)))


def test_differentiation_fp32_mixed_vs_fp64(ca):
    from cabana_amd import autograd as AG
    cells, n, lg, lcl, x, v, gf = _differentiation_case(ca, seed=777)
    grads = {}
    for dt in (torch.float64, torch.float32):
        xr = x.clone().requires_grad_()
        vr = v.detach().to(dt).clone().requires_grad_()
        loss, _ = _loss(ca, AG, lg, xr, vr, gf.to(dt), lcl)
        loss.backward()
        grads[dt] = (xr.grad.double(), vr.grad.double())
    ex = (torch.linalg.norm(grads[torch.float32][0] - grads[torch.float64][0]) /
          torch.linalg.norm(grads[torch.float64][0])).item()
    ev = (torch.linalg.norm(grads[torch.float32][1] - grads[torch.float64][1]) /
          torch.linalg.norm(grads[torch.float64][1])).item()
    print("FP32-mixed vs FP64 gradient rel-L2: positions %.3e, values %.3e" % (ex, ev))
    assert ex <= 1e-5 and ev <= 1e-5


def _loop_case(ca, n=3000, cells=12, seed=5):
    from cabana_amd import mpm
    box = mpm.PeriodicBox(cells)
    x = mpm.clustered_positions(n, cells, num_cluster=4, sigma=(1.0, 3.0), seed=seed)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed + 1)
    v = torch.randn(n, 3, dtype=torch.float64, device="cuda", generator=g) * 0.2
    m = torch.rand(n, dtype=torch.float64, device="cuda", generator=g) + 0.5
    return mpm, box, x, v, m


def _loop_loss(x, v):
    return (torch.sin(x) ** 2).sum() + (v ** 2).sum()


def test_checkpointed_loop_matches_plain(ca):
    mpm, box, x0, v0, m = _loop_case(ca)
    grads = []
    for every in (0, 2):
        x = x0.clone().requires_grad_()
        v = v0.clone().requires_grad_()
        xs, vs, _, _ = mpm.run_loop(box, x, v, m, num_steps=6, checkpoint_every=every)
        _loop_loss(xs, vs).backward()
        grads.append((x.grad, v.grad))
    # Atomic accumulation order may differ between the first run and the
    # recomputation, so equality is to rounding, not bitwise.
    for a, b in zip(grads[0], grads[1]):
        assert torch.allclose(a, b, rtol=1e-9, atol=1e-11)


def _knot_distance(x):
    """Distance of every coordinate to the nearest knot of the quadratic
    spline on the unit-cell box (the half-integers: node 0 of the ghosted
    grid is at -1 and Spline<2> adds 0.5), where the second derivative of
    the weights jumps."""
    return (torch.remainder(x - 0.5, 1.0) - 0.5).abs().neg().add(0.5)


def test_loop_gradient_fd(ca):
    """Whole-loop gradient (re-binning gather, P2G, grid update, G2P,
    advection; two steps, checkpointed) against central differences.  The
    spline weights are C1 with second derivatives that jump at the knots, so
    the perturbed particles are chosen at least 1e-3 from every knot at both
    steps: a central difference across a jump in f'' is off by O(h), which
    would hide a real error of that size."""
    mpm, box, x0, v0, m = _loop_case(ca, n=2000, seed=11)
    x = x0.clone().requires_grad_()
    v = v0.clone().requires_grad_()
    xs, vs, _, _ = mpm.run_loop(box, x, v, m, num_steps=2, checkpoint_every=2)
    _loop_loss(xs, vs).backward()
    rng = np.random.default_rng(3)
    h = 1e-6
    worst = 0.0
    with torch.no_grad():
        x1, _, _, _ = mpm.run_loop(box, x0, v0, m, num_steps=1, checkpoint_every=0)
        clear = ((_knot_distance(x0).min(dim=1).values > 1e-3) &
                 (_knot_distance(x1).min(dim=1).values > 1e-3)).cpu().numpy()
        candidates = np.flatnonzero(clear)
        assert candidates.size >= 12

        def loss_at(xx, vv):
            xs, vs, _, _ = mpm.run_loop(box, xx, vv, m, num_steps=2, checkpoint_every=0)
            return _loop_loss(xs, vs).item()
        for p in rng.choice(candidates, size=12, replace=False):
            for d in range(3):
                for which, grad in (("x", x.grad), ("v", v.grad)):
                    xp, xm = x0.clone(), x0.clone()
                    vp, vm = v0.clone(), v0.clone()
                    if which == "x":
                        xp[p, d] += h
                        xm[p, d] -= h
                    else:
                        vp[p, d] += h
                        vm[p, d] -= h
                    fd = (loss_at(xp, vp) - loss_at(xm, vm)) / (2 * h)
                    ad = grad[p, d].item()
                    err = abs(ad - fd) / max(1.0, abs(fd))
                    worst = max(worst, err)
    print("loop gradient vs FD: worst relative error %.3e" % worst)
    assert worst <= 1e-6
