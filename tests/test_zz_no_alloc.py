"""cabana_amd_p2g_no_alloc (no reference counterpart) against cabana_amd_p2g,
both called through the C ABI on the same inputs: the fused MPM set
(ScalarValue, VectorValue, TensorDivergence) and a single VectorValue term,
cell-sorted (tiled scatter, part table in the fixed scratch) and unsorted,
FP64 (rel-L2 <= 1e-12) and FP32-mixed (rel-L2 <= 1e-5: both sum the same
contributions in different atomic orders).  Two calls in a row, the second
on a smaller input, so the second reuses scratch the first wrote."""
import ctypes

import numpy as np
import pytest
import torch

from cabana_amd import _native as N
from cabana_amd import grid as G

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-12, torch.float32: 1e-5}


def _setup(ca, cells, ppc, dtype, seed):
    L = float(cells)
    gm = ca.createUniformGlobalMesh([0.0] * 3, [L] * 3, 1.0)
    lg = ca.createLocalGrid(ca.createGlobalGrid(gm, [True] * 3), 1)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    n = cells ** 3 * ppc
    pts = torch.rand(n, 3, dtype=torch.float64, device="cuda", generator=g) * L
    lcl = ca.createLinkedCellList(pts, [1.0] * 3, [0.0] * 3, [L] * 3)
    ca.permute(lcl, pts)
    m = (torch.rand(n, dtype=torch.float64, device="cuda", generator=g) + 0.5).to(dtype)
    q = torch.randn(n, 3, dtype=torch.float64, device="cuda", generator=g).to(dtype)
    sig = torch.randn(n, 3, 3, dtype=torch.float64, device="cuda", generator=g).to(dtype)
    return lg, pts, lcl, n, [(0, m, 1), (1, q, 3), (4, sig, 3)]


def _call(ca, entry, lg, pts, n, order, terms_in, bins):
    """One p2g call through `entry` into fresh zeroed grids; returns them."""
    dtype = terms_in[0][1].dtype
    grids = [ca.createArray("g", ca.createArrayLayout(lg, d, ca.Node()), dtype=dtype)
             for _, _, d in terms_in]
    m = lg.mesh_descriptor()
    p = G._points_descriptor(pts, n)
    terms = (N.P2GTerm * len(terms_in))()
    for i, ((op, vals, _), a) in enumerate(zip(terms_in, grids)):
        terms[i].op = op
        terms[i].multiplier = -1.0 if op == 4 else 1.0
        terms[i].values = G._values_descriptor(vals, n)
        terms[i].grid = a.descriptor()
    b = bins.binning_descriptor() if bins is not None else None
    ctx = N.context()
    N.check(getattr(N.load(), entry)(ctx, ctypes.byref(m), order, ctypes.byref(p),
                                     len(terms_in), terms, -1,
                                     ctypes.byref(b) if b is not None else None,
                                     N.current_stream()), ctx)
    torch.cuda.synchronize()
    return [a.view().double().cpu().numpy() for a in grids]


def _rel(a, b):
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_no_alloc_matches_p2g_two_calls(ca, dtype, order):
    for call, (cells, ppc, seed) in enumerate(((24, 8, 7), (12, 5, 8))):
        lg, pts, lcl, n, terms = _setup(ca, cells, ppc, dtype, seed)
        print("no-alloc call %d: %d^3 cells, %d particles, order %d, %s"
              % (call, cells, n, order, dtype), flush=True)
        for subset in (terms, terms[1:2]):
            for bins in (lcl, None):
                want = _call(ca, "cabana_amd_p2g", lg, pts, n, order, subset, bins)
                got = _call(ca, "cabana_amd_p2g_no_alloc", lg, pts, n, order, subset, bins)
                for a, b in zip(got, want):
                    e = _rel(a, b)
                    assert e <= TOL[dtype], (call, len(subset), bins is not None, e)


def test_no_alloc_rejects_a_part_table_beyond_the_scratch(ca):
    """A mesh whose brick part table exceeds the fixed scratch (1000^3 cells:
    about 2 M bricks of 8^3 cells, 39 MB of table) is refused with an error
    before anything is launched."""
    cells = 1000
    L = float(cells)
    gm = ca.createUniformGlobalMesh([0.0] * 3, [L] * 3, 1.0)
    lg = ca.createLocalGrid(ca.createGlobalGrid(gm, [True] * 3), 1)
    n = 1 << 16
    pts = torch.rand(n, 3, dtype=torch.float64, device="cuda") * L
    lcl = ca.createLinkedCellList(pts, [1.0] * 3, [0.0] * 3, [L] * 3)
    ca.permute(lcl, pts)
    m = torch.ones(n, dtype=torch.float32, device="cuda")
    print("no-alloc rejection: %d^3 cells, %d particles" % (cells, n), flush=True)
    with pytest.raises(RuntimeError, match="no-allocation scratch"):
        _call(ca, "cabana_amd_p2g_no_alloc", lg, pts, n, 2, [(0, m, 1)], lcl)
    del lcl, pts
    torch.cuda.empty_cache()
