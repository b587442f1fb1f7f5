"""Differentiable P2G / G2P: torch.autograd.Function wrappers over the C ABI.

Forward calls cabana_amd_p2g / cabana_amd_g2p; backward calls the
hand-written adjoint kernels (cabana_amd_p2g_adjoint, a G2P-shaped gather,
and cabana_amd_g2p_adjoint, a P2G-shaped scatter), which recompute the spline
weights instead of storing them.  Gradients flow to the particle values, to
the grid fields and to the particle positions (through the spline weight
derivatives).

Grids cross this interface as their OWNED nodes only: a P2G output is the
folded owned node array (Cabana's p2g followed by Halo::scatter(Sum),
Cabana_Grid_Interpolation.hpp:1217-1268), and a G2P input is an owned node
array whose ghosts are filled from the periodic owners before the gather
(Halo::gather then the point loop, :831-881).  Shapes are
(n_i, n_j, n_k, dofs) over LocalGrid.indexSpace(Own(), Node()).

Operators are named by the reference factories without the create/P2G/G2P
affixes: "ScalarValue", "VectorValue", "ScalarGradient", "VectorDivergence",
"TensorDivergence" (P2G) and "ScalarValue", "VectorValue", "ScalarGradient",
"VectorGradient", "VectorDivergence" (G2P).
"""
import torch

from . import grid as G

_P2G = {
    "ScalarValue": (G.ScalarValueP2G, 1),
    "VectorValue": (G.VectorValueP2G, 3),
    "ScalarGradient": (G.ScalarGradientP2G, 3),
    "VectorDivergence": (G.VectorDivergenceP2G, 1),
    "TensorDivergence": (G.TensorDivergenceP2G, 3),
}
# G2P operator -> (functor class, grid dofs read, particle output shape)
_G2P = {
    "ScalarValue": (G.ScalarValueG2P, 1, ()),
    "VectorValue": (G.VectorValueG2P, 3, (3,)),
    "ScalarGradient": (G.ScalarGradientG2P, 1, (3,)),
    "VectorGradient": (G.VectorGradientG2P, 3, (3, 3)),
    "VectorDivergence": (G.VectorDivergenceG2P, 3, ()),
}


def _own_slices(local_grid):
    s = local_grid.indexSpace(G.Own(), G.Node())
    return tuple(slice(s.min(d), s.max(d)) for d in range(3))


def _new_array(local_grid, dofs, dtype, device):
    layout = G.createArrayLayout(local_grid, dofs, G.Node())
    return G.createArray("autograd", layout, dtype=dtype, device=device)


class _Spec:
    """Non-tensor arguments of one transfer (held by the autograd node)."""

    def __init__(self, local_grid, ops, multipliers, order, bins, halo_width):
        self.local_grid = local_grid
        self.ops = list(ops)
        self.multipliers = [1.0] * len(self.ops) if multipliers is None \
            else [float(m) for m in multipliers]
        if len(self.multipliers) != len(self.ops):
            raise ValueError("one multiplier per operator")
        self.spline = G.Spline(order)
        self.bins = bins
        self.halo_width = halo_width


class P2GFunction(torch.autograd.Function):
    """(points, values...) -> owned node grids, one per operator."""

    @staticmethod
    def forward(ctx, spec, points, *values):
        lg = spec.local_grid
        own = _own_slices(lg)
        dtype = values[0].dtype
        arrays, functors = [], []
        for op, mult, v in zip(spec.ops, spec.multipliers, values):
            cls, dofs = _P2G[op]
            arrays.append(_new_array(lg, dofs, dtype, points.device))
            functors.append(cls(v.detach(), mult))
        halo = G.createHalo(G.NodeHaloPattern(), spec.halo_width, *arrays)
        G.p2g(functors, points.detach(), points.shape[0], spec.spline, halo, arrays,
              bins=spec.bins)
        ctx.spec = spec
        ctx.save_for_backward(points, *values)
        return tuple(a.view()[own] for a in arrays)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, *grad_grids):
        spec = ctx.spec
        points, *values = ctx.saved_tensors
        lg = spec.local_grid
        own = _own_slices(lg)
        need_x = ctx.needs_input_grad[1]
        need_v = ctx.needs_input_grad[2:]
        terms, gbars, vbars = [], [], []
        for i, (op, mult, v) in enumerate(zip(spec.ops, spec.multipliers, values)):
            if grad_grids[i] is None or not (need_x or need_v[i]):
                continue
            cls, dofs = _P2G[op]
            gb = _new_array(lg, dofs, v.dtype, points.device)
            gb.view()[own] = grad_grids[i]
            terms.append((i, cls(v.detach(), mult)))
            gbars.append(gb)
            vbars.append(torch.zeros_like(v) if need_v[i] else None)
        xbar = torch.zeros_like(points) if need_x else None
        if terms:
            halo = G.createHalo(G.NodeHaloPattern(), spec.halo_width, *gbars)
            G.p2g_adjoint([f for _, f in terms], points.detach(), points.shape[0], spec.spline,
                          halo, gbars, vbars, points_bar=xbar, bins=spec.bins)
        vgrads = [None] * len(values)
        for (i, _), vb in zip(terms, vbars):
            vgrads[i] = vb
        return (None, xbar, *vgrads)


class G2PFunction(torch.autograd.Function):
    """(points, grids...) -> particle fields, one per operator.  The same
    grid tensor may feed several operators; it is then gathered once."""

    @staticmethod
    def forward(ctx, spec, points, *grids):
        lg = spec.local_grid
        own = _own_slices(lg)
        n = points.shape[0]
        # One ghosted array per distinct input tensor.
        distinct, slot_of = [], []
        for g in grids:
            for j, h in enumerate(distinct):
                if h is g:
                    slot_of.append(j)
                    break
            else:
                slot_of.append(len(distinct))
                distinct.append(g)
        arrays = []
        for g in distinct:
            a = _new_array(lg, g.shape[-1], g.dtype, points.device)
            a.view()[own] = g.detach()
            arrays.append(a)
        outs, functors = [], []
        for op, mult, j in zip(spec.ops, spec.multipliers, slot_of):
            cls, dofs, shape = _G2P[op]
            if arrays[j].view().shape[-1] != dofs:
                raise ValueError("%sG2P reads a %d-dof grid" % (op, dofs))
            out = torch.zeros((n,) + shape, dtype=arrays[j].view().dtype, device=points.device)
            outs.append(out)
            functors.append(cls(out, mult))
        halo = G.createHalo(G.NodeHaloPattern(), spec.halo_width, *arrays)
        G.g2p([arrays[j] for j in slot_of], halo, points.detach(), n, spec.spline, functors,
              bins=spec.bins)
        ctx.spec = spec
        ctx.slot_of = slot_of
        ctx.save_for_backward(points, *distinct)
        return tuple(outs)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, *grad_outs):
        spec = ctx.spec
        slot_of = ctx.slot_of
        points, *distinct = ctx.saved_tensors
        lg = spec.local_grid
        own = _own_slices(lg)
        n = points.shape[0]
        need_x = ctx.needs_input_grad[1]
        need_g = ctx.needs_input_grad[2:]
        arrays = []
        for g in distinct:
            a = _new_array(lg, g.shape[-1], g.dtype, points.device)
            a.view()[own] = g.detach()
            arrays.append(a)
        gbars = [_new_array(lg, g.shape[-1], g.dtype, points.device) for g in distinct]
        want_g = [False] * len(distinct)
        for i, j in enumerate(slot_of):
            want_g[j] = want_g[j] or need_g[i]
        terms, grids, vbars, gb_terms = [], [], [], []
        for i, (op, mult, j) in enumerate(zip(spec.ops, spec.multipliers, slot_of)):
            if grad_outs[i] is None or not (need_x or want_g[j]):
                continue
            cls, _, shape = _G2P[op]
            vb = grad_outs[i].contiguous()
            terms.append(cls(vb, mult))
            grids.append(arrays[j])
            vbars.append(vb)
            gb_terms.append(gbars[j] if want_g[j] else None)
        xbar = torch.zeros_like(points) if need_x else None
        if terms:
            halo = G.createHalo(G.NodeHaloPattern(), spec.halo_width, *arrays)
            G.g2p_adjoint(grids, halo, points.detach(), n, spec.spline, terms, vbars, gb_terms,
                          points_bar=xbar, bins=spec.bins)
        ggrads = [None] * len(slot_of)
        seen = set()
        for i, j in enumerate(slot_of):
            if want_g[j] and j not in seen and need_g[i]:
                ggrads[i] = gbars[j].view()[own]
                seen.add(j)
        return (None, xbar, *ggrads)


def p2g(points, values, local_grid, ops, multipliers=None, order=2, bins=None, halo_width=-1):
    """Differentiable fused P2G.  ``values`` and ``ops`` are lists (or single
    items); returns one owned node grid per operator."""
    single = not isinstance(ops, (list, tuple))
    ops = [ops] if single else list(ops)
    values = [values] if single else list(values)
    spec = _Spec(local_grid, ops, multipliers, order, bins, halo_width)
    out = P2GFunction.apply(spec, points, *values)
    return out[0] if single else out


def g2p(points, grids, local_grid, ops, multipliers=None, order=2, bins=None, halo_width=-1):
    """Differentiable fused G2P.  ``grids`` are owned node arrays (one per
    operator; pass the same tensor to share it); returns one particle field
    per operator (fresh tensors: result = multiplier * interpolation)."""
    single = not isinstance(ops, (list, tuple))
    ops = [ops] if single else list(ops)
    grids = [grids] if single else list(grids)
    spec = _Spec(local_grid, ops, multipliers, order, bins, halo_width)
    out = G2PFunction.apply(spec, points, *grids)
    return out[0] if single else out
