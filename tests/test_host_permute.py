"""The re-binning permutation used by the MPM loop: forward is the gather
new[i] = old[perm[i]] (Cabana_Sort.hpp:600-656), backward its transpose.
Host-only: runs on CPU tensors."""
import torch

from cabana_amd.mpm import permute_tensor


def test_permute_forward_and_adjoint():
    g = torch.Generator().manual_seed(5)
    n = 257
    perm = torch.randperm(n, generator=g)
    x = torch.randn(n, 3, dtype=torch.float64, generator=g, requires_grad=True)
    y = permute_tensor(x, perm)
    assert torch.equal(y, x.detach()[perm])
    w = torch.randn(n, 3, dtype=torch.float64, generator=g)
    (y * w).sum().backward()
    # <P x, w> = <x, P^T w>: the gradient is w scattered back through perm.
    inv = torch.empty_like(perm)
    inv[perm] = torch.arange(n)
    assert torch.equal(x.grad, w[inv])


def test_permute_gradcheck():
    g = torch.Generator().manual_seed(6)
    perm = torch.randperm(11, generator=g)
    x = torch.randn(11, 2, dtype=torch.float64, generator=g, requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: permute_tensor(t, perm), (x,))
