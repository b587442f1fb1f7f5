"""Textbook cardinal B-splines, written independently of the oracle, for
pinning it: the centred B-spline of degree p by the truncated-power formula

    B_p(t) = 1/p! * sum_{k=0}^{p+1} (-1)^k C(p+1, k) (t + (p+1)/2 - k)_+^p

with B_p' (t) = B_{p-1}(t + 1/2) - B_{p-1}(t - 1/2) and
B_p''(t) = B_{p-2}(t + 1) - 2 B_{p-2}(t) + B_{p-2}(t - 1).  A particle at
logical node coordinate x (node units) gives node n the weight B_p(x - n); on a
periodic mesh the folded weight of owned node o sums the periodic images.
"""
import math

import numpy as np


def bspline(p, t):
    t = np.asarray(t, dtype=np.float64)
    if p == 0:
        return ((t >= -0.5) & (t < 0.5)).astype(np.float64)
    acc = np.zeros_like(t)
    for k in range(p + 2):
        u = t + (p + 1) / 2.0 - k
        acc += (-1) ** k * math.comb(p + 1, k) * np.where(u > 0, u, 0.0) ** p
    # B_p vanishes for |t| >= (p + 1) / 2; there the alternating sum of large
    # powers would only leave its rounding error.
    return np.where(np.abs(t) < (p + 1) / 2.0, acc / math.factorial(p), 0.0)


def dbspline(p, t):
    if p == 0:
        return np.zeros_like(np.asarray(t, dtype=np.float64))
    return bspline(p - 1, np.asarray(t) + 0.5) - bspline(p - 1, np.asarray(t) - 0.5)


def ddbspline(p, t):
    if p < 2:
        return np.zeros_like(np.asarray(t, dtype=np.float64))
    t = np.asarray(t)
    return bspline(p - 2, t + 1.0) - 2.0 * bspline(p - 2, t) + bspline(p - 2, t - 1.0)


def folded_factors(order, x_node, num_cell, halo, rdx, deriv=0):
    """(np, N) matrix of 1-D folded weights (deriv 0), gradients (1) or second
    derivatives (2) of particles at logical node coordinates x_node onto the
    owned nodes halo .. halo + num_cell - 1 of a periodic dimension."""
    fn = {0: bspline, 1: dbspline, 2: ddbspline}[deriv]
    scale = rdx ** deriv
    nodes = halo + np.arange(num_cell)
    out = np.zeros((len(x_node), num_cell))
    for img in (-2, -1, 0, 1, 2):
        out += fn(order, x_node[:, None] - (nodes[None, :] + img * num_cell)) * scale
    return out


def dense_weights(order, x_node3, num_cell, halo, rdx, derivs):
    """(np, Nx*Ny*Nz) folded weight matrix with derivs[d] in {0, 1, 2} per
    dimension (owned nodes in (i, j, k) row-major order)."""
    f = [folded_factors(order, x_node3[:, d], num_cell[d], halo, rdx[d], derivs[d])
         for d in range(3)]
    w = f[0][:, :, None, None] * f[1][:, None, :, None] * f[2][:, None, None, :]
    return w.reshape(len(x_node3), -1)
