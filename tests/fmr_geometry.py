"""Nested cell-centred box geometry for FMR tests (BAM's 2:1 staggering).

A coarse point sits at X0 + I*H; its children at (I -/+ 1/4)*H, so a fine
box starting at coarse cell c0 has x0 = X0 + (c0 - 1/4)*H and h = H/2.
"""
import numpy as np


def coarse_box(mk, m, n, o, H, x0=None):
    """coarse box centred on the origin unless x0 is given"""
    if x0 is None:
        x0 = (-(m - 1) / 2 * H, -(n - 1) / 2 * H, -(o - 1) / 2 * H)
    return mk(m, n, o, x0[0], x0[1], x0[2], H)


def fine_box(mk, cb, c0, size):
    """fine box of size (m, n, o) whose first point is the left child of
    coarse point c0 of box cb"""
    H = cb.dx
    x0 = [cb.x0 + (c0[0] - 0.25) * H, cb.y0 + (c0[1] - 0.25) * H, cb.z0 + (c0[2] - 0.25) * H]
    return mk(size[0], size[1], size[2], x0[0], x0[1], x0[2], H / 2)


def coords(b):
    x = b.x0 + np.arange(b.m) * b.dx
    y = b.y0 + np.arange(b.n) * b.dy
    z = b.z0 + np.arange(b.o) * b.dz
    return x, y, z


def lagrange_weights(n, x):
    w = np.ones(n)
    for i in range(n):
        for j in range(n):
            if j != i:
                w[i] *= (x - j) / (i - j)
    return w


def centred_window(xi, n):
    """first index of the n coarse points centred on coarse coordinate xi and
    the Lagrange evaluation point inside that window"""
    s = int(np.floor(xi - (n - 1) / 2 + 0.5))
    return s, xi - s


def small_hierarchy(mk0, mkf):
    """three levels around a Brill-Lindquist binary at x = -3, +3: a 32^3
    coarse box (H = 0.5, made by mk0), two uneven boxes on level 1 and one
    box inside each of them on level 2 (made by mkf).  Returns (boxes per
    level, parent index of every box per level)."""
    L0 = [mk0(32, 32, 32, -7.75, -7.75, -7.75, 0.5)]
    a1 = fine_box(mkf, L0[0], (4, 12, 12), (16, 16, 16))
    b1 = fine_box(mkf, L0[0], (18, 12, 11), (20, 16, 18))
    a2 = fine_box(mkf, a1, (4, 4, 4), (16, 16, 16))
    b2 = fine_box(mkf, b1, (6, 5, 5), (16, 12, 14))
    return [L0, [a1, b1], [a2, b2]], [[0], [0, 0], [0, 1]]
