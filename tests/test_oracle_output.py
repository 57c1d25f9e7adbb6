"""Pins the CPU oracle of the 0d reductions and the 1d point lists of BAM's
output: constant fields (max = min = c, both norms |c|, integral c N dx^3),
random fields against numpy over the inner points (the l2 norm of each
variable is sqrt(mean(u^2)) once the count over all variables is undone), and
the 1d lines of cell-centred boxes: the origin snaps to the next grid point,
the points are the grid points of the line, and points no box holds (a gap
between two boxes) are reported as missing."""
import numpy as np

import oracle as O


def _boxes():
    h = 0.25
    a = O.make_box(12, 10, 8, -1.375, -1.125, -0.875, h)
    b = O.make_box(6, 10, 8, 2.125, -1.125, -0.875, h)     # gap of 2 points in x after box a
    return [a, b], h


def _inner(box, u):
    s = box.skip
    return u.reshape(box.o, box.n, box.m)[s[4]:box.o - s[5], s[2]:box.n - s[3], s[0]:box.m - s[1]]


def test_reductions_constant_and_random():
    boxes, h = _boxes()
    c = -0.37
    const = [[np.full(b.m * b.n * b.o, c) for b in boxes]]
    ninner = sum(_inner(b, f).size for b, f in zip(boxes, const[0]))
    assert O.reduce_0d(boxes, const, "max")[0] == c and O.reduce_0d(boxes, const, "min")[0] == c
    assert abs(O.reduce_0d(boxes, const, "norm")[0] - abs(c)) < 1e-14
    assert O.reduce_0d(boxes, const, "norminf")[0] == abs(c)
    assert abs(O.reduce_0d(boxes, const, "integral")[0] - c * ninner * h ** 3) < 1e-13
    rng = np.random.default_rng(1)
    fields = [[rng.standard_normal(b.m * b.n * b.o) for b in boxes] for _ in range(3)]
    for v, f in enumerate(fields):
        inner = np.concatenate([_inner(b, u).reshape(-1) for b, u in zip(boxes, f)])
        assert O.reduce_0d(boxes, fields, "max")[v] == inner.max()
        assert O.reduce_0d(boxes, fields, "min")[v] == inner.min()
        assert O.reduce_0d(boxes, fields, "norminf")[v] == np.abs(inner).max()
        np.testing.assert_allclose(O.reduce_0d(boxes, fields, "norm")[v], np.sqrt(np.mean(inner ** 2)), rtol=1e-14)
        np.testing.assert_allclose(O.reduce_0d(boxes, fields, "integral")[v], inner.sum() * h ** 3, rtol=1e-12)
    # boundary points do not count
    f = [[np.zeros(b.m * b.n * b.o) for b in boxes]]
    f[0][0][0] = 100.0                                  # a corner point of box a
    assert O.reduce_0d(boxes, f, "max")[0] == 0.0


def test_lines_snap_to_grid_points_and_mark_gaps():
    h = 0.25
    box = O.make_box(16, 16, 16, -1.875, -1.875, -1.875, h)   # cell centred around the origin
    crd, bo, ix = O.line_1d([box], "x")
    np.testing.assert_allclose(crd, -1.875 + h * np.arange(16), rtol=0, atol=1e-12)
    assert np.all(bo == 0)
    np.testing.assert_array_equal(ix, np.arange(16) + 16 * 8 + 256 * 8)   # y = z = +h/2: j = k = 8
    crd, bo, ix = O.line_1d([box], "z", origin=(0.3, -0.3, 0.0))
    # x snaps to 0.375 (i = 9), y to -0.125 (j = 7)
    np.testing.assert_array_equal(ix, 9 + 16 * 7 + 256 * np.arange(16))
    boxes, h = _boxes()
    crd, bo, ix = O.line_1d(boxes, "x")
    np.testing.assert_allclose(crd, -1.375 + h * np.arange(20), rtol=0, atol=1e-12)
    np.testing.assert_array_equal(bo, [0] * 12 + [-1] * 2 + [1] * 6)
    assert np.all(ix[12:14] == -1)
