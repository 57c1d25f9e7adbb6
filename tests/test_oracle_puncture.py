"""Pins the CPU oracle of BAM's puncture tracker (point interpolation and the
three-step ICN) against what the mathematics fixes: Lagrange interpolation
of `order` points per direction is exact for tensor polynomials of degree
order - 1, a cube that no single box holds gives 0 (interpolate.c:458-461),
and the ICN iteration has closed forms for constant and linear shifts."""
import numpy as np

import oracle as O


def _level():
    h = 0.25
    b0 = O.make_box(20, 18, 16, -2.375, -2.125, -1.875, h)
    b1 = O.make_box(12, 14, 10, -2.375 + 20 * h, -2.125, -1.875, h)   # the next 12 points in x
    return [b0, b1], h


def _coords(b):
    x = b.x0 + b.dx * np.arange(b.m)
    y = b.y0 + b.dy * np.arange(b.n)
    z = b.z0 + b.dz * np.arange(b.o)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def _poly(X, Y, Z):
    return 0.3 - 0.2 * X + 0.7 * X ** 3 + 0.1 * Y ** 2 * Z - 0.05 * X * Y * Z + 0.4 * Z ** 3 + 0.02 * X ** 2 * Y


def test_interpolation_exact_for_cubic_tensor_polynomials():
    boxes, h = _level()
    var = [_poly(*_coords(b)).reshape(-1) for b in boxes]
    rng = np.random.default_rng(3)
    hits = 0
    for _ in range(40):
        x = rng.uniform(-2.0, 4.5)
        y = rng.uniform(-1.6, 1.6)
        z = rng.uniform(-1.4, 1.4)
        v, found = O.interp_point(boxes, var, x, y, z, order=4)
        if found:
            hits += 1
            assert abs(v - _poly(x, y, z)) < 1e-12, (x, y, z)
        else:
            assert v == 0.0
    assert hits >= 20


def test_cube_across_two_boxes_is_not_found():
    boxes, h = _level()
    var = [np.ones(b.m * b.n * b.o) for b in boxes]
    # x between the last point of box 0 and the first of box 1: the cube spans both
    x = boxes[0].x0 + (boxes[0].m - 1) * h + 0.4 * h
    v, found = O.interp_point(boxes, var, x, 0.01, 0.02, order=4)
    assert not found and v == 0.0
    v, found = O.interp_point(boxes, var, x - 3 * h, 0.01, 0.02, order=4)
    assert found and abs(v - 1.0) < 1e-14


def test_icn_constant_and_linear_shift():
    boxes, h = _level()
    dt = 0.1
    p = np.array([0.37, -0.21, 0.13])
    # constant shift: x_new = x - dt beta exactly
    beta = [[np.full(b.m * b.n * b.o, c) for b in boxes] for c in (0.3, -0.2, 0.1)]
    got = O.track_puncture(boxes, beta, beta, dt, p)
    np.testing.assert_allclose(got, p - dt * np.array([0.3, -0.2, 0.1]), rtol=0, atol=1e-15)
    # linear shift beta^i = a_i + b_i x^i, previous level with other constants:
    # three fixed-point iterations of x0 = x1 - dt (beta(x0) + beta_p(x1)) / 2
    a, bb = np.array([0.2, -0.1, 0.05]), np.array([0.5, -0.3, 0.8])
    ap, bp = np.array([0.1, 0.0, -0.05]), np.array([0.4, 0.2, 0.6])
    crd = [_coords(b) for b in boxes]
    beta = [[(a[i] + bb[i] * crd[k][i]).reshape(-1) for k in range(2)] for i in range(3)]
    beta_p = [[(ap[i] + bp[i] * crd[k][i]).reshape(-1) for k in range(2)] for i in range(3)]
    got = O.track_puncture(boxes, beta, beta_p, dt, p)
    x1 = p.copy()
    b1 = ap + bp * x1
    x0 = x1.copy()
    for _ in range(3):
        x0 = x1 - dt * 0.5 * ((a + bb * x0) + b1)
    np.testing.assert_allclose(got, x0, rtol=0, atol=1e-14)
    # two iterations would differ: the count is pinned
    x2 = x1.copy()
    for _ in range(2):
        x2 = x1 - dt * 0.5 * ((a + bb * x2) + b1)
    assert np.abs(x2 - got).max() > 1e-6


def test_pin_z_and_step_out_of_the_level():
    """track_puncture around the integrator (moving_punctures.c:174-185): with
    moving_puncture_fixz BH the z coordinate stays; a new position in no box's
    bounding box is dropped for the old one (constant shift: closed form)"""
    boxes, h = _level()
    dt = 0.1
    p = np.array([0.37, -0.21, 0.13])
    beta = [[np.full(b.m * b.n * b.o, c) for b in boxes] for c in (0.3, -0.2, 0.1)]
    got = O.track_puncture(boxes, beta, beta, dt, p, pin_z=True)
    np.testing.assert_allclose(got, [p[0] - dt * 0.3, p[1] + dt * 0.2, p[2]], rtol=0, atol=1e-15)
    assert got[2] == p[2]
    # the level spans x in [-2.375, 5.375]; where no box holds the cube the
    # shift is 0, so with dt = 4 from x = -1.9 the ICN trial points are -3.1,
    # then -1.9 - dt * 0.3 / 2 = -2.5 twice: outside, the old position stays
    out = O.track_puncture(boxes, beta, beta, 4.0, np.array([-1.9, 0.1, 0.0]))
    np.testing.assert_array_equal(out, [-1.9, 0.1, 0.0])
    # from x = -1.2 with dt = 2 every trial point keeps its cube: x - dt beta
    inside = O.track_puncture(boxes, beta, beta, 2.0, np.array([-1.2, 0.1, 0.0]))
    np.testing.assert_allclose(inside, [-1.2 - 0.6, 0.1 + 0.4, -0.2], rtol=0, atol=1e-14)
