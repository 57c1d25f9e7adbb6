"""Pins the CPU oracle of the puncture initial data and of the ADM -> Z4c
conversion (z4_init with z4_setbound_G) against closed forms:
- Brill-Lindquist: psi = 1 + sum m_i / (2 r_i), g_ij = psi^4 delta_ij, K_ij = 0,
  and z4_init of it: chi = psi^-4, conformal metric delta, Khat = A_ij = 0;
- a conformal metric whose inverse is quadratic, diag(1 + c x^2, 1 + d y^2,
  1 / ((1 + c x^2)(1 + d y^2))) (unit determinant), times a smooth psi^4:
  the 2nd-order stencil is exact, so G^i = -d_j gt^{ij} = (-2 c x, -2 d y, 0)
  at the inner points, and the boundary extrapolation, being linear, must
  reproduce this linear field on faces, edges and corners."""
import numpy as np

import oracle as O


def _grid(box):
    x = box.x0 + box.dx * np.arange(box.m)
    y = box.y0 + box.dy * np.arange(box.n)
    z = box.z0 + box.dz * np.arange(box.o)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def test_brill_lindquist_and_z4_init():
    box = O.make_box(14, 12, 10, -1.625, -1.375, -1.125, 0.25)
    masses, pos = [0.5, 0.3], [[-0.6, 0.1, 0.0], [0.7, -0.2, 0.1]]
    g, k, psi = O.punctures_bl(box, masses, pos)
    X, Y, Z = _grid(box)
    want = 1 + sum(m / (2 * np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)) for m, c in zip(masses, pos))
    np.testing.assert_allclose(psi, want, rtol=1e-15)
    np.testing.assert_allclose(g[0], want ** 4, rtol=1e-14)
    assert np.all(g[[1, 2, 4]] == 0) and np.all(k == 0)
    np.testing.assert_array_equal(g[0], g[3])
    u = O.adm_to_z4c(box, g, k, phybound=True)
    np.testing.assert_allclose(u[6], want ** -4, rtol=1e-13)          # chi
    np.testing.assert_allclose(u[[0, 3, 5]], 1.0, rtol=0, atol=1e-14)  # conformal metric
    assert np.abs(u[[1, 2, 4]]).max() < 1e-14
    assert np.abs(u[7:14]).max() == 0.0 and np.all(u[17] == 0)


def test_gamma_exact_for_quadratic_inverse_metric_with_extrapolated_boundary():
    box = O.make_box(12, 11, 10, -0.55, -0.5, -0.45, 0.1)
    X, Y, Z = _grid(box)
    c, d = 0.4, -0.3
    psi4 = (1 + 0.1 * np.sin(X + 2 * Y) * np.cos(Z)) ** 4
    g = np.zeros((6,) + X.shape)
    g[0] = psi4 / (1 + c * X * X)
    g[3] = psi4 / (1 + d * Y * Y)
    g[5] = psi4 * (1 + c * X * X) * (1 + d * Y * Y)
    k = np.zeros_like(g)
    u = O.adm_to_z4c(box, g, k, phybound=True)
    want = [-2 * c * X, -2 * d * Y, 0 * X]
    for a in range(3):
        np.testing.assert_allclose(u[14 + a][1:-1, 1:-1, 1:-1], want[a][1:-1, 1:-1, 1:-1], rtol=0, atol=1e-12)
        np.testing.assert_allclose(u[14 + a], want[a], rtol=0, atol=1e-11)   # faces, edges, corners too
    # without the extrapolation the outer layer keeps no value
    u0 = O.adm_to_z4c(box, g, k, phybound=False)
    assert np.all(u0[14][0] == 0) and np.abs(u0[14][1:-1, 1:-1, 1:-1] - u[14][1:-1, 1:-1, 1:-1]).max() == 0
