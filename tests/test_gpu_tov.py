"""A TOV star on a level (bamhip_tov_set, set_TOV of src/projects/
BBIDdataReader/BBIDdataReader.c:483-690).  With a table whose Schwarzschild
radius, potential and enthalpy are cubics of the (non-uniform) isotropic
radius, the 4-point Lagrange read of interp_lag4 is exact, so every interior
point has a closed form; outside the surface the exterior Schwarzschild
solution.  Then the default star of tov_h on two boxes: central density,
continuity of the metric at the surface, matter only inside."""
import numpy as np
import pytest

import baum_amd as B

pytestmark = pytest.mark.gpu

G6, K6, ALPHA, PRIM = list(range(6)), list(range(6, 12)), 12, list(range(13, 19))
NVAR = 19
CENTER = (0.1, -0.2, 0.05)


def level(boxes):
    lv = B.Level(boxes, NVAR)
    lv.enable(range(NVAR))
    for v in range(NVAR):
        lv.fill(v, 7.0)
    return lv


def radius(lv):
    out = []
    for b in lv.boxes:
        Z, Y, X = np.meshgrid(b.z0 + b.dz * np.arange(b.o), b.y0 + b.dy * np.arange(b.n),
                              b.x0 + b.dx * np.arange(b.m), indexing="ij")
        out.append(np.sqrt((X - CENTER[0]) ** 2 + (Y - CENTER[1]) ** 2 + (Z - CENTER[2]) ** 2).ravel())
    return np.concatenate(out)


def eos_h(h, K, G):
    n = ((G - 1) / (G * K) * np.expm1(np.maximum(h, 0))) ** (1 / (G - 1))
    return n, K * n ** (G - 1) / (G - 1), K * n ** G


def test_cubic_table_is_read_exactly_and_exterior_is_schwarzschild():
    x = 6.0 * np.linspace(0, 1, 300) ** 1.3
    rs = lambda s: 0.02 + 0.9 * s + 0.01 * s * s - 0.001 * s ** 3
    ph = lambda s: -0.3 + 0.01 * s * s - 0.0005 * s ** 3
    hh = lambda s: 0.2 - 0.005 * s * s + 0.0001 * s ** 3
    star = dict(M=1.0, K=100.0, Gamma=2.0)
    lv = level([B.make_box(32, 30, 28, -7.75, -7.25, -6.75, 0.5)])
    lv.tov_set(star, G6, K6, ALPHA, PRIM, center=CENTER, riso=x, rsch=rs(x), phi=ph(x), h=hh(x))
    r = radius(lv)
    g = lv.get_state(G6)
    alpha = lv.get(ALPHA)
    prim = lv.get_state(PRIM)
    kij = lv.get_state(K6)
    lv.free()
    inside = r < x[-1]
    assert 0 < inside.sum() < r.size
    rho, eps, p = eos_h(hh(r), 100.0, 2.0)
    psi4_in = (rs(r) / r) ** 2
    psi4_out = (1 + 0.5 / r) ** 4
    a_out = (1 - 0.5 / r) / (1 + 0.5 / r)
    psi4 = np.where(inside, psi4_in, psi4_out)
    np.testing.assert_allclose(g[0], psi4, rtol=1e-12)
    np.testing.assert_allclose(g[3], psi4, rtol=1e-12)
    np.testing.assert_allclose(g[5], psi4, rtol=1e-12)
    np.testing.assert_array_equal(g[[1, 2, 4]], 0.0)
    np.testing.assert_array_equal(kij, 0.0)
    np.testing.assert_allclose(alpha, np.where(inside, np.exp(ph(r)), a_out), rtol=1e-12)
    np.testing.assert_allclose(prim[0], np.where(inside, rho, 0.0), rtol=1e-11, atol=1e-300)
    np.testing.assert_allclose(prim[1], np.where(inside, eps, 0.0), rtol=1e-11, atol=1e-300)
    np.testing.assert_allclose(prim[5], np.where(inside, p, 0.0), rtol=1e-11, atol=1e-300)
    np.testing.assert_array_equal(prim[2:5], 0.0)


def test_default_star_on_two_boxes():
    star = B.tov_h(0.227869, 100.0, 2.0, 100000)
    boxes = [B.make_box(24, 48, 48, -11.75, -11.75, -11.75, 0.5), B.make_box(24, 48, 48, 0.25, -11.75, -11.75, 0.5)]
    lv = level(boxes)
    lv.tov_set(star, G6, K6, ALPHA, PRIM, center=CENTER)
    r = radius(lv)
    g = lv.get_state(G6)
    alpha = lv.get(ALPHA)
    prim = lv.get_state(PRIM)
    lv.free()
    R, M = star["Riso"], star["M"]
    inside = r < R
    assert np.all(prim[0][~inside] == 0) and np.all(prim[0][inside] > 0)
    c = np.argmin(r)
    assert abs(prim[0][c] - 1.28e-3) < 2e-5                      # r = 0.17 from the centre
    # metric and lapse continuous across the surface: compare the last points inside with the exterior formula
    band = inside & (r > R - 0.3)
    np.testing.assert_allclose(g[0][band], (1 + 0.5 * M / r[band]) ** 4, rtol=2e-3)
    np.testing.assert_allclose(alpha[band], (1 - 0.5 * M / r[band]) / (1 + 0.5 * M / r[band]), rtol=2e-3)
    assert alpha.min() > 0.5 and alpha.max() < 1 and g[0].max() < 4
    np.testing.assert_allclose(prim[5], 100.0 * prim[0] ** 2, rtol=1e-12, atol=1e-300)
