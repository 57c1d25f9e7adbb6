"""Pins the CPU oracle of BAM's GRHD source terms (grhd_sources) against what
the mathematics fixes.  The oracle follows the reference's T^{mu nu} form
(T^00, T^0i, T^ij, T^0_j contracted with K, d alpha, d beta, d g).  Expanding
T^{mu nu} in the Eulerian energy E, momentum S^i and stress S^ij gives the
3+1 form, written here independently:
    s_tau = sqrt(g) (alpha S^ij K_ij - S^i d_i alpha)
    s_S_c = sqrt(g) (-E d_c alpha + S_k d_c beta^k + alpha S^ij d_c g_ij / 2).
With quadratic metric fields the centred stencils are exact, so both forms
must agree to round-off at every inner point; a quartic lapse separates the
4th-order interior from the 2nd-order points next to a face.  Also: the trace
of the matter stress, and the atmosphere, cold-EOS, non-finite and c rules."""
import numpy as np

import oracle as O

S6 = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]


def _grid(box):
    x = box.x0 + box.dx * np.arange(box.m)
    y = box.y0 + box.dy * np.arange(box.n)
    z = box.z0 + box.dz * np.arange(box.o)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def _quadratic_metric(X, Y, Z):
    """alpha, beta^i, g_ij quadratic in x, y, z, and their exact gradients"""
    coef = np.random.default_rng(1).uniform(-0.02, 0.02, (10, 10))
    base = np.array([0.9, 0.03, -0.02, 0.01, 1.1, 0.02, -0.01, 1.05, 0.015, 1.2])
    mons = [np.ones_like(X), X, Y, Z, X * X, Y * Y, Z * Z, X * Y, X * Z, Y * Z]
    dmons = [[0 * X, 1 + 0 * X, 0 * X, 0 * X, 2 * X, 0 * X, 0 * X, Y, Z, 0 * X],
             [0 * X, 0 * X, 1 + 0 * X, 0 * X, 0 * X, 2 * Y, 0 * X, X, 0 * X, Z],
             [0 * X, 0 * X, 0 * X, 1 + 0 * X, 0 * X, 0 * X, 2 * Z, 0 * X, X, Y]]
    f = np.array([base[v] + sum(coef[v, k] * mons[k] for k in range(1, 10)) for v in range(10)])
    df = np.array([[sum(coef[v, k] * dmons[d][k] for k in range(1, 10)) for v in range(10)] for d in range(3)])
    return f, df


def _prims(shape, seed=2):
    rng = np.random.default_rng(seed)
    rho = 1e-3 * (1 + rng.random(shape))
    eps = 0.1 + 0.2 * rng.random(shape)
    v = 0.3 * (rng.random((3,) + shape) - 0.5)
    p = rho * eps                                # ideal gas, Gamma = 2
    return np.array([rho, eps, v[0], v[1], v[2], p])


def _sym(a6):
    return np.array([[a6[S6[a][b]] for b in range(3)] for a in range(3)])


def test_sources_equal_the_3p1_valencia_form():
    box = O.make_box(12, 11, 10, -0.6, -0.5, -0.45, 0.1)
    X, Y, Z = _grid(box)
    met, dmet = _quadratic_metric(X, Y, Z)
    K = np.random.default_rng(3).uniform(-0.1, 0.1, (6,) + X.shape)
    prim = _prims(X.shape)
    rhs = np.zeros((5,) + X.shape)
    madm = np.zeros((11,) + X.shape)
    O.grhd_sources(box, O.hrsc_params(), met, K, prim, rhs, madm)
    n_checked = 0
    for k in range(1, box.o - 1):
        for j in range(1, box.n - 1):
            for i in range(1, box.m - 1):
                q = (slice(None), k, j, i)
                al, be, g = met[0][k, j, i], met[1:4][q[:1] + q[1:]], _sym(met[4:10][q])
                gi = np.linalg.inv(g)
                rho, eps, vu, p = prim[0][k, j, i], prim[1][k, j, i], prim[2:5][q], prim[5][k, j, i]
                vl = g @ vu
                W = 1 / np.sqrt(1 - vu @ vl)
                hrho = W * W * (rho * (1 + eps) + p)
                E, Su, Sl = hrho - p, hrho * vu, hrho * vl
                Sij = hrho * np.outer(vu, vu) + p * gi
                Kq = _sym(K[q])
                dal = dmet[:, 0, k, j, i]
                dbe = dmet[:, 1:4, k, j, i]
                dg = [_sym(dmet[d, 4:10, k, j, i]) for d in range(3)]
                sg = np.sqrt(np.linalg.det(g))
                s_tau = sg * (al * np.sum(Sij * Kq) - Su @ dal)
                s_S = [sg * (-E * dal[c] + Sl @ dbe[c] + 0.5 * al * np.sum(Sij * dg[c])) for c in range(3)]
                # tolerances relative to the size of the terms (they cancel in part)
                t_tau = sg * (al * np.sum(np.abs(Sij * Kq)) + np.abs(Su) @ np.abs(dal))
                t_S = [sg * (abs(E * dal[c]) + np.abs(Sl) @ np.abs(dbe[c]) + 0.5 * al * np.sum(np.abs(Sij * dg[c])))
                       for c in range(3)]
                assert rhs[0][k, j, i] == 0.0
                assert abs(rhs[1][k, j, i] - s_tau) <= 1e-12 * t_tau
                for c in range(3):
                    assert abs(rhs[2 + c][k, j, i] - s_S[c]) <= 1e-12 * t_S[c], (c, i, j, k)
                # matter ADM variables: E, S^i, S_ij, and S = g^ij S_ij
                assert abs(madm[0][k, j, i] - E) <= 1e-15 * E
                np.testing.assert_allclose(madm[1:4][q], Su, rtol=1e-14)
                Sdown = _sym(madm[4:10][q])
                np.testing.assert_allclose(Sdown, hrho * np.outer(vl, vl) + p * g, rtol=1e-14,
                                           atol=1e-14 * np.abs(Sdown).max())
                assert abs(madm[10][k, j, i] - np.sum(gi * Sdown)) <= 1e-13 * madm[10][k, j, i]
                n_checked += 1
    assert n_checked == (box.m - 2) * (box.n - 2) * (box.o - 2)
    # boundary layer untouched
    assert np.all(rhs[:, 0] == 0) and np.all(madm[:, :, :, -1] == 0)


def test_fourth_order_inside_second_order_next_to_faces():
    box = O.make_box(12, 9, 9, -0.55, -0.4, -0.4, 0.1)
    X, Y, Z = _grid(box)
    met = np.zeros((10,) + X.shape)
    met[0] = 0.9 + 3.0 * X ** 4
    met[4] = met[7] = met[9] = 1.0
    prim = np.zeros((6,) + X.shape)
    prim[0], prim[1] = 1e-3, 0.2
    prim[5] = prim[0] * prim[1]
    K = np.zeros((6,) + X.shape)
    for order_centered, exact_edge in ((4, False), (2, False)):
        rhs = np.zeros((5,) + X.shape)
        O.grhd_sources(box, O.hrsc_params(), met, K, prim, rhs, np.zeros((11,) + X.shape), order_centered=order_centered)
        want = -(prim[0] * (1 + prim[1])) * 12.0 * X ** 3          # -E d_x alpha, sqrt(g) = 1
        err = np.abs(rhs[2] - want)[1:-1, 1:-1, 1:-1]
        interior = err[1:-1, 1:-1, 1:-1][:, :, 1:-1]
        if order_centered == 4:
            assert interior.max() <= 1e-12 * np.abs(want).max()
        else:
            assert interior.max() > 1e-6 * np.abs(want).max()
        assert err[:, :, 0].max() > 1e-6 * np.abs(want).max()        # i = 1: second order
    # the gradient of the quartic at i = 1 is not exact at 2nd order, but is at i >= 2 with 4th


def test_atmosphere_cold_nonfinite_and_c_rules():
    box = O.make_box(8, 8, 8, -0.35, -0.35, -0.35, 0.1)
    X, Y, Z = _grid(box)
    met, _ = _quadratic_metric(X, Y, Z)
    K = np.random.default_rng(4).uniform(-0.1, 0.1, (6,) + X.shape)
    prim = _prims(X.shape, seed=5)
    base = np.random.default_rng(6).standard_normal((5,) + X.shape)
    madm = np.zeros((11,) + X.shape)
    src = O.grhd_sources(box, O.hrsc_params(), met, K, prim, np.zeros_like(base), madm)[0]
    got = O.grhd_sources(box, O.hrsc_params(), met, K, prim, base.copy(), madm, c=0.25)[0]
    np.testing.assert_allclose(got, base + 0.25 * src, rtol=1e-14, atol=1e-16)
    cold = O.grhd_sources(box, O.hrsc_params(eos=1), met, K, prim, base.copy(), madm)[0]
    np.testing.assert_array_equal(cold[1], base[1])
    atm = O.grhd_sources(box, O.hrsc_params(), met, K, prim, base.copy(), madm, rhoatm=1.5e-3)[0]
    below = prim[0] < 1.5e-3
    inner = np.zeros(X.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    for c in (2, 3, 4):
        np.testing.assert_array_equal(atm[c][below & inner], base[c][below & inner])
        assert np.all(atm[c][~below & inner] != base[c][~below & inner])
    bad = prim.copy()
    bad[0][4, 4, 4] = np.nan
    badmet = met.copy()
    badmet[0][3, 3, 3] = -0.5                       # alpha <= 0
    r = O.grhd_sources(box, O.hrsc_params(), badmet, K, bad, np.zeros_like(base), madm)[0]
    assert np.all(np.isfinite(r)) and np.all(np.isfinite(madm))
    assert np.all(r[:, 4, 4, 4] == 0) and np.all(madm[:, 4, 4, 4] == 0)
    assert np.all(r[:, 3, 3, 3] == 0) and np.all(madm[:, 3, 3, 3] == 0)
