"""Pins the CPU oracle's Z4c path against what the mathematics fixes.

The reference Z4c RHS cannot be compiled standalone (every file includes
bam.h -> the build-generated bam_automatic_include.h, and links MPI), and
the reference holds no Z4c golden data.  So the oracle is pinned by:

* an exact stationary solution (rotated, displaced Kerr-Schild Kerr): with
  its own lapse and shift, the RHS of the 18 non-gauge fields must vanish
  at 4th order; a dropped term, wrong sign or index leaves an O(1)
  residual that does not converge;
* the gauge RHS on the same data against closed forms (1+log, Gamma driver);
* the Hamiltonian relation rTheta = alpha * H / 2 on generic smooth data,
  against the independent ADM-constraint code path;
* flat space in polynomial gauge: finite differences are exact, so the RHS
  equals closed forms at every point, including the box-edge points where
  the reference lowers the stencil order;
* RK4 on the linear decay dB/dt = -eta B (Gamma driver with B):
  amplification 1 + z + z^2/2 + z^3/6 + z^4/24 exactly;
* algebraic constraints: det(gt) = 1 and tr(At) = 0 after projection;
* the radiative condition against an exact outgoing spherical wave.
"""
import numpy as np
import pytest

import oracle as O
from spacetimes import KerrSchild, flat_z4c, grid

INNER = (slice(3, -3),) * 3


def _ks_box(h, center=(2.5, 1.5, -2.0), half=0.4):
    n = int(round(2 * half / h)) + 1 + 6
    c = np.array(center)
    return O.make_box(n, n, n, *(c - (n - 1) / 2 * h), h, skip=(3,) * 6)


@pytest.fixture(scope="module")
def ks_runs():
    ks = KerrSchild()
    p = O.z4c_params(order_dissipation=0, z4_shiftdriver=0.0)
    out = {}
    for h in (0.08, 0.04):
        b = _ks_box(h)
        X = grid(b)
        u = ks.z4c(X)
        r, nbad = O.z4c_rhs(b, p, u)
        assert nbad == 0
        out[h] = (b, X, u, r)
    return ks, out


def test_ks_stationary_nongauge_rhs_vanishes_at_4th_order(ks_runs):
    _, runs = ks_runs
    e1 = np.array([np.abs(runs[0.08][3][v][INNER]).max() for v in range(18)])
    e2 = np.array([np.abs(runs[0.04][3][v][INNER]).max() for v in range(18)])
    # terms of the RHS are O(0.1) here; residuals are truncation errors
    assert e1.max() < 1e-5
    ratio = e1 / e2
    assert np.all(ratio > 11.0), ratio   # 16 for exact 4th order


def test_ks_constraints_vanish_at_4th_order(ks_runs):
    """Kerr-Schild data satisfy the Hamiltonian and momentum constraints; with
    nonzero K_ij and Christoffel symbols a dropped connection term, a wrong
    sign or a wrong contraction leaves an O(1) residual."""
    _, runs = ks_runs
    errs = []
    for h in (0.08, 0.04):
        b, X, u, r = runs[h]
        l2, ham, mom = O.z4c_constraints_l2(b, u, want_fields=True)
        e = [np.abs(ham[INNER]).max()] + [np.abs(mom[a][INNER]).max() for a in range(3)]
        errs.append(np.array(e))
        # the norms are over the same inner points (skip layers of 3)
        assert np.allclose(l2, [np.sqrt(np.mean(f[INNER] ** 2)) for f in (ham, *mom)], rtol=1e-12)
    assert errs[0].max() < 1e-4, errs[0]
    ratio = errs[0] / errs[1]
    assert np.all(ratio > 11.0), ratio   # 16 for exact 4th order


def test_momentum_constraint_flat_polynomial_exact():
    """flat metric with quadratic K_ij (At_ij and Khat): Christoffels vanish
    and the finite differences are exact, so mom_i = d_j K_ij - d_i K exactly
    at every inner point, including those one point from the edge (order 2)."""
    h = 0.1
    n = 12
    b = O.make_box(n, n + 1, n + 2, -0.55, -0.6, -0.65, h, skip=(1,) * 6)
    X = grid(b)
    x, y, z = X
    u = flat_z4c(x.shape)
    SYM = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    rng = np.random.default_rng(7)
    coef = rng.normal(size=(7, 10)) * 0.1

    def quad(cf):
        return (cf[0] + cf[1] * x + cf[2] * y + cf[3] * z + cf[4] * x * x + cf[5] * y * y + cf[6] * z * z
                + cf[7] * x * y + cf[8] * x * z + cf[9] * y * z)

    def dquad(cf):
        return np.array([cf[1] + 2 * cf[4] * x + cf[7] * y + cf[8] * z,
                         cf[2] + 2 * cf[5] * y + cf[7] * x + cf[9] * z,
                         cf[3] + 2 * cf[6] * z + cf[8] * x + cf[9] * y])
    for c in range(6):
        u[7 + c] = quad(coef[c])
    u[13] = quad(coef[6])
    dK = np.zeros((3, 3, 3) + x.shape)     # dK[e, i, j] = d_e K_ij
    dKh = dquad(coef[6])
    for c, (i, j) in enumerate(SYM):
        d = dquad(coef[c]) + (i == j) * dKh / 3.0
        dK[:, i, j] = d
        dK[:, j, i] = d
    l2, ham, mom = O.z4c_constraints_l2(b, u, want_fields=True)
    sl = (slice(1, -1),) * 3
    for i in range(3):
        expect = sum(dK[j, i, j] for j in range(3)) - sum(dK[i, j, j] for j in range(3))
        assert np.abs(expect[sl]).max() > 1e-2
        np.testing.assert_allclose(mom[i][sl], expect[sl], rtol=0, atol=1e-12)
    # H = K^2 - K_ij K^ij for flat space
    Kij = np.zeros((3, 3) + x.shape)
    for c, (i, j) in enumerate(SYM):
        Kij[i, j] = Kij[j, i] = u[7 + c] + (i == j) * u[13] / 3.0
    trK = Kij[0, 0] + Kij[1, 1] + Kij[2, 2]
    H = trK ** 2 - np.einsum("ij...,ij...->...", Kij, Kij)
    np.testing.assert_allclose(ham[sl], H[sl], rtol=0, atol=1e-12)


def test_ks_gauge_rhs_matches_closed_form(ks_runs):
    ks, runs = ks_runs
    errs = []
    for h in (0.08, 0.04):
        b, X, u, r = runs[h]
        alp = u[18]
        beta = u[19:22]
        K = u[13]
        dalp = np.array([np.gradient(alp, h, axis=2 - k, edge_order=2) for k in range(3)])
        # exact derivatives from the analytic lapse instead of np.gradient
        from spacetimes import fd_grad
        dalp = fd_grad(ks.alpha, X)
        dbeta = fd_grad(lambda Xs: ks.adm(Xs)[3], X)      # [k, i]
        ralpha = np.einsum("k...,k...->...", beta, dalp) - 2 * alp * K
        rbeta = u[14:17] + np.einsum("k...,ki...->i...", beta, dbeta)
        e = max(np.abs(r[18][INNER] - ralpha[INNER]).max(),
                max(np.abs(r[19 + i][INNER] - rbeta[i][INNER]).max() for i in range(3)))
        errs.append(e)
    assert errs[0] < 1e-5
    assert errs[0] / errs[1] > 11.0, errs


def _warped(X):
    """smooth generic non-solution with Gamma consistent with gt"""
    x, y, z = X
    u = flat_z4c(X.shape[1:])
    u[0] = 1 + 0.05 * np.sin(x + 0.3 * y)
    u[1] = 0.03 * np.cos(0.7 * y - z)
    u[2] = 0.02 * np.sin(0.5 * x * z)
    u[3] = 1 + 0.04 * np.cos(z + 0.2 * x)
    u[4] = 0.025 * np.sin(x - y + 0.3 * z)
    u[5] = 1 + 0.03 * np.sin(0.6 * y + 0.4 * z)
    u[6] = 0.8 + 0.1 * np.cos(0.5 * x + 0.2 * y * z)
    for c in range(6):
        u[7 + c] = 0.05 * np.sin((c + 1) * 0.3 * x + 0.2 * y - 0.1 * c * z)
    u[13] = 0.1 * np.cos(0.4 * x * y + z)
    u[18] = 0.9 + 0.05 * np.sin(x * y * 0.3)
    # Z4c data satisfy det(gt) = 1 and tr(At) = 0 (the algebraic constraints)
    g = np.array([[u[0], u[1], u[2]], [u[1], u[3], u[4]], [u[2], u[4], u[5]]])
    gm = np.moveaxis(g, (0, 1), (-2, -1))
    s = np.linalg.det(gm) ** (-1.0 / 3.0)
    u[0:6] *= s
    gi = np.linalg.inv(gm) / s[..., None, None]
    A = np.array([[u[7], u[8], u[9]], [u[8], u[10], u[11]], [u[9], u[11], u[12]]])
    tr = np.einsum("...ij,ij...->...", gi, A)
    for c in range(6):
        u[7 + c] -= tr * u[c] / 3.0
    return u


def _gamma_from_gt(gt_fn, X):
    from spacetimes import fd_grad

    def gtinv(Xs):
        g6 = gt_fn(Xs)
        g = np.array([[g6[0], g6[1], g6[2]], [g6[1], g6[3], g6[4]], [g6[2], g6[4], g6[5]]])
        return np.moveaxis(np.linalg.inv(np.moveaxis(g, (0, 1), (-2, -1))), (-2, -1), (0, 1))

    d = fd_grad(gtinv, X)
    return -np.einsum("jij...->i...", d)


def test_theta_rhs_equals_half_alpha_hamiltonian():
    """rTheta = alpha * Hhat / 2 where Hhat -> ADM Hamiltonian constraint when
    Gamma^i = -d_j gt^ij (z4_rhs_movpunc_N.m:271-276,315-316)."""
    p = O.z4c_params(order_dissipation=0)
    errs = []
    for h in (0.1, 0.05):
        n = int(round(1.2 / h)) + 1 + 8
        b = O.make_box(n, n, n, -0.6 - 4 * h, -0.6 - 4 * h, -0.6 - 4 * h, h, skip=(4,) * 6)
        X = grid(b)
        u = _warped(X)
        u[14:17] = _gamma_from_gt(lambda Xs: _warped(Xs)[0:6], X)
        r, nbad = O.z4c_rhs(b, p, u)
        _, ham = O.z4c_ham_l2(b, u, want_field=True)
        sl = (slice(4, -4),) * 3
        H = ham[sl]
        assert np.abs(H).max() > 1e-2          # genuinely constraint violating
        errs.append(np.abs(r[17][sl] - 0.5 * u[18][sl] * H).max())
    assert errs[0] / errs[1] > 11.0, errs


def _poly_gauge(X):
    x, y, z = X
    u = flat_z4c(X.shape[1:])
    u[18] = 1 + 0.1 * x - 0.05 * y + 0.02 * z + 0.03 * x * y - 0.02 * z * z + 0.01 * x * x
    u[19] = 0.01 + 0.02 * y + 0.01 * x * z
    u[20] = -0.02 * x + 0.015 * z * z
    u[21] = 0.005 * x * y - 0.01 * z
    return u


def test_flat_polynomial_gauge_exact_everywhere():
    """flat space with quadratic lapse and shift: all stencils (2nd and 4th
    order, lop-sided and shifted near the edge) are exact, so the RHS equals
    the closed forms at every computed point, including 1 and 2 points from
    the box edge."""
    p = O.z4c_params()
    h = 0.1
    n = 14
    b = O.make_box(n, n + 1, n + 2, -0.7, -0.6, -0.65, h, skip=(1,) * 6)
    X = grid(b)
    x, y, z = X
    u = _poly_gauge(X)
    r, nbad = O.z4c_rhs(b, p, u)
    assert nbad == 0
    alp = u[18]
    dalp = np.array([0.1 + 0.03 * y + 0.02 * x, -0.05 + 0.03 * x, 0.02 - 0.04 * z])
    dda = np.zeros((3, 3) + x.shape)
    dda[0, 0] = 0.02
    dda[0, 1] = dda[1, 0] = 0.03
    dda[2, 2] = -0.04
    beta = u[19:22]
    db = np.zeros((3, 3) + x.shape)        # db[k, i] = d_k beta^i
    db[1, 0] = 0.02
    db[0, 0] = 0.01 * z
    db[2, 0] = 0.01 * x
    db[0, 1] = -0.02
    db[2, 1] = 0.03 * z
    db[0, 2] = 0.005 * y
    db[1, 2] = 0.005 * x
    db[2, 2] = -0.01
    ddb = np.zeros((3, 3, 3) + x.shape)    # ddb[k, l, i]
    ddb[0, 2, 0] = ddb[2, 0, 0] = 0.01
    ddb[2, 2, 1] = 0.03
    ddb[0, 1, 2] = ddb[1, 0, 2] = 0.005
    divb = db[0, 0] + db[1, 1] + db[2, 2]
    sl = (slice(1, -1),) * 3
    SYM = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    lap = dda[0, 0] + dda[1, 1] + dda[2, 2]
    for c, (i, j) in enumerate(SYM):
        rg = db[i, j] + db[j, i] - (2.0 / 3.0) * (i == j) * divb
        np.testing.assert_allclose(r[c][sl], rg[sl], rtol=0, atol=1e-13)
        rA = -(dda[i, j] - (i == j) * lap / 3.0)
        np.testing.assert_allclose(r[7 + c][sl], rA[sl], rtol=0, atol=1e-12)
    np.testing.assert_allclose(r[6][sl], (-4.0 / 6.0) * divb[sl], rtol=0, atol=1e-13)
    np.testing.assert_allclose(r[13][sl], -lap[sl], rtol=0, atol=1e-12)
    for i in range(3):
        rG = ddb[0, 0, i] + ddb[1, 1, i] + ddb[2, 2, i] + (ddb[i, 0, 0] + ddb[i, 1, 1] + ddb[i, 2, 2]) / 3.0
        np.testing.assert_allclose(r[14 + i][sl], rG[sl], rtol=0, atol=1e-12)
        rb = np.einsum("k...,k...->...", beta, db[:, i]) - 2.0 * beta[i]
        np.testing.assert_allclose(r[19 + i][sl], rb[sl], rtol=0, atol=1e-13)
    ra = np.einsum("k...,k...->...", beta, dalp) - 2.0 * alp * 0.0
    np.testing.assert_allclose(r[18][sl], ra[sl], rtol=0, atol=1e-13)
    # untouched skip layer
    assert np.all(r[:, 0] == 0) and np.all(r[:, :, :, -1] == 0)


def test_minkowski_rhs_zero_with_dissipation():
    p = O.z4c_params()
    b = O.make_box(12, 12, 12, -0.55, -0.55, -0.55, 0.1, skip=(1,) * 6)
    u = flat_z4c((12, 12, 12))
    r, nbad = O.z4c_rhs(b, p, u)
    assert np.abs(r).max() == 0.0


def test_dissipation_annihilates_cubics_and_matches_fourth_difference():
    """KO term: -eps * sum_d h^3 d^4u/dx^4 (discretize.m:131-135) is zero for
    cubic data and equals -eps*24*c4*h^3 for u = c4 x^4 in the interior."""
    h = 0.1
    eps_sigma = 0.1
    p = O.z4c_params(dissipation_factor=eps_sigma, z4_shiftdriver=0.0, z4_lapse="constant")
    b = O.make_box(12, 12, 12, -0.55, -0.55, -0.55, h, skip=(2,) * 6)
    X = grid(b)
    u = flat_z4c((12, 12, 12))
    # constant-lapse gauge: the lapse RHS is the KO term alone
    u[18] = 1 + 0.3 * X[0] ** 3 - 0.2 * X[1] ** 2 * X[2]
    r, _ = O.z4c_rhs(b, p, u)
    sl = (slice(2, -2),) * 3
    assert np.abs(r[18][sl]).max() < 1e-12
    u[18] = 1 + 0.5 * X[0] ** 4
    r, _ = O.z4c_rhs(b, p, u)
    expect = -(eps_sigma / 16.0) * 24 * 0.5 * h ** 3
    np.testing.assert_allclose(r[18][sl], expect, rtol=1e-9)


def test_addlinear_projects_det_and_trace():
    ks = KerrSchild()
    p = O.z4c_params()
    b = _ks_box(0.08)
    X = grid(b)
    u = ks.z4c(X)
    rng = np.random.default_rng(1)
    upre = u + 1e-3 * rng.standard_normal(u.shape)
    unew = np.zeros_like(u)
    O.z4c_rhs(b, p, u, unew=unew, upre=upre, c=0.05)
    g = unew[0:6][(slice(None),) + INNER]
    A = unew[7:13][(slice(None),) + INNER]
    G = np.array([[g[0], g[1], g[2]], [g[1], g[3], g[4]], [g[2], g[4], g[5]]])
    Am = np.array([[A[0], A[1], A[2]], [A[1], A[3], A[4]], [A[2], A[4], A[5]]])
    Gm = np.moveaxis(G, (0, 1), (-2, -1))
    det = np.linalg.det(Gm)
    np.testing.assert_allclose(det, 1.0, atol=1e-13)
    Gi = np.linalg.inv(Gm)
    tr = np.einsum("...ij,...ij->...", Gi, np.moveaxis(Am, (0, 1), (-2, -1)))
    assert np.abs(tr).max() < 1e-14 * max(1.0, np.abs(A).max()) * 50
    # other fields: exactly upre + c rhs
    r, _ = O.z4c_rhs(b, p, u)
    np.testing.assert_allclose(unew[13][INNER], (upre[13] + 0.05 * r[13])[INNER], rtol=1e-15, atol=1e-16)


def test_algcon_projection():
    rng = np.random.default_rng(7)
    n = 6
    b = O.make_box(n, n, n, 0, 0, 0, 0.1)
    u = flat_z4c((n, n, n))
    u[0:6] += 0.05 * rng.standard_normal((6, n, n, n))
    u[7:13] = 0.1 * rng.standard_normal((6, n, n, n))
    O.z4c_algcon(b, u)
    G = np.array([[u[0], u[1], u[2]], [u[1], u[3], u[4]], [u[2], u[4], u[5]]])
    Gm = np.moveaxis(G, (0, 1), (-2, -1))
    np.testing.assert_allclose(np.linalg.det(Gm), 1.0, atol=2e-15)
    A = np.array([[u[7], u[8], u[9]], [u[8], u[10], u[11]], [u[9], u[11], u[12]]])
    tr = np.einsum("...ij,...ij->...", np.linalg.inv(Gm), np.moveaxis(A, (0, 1), (-2, -1)))
    assert np.abs(tr).max() < 1e-15


def test_rk4_amplification_linear_decay():
    """withB Gamma driver in flat space: dB/dt = -eta B at every point, all
    other RHS zero -> one RK4 step multiplies B by the 4th-order Taylor
    polynomial of exp(-eta dt) (rk.c:98-103,220-257)."""
    p = O.z4c_params(z4_shift="withB", z4_shiftdriver=2.0, order_dissipation=0)
    n = 10
    b = O.make_box(n, n, n, 0, 0, 0, 0.1, skip=(2,) * 6)
    u = flat_z4c((n, n, n))
    u[22] = 0.3
    u[23] = -0.7
    u[24] = 1.1
    dt = 0.3
    O.z4c_rk4_step(b, p, dt, u, bc=2, pgh=2, algcon=True)
    z = -2.0 * dt
    amp = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    np.testing.assert_allclose(u[22], 0.3 * amp, rtol=1e-15)
    np.testing.assert_allclose(u[24], 1.1 * amp, rtol=1e-15)
    assert np.all(u[19:22] == 0)


def test_radiative_bc_outgoing_wave():
    """u = u0 + f(r - v t)/r solves d_t u = -v (u-u0)/r - v d_r u exactly;
    the BC (2nd-order one-sided/centred differences) converges to it."""
    p = O.z4c_params()
    errs = []
    for h in (0.1, 0.05):
        n = int(round(1.0 / h)) + 1
        b = O.make_box(n, n, n, 2.0, 1.0, 1.5, h)
        X = grid(b)
        r = np.sqrt((X ** 2).sum(0))
        v = np.sqrt(2.0)
        f = np.sin(r)
        u = flat_z4c((n, n, n))
        u[13] = f / r                      # Khat: far limit 0, speed sqrt(2)
        rhs = np.zeros_like(u)
        O.z4c_radiative_bc(b, p, rhs, u)
        exact = -v * np.cos(r) / r
        face = np.zeros((n, n, n), bool)
        face[0, :, :] = face[-1] = face[:, 0] = face[:, -1] = face[:, :, 0] = face[:, :, -1] = True
        errs.append(np.abs(rhs[13][face] - exact[face]).max())
        assert np.all(rhs[13][~face] == 0)
        # constant far state preserved for gxx (far limit 1)
        assert np.abs(rhs[0][face]).max() == 0.0
    assert errs[0] / errs[1] > 3.0, errs


def test_periodic_fill():
    n = 10
    b = O.make_box(n, n, n, 0, 0, 0, 0.1)
    rng = np.random.default_rng(3)
    u = rng.standard_normal((1, n, n, n))
    O.periodic_fill(b, u, 3)
    per = n - 6
    for k in range(n):
        kk = (k - 3) % per + 3
        np.testing.assert_array_equal(u[0, k, :, :][3:-3, 3:-3], u[0, kk, :, :][3:-3, 3:-3])


def _fd_weights(pos, deriv=1):
    """weights of the finite-difference approximant of the deriv-th
    derivative at 0 from the points pos (exact for polynomials of degree
    < len(pos)): the Vandermonde system sum_k w_k pos_k^p = p! delta_{p,deriv}"""
    pos = np.asarray(pos, dtype=float)
    n = len(pos)
    A = np.array([pos ** p for p in range(n)])
    b = np.zeros(n)
    b[deriv] = float(np.prod(np.arange(1, deriv + 1)))
    return np.linalg.solve(A, b)


def test_advection_rows_near_box_edges():
    """The lop-sided advection row the oracle uses at each point of a line
    through a box (set_advection_wsign / set_advection4_onedir,
    bam_amr_advection.h:90-146, with the order-2 rows one point from the
    edge, bam_amr_loops_box.h:186-215), pinned without the reference's
    tables: the expected row is the least-shifted upwind stencil (offset at
    least advection_lopsided = 1 for order 4, 0 for order 2) that stays
    inside the box [0, imax], with weights derived here from exactness for
    degree <= 4 (<= 2).  Khat has quintic and sextic parts in x, so every
    candidate row gives a different value (rows differ in their fifth
    moments, mirror rows in their sixth); flat space, unit lapse,
    no dissipation: rhs[Khat] = beta^x d_x Khat + Khat^2 / 3."""
    m, n = 14, 7
    h = 0.25
    p = O.z4c_params(order_dissipation=0)
    b = O.make_box(m, n, n, -1.3, -0.75, -0.75, h)
    X = grid(b)
    imax = m - 1
    j = k = 3
    x = X[0][k, j, :]
    c5 = 1e-3

    def f(xx):
        t = (xx + 0.37) / 1.7
        return c5 * (t ** 5 + 0.7 * t ** 6 + 0.3 * ((xx - 0.2) / 1.7) ** 4)

    for bx in (-0.6, 0.45):
        u = flat_z4c((n, n, m))
        u[19] = bx
        u[13] = f(X[0])
        r, nbad = O.z4c_rhs(b, p, u)
        assert nbad == 0
        adv = r[13][k, j, :] - u[13][k, j, :] ** 2 / 3
        for i in range(1, imax):
            order = 2 if i in (1, imax - 1) else 4
            npt = order + 1
            base = 0 if order == 2 else 1
            cands = {}
            for off in range(npt):
                # upwind: beta < 0 uses points on the low side of i, shifted up by off
                lo = i + off - order if bx < 0 else i - off
                pos = np.arange(lo, lo + npt)
                if pos.min() < 0 or pos.max() > imax:
                    continue
                w = _fd_weights(pos - i)
                cands[off] = bx * np.dot(w, f(x[pos])) / h
            inside = [o for o in cands if o >= base]
            want = cands[min(inside)]
            assert abs(adv[i] - want) <= 1e-12 * abs(want) + 1e-18, (bx, i, adv[i], want)
            for off, val in cands.items():
                if off != min(inside):
                    assert abs(val - want) > 1e-6 * abs(want), (bx, i, off)


def test_radiative_bc_neighbour_choice_at_faces_edges_and_corners():
    """set_boundary_radiative's difference choice (computedel, radiative.c:
    18-40, neighbours 1 and 2 away that exist in the box, set_nbinds_check_13)
    at every PHYBOUND point of a box that contains the origin's plane x = 0
    (so v^x > 0, < 0 and = 0 all occur on the faces, edges and corners),
    restated by enumerating the existing neighbours of each point: upwind
    one-sided 2nd-order difference where both upwind neighbours exist, the
    centred one where only the near one does, else the downwind one-sided;
    with v^x = 0 the centred one where both near neighbours exist.  Quartic
    data separate the three differences at every point."""
    m, nn, o = 9, 6, 7
    h = 0.5
    b = O.make_box(m, nn, o, -2.0, 0.3, -1.6, h)
    X = grid(b)
    p = O.z4c_params()
    rng = np.random.default_rng(12)
    coef = rng.uniform(0.5, 1.5, 9)
    x, y, z = X
    Kh = (coef[0] * x ** 4 + coef[1] * y ** 4 + coef[2] * z ** 4 + coef[3] * x ** 3 * y + coef[4] * y * z ** 3
          + coef[5] * x * y * z + coef[6] * x ** 2 + coef[7] * y + coef[8])
    u = flat_z4c((o, nn, m))
    u[13] = 1e-2 * Kh
    rhs = np.zeros_like(u)
    O.z4c_radiative_bc(b, p, rhs, u)
    v = np.sqrt(2.0)                        # 1+log with f = 2: speed sqrt(2) for Khat, far limit 0
    U = u[13]
    dims = (m, nn, o)
    nface = 0
    for kk in range(o):
        for jj in range(nn):
            for ii in range(m):
                idx = (ii, jj, kk)
                if not any(t in (0, d - 1) for t, d in zip(idx, dims)):
                    continue
                xs = (x[kk, jj, ii], y[kk, jj, ii], z[kk, jj, ii])
                rr = np.sqrt(sum(c * c for c in xs))
                out = -v * (U[kk, jj, ii] - 0.0) / rr
                acc = 0.0
                for d in range(3):
                    def at(s):
                        q = list(idx)
                        q[d] += s
                        return U[q[2], q[1], q[0]]
                    have = {s for s in (-2, -1, 1, 2) if 0 <= idx[d] + s < dims[d]}
                    cen = (at(1) if 1 in have else at(0)) - (at(-1) if -1 in have else at(0))
                    rgt = 4 * at(1) - at(2) - 3 * at(0) if {1, 2} <= have else None
                    lft = 3 * at(0) + at(-2) - 4 * at(-1) if {-1, -2} <= have else None
                    vd = v * xs[d] / rr
                    if vd > 0:
                        dl = lft if -2 in have else (cen if -1 in have else rgt)
                    elif vd < 0:
                        dl = rgt if 2 in have else (cen if -1 in have else lft)
                    else:
                        dl = cen if (1 in have and -1 in have) else (rgt if -1 not in have else lft)
                    if lft is not None and rgt is not None:
                        assert abs(lft - rgt) > 1e-9 and abs(lft - cen) > 1e-9
                    acc += vd * dl / (2 * h)
                out -= acc
                assert abs(rhs[13][kk, jj, ii] - out) <= 1e-13 * max(abs(out), 1e-12), (idx, rhs[13][kk, jj, ii], out)
                nface += 1
    assert nface == m * nn * o - (m - 2) * (nn - 2) * (o - 2)
