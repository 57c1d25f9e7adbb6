"""Pins the CPU oracle of the Z4c matter terms (z4_rhs_add_source).  The
oracle follows the reference through psi^4 = exp(4 log(chi) / chipsipower);
the checks below are written directly in chi (chipsipower = -4, so
psi^4 = 1/chi) from the Z4c equations with matter:
    A_ij  += -8 pi alpha chi (S_ij - gamma_ij S / 3),  gamma_ij = g~_ij / chi
    Khat  += 4 pi alpha (rho + S),  Theta += -8 pi alpha rho
    G^i   += -16 pi alpha gamma~^ij S_j = -16 pi alpha S^i / chi
    B^i   += (gamma0 + gamma2 shiftgammacoeff alpha^p) x the G^i term (withB)
with the trace of the A_ij term zero for a consistent S = gamma^ij S_ij, the
chi floor, the non-finite rule, inner points only and the c form."""
import numpy as np

import oracle as O

S6 = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]


def _state(box, seed):
    rng = np.random.default_rng(seed)
    shape = (box.o, box.n, box.m)
    u = np.zeros((25,) + shape)
    a = 0.05 * rng.standard_normal((6,) + shape)
    u[0:6] = a + np.array([1, 0, 0, 1, 0, 1.0])[:, None, None, None]
    u[6] = 0.3 + 0.6 * rng.random(shape)            # chi
    u[18] = 0.5 + 0.5 * rng.random(shape)           # alpha
    madm = np.zeros((11,) + shape)
    madm[0] = 1e-3 * rng.random(shape)
    madm[1:4] = 1e-4 * rng.standard_normal((3,) + shape)
    madm[4:10] = 1e-4 * rng.standard_normal((6,) + shape)
    # consistent trace S = gamma^ij S_ij with gamma_ij = g~_ij / chi
    for k in range(box.o):
        for j in range(box.n):
            for i in range(box.m):
                g = np.array([[u[S6[p][q]][k, j, i] for q in range(3)] for p in range(3)]) / u[6][k, j, i]
                Sij = np.array([[madm[4 + S6[p][q]][k, j, i] for q in range(3)] for p in range(3)])
                madm[10][k, j, i] = np.sum(np.linalg.inv(g) * Sij)
    return u, madm


def test_matter_terms_closed_form():
    box = O.make_box(7, 6, 5, 0, 0, 0, 0.1)
    u, madm = _state(box, 1)
    P = O.z4c_params(z4_shift="withB gamma2 withShiftadv", z4_shiftalphapower=2.0, z4_shiftgammacoeff=0.75)
    r = O.z4c_matter_source(box, P, u, madm, np.zeros((25,) + u.shape[1:]))
    pi = 3.1415926535897932
    inner = (slice(1, -1), slice(1, -1), slice(1, -1))
    chi, al, rho, S = u[6], u[18], madm[0], madm[10]
    for a in range(6):
        want = -8 * pi * al * chi * (madm[4 + a] - u[a] / chi * S / 3)
        np.testing.assert_allclose(r[7 + a][inner], want[inner], rtol=1e-12, atol=1e-12 * np.abs(want).max())
    np.testing.assert_allclose(r[13][inner], (4 * pi * al * (rho + S))[inner], rtol=1e-14)
    np.testing.assert_allclose(r[17][inner], (-8 * pi * al * rho)[inner], rtol=1e-14)
    for a in range(3):
        wantG = -16 * pi * al * madm[1 + a] / chi
        np.testing.assert_allclose(r[14 + a][inner], wantG[inner], rtol=1e-12)
        np.testing.assert_allclose(r[22 + a][inner], (0.75 * al ** 2 * wantG)[inner], rtol=1e-12)
    # trace-free A_ij term: g~^ij dA_ij = 0
    for k, j, i in [(1, 1, 1), (2, 3, 4), (3, 4, 5)]:
        g = np.array([[u[S6[p][q]][k, j, i] for q in range(3)] for p in range(3)])
        dA = np.array([[r[7 + S6[p][q]][k, j, i] for q in range(3)] for p in range(3)])
        assert abs(np.sum(np.linalg.inv(g) * dA)) <= 1e-12 * np.abs(dA).max()
    # inner points only; untouched fields stay zero
    assert np.all(r[:, 0] == 0) and np.all(r[:, :, :, -1] == 0)
    for f in list(range(0, 7)) + [18, 19, 20, 21]:
        assert np.all(r[f] == 0)


def test_chi_floor_nonfinite_and_c():
    box = O.make_box(6, 6, 6, 0, 0, 0, 0.1)
    u, madm = _state(box, 2)
    u[6][2, 2, 2] = 1e-9
    P = O.z4c_params(z4_chi_div_floor=1e-3)
    r = O.z4c_matter_source(box, P, u, madm, np.zeros((25,) + u.shape[1:]))
    pi = 3.1415926535897932
    want = -16 * pi * u[18][2, 2, 2] * madm[1][2, 2, 2] / 1e-3       # chi floored
    assert abs(r[14][2, 2, 2] - want) <= 1e-12 * abs(want)
    base = np.random.default_rng(3).standard_normal((25,) + u.shape[1:])
    full = O.z4c_matter_source(box, P, u, madm, np.zeros_like(base))
    got = O.z4c_matter_source(box, P, u, madm, base.copy(), c=0.3)
    np.testing.assert_allclose(got, base + 0.3 * full, rtol=1e-14, atol=1e-16)
    madm[0][3, 3, 3] = np.inf
    r = O.z4c_matter_source(box, P, u, madm, np.zeros((25,) + u.shape[1:]))
    assert np.all(r[:, 3, 3, 3] == 0) and np.all(np.isfinite(r))
