"""Puncture initial data and the ADM -> Z4c conversion on the device
(bamhip_punctures_bl, bamhip_adm_to_z4c) against the CPU oracle on a level
of two boxes: one with PHYBOUND faces (G extrapolated on its outer layer,
three passes in point order), one with refinement faces.  pow and sqrt of
the conversion round differently on the device, so fields are compared at
1e-13 of their scale; the lapse is the precollapsed 1/sqrt(g_xx)."""
import numpy as np
import pytest

import baum_amd as B
import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(20, 18, 16, (-2.375, -2.125, -1.875), B.PHYBOUND), (12, 10, 14, (3.125, -1.125, -1.625), B.REFBOUND)]
H = 0.25
MASSES, POS = [0.5, 0.4], [[-0.6, 0.1, 0.0], [3.9, -0.3, 0.2]]


def test_bl_data_and_z4_init_match_oracle():
    oboxes = [O.make_box(m, n, o, *x0, H) for m, n, o, x0, _ in SHAPES]
    gboxes = [B.make_box(m, n, o, *x0, H, bflag=(f,) * 6) for m, n, o, x0, f in SHAPES]
    G6, K6, U = list(range(6)), list(range(6, 12)), list(range(12, 37))
    lev = B.Level(gboxes, 37)
    lev.enable(range(37))
    for v in range(37):
        lev.fill(v, 7.0)                       # stale values the conversion must overwrite
    lev.punctures_bl(MASSES, POS, G6, K6)
    lev.adm_to_z4c(G6, K6, U, initial_lapse=2, zero_shift=True)
    got_g = lev.get_state(G6)
    got_u = lev.get_state(U)
    lev.free()
    want_g, want_u = [], []
    for b, (m, n, o, x0, f) in zip(oboxes, SHAPES):
        g, k, psi = O.punctures_bl(b, MASSES, POS)
        u = O.adm_to_z4c(b, g, k, phybound=(f == B.PHYBOUND))
        u[18] = 1 / np.sqrt(g[0])
        if f != B.PHYBOUND:                     # G keeps the stale value on the outer layer
            inner = np.zeros(u.shape[1:], bool)
            inner[1:-1, 1:-1, 1:-1] = True
            for a in range(3):
                u[14 + a][~inner] = 7.0
        want_g.append(g.reshape(6, -1))
        want_u.append(u.reshape(25, -1))
    want_g, want_u = np.concatenate(want_g, axis=1), np.concatenate(want_u, axis=1)
    np.testing.assert_allclose(got_g, want_g, rtol=1e-14, atol=0)
    for f in range(25):
        scale = max(np.abs(want_u[f]).max(), 1e-300)
        err = np.abs(got_u[f] - want_u[f]).max()
        assert err <= 1e-13 * scale + 1e-14, (f, B.VAR_NAMES[f], err, scale)
    # G is not trivially zero: the BL conformal metric is flat only up to round-off, so check chi and alpha shape
    assert np.abs(got_u[6] - 1).max() > 0.1


def test_gamma_and_boundary_extrapolation_match_oracle():
    """a conformal metric with a nonzero G: diag(1/(1 + c x^2), 1/(1 + d y^2),
    (1 + c x^2)(1 + d y^2)) times psi^4, on the PHYBOUND box"""
    m, n, o, x0, f = SHAPES[0]
    ob = O.make_box(m, n, o, *x0, H)
    gb = B.make_box(m, n, o, *x0, H, bflag=(B.PHYBOUND,) * 6)
    x = x0[0] + H * np.arange(m)
    y = x0[1] + H * np.arange(n)
    z = x0[2] + H * np.arange(o)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    psi4 = (1 + 0.1 * np.sin(X + 2 * Y) * np.cos(Z)) ** 4
    g = np.zeros((6,) + X.shape)
    g[0] = psi4 / (1 + 0.4 * X * X)
    g[3] = psi4 / (1 - 0.03 * Y * Y)
    g[5] = psi4 * (1 + 0.4 * X * X) * (1 - 0.03 * Y * Y)
    g[1] = 0.01 * np.cos(Z)
    k = 0.05 * np.random.default_rng(3).standard_normal((6,) + X.shape)
    G6, K6, U = list(range(6)), list(range(6, 12)), list(range(12, 37))
    lev = B.Level([gb], 37)
    lev.enable(range(37))
    for v in range(37):
        lev.fill(v, 0.0)
    lev.put_state(G6, g.reshape(6, -1))
    lev.put_state(K6, k.reshape(6, -1))
    lev.adm_to_z4c(G6, K6, U, initial_lapse=0, zero_shift=False)
    got = lev.get_state(U).reshape((25,) + X.shape)
    lev.free()
    want = O.adm_to_z4c(ob, g, k, phybound=True)
    for fld in list(range(0, 18)):
        # G^i: round-off of the differenced inverse metric scales with |G|, not with one component
        scale = np.abs(want[14:17]).max() if 14 <= fld <= 16 else max(np.abs(want[fld]).max(), 1e-300)
        err = np.abs(got[fld] - want[fld]).max()
        assert err <= 1e-12 * scale, (fld, B.VAR_NAMES[fld], err, scale)
    assert np.abs(want[14]).max() > 0.5                      # G^x = -2 c x + ...
    assert np.abs(want[14][0] - want[14][1]).max() > 0       # the outer layer was extrapolated
