"""The matter terms of the Z4c right-hand side on the device
(bamhip_z4c_matter_source) against the CPU oracle of z4_rhs_add_source on a
two-box level, with and without withB / gamma2 (alpha^p), the c form, the chi
floor and a non-finite matter value.  log, exp and pow round differently on
the device, so each right-hand side is compared at 1e-13 of its largest
change; the fields the matter terms do not touch stay bitwise equal."""
import numpy as np
import pytest

import baum_amd as B
import oracle as O

pytestmark = pytest.mark.gpu

U = list(range(25))
R = list(range(25, 50))
MADM = list(range(50, 61))
NV = 61
SHAPES = [(12, 9, 10, (0.0, 0.0, 0.0)), (7, 15, 8, (1.5, 0.0, 0.0))]


def _state(box, seed):
    rng = np.random.default_rng(seed)
    shape = (box.o, box.n, box.m)
    u = np.zeros((25,) + shape)
    u[0:6] = 0.05 * rng.standard_normal((6,) + shape) + np.array([1, 0, 0, 1, 0, 1.0])[:, None, None, None]
    u[6] = 0.2 + 0.7 * rng.random(shape)
    u[18] = 0.4 + 0.6 * rng.random(shape)
    madm = np.zeros((11,) + shape)
    madm[0] = 1e-3 * rng.random(shape)
    madm[1:4] = 1e-4 * rng.standard_normal((3,) + shape)
    madm[4:10] = 1e-4 * rng.standard_normal((6,) + shape)
    madm[10] = 3e-4 * rng.random(shape)
    return u, madm


@pytest.mark.parametrize("shift,c", [("withShiftadv", 0.0), ("withB gamma2 gamma0 withShiftadv", 0.25)])
def test_matter_terms_match_oracle(shift, c):
    over = dict(z4_shift=shift, z4_shiftalphapower=1.5, z4_shiftgammacoeff=0.75, z4_chi_div_floor=0.3)
    gpar = B.z4c_params(**over)
    opar = O.z4c_params(**over)
    oboxes = [O.make_box(m, n, o, *x0, 0.1) for m, n, o, x0 in SHAPES]
    gboxes = [B.make_box(m, n, o, *x0, 0.1) for m, n, o, x0 in SHAPES]
    data = [_state(b, 5 + k) for k, b in enumerate(oboxes)]
    data[1][1][4][3, 5, 4] = np.nan
    r0 = [np.random.default_rng(9 + k).standard_normal((25, b.o, b.n, b.m)) for k, b in enumerate(oboxes)]
    lev = B.Level(gboxes, NV)
    lev.enable(range(NV))
    cat = lambda arrs, a: np.concatenate([x[a].reshape(-1) for x in arrs])
    for a in range(25):
        lev.put(U[a], cat([d[0] for d in data], a))
        lev.put(R[a], cat(r0, a))
    for a in range(11):
        lev.put(MADM[a], cat([d[1] for d in data], a))
    lev.z4c_matter_source(gpar, R, U, MADM, c=c)
    got = np.stack([lev.get(v) for v in R])
    lev.free()
    want = np.concatenate([O.z4c_matter_source(b, opar, u, madm, r.copy(), c=c).reshape(25, -1)
                           for b, (u, madm), r in zip(oboxes, data, r0)], axis=1)
    base = np.concatenate([r.reshape(25, -1) for r in r0], axis=1)
    assert np.all(np.isfinite(want))
    touched = [7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17] + ([22, 23, 24] if "withB" in shift else [])
    for f in range(25):
        if f in touched:
            scale = np.abs(want[f] - base[f]).max()
            assert scale > 0, f
            assert np.abs(got[f] - want[f]).max() <= 1e-13 * scale, (f, B.VAR_NAMES[f])
        else:
            np.testing.assert_array_equal(got[f], base[f])
