"""The moving-puncture tracker on the device (bamhip_puncture_track) against
the CPU oracle of track_moving_puncture_integrate: a two-box level, a cubic
shift at the new and a different one at the previous time level, punctures
inside one box, near the box seam and outside the level (no cube: shift 0).
FP contraction is off on both sides, so the positions agree bitwise."""
import numpy as np
import pytest

import baum_amd as B
import oracle as O

pytestmark = pytest.mark.gpu

BETA, BETA_P = [19, 20, 21], [22, 23, 24]


def _coords(b):
    x = b.x0 + b.dx * np.arange(b.m)
    y = b.y0 + b.dy * np.arange(b.n)
    z = b.z0 + b.dz * np.arange(b.o)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def _shift(X, Y, Z, s):
    return [s * (0.2 + 0.1 * X - 0.05 * Y * Z + 0.03 * X ** 3),
            s * (-0.1 + 0.07 * Y + 0.02 * X * Z - 0.01 * Y ** 3),
            s * (0.05 - 0.04 * Z + 0.06 * X * Y + 0.02 * Z ** 3)]


@pytest.mark.parametrize("order", [4, 6])
def test_puncture_track_matches_oracle_bitwise(order):
    h = 0.25
    shapes = [(20, 18, 16, (-2.375, -2.125, -1.875)), (12, 14, 10, (2.625, -2.125, -1.875))]
    oboxes = [O.make_box(m, n, o, *x0, h) for m, n, o, x0 in shapes]
    gboxes = [B.make_box(m, n, o, *x0, h, bflag=(B.REFBOUND,) * 6) for m, n, o, x0 in shapes]
    crd = [_coords(b) for b in oboxes]
    beta = [[f.reshape(-1) for f in _shift(*crd[k], 1.0)] for k in range(2)]
    beta_p = [[f.reshape(-1) for f in _shift(*crd[k], 0.9)] for k in range(2)]
    lev = B.Level(gboxes, 25)
    lev.enable(range(25))
    for a in range(3):
        lev.put(BETA[a], np.concatenate([beta[k][a] for k in range(2)]))
        lev.put(BETA_P[a], np.concatenate([beta_p[k][a] for k in range(2)]))
    pos0 = np.array([[0.31, -0.17, 0.09], [-1.1, 0.6, -0.4], [2.47, 0.12, 0.2], [9.0, 0.0, 0.0]])
    dt = 0.0625
    dpos = B.DeviceArray.from_host(pos0)
    want = pos0.copy()
    for step in range(5):
        lev.puncture_track(BETA, BETA_P, dt, dpos, order=order)
        for p in range(want.shape[0]):
            want[p] = O.track_puncture(oboxes, [[beta[k][a] for k in range(2)] for a in range(3)],
                                       [[beta_p[k][a] for k in range(2)] for a in range(3)], dt, want[p], order=order)
    got = dpos.get()
    np.testing.assert_array_equal(got, want)
    assert np.abs(got[0] - pos0[0]).max() > 1e-3      # it moved
    np.testing.assert_array_equal(got[3], pos0[3])    # outside the level: shift 0
    dpos.free()
    lev.free()


@pytest.mark.parametrize("pin_z", [False, True])
def test_puncture_track_pin_z_and_leaving_the_level(pin_z):
    """large steps that carry some punctures out of the level (their old
    positions stay) and pin_z (moving_puncture_fixz BH) against the oracle"""
    h = 0.25
    shapes = [(20, 18, 16, (-2.375, -2.125, -1.875)), (12, 14, 10, (2.625, -2.125, -1.875))]
    oboxes = [O.make_box(m, n, o, *x0, h) for m, n, o, x0 in shapes]
    gboxes = [B.make_box(m, n, o, *x0, h, bflag=(B.REFBOUND,) * 6) for m, n, o, x0 in shapes]
    crd = [_coords(b) for b in oboxes]
    beta = [[f.reshape(-1) for f in _shift(*crd[k], 1.0)] for k in range(2)]
    beta_p = [[f.reshape(-1) for f in _shift(*crd[k], 0.9)] for k in range(2)]
    lev = B.Level(gboxes, 25)
    lev.enable(range(25))
    for a in range(3):
        lev.put(BETA[a], np.concatenate([beta[k][a] for k in range(2)]))
        lev.put(BETA_P[a], np.concatenate([beta_p[k][a] for k in range(2)]))
    rng = np.random.default_rng(11)
    pos0 = np.column_stack([rng.uniform(-2.0, 4.8, 24), rng.uniform(-1.6, 1.6, 24), rng.uniform(-1.4, 1.4, 24)])
    dpos = B.DeviceArray.from_host(pos0)
    lev.puncture_track(BETA, BETA_P, 1.5, dpos, order=4, pin_z=pin_z)
    got = dpos.get()
    want = np.array([O.track_puncture(oboxes, [[beta[k][a] for k in range(2)] for a in range(3)],
                                      [[beta_p[k][a] for k in range(2)] for a in range(3)], 1.5, p, order=4,
                                      pin_z=pin_z) for p in pos0])
    np.testing.assert_array_equal(got, want)
    kept = np.all(got == pos0, axis=1)
    assert 0 < kept.sum() < len(pos0)          # some left the level and stayed put, some moved
    if pin_z:
        np.testing.assert_array_equal(got[:, 2], pos0[:, 2])
    dpos.free()
    lev.free()
