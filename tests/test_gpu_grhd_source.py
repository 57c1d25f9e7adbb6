"""GRHD source terms and matter ADM variables on the device
(bamhip_grhd_sources) against the CPU oracle of grhd_sources on a level of
three boxes (radiative and refinement faces), with a general metric,
extrinsic curvature and shift, points below the atmosphere density, a point
with alpha <= 0 and a non-finite primitive.  FP contraction is off and the
sums run in the same order on both sides, so the results agree bitwise."""
import numpy as np
import pytest

import baum_amd as B
import oracle as O

pytestmark = pytest.mark.gpu

RHS = list(range(0, 5))
MET = list(range(5, 15))
KIJ = list(range(15, 21))
PRIM = list(range(21, 27))
MADM = list(range(27, 38))
NV = 38
H = 0.1
SHAPES = [(21, 13, 11, (-1.0, -0.6, -0.5)), (10, 24, 9, (1.2, -0.6, -0.5)), (9, 9, 30, (-1.0, 1.0, 1.0))]


def _fields(box, seed):
    rng = np.random.default_rng(seed)
    shape = (box.o, box.n, box.m)
    x = (box.x0 + H * np.arange(box.m))[None, None, :]
    y = (box.y0 + H * np.arange(box.n))[None, :, None]
    z = (box.z0 + H * np.arange(box.o))[:, None, None]
    met = np.zeros((10,) + shape)
    met[0] = 0.8 + 0.1 * np.sin(x - y) * np.cos(z) + 0.001 * rng.random(shape)
    one = np.ones(shape)
    met[1:4] = 0.05 * np.array([np.sin(y + z) * one, np.cos(x * z) * one, np.sin(x - y) * one])
    wob = [np.sin(x), np.cos(y), np.sin(z), np.cos(x + y), np.sin(y + z), np.cos(x * y)]
    met[4:10] = np.array([1.1, 0.02, -0.01, 1.05, 0.015, 1.2])[:, None, None, None] + \
        0.02 * np.array([w * one for w in wob])
    K = 0.1 * rng.standard_normal((6,) + shape)
    prim = np.zeros((6,) + shape)
    prim[0] = 1e-3 * (0.5 + rng.random(shape))
    prim[1] = 0.1 + 0.2 * rng.random(shape)
    prim[2:5] = 0.3 * (rng.random((3,) + shape) - 0.5)
    prim[5] = prim[0] * prim[1]
    return met, K, prim


@pytest.mark.parametrize("eos,c,order", [("ideal", 0.0, 4), ("poly", 0.125, 4), ("ideal", 0.0, 2)])
def test_sources_bitwise(eos, c, order):
    oboxes = [O.make_box(m, n, o, *x0, H) for m, n, o, x0 in SHAPES]
    flags = [(B.PHYBOUND,) * 6, (B.REFBOUND,) * 6, (B.PHYBOUND,) * 6]
    gboxes = [B.make_box(m, n, o, *x0, H, bflag=f) for (m, n, o, x0), f in zip(SHAPES, flags)]
    data = [_fields(b, 20 + k) for k, b in enumerate(oboxes)]
    data[0][2][0][5, 6, 7] = np.nan
    data[1][0][0][4, 10, 5] = -0.2
    rhs0 = [np.random.default_rng(40 + k).standard_normal((5, b.o, b.n, b.m)) for k, b in enumerate(oboxes)]
    lev = B.Level(gboxes, NV)
    lev.enable(range(NV))
    for v in range(NV):
        lev.fill(v, 0.0)
    cat = lambda arrs, a: np.concatenate([x[a].reshape(-1) for x in arrs])
    for a in range(10):
        lev.put(MET[a], cat([d[0] for d in data], a))
    for a in range(6):
        lev.put(KIJ[a], cat([d[1] for d in data], a))
        lev.put(PRIM[a], cat([d[2] for d in data], a))
    for a in range(5):
        lev.put(RHS[a], cat(rhs0, a))
    rhoatm = 7e-4
    lev.grhd_sources(B.hrsc_params(eos=eos), RHS, MET, KIJ, PRIM, MADM, c=c, rhoatm=rhoatm, order_centered=order)
    got_rhs = np.stack([lev.get(v) for v in RHS])
    got_madm = np.stack([lev.get(v) for v in MADM])
    lev.free()
    want_rhs, want_madm = [], []
    op = O.hrsc_params(eos=0 if eos == "ideal" else 1)
    for b, (met, K, prim), r in zip(oboxes, data, rhs0):
        madm = np.zeros((11, b.o, b.n, b.m))
        rr = r.copy()
        O.grhd_sources(b, op, met, K, prim, rr, madm, c=c, rhoatm=rhoatm, order_centered=order)
        want_rhs.append(rr.reshape(5, -1))
        want_madm.append(madm.reshape(11, -1))
    want_rhs, want_madm = np.concatenate(want_rhs, axis=1), np.concatenate(want_madm, axis=1)
    assert np.all(np.isfinite(want_rhs)) and np.all(np.isfinite(want_madm))
    np.testing.assert_array_equal(got_rhs, want_rhs)
    np.testing.assert_array_equal(got_madm, want_madm)
    rhs0c = np.concatenate([r.reshape(5, -1) for r in rhs0], axis=1)
    assert np.abs(got_rhs[2:] - rhs0c[2:]).max() > 1e-6      # momentum sources act
