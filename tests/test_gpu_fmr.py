"""GPU parity of the FMR coupling (HIP kernels through the C ABI) with the
CPU oracle: R/P flags, prolongation and restriction are integer / fixed
sequences of IEEE operations evaluated in the reference's order with FP
contraction off on both sides, so every value must agree bitwise.
"""
import numpy as np
import pytest

import baum_amd as B
import oracle as O
from fmr_geometry import coarse_box, fine_box

pytestmark = pytest.mark.gpu

FP, FR = 0, 1          # flag variables
U0 = 2                 # first data variable


def _mk(m, n, o, x0, y0, z0, dx):
    return B.make_box(m, n, o, x0, y0, z0, dx, bflag=(B.REFBOUND,) * 6)


def _omk(m, n, o, x0, y0, z0, dx):
    return O.make_box(m, n, o, x0, y0, z0, dx)


def _hierarchy(mk):
    """coarse level of two boxes; fine level of three boxes (two in the first
    coarse box, one in the second), uneven sizes"""
    cb0 = coarse_box(mk, 44, 28, 26, 0.5)
    cb1 = coarse_box(mk, 24, 22, 24, 0.5, x0=(40.0, 0.25, -3.0))
    f1 = fine_box(mk, cb0, (4, 5, 4), (20, 16, 16))
    f2 = fine_box(mk, cb0, (24, 4, 5), (28, 18, 14))
    f3 = fine_box(mk, cb1, (5, 6, 4), (24, 20, 28))
    return [cb0, cb1], [f1, f2, f3], [0, 0, 1]


def _levels(nvar, nbuffer):
    cbs, fbs, parent = _hierarchy(_mk)
    lc = B.Level(cbs, U0 + nvar, l=0)
    lf = B.Level(fbs, U0 + nvar, l=1)
    lc.enable(range(U0 + nvar))
    lf.enable(range(U0 + nvar))
    B.set_rp_flags(lc, lf, parent, nbuffer, FP, FR)
    ocbs, ofbs, _ = _hierarchy(_omk)
    return lc, lf, ocbs, ofbs, parent


@pytest.mark.parametrize("nbuffer", [6, 4, 3])
def test_rp_flags_match_oracle(nbuffer):
    lc, lf, ocbs, ofbs, parent = _levels(1, nbuffer)
    fp, fr = O.rp_flags(ofbs, ocbs, parent, nbuffer)
    np.testing.assert_array_equal(lf.get(FP), fp)
    np.testing.assert_array_equal(lc.get(FR), fr)
    assert fp.max() == nbuffer and (fr == 3).sum() > 0


def _random_levels(lc, lf, nvar, seed):
    rng = np.random.default_rng(seed)
    cv = rng.standard_normal((nvar, lc.npoints))
    fv = rng.standard_normal((nvar, lf.npoints))
    for v in range(nvar):
        lc.put(U0 + v, cv[v])
        lf.put(U0 + v, fv[v])
    return cv, fv


@pytest.mark.parametrize("order", [2, 4, 6, 8])
@pytest.mark.parametrize("nbuffer", [3, 1, -1])
def test_prolong_matches_oracle_bitwise(order, nbuffer):
    nb, nvar = 6, 3
    lc, lf, ocbs, ofbs, parent = _levels(nvar, nb)
    cv, fv = _random_levels(lc, lf, nvar, 100 + order)
    idx = list(range(U0, U0 + nvar))
    B.prolong_varlist(lc, lf, idx, idx, nbuffer=nbuffer, order=order)
    fp, _ = O.rp_flags(ofbs, ocbs, parent, nb)
    expect = fv.copy()
    assert O.prolong_level(ofbs, ocbs, parent, fp, nbuffer, nb, order, cv, expect) == 0
    got = np.array([lf.get(v) for v in idx])
    assert np.count_nonzero(got != fv) > 0
    np.testing.assert_array_equal(got, expect)
    np.testing.assert_array_equal(np.array([lc.get(v) for v in idx]), cv)


@pytest.mark.parametrize("order", [2, 4, 6, 8])
@pytest.mark.parametrize("nbuffer", [3, 1])
def test_restrict_matches_oracle_bitwise(order, nbuffer):
    nb, nvar = 6, 3
    lc, lf, ocbs, ofbs, parent = _levels(nvar, nb)
    cv, fv = _random_levels(lc, lf, nvar, 200 + order)
    idx = list(range(U0, U0 + nvar))
    _, fr = O.rp_flags(ofbs, ocbs, parent, nb)
    expect = cv.copy()
    if O.restrict_level(ofbs, ocbs, parent, fr, nbuffer, order, fv, expect) != 0:
        # the flag-2 region with an 8-point stencil leaves these fine boxes:
        # the reference stops with "restriction: boxes are too small"
        with pytest.raises(B.BamhipError, match="too small"):
            B.restrict_varlist(lf, lc, idx, idx, nbuffer=nbuffer, order=order)
        return
    B.restrict_varlist(lf, lc, idx, idx, nbuffer=nbuffer, order=order)
    got = np.array([lc.get(v) for v in idx])
    assert np.count_nonzero(got != cv) > 0
    np.testing.assert_array_equal(got, expect)


def test_restrict_prolong_many_variables():
    """40 variables: more than one launch chunk; restriction precedes prolongation"""
    nb, nvar, order = 6, 40, 4
    lc, lf, ocbs, ofbs, parent = _levels(nvar, nb)
    cv, fv = _random_levels(lc, lf, nvar, 7)
    idx = list(range(U0, U0 + nvar))
    B.restrict_prolong_varlist(lc, lf, idx, idx, nbuffer=3, order=order)
    fp, fr = O.rp_flags(ofbs, ocbs, parent, nb)
    ec, ef = cv.copy(), fv.copy()
    assert O.restrict_level(ofbs, ocbs, parent, fr, 3, order, ef, ec) == 0
    assert O.prolong_level(ofbs, ocbs, parent, fp, 3, nb, order, ec, ef) == 0
    np.testing.assert_array_equal(np.array([lc.get(v) for v in idx]), ec)
    np.testing.assert_array_equal(np.array([lf.get(v) for v in idx]), ef)


def test_large_box_pair_bitwise():
    """a 96^3 fine box in a 128^3 parent, 25 variables, order 4, nbuffer 6"""
    nb, nvar, order = 6, 25, 4
    cb = coarse_box(_mk, 128, 128, 128, 1.0)
    fb = fine_box(_mk, cb, (40, 40, 40), (96, 96, 96))
    ocb = coarse_box(_omk, 128, 128, 128, 1.0)
    ofb = fine_box(_omk, ocb, (40, 40, 40), (96, 96, 96))
    lc = B.Level([cb], U0 + nvar)
    lf = B.Level([fb], U0 + nvar)
    lc.enable(range(U0 + nvar))
    lf.enable(range(U0 + nvar))
    B.set_rp_flags(lc, lf, [0], nb, FP, FR)
    cv, fv = _random_levels(lc, lf, nvar, 3)
    idx = list(range(U0, U0 + nvar))
    B.restrict_prolong_varlist(lc, lf, idx, idx, nbuffer=3, order=order)
    fp, fr = O.rp_flags([ofb], [ocb], [0], nb)
    ec, ef = cv.copy(), fv.copy()
    O.restrict_level([ofb], [ocb], [0], fr, 3, order, ef, ec)
    O.prolong_level([ofb], [ocb], [0], fp, 3, nb, order, ec, ef)
    for v in (0, 12, 24):
        np.testing.assert_array_equal(lc.get(idx[v]), ec[v])
        np.testing.assert_array_equal(lf.get(idx[v]), ef[v])


@pytest.mark.parametrize("ti", [2, 3, 4, 6])
def test_bo_time_interpolation(ti):
    """evolve.c:298-340, 367-380 on device lists vs the same vladd sequence in numpy"""
    cbs, _, _ = _hierarchy(_mk)
    lev = B.Level(cbs[:1], 8)
    lev.enable(range(8))
    rng = np.random.default_rng(ti)
    upp, uc, up, up2, up3, up4, unew = 0, 1, 2, 3, 4, 5, 6
    data = {k: rng.standard_normal(lev.npoints) for k in (uc, up, up2, up3, up4, unew)}
    for k, a in data.items():
        lev.put(k, a)
    for it in (0, 5):
        o = min(ti, it + 2)
        lev.bo_interp_begin(ti, it, [upp], [uc], [up], [up2], [up3], [up4])
        lev.put(uc, data[unew])             # the step
        lev.bo_interp_end(ti, it, [upp], [uc])
        w = {2: [0.5], 3: [0.75, -0.125], 4: [15 / 16, -5 / 16, 1 / 16],
             6: [315 / 256, -210 / 256, 126 / 256, -45 / 256, 7 / 256]}[o]
        hist = [data[uc], data[up], data[up2], data[up3], data[up4]]
        exp = w[0] * hist[0] + (w[1] * hist[1] if len(w) > 1 else 0)
        for k in range(2, len(w)):
            exp = exp + w[k] * hist[k]
        exp = exp + {2: 0.5, 3: 3 / 8, 4: 5 / 16, 6: 63 / 256}[o] * data[unew]
        np.testing.assert_allclose(lev.get(upp), exp, rtol=1e-14, atol=1e-14)
        # history shift for o >= 4: up2 <- up
        if o >= 4:
            np.testing.assert_array_equal(lev.get(up2), data[up])
        lev.put(uc, data[uc])
        for k in (up2, up3, up4):
            lev.put(k, data[k])


def test_rejects_bad_calls():
    lc, lf, _, _, parent = _levels(1, 6)
    with pytest.raises(B.BamhipError, match="order_RP"):
        B.prolong_varlist(lc, lf, [U0], [U0], order=5)
    with pytest.raises(B.BamhipError, match="P needs"):
        B.prolong_varlist(lc, lf, [U0], [U0], nbuffer=2)
    with pytest.raises(B.BamhipError, match="R needs"):
        B.restrict_varlist(lf, lc, [U0], [U0], nbuffer=0)
    with pytest.raises(B.BamhipError, match="not linked"):
        B.prolong_varlist(lf, lc, [U0], [U0])
    # stencil leaving the parent box: fine box starting at coarse cell 1
    cb = coarse_box(_mk, 20, 20, 20, 0.5)
    fb = fine_box(_mk, cb, (1, 6, 6), (16, 16, 16))
    lc2, lf2 = B.Level([cb], 3), B.Level([fb], 3)
    lc2.enable(range(3))
    lf2.enable(range(3))
    B.set_rp_flags(lc2, lf2, [0], 6, FP, FR)
    with pytest.raises(B.BamhipError, match="too small"):
        B.prolong_varlist(lc2, lf2, [U0], [U0], order=4)
