"""Pins the CPU oracle's FMR coupling (R/P flags, regions, prolongation,
restriction) against the reference's own data and what the mathematics fixes.

* golden vectors: src/utility/interpolate/test/lagrange_{2,4,6,8}_fine.dat
  (the reference's 1-D prolongation of ref.dat) reproduced by the 3-D
  box-to-box prolongation along each of x, y and z;
* known answers: the R/P flag rows quoted in flag.c:104-118 for nbuffer 6
  and 3;
* brute force: every prolonged and restricted value against an independent
  tensor-product Lagrange sum over the centred window (catches window and
  index offsets, which a polynomial test cannot see);
* closed form: polynomials of degree N-1 per direction are reproduced;
* the reference's control flow: the fbox0 interior is left untouched, boxes
  whose stencil leaves the parent are rejected.
"""
import os

import numpy as np
import pytest

import oracle as O
from fmr_geometry import centred_window, coarse_box, coords, fine_box, lagrange_weights

GOLD = os.path.join(os.path.dirname(__file__), "golden", "interpolate")
EMPTY = [2 ** 31 - 1, -2 ** 31, 2 ** 31 - 1, -2 ** 31, 2 ** 31 - 1, -2 ** 31]


def _gold(name):
    return np.loadtxt(os.path.join(GOLD, name), comments="#")


@pytest.mark.parametrize("N", [2, 4, 6, 8])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_prolong_box_reproduces_reference_1d_golden(N, axis):
    ref = _gold("ref.dat")[:, 1]
    gold = _gold("lagrange_%d_fine.dat" % N)[:, 1]
    dims = [N + 2] * 3                       # (x, y, z) extents
    dims[axis] = ref.size
    shape = (dims[2], dims[1], dims[0])
    bshape = [1, 1, 1]
    bshape[2 - axis] = ref.size
    cv = np.ascontiguousarray(np.broadcast_to(ref.reshape(bshape), shape))
    nf = [(d - 1 - N) * 2 + 2 for d in dims]
    cbox = [0, dims[0] - 1, 0, dims[1] - 1, 0, dims[2] - 1]
    fbox = [0, nf[0] - 1, 0, nf[1] - 1, 0, nf[2] - 1]
    fv = np.zeros((nf[2], nf[1], nf[0]))
    O.prolong_box(N, cbox, fbox, EMPTY, cv, fv)
    line = np.moveaxis(fv, 2 - axis, 0)[:, 0, 0]
    # fine point f sits at coarse coordinate N/2 - 1/4 + f/2: golden index f + N - 1.
    # The golden run leaves its first point with a full stencil unset (0, or
    # the boundary value 1 for N = 2), so the comparison starts at f = 1.
    # The golden values carry their own rounding (up to ~9e-15 from the
    # harness's weights); a wrong window or weight shows up at O(h) ~ 1e-2.
    expect = gold[N - 1:N - 1 + nf[axis]]
    assert np.all(expect[1:] != 0)
    np.testing.assert_allclose(line[1:], expect[1:], rtol=0, atol=2e-14)


def _flag_line(nbuffer, m_f=24):
    cb = coarse_box(O.make_box, 32, 32, 32, 1.0)
    fb = fine_box(O.make_box, cb, (10, 10, 10), (m_f, m_f, m_f))
    fp, fr = O.rp_flags([fb], [cb], [0], nbuffer)
    fp = fp.reshape(m_f, m_f, m_f)
    fr = fr.reshape(32, 32, 32)
    return fp[m_f // 2, m_f // 2, :], fr[16, 16, :]


def test_rp_flags_nbuffer6_match_flag_c_comment():
    fline, cline = _flag_line(6)
    # flag.c:114-116: coarse 3 3 3 2 2 1 0 0 over fine 0 0 0 0 0 0 6 5 4 3 2 1
    np.testing.assert_array_equal(fline[-12:], [0, 0, 0, 0, 0, 0, 6, 5, 4, 3, 2, 1])
    np.testing.assert_array_equal(fline[:6], [1, 2, 3, 4, 5, 6])
    # fine box covers coarse cells 10..21; to its right the coarse level is 0
    np.testing.assert_array_equal(cline[16:24], [3, 3, 3, 2, 2, 1, 0, 0])
    np.testing.assert_array_equal(cline[8:13], [0, 0, 1, 2, 2])


def test_rp_flags_nbuffer3_match_flag_c_comment():
    fline, cline = _flag_line(3)
    # flag.c:111-112: coarse 3 3 3 3 2 1 0 0 over fine 0 0 0 0 0 0 0 0 0 3 2 1
    np.testing.assert_array_equal(fline[-12:], [0] * 9 + [3, 2, 1])
    np.testing.assert_array_equal(cline[16:24], [3, 3, 3, 3, 2, 1, 0, 0])


def test_rp_flags_three_d_shell():
    m, nb = 20, 6
    cb = coarse_box(O.make_box, 30, 30, 30, 0.5)
    fb = fine_box(O.make_box, cb, (10, 10, 10), (m, m, m))
    fp, fr = O.rp_flags([fb], [cb], [0], nb)
    i = np.arange(m)
    d = np.minimum(i + 1, m - i)
    dist = np.minimum(np.minimum(d[:, None, None], d[None, :, None]), d[None, None, :])
    np.testing.assert_array_equal(fp.reshape(m, m, m), np.where(dist <= nb, dist, 0))
    # restriction flags live on the coarse footprint only
    fr = fr.reshape(30, 30, 30)
    assert fr[:10].max() == 0 and fr[20:].max() == 0
    assert np.count_nonzero(fr == 3) == (m // 2 - nb) ** 3


def _nested(N, sizes=((24, 20, 16),), c0s=((6, 7, 8),), cdims=(30, 28, 26), H=0.5):
    cb = coarse_box(O.make_box, *cdims, H)
    fbs = [fine_box(O.make_box, cb, c0, s) for c0, s in zip(c0s, sizes)]
    return cb, fbs


def _brute_prolong(cb, fb, cv, N):
    """tensor-product Lagrange over the centred coarse window at every fine point"""
    xs = [(c - x0) / cb.dx for c, x0 in zip(coords(fb), (cb.x0, cb.y0, cb.z0))]
    W, S = [], []
    for xi in xs:
        s_w = [centred_window(v, N) for v in xi]
        S.append(np.array([s for s, _ in s_w]))
        W.append(np.array([lagrange_weights(N, t) for _, t in s_w]))
    out = np.zeros((fb.o, fb.n, fb.m))
    for c in range(N):
        for b in range(N):
            for a in range(N):
                g = cv[(S[2] + c)[:, None, None], (S[1] + b)[None, :, None], (S[0] + a)[None, None, :]]
                out += W[2][:, c][:, None, None] * W[1][:, b][None, :, None] * W[0][:, a][None, None, :] * g
    return out


def _brute_restrict(cb, fb, fv, N, region):
    xs = [(c - x0) / fb.dx for c, x0 in zip(coords(cb), (fb.x0, fb.y0, fb.z0))]
    out = np.full((cb.o, cb.n, cb.m), np.nan)
    i0, i1, j0, j1, k0, k1 = region
    sw = [[centred_window(v, N) for v in xi] for xi in xs]
    for k in range(k0, k1 + 1):
        sk, tk = sw[2][k]
        wk = lagrange_weights(N, tk)
        for j in range(j0, j1 + 1):
            sj, tj = sw[1][j]
            wj = lagrange_weights(N, tj)
            for i in range(i0, i1 + 1):
                si, ti = sw[0][i]
                wi = lagrange_weights(N, ti)
                blk = fv[sk:sk + N, sj:sj + N, si:si + N]
                out[k, j, i] = np.einsum("c,b,a,cba->", wk, wj, wi, blk)
    return out


@pytest.mark.parametrize("N", [2, 4, 6, 8])
def test_prolong_level_matches_brute_force(N):
    nb = 6
    cb, (fb,) = _nested(N)
    rng = np.random.default_rng(N)
    cv = rng.standard_normal((1, cb.o, cb.n, cb.m))
    fv0 = rng.standard_normal((1, fb.o, fb.n, fb.m))
    fp, _ = O.rp_flags([fb], [cb], [0], nb)
    fv = fv0.copy()
    assert O.prolong_level([fb], [cb], [0], fp, 3, nb, N, cv.reshape(1, -1), fv.reshape(1, -1)) == 0
    shell = (fp > 0).reshape(fb.o, fb.n, fb.m)
    expect = np.where(shell, _brute_prolong(cb, fb, cv[0], N), fv0[0])
    np.testing.assert_array_equal(fv[0][~shell], fv0[0][~shell])
    np.testing.assert_allclose(fv[0], expect, rtol=0, atol=1e-12 * 8 ** 3)


@pytest.mark.parametrize("N", [2, 4, 6, 8])
def test_restrict_level_matches_brute_force(N):
    nb = 6
    cb, (fb,) = _nested(N)
    rng = np.random.default_rng(10 + N)
    fv = rng.standard_normal((1, fb.o, fb.n, fb.m))
    cv0 = rng.standard_normal((1, cb.o, cb.n, cb.m))
    fp, fr = O.rp_flags([fb], [cb], [0], nb)
    n, icbox, ifbox = O.restrict_region(fb, cb, fr, 3, N)
    assert n == (12 - nb) * (10 - nb) * (8 - nb)
    # restricted parents: coarse cells c0 + nb/2 ... of the fine box
    assert icbox == [6 + 3, 6 + 11 - 3, 7 + 3, 7 + 9 - 3, 8 + 3, 8 + 7 - 3]
    cv = cv0.copy()
    assert O.restrict_level([fb], [cb], [0], fr, 3, N, fv.reshape(1, -1), cv.reshape(1, -1)) == 0
    expect = _brute_restrict(cb, fb, fv[0], N, icbox)
    inreg = ~np.isnan(expect)
    np.testing.assert_array_equal(cv[0][~inreg], cv0[0][~inreg])
    np.testing.assert_allclose(cv[0][inreg], expect[inreg], rtol=0, atol=1e-12)


def _poly(N, rng):
    c = rng.standard_normal((N, N, N))

    def f(x, y, z):
        X = np.stack([x ** a for a in range(N)])
        Y = np.stack([y ** a for a in range(N)])
        Z = np.stack([z ** a for a in range(N)])
        return np.einsum("abc,cz,by,ax->zyx", c, Z, Y, X)
    return f


@pytest.mark.parametrize("N", [2, 4, 6])
def test_polynomials_of_degree_N_minus_1_are_exact(N):
    nb = 6
    cb, (fb,) = _nested(N, H=0.25)
    f = _poly(N, np.random.default_rng(3))
    cv = f(*coords(cb))[None].copy()
    fv = np.zeros((1, fb.o, fb.n, fb.m))
    fp, fr = O.rp_flags([fb], [cb], [0], nb)
    O.prolong_level([fb], [cb], [0], fp, 3, nb, N, cv.reshape(1, -1), fv.reshape(1, -1))
    shell = (fp > 0).reshape(fb.o, fb.n, fb.m)
    exact_f = f(*coords(fb))
    scale = np.abs(cv).max()
    np.testing.assert_allclose(fv[0][shell], exact_f[shell], rtol=0, atol=1e-13 * scale)
    cv2 = np.zeros_like(cv)
    fv_exact = exact_f[None].copy()
    O.restrict_level([fb], [cb], [0], fr, 3, N, fv_exact.reshape(1, -1), cv2.reshape(1, -1))
    sel = cv2[0] != 0
    assert sel.sum() == (12 - nb) * (10 - nb) * (8 - nb)
    np.testing.assert_allclose(cv2[0][sel], cv[0][sel], rtol=0, atol=1e-13 * scale)


def test_prolong_regions_and_interior_skip():
    N, nb = 4, 6
    cb, (fb,) = _nested(N)
    fp, _ = O.rp_flags([fb], [cb], [0], nb)
    n, ifbox, ifbox0, icbox = O.prolong_region(fb, cb, fp, 3, nb, N)
    assert ifbox == [0, 23, 0, 19, 0, 15]
    assert ifbox0 == [6, 17, 6, 13, 6, 9]
    assert n == 24 * 20 * 16 - 12 * 8 * 4
    # first fine point is the left child of coarse c0: snap down gives c0-1
    assert icbox == [6 - 1 - 1, 6 + 11 + 2, 7 - 2, 7 + 9 + 2, 8 - 2, 8 + 7 + 2]
    # nbuffer == 1: only flag-1 points; nbuffer < 0: every point
    n1, ifb1, _, _ = O.prolong_region(fb, cb, fp, 1, nb, N)
    assert n1 == 24 * 20 * 16 - 22 * 18 * 14 and ifb1 == [0, 23, 0, 19, 0, 15]
    na, _, ifb0a, _ = O.prolong_region(fb, cb, fp, -1, nb, N)
    assert na == 24 * 20 * 16 and ifb0a == EMPTY


def test_two_fine_boxes_in_one_parent_are_independent():
    N, nb = 4, 6
    cb = coarse_box(O.make_box, 40, 24, 24, 0.5)
    f1 = fine_box(O.make_box, cb, (4, 5, 4), (20, 16, 16))
    f2 = fine_box(O.make_box, cb, (22, 4, 5), (24, 16, 14))
    rng = np.random.default_rng(1)
    cv = rng.standard_normal((2, O.level_points([cb])))
    fv = rng.standard_normal((2, O.level_points([f1, f2])))
    fp, fr = O.rp_flags([f1, f2], [cb], [0, 0], nb)
    out = fv.copy()
    assert O.prolong_level([f1, f2], [cb], [0, 0], fp, 3, nb, N, cv, out) == 0
    n1 = O.level_points([f1])
    for b, (fb, sl) in enumerate(((f1, slice(0, n1)), (f2, slice(n1, None)))):
        fp1, fr1 = O.rp_flags([fb], [cb], [0], nb)
        one = np.ascontiguousarray(fv[:, sl])
        O.prolong_level([fb], [cb], [0], fp1, 3, nb, N, cv, one)
        np.testing.assert_array_equal(out[:, sl], one)
    cout = cv.copy()
    assert O.restrict_level([f1, f2], [cb], [0, 0], fr, 3, N, out, cout) == 0
    # parents with flag 3: the coarse cells of each box minus nb/2 on every side
    assert np.count_nonzero(cout != cv) == 2 * ((10 - 6) * (8 - 6) * (8 - 6) + (12 - 6) * (8 - 6) * (7 - 6))


def test_stencil_outside_parent_is_rejected():
    N, nb = 4, 6
    cb = coarse_box(O.make_box, 20, 20, 20, 0.5)
    fb = fine_box(O.make_box, cb, (1, 6, 6), (16, 16, 16))   # needs coarse index -1
    fp, _ = O.rp_flags([fb], [cb], [0], nb)
    cv = np.zeros((1, O.level_points([cb])))
    fv = np.zeros((1, O.level_points([fb])))
    assert O.prolong_level([fb], [cb], [0], fp, 3, nb, N, cv, fv) == -1
