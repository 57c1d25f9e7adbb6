"""Moving boxes on the device (bamhip_move_box, bamhip_level_copy_overlap):
the fine levels of a hierarchy follow the punctures.  The new boxes are
those of bamhip_find_nested_bboxes (pinned on the host); the points a new
box shares with the old one hold the old values bit for bit, every other
point holds the order-4 prolongation from the parent, which reproduces the
cubic data below to round-off.  The old fine data carry a marker (+1000) so
that copied and prolonged points cannot be confused.  Then the move inside
advance_levelandsublevels on puncture data, with the Python level handles
following the library's replacements."""
import numpy as np
import pytest

import baum_amd as B
from spacetimes import grid, puncture_z4c

pytestmark = pytest.mark.gpu

U, UP, UPP, R, V = (list(range(25 * k, 25 * k + 25)) for k in range(5))
FP, FR = 125, 126
NVAR, NB, ORDER, NXYZ = 127, 6, 4, 24
H0 = 0.5
# 24 points per direction on the moving levels, without the extra buffer
# points (the size rules themselves are pinned on the host)
SIZES = B.NestedSizes(nx=NXYZ, move_nxyz=NXYZ, nbuffer=NB, addextra=False)


def coarse_level():
    b = B.make_box(40, 40, 40, -9.75, -9.75, -9.75, H0, bflag=(B.PHYBOUND,) * 6)
    return B.Level([b], NVAR, l=0)


def nested_boxes(points, h):
    """the fine boxes move_box makes on a parent of spacing h"""
    bb, ib = B.find_nested_bboxes(SIZES, 1, points, h)
    out = []
    for b, i in zip(bb, ib):
        d = h / 2
        out.append(B.make_box(2 * (i[1] + 1), 2 * (i[3] + 1), 2 * (i[5] + 1), b[0] - d / 2, b[2] - d / 2,
                              b[4] - d / 2, d, skip=(1,) * 6, bflag=(B.REFBOUND,) * 6))
    return out


def coords(lv):
    xs = []
    for b in lv.boxes:
        Z, Y, X = np.meshgrid(b.z0 + b.dz * np.arange(b.o), b.y0 + b.dy * np.arange(b.n),
                              b.x0 + b.dx * np.arange(b.m), indexing="ij")
        xs.append(np.stack([X.ravel(), Y.ravel(), Z.ravel()]))
    return np.concatenate(xs, axis=1)


def cubic(v, x, y, z):
    return (v + 1) * 0.01 + 0.3 * x - 0.2 * y * y + 0.05 * x * y * z + 0.02 * z ** 3 + 0.001 * v * x ** 3 * y


def fill_state(lv, marker):
    x, y, z = coords(lv)
    for v in range(NVAR):
        lv.fill(v, 0.0)
    for v in U + UP + UPP:
        lv.put(v, cubic(v, x, y, z) + marker)


def parents(lc, boxes):
    out = []
    for f in boxes:
        for k, c in enumerate(lc.boxes):
            if (c.x0 <= f.x0 and f.x0 + (f.m - 1) * f.dx <= c.x0 + (c.m - 1) * c.dx and c.y0 <= f.y0
                    and f.y0 + (f.n - 1) * f.dy <= c.y0 + (c.n - 1) * c.dy and c.z0 <= f.z0
                    and f.z0 + (f.o - 1) * f.dz <= c.z0 + (c.o - 1) * c.dz):
                out.append(k)
                break
    assert len(out) == len(boxes)
    return out


def build(points, nlev, markers):
    levels = [coarse_level()]
    for l in range(1, nlev):
        h = levels[-1].boxes[0].dx
        lv = B.Level(nested_boxes(points, h), NVAR, l=l)
        levels.append(lv)
    for lv, mk in zip(levels, markers):
        lv.enable(range(NVAR))
        fill_state(lv, mk)
    for l in range(1, nlev):
        B.set_rp_flags(levels[l - 1], levels[l], parents(levels[l - 1], levels[l].boxes), NB, FP, FR)
    return levels


def shared(old, new):
    """for every point of the new level the old level index at the same
    coordinates (or -1)"""
    xo = coords(old)
    xn = coords(new)
    idx = np.full(xn.shape[1], -1)
    off = 0
    for b in old.boxes:
        i = np.rint((xn[0] - b.x0) / b.dx).astype(int)
        j = np.rint((xn[1] - b.y0) / b.dy).astype(int)
        k = np.rint((xn[2] - b.z0) / b.dz).astype(int)
        ok = (i >= 0) & (i < b.m) & (j >= 0) & (j < b.n) & (k >= 0) & (k < b.o)
        q = off + i + b.m * j + b.m * b.n * k
        ok &= np.all(np.abs(xo[:, np.where(ok, q, 0)] - xn) < 1e-10, axis=0)
        idx[ok] = q[ok]
        off += b.m * b.n * b.o
    return idx


def check_level(old_vals, new, idx, marker_new_region):
    x, y, z = coords(new)
    hit = idx >= 0
    assert 0 < hit.sum() < idx.size
    for v in U + UP + UPP:
        got = new.get(v)
        np.testing.assert_array_equal(got[hit], old_vals[v][idx[hit]])
        want = cubic(v, x, y, z)[~hit] + marker_new_region
        err = np.abs(got[~hit] - want).max()
        assert err <= 1e-12 * np.abs(want).max(), (v, err)
    for v in R + V:
        np.testing.assert_array_equal(new.get(v), 0.0)


def hierarchy(levels, points, dt0=0.125):
    punc = B.DeviceArray.from_host(np.array(points, dtype=np.float64))
    H = B.Hierarchy(levels, [B.z4c_params()] * len(levels), dt0, U, UP, UPP, R, V, order_RP=ORDER,
                    timeinterpolation=3, punctures=punc, sizes=SIZES, rp_flags=(NB, FP, FR))
    H.set_check(0)
    return H


def test_move_one_level_copies_overlap_and_prolongs_the_rest():
    levels = build([[0.0, 0.0, 0.0]], 2, [0.0, 1000.0])
    old_vals = {v: levels[1].get(v) for v in U + UP + UPP}
    old_l1 = levels[1]
    new_pos = [[0.9, -0.6, 0.3]]
    H = hierarchy(levels, new_pos)
    old_h = old_l1.h.value
    assert H.move_box(1)
    new = H.levels[1]
    assert new.h.value != old_h and old_l1.h is None
    # the box of find_nested_bboxes: p = (0.5, -0.5, 0), half width 2.75
    b = new.boxes[0]
    assert (b.m, b.n, b.o) == (24, 24, 24)
    np.testing.assert_allclose([b.x0, b.y0, b.z0], [-2.375, -3.375, -2.875], rtol=0, atol=1e-14)
    # the old level is gone: compare with a host copy of its geometry
    old_geom = B.Level.__new__(B.Level)
    old_geom.boxes = nested_boxes([[0.0, 0.0, 0.0]], H0)
    idx = shared(old_geom, new)
    check_level(old_vals, new, idx, 0.0)
    # the same punctures again: nothing moves
    assert not H.move_box(1)
    assert H.levels[1] is new
    # the level can be prolonged into and restricted from again
    B.restrict_prolong_varlist(H.levels[0], new, U, U, NB, ORDER)


def test_move_two_levels_relinks_children():
    levels = build([[0.0, 0.0, 0.0]], 3, [0.0, 0.0, 1000.0])
    old_vals = {v: levels[2].get(v) for v in U + UP + UPP}
    old_geom = B.Level.__new__(B.Level)
    old_geom.boxes = list(levels[2].boxes)
    H = hierarchy(levels, [[0.9, -0.6, 0.3]])
    assert H.move_box(1)
    l1, l2 = H.levels[1], H.levels[2]
    np.testing.assert_allclose([l1.boxes[0].x0, l2.boxes[0].x0], [-2.375, 0.75 - 1.375 - 0.0625], atol=1e-14)
    # level 1: the cubic everywhere (copied or prolonged)
    x, y, z = coords(l1)
    for v in U + UP + UPP:
        want = cubic(v, x, y, z)
        assert np.abs(l1.get(v) - want).max() <= 1e-12 * np.abs(want).max()
    check_level(old_vals, l2, shared(old_geom, l2), 0.0)
    # the time condition: level 1 at the time of level 0 keeps level 2 in place
    H.iteration[0], H.iteration[1], H.iteration[2] = 1, 2, 4
    H.punctures.put(np.array([[1.0, 0.0, 0.0]]))
    assert not H.move_box(2)
    H.iteration[0], H.iteration[1], H.iteration[2] = 1, 1, 2
    assert H.move_box(2)
    assert H.levels[1] is l1 and H.levels[2] is not l2


def test_boxes_follow_punctures_during_evolution():
    pts = [[0.0, 0.0, 0.0]]
    levels = build(pts, 3, [0.0, 0.0, 0.0])
    for lv in levels:
        u = np.concatenate([puncture_z4c(grid(b), ((0.5, (0.0, 0.0, 0.0)),)).reshape(25, -1) for b in lv.boxes],
                           axis=1)
        lv.put_state(U, u)
    H = hierarchy(levels, [[0.9, -0.6, 0.3]])
    H.advance(0)
    assert H.nonfinite() == 0
    pos = H.punctures.get()
    for l in (1, 2):
        h = H.levels[l - 1].boxes[0].dx
        want = nested_boxes(pos, h)[0]
        got = H.levels[l].boxes[0]
        np.testing.assert_allclose([got.x0, got.y0, got.z0], [want.x0, want.y0, want.z0], atol=1e-12)
        for v in U:
            assert np.isfinite(H.levels[l].get(v)).all()
    assert abs(H.levels[1].boxes[0].x0 - (-2.875)) > 0.1      # level 1 left the origin
    H.advance(0)
    assert H.nonfinite() == 0


def test_move_keeps_a_deferred_nonfinite_flag():
    """a non-finite RHS recorded on a fine level in check mode 0 is still
    reported after move_box replaced that level"""
    levels = build([[0.0, 0.0, 0.0]], 2, [0.0, 0.0])
    H = hierarchy(levels, [[0.9, -0.6, 0.3]])
    assert H.nonfinite(reset=True) == 0
    bad = levels[1].get(U[18])
    bad[bad.size // 2] = np.nan
    levels[1].put(U[18], bad)
    levels[1].z4c_rhs(B.z4c_params(), R, U)          # check mode 0: the flag is only recorded
    old = levels[1]
    assert H.move_box(1)
    assert H.levels[1] is not old
    assert H.nonfinite(reset=True) == 1
    assert H.nonfinite(reset=True) == 0