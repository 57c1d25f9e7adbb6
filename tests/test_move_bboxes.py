"""The nested boxes of find_nested_bboxes (src/main/amr/move.c:121-465),
bamhip_find_nested_bboxes, on the host: every expected box below is worked
out by hand from the rules (points snapped to the coarse grid by truncation
toward zero, the size of move.c:233-267 halved and rounded up to even coarse
points, half width (n - 1) h / 2, touching or overlapping boxes joined, the
joined box compared again with all others).  The size rules: nx (nz) of the
coarsest grid, amr_move_nxyz above amr_move_lcube, the amr_nxyz entry
min(nlist - 1, lfine), plus 2 amr_nbuffer with amr_buffer_addextra."""
import numpy as np
import pytest

import baum_amd as B


def boxes(points, nxyz, h=0.5):
    # the size rule alone: move_nxyz on a moving level, no extra buffer
    sz = B.NestedSizes(nx=nxyz, move_nxyz=nxyz, nbuffer=0, addextra=False)
    return B.find_nested_bboxes(sz, 1, points, h)


def test_single_point_snapped_by_truncation():
    bb, ib = boxes([[0.3, -0.7, 1.2]], 32)       # n = 16, half width 3.75, p = (0, -0.5, 1.0)
    np.testing.assert_array_equal(bb, [[-3.75, 3.75, -4.25, 3.25, -2.75, 4.75]])
    np.testing.assert_array_equal(ib, [[0, 15, 0, 15, 0, 15]])
    bb, _ = boxes([[-0.74, 0.74, -0.2]], 32)     # truncation, not floor: -1.48 -> -1
    np.testing.assert_array_equal(bb[0, ::2], [-4.25, -3.25, -3.75])


def test_box_size_rounded_up_to_even():
    bb, ib = boxes([[0.0, 0.0, 0.0]], 34)         # 34/2 = 17 -> 18 coarse points
    np.testing.assert_array_equal(bb, [[-4.25, 4.25] * 3])
    np.testing.assert_array_equal(ib, [[0, 17] * 3])


def test_separate_touching_and_overlapping_boxes():
    bb, ib = boxes([[-6.0, 0.0, 0.0], [6.0, 0.0, 0.0]], 16)   # n = 8, half width 1.75
    np.testing.assert_array_equal(bb, [[-7.75, -4.25, -1.75, 1.75, -1.75, 1.75],
                                       [4.25, 7.75, -1.75, 1.75, -1.75, 1.75]])
    bb, ib = boxes([[-2.0, 0.0, 0.0], [2.0, 0.0, 0.0]], 16)   # one coarse cell apart: two boxes
    assert bb.shape[0] == 2 and bb[0, 1] == -0.25 and bb[1, 0] == 0.25
    bb, ib = boxes([[-1.5, 0.0, 0.0], [2.0, 0.0, 0.0]], 16)   # touching at x = 0.25: joined
    np.testing.assert_array_equal(bb, [[-3.25, 3.75, -1.75, 1.75, -1.75, 1.75]])
    np.testing.assert_array_equal(ib, [[0, 14, 0, 7, 0, 7]])
    bb, ib = boxes([[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0]], 16)
    np.testing.assert_array_equal(bb, [[-2.75, 2.75, -1.75, 1.75, -1.75, 1.75]])
    np.testing.assert_array_equal(ib, [[0, 11, 0, 7, 0, 7]])


def test_joined_box_is_compared_again():
    # box 0 and box 1 are apart; box 2 overlaps box 0; the join reaches box 1
    bb, ib = boxes([[-3.0, 0.0, 0.0], [3.0, 0.0, 0.0], [0.0, 0.0, 0.0]], 16)
    np.testing.assert_array_equal(bb, [[-4.75, 4.75, -1.75, 1.75, -1.75, 1.75]])
    np.testing.assert_array_equal(ib, [[0, 19, 0, 7, 0, 7]])
    # three apart boxes keep their order
    bb, _ = boxes([[-6.0, 0.0, 0.0], [0.0, 6.0, 0.0], [0.0, 0.0, 6.0]], 16)
    np.testing.assert_array_equal(bb[:, ::2], [[-7.75, -1.75, -1.75], [-1.75, 4.25, -1.75], [-1.75, -1.75, 4.25]])


def test_bad_arguments():
    with pytest.raises(Exception):
        boxes([[0.0, 0.0, 0.0]], 15)
    with pytest.raises(Exception):
        boxes([[0.0, 0.0, 0.0]], 0)
    with pytest.raises(Exception):
        boxes(np.zeros((11, 3)), 16)


def test_buffer_added_as_extra_points():
    """amr_buffer_addextra = yes (the reference default) with amr_nbuffer 4:
    n = (32 + 8) / 2 = 20 coarse points, half width 19 h / 2"""
    sz = B.NestedSizes(nx=64, move_nxyz=32, nbuffer=4, addextra=True)
    assert B.nested_box_size(sz, 2) == [20, 20, 20]
    bb, ib = B.find_nested_bboxes(sz, 2, [[0.0, 0.0, 0.0]], 0.5)
    np.testing.assert_array_equal(bb, [[-4.75, 4.75] * 3])
    np.testing.assert_array_equal(ib, [[0, 19] * 3])
    # without it: 32 / 2 = 16
    assert B.nested_box_size(B.NestedSizes(nx=64, move_nxyz=32, nbuffer=4, addextra=False), 2) == [16] * 3
    # 6 buffer points: (32 + 12) / 2 = 22; (30 + 12) / 2 = 21 -> 22; (28 + 12) / 2 = 20
    for mv, want in ((32, 22), (30, 22), (28, 20)):
        assert B.nested_box_size(B.NestedSizes(nx=64, move_nxyz=mv, nbuffer=6), 3) == [want] * 3


def test_cubes_up_to_lcube_use_nx_around_the_origin():
    """levels up to amr_move_lcube: the coarsest grid's nx (not amr_move_nxyz)
    and one cube centred on the origin whatever the punctures; above it
    amr_move_nxyz around each puncture; nz sets the z size of the cubes"""
    sz = B.NestedSizes(nx=64, nz=48, move_nxyz=32, move_lcube=2, nbuffer=6, addextra=True)
    assert B.nested_box_size(sz, 1) == [38, 38, 30]          # (64 + 12) / 2, (48 + 12) / 2 = 30
    assert B.nested_box_size(sz, 2) == [38, 38, 30]
    assert B.nested_box_size(sz, 3) == [22, 22, 22]          # (32 + 12) / 2
    punc = [[-3.0, 0.4, 0.0], [3.0, 0.4, 0.0]]
    bb, ib = B.find_nested_bboxes(sz, 2, punc, 0.5)
    np.testing.assert_array_equal(bb, [[-9.25, 9.25, -9.25, 9.25, -7.25, 7.25]])
    np.testing.assert_array_equal(ib, [[0, 37, 0, 37, 0, 29]])
    bb, ib = B.find_nested_bboxes(sz, 3, punc, 0.25)         # half width 21 * 0.25 / 2 = 2.625
    np.testing.assert_array_equal(bb, [[-5.625, -0.375, -2.375, 2.875, -2.625, 2.625],
                                       [0.375, 5.625, -2.375, 2.875, -2.625, 2.625]])
    np.testing.assert_array_equal(ib, [[0, 21] * 3] * 2)


def test_amr_nxyz_list():
    """amr_nxyz: entry min(nlist - 1, lfine) for every level, multiples of 4"""
    sz = B.NestedSizes(nx=64, move_nxyz=32, nxyz_list=(64, 48, 40), nbuffer=6, addextra=True)
    assert B.nested_box_size(sz, 1) == [30] * 3                # (48 + 12) / 2
    assert B.nested_box_size(sz, 2) == [26] * 3                # (40 + 12) / 2
    assert B.nested_box_size(sz, 7) == [26] * 3
    bad = B.NestedSizes(nx=64, move_nxyz=32, nxyz_list=(64, 50), nbuffer=6)
    with pytest.raises(B.BamhipError, match="multiples of 4"):
        B.nested_box_size(bad, 1)


def test_odd_coarse_grid_is_rejected_even_with_move_nxyz():
    sz = B.NestedSizes(nx=63, move_nxyz=32, nbuffer=6)
    with pytest.raises(B.BamhipError, match="even nx"):
        B.nested_box_size(sz, 3)
