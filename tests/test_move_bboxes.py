"""The nested boxes of find_nested_bboxes (src/main/amr/move.c:121-465),
bamhip_find_nested_bboxes, on the host: every expected box below is worked
out by hand from the rules (points snapped to the coarse grid by truncation
toward zero, n = nxyz/2 rounded up to even coarse points, half width
(n - 1) h / 2, touching or overlapping boxes joined, the joined box compared
again with all others)."""
import numpy as np
import pytest

import baum_amd as B


def boxes(points, nxyz, h=0.5):
    bb, ib = B.find_nested_bboxes(points, h, nxyz)
    return bb, ib


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
