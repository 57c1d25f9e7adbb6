"""get_dissipation_factor (src/utility/evolve/evolve.c:486-530) as
bamhip_dissipation_factor, on the host.  Every expected value is worked out
by hand: factor / 2^order with, in order of precedence, the shells factor on
shell levels, dissipation_factor_move above amr_move_lcube, 0.2 times the
factor above amr_move_lcube unless physics contains Z4d, the level-0 factor
on level 0, else the plain factor; -1 means an optional factor is unset.
The parameter sets are those of the reference's parfiles
(BBIDdataReader/par/TOV_binary_boost.par:61-74: order 6, factor 0.5,
shells 0.1, amr_move_lcube 1)."""
import pytest

import baum_amd as B


def test_plain_factor_on_levels_up_to_lcube():
    # order 4, factor 0.1: 0.1 / 16 on every level <= amr_move_lcube
    for l in range(0, 3):
        assert B.dissipation_factor(l, 0.1, 4, amr_move_lcube=2) == 0.1 / 16


def test_moving_levels_get_a_fifth_unless_z4d():
    assert B.dissipation_factor(3, 0.1, 4, amr_move_lcube=2) == 0.2 * 0.1 / 16
    assert B.dissipation_factor(3, 0.1, 4, amr_move_lcube=2, physics="Z4d") == 0.1 / 16
    # Getv is a whole-word test: "Z4dx" does not contain the word Z4d
    assert B.dissipation_factor(3, 0.1, 4, amr_move_lcube=2, physics="Z4dx") == 0.2 * 0.1 / 16
    assert B.dissipation_factor(3, 0.1, 4, amr_move_lcube=2, physics="grhd Z4d") == 0.1 / 16


def test_move_factor_overrides_the_fifth_and_z4d():
    kw = dict(amr_move_lcube=1, dissipation_factor_move=0.1)
    assert B.dissipation_factor(2, 0.5, 4, **kw) == 0.1 / 16
    assert B.dissipation_factor(2, 0.5, 4, physics="Z4d", **kw) == 0.1 / 16
    assert B.dissipation_factor(1, 0.5, 4, **kw) == 0.5 / 16      # not above lcube


def test_level0_factor_only_on_level_zero():
    kw = dict(amr_move_lcube=0, dissipation_factor_level0=0.3)
    assert B.dissipation_factor(0, 0.1, 4, **kw) == 0.3 / 16
    assert B.dissipation_factor(1, 0.1, 4, **kw) == 0.2 * 0.1 / 16
    # lcube < 0 puts level 0 above it: the move rule wins over level0
    assert B.dissipation_factor(0, 0.1, 4, amr_move_lcube=-1, dissipation_factor_level0=0.3) == 0.2 * 0.1 / 16


def test_shells_win_over_everything():
    kw = dict(amr_move_lcube=1, dissipation_factor_shells=0.1, dissipation_factor_move=0.7,
              dissipation_factor_level0=0.9)
    assert B.dissipation_factor(4, 0.5, 6, shells=True, **kw) == 0.1 / 64
    assert B.dissipation_factor(0, 0.5, 6, shells=True, **kw) == 0.1 / 64
    # not a shell level: the shells factor plays no part
    assert B.dissipation_factor(4, 0.5, 6, **kw) == 0.7 / 64
    # shells level with the shells factor unset falls through to the other rules
    assert B.dissipation_factor(4, 0.5, 6, shells=True, amr_move_lcube=1) == 0.2 * 0.5 / 64


def test_tov_binary_boost_parfile():
    # order 6, factor 0.5, shells 0.1, lcube 1, physics without Z4d
    kw = dict(dissipation_factor=0.5, order_dissipation=6, amr_move_lcube=1, dissipation_factor_shells=0.1)
    assert [B.dissipation_factor(l, **kw) for l in range(4)] == [0.5 / 64, 0.5 / 64, 0.1 / 64, 0.1 / 64]


def test_bad_level_refused():
    with pytest.raises(B.BamhipError):
        B.dissipation_factor(-1)
