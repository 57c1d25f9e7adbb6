"""bamhip_level_reduce_noalloc (the reduction that allocates nothing) called
through the C ABI against bamhip_level_reduce: bitwise equal results on two
calls in a row, the second on a smaller level, and a clear error for a level
beyond the fixed scratch."""
import ctypes as C

import numpy as np
import pytest

import baum_amd as B

pytestmark = pytest.mark.gpu

DP = C.POINTER(C.c_double)


def _bind():
    L = B.lib()
    for name in ("bamhip_level_reduce", "bamhip_level_reduce_noalloc"):
        f = getattr(L, name)
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_int, DP]
    return L


def _level(sizes, nvars, seed):
    boxes = [B.make_box(m, n, o, 10.0 * b, 0.0, 0.0, 0.5, skip=(1, 2, 1, 1, 3, 1)) for b, (m, n, o) in
             enumerate(sizes)]
    lev = B.Level(boxes, nvars)
    lev.enable(range(nvars))
    rng = np.random.default_rng(seed)
    for v in range(nvars):
        lev.put(v, rng.standard_normal(lev.npoints) * (1 + v))
    return lev


def _both(L, lev, vars_, kind):
    idx = (C.c_int * len(vars_))(*vars_)
    a = np.zeros(len(vars_))
    b = np.full(len(vars_), np.nan)
    assert L.bamhip_level_reduce(lev.h, len(vars_), idx, kind, a.ctypes.data_as(DP)) == 0, B.lib().bamhip_last_error()
    rc = L.bamhip_level_reduce_noalloc(lev.h, len(vars_), idx, kind, b.ctypes.data_as(DP))
    assert rc == 0, B.lib().bamhip_last_error()
    return a, b


def test_noalloc_reduce_matches_allocating_reduce_twice():
    B.init()
    L = _bind()
    big = _level([(70, 64, 58), (40, 52, 36)], 5, seed=3)
    small = _level([(21, 17, 19)], 3, seed=4)
    for lev, vars_ in ((big, [0, 2, 4, 1]), (small, [2, 0])):
        for kind in range(5):
            a, b = _both(L, lev, vars_, kind)
            np.testing.assert_array_equal(b, a)
            assert np.all(np.isfinite(b))
    # the smaller second call leaves no trace of the first one's partials
    a, b = _both(L, small, [1], 2)
    np.testing.assert_array_equal(b, a)
    big.free()
    small.free()


def test_noalloc_reduce_rejects_levels_beyond_the_scratch():
    B.init()
    L = _bind()
    lev = _level([(66, 66, 66)] * 5, 1, seed=5)     # 5 x 1024 reduction blocks > 4096
    out = np.zeros(1)
    idx = (C.c_int * 1)(0)
    rc = L.bamhip_level_reduce_noalloc(lev.h, 1, idx, 0, out.ctypes.data_as(DP))
    assert rc == -1                                  # BAMHIP_EINVAL
    assert b"reduction blocks" in B.lib().bamhip_last_error()
    lev.free()
