"""Checkpoint and restart of device levels in the reference's file layout
(bamhip_checkpoint_write / _read, after checkpoint_write_local and
checkpoint_read): the file is parsed here independently (text header, level
lines, parameter database, "$variable = ..." lines each followed by the
level's doubles in binary and a newline), the data bitwise equal to the
device state; a restart into fresh levels restores every listed variable
bitwise with the iterations and times, skips what it is not asked for, and
refuses a level whose point count differs."""
import numpy as np
import pytest

import baum_amd as B

pytestmark = pytest.mark.gpu


def _levels():
    l0 = B.Level([B.make_box(10, 9, 8, -1.0, -1.0, -1.0, 0.25)], 4, l=0)
    l1 = B.Level([B.make_box(6, 7, 5, -0.6, -0.6, -0.6, 0.125), B.make_box(5, 5, 5, 0.4, 0.4, 0.4, 0.125)], 4, l=1)
    return [l0, l1]


def _parse(path):
    data = open(path, "rb").read()
    head, rest = data.split(b"$END_checkpoint_header\n\n", 1)
    params, rest = rest.split(b"$END_parameter_database\n\n", 1)
    assert rest.startswith(b"$BEGIN_variables:\n")
    rest = rest[len(b"$BEGIN_variables:\n"):]
    blocks = []
    while not rest.startswith(b"$END_variables\n\n"):
        line, rest = rest.split(b"\n", 1)
        name = line.split(b"$variable = ")[1].split(b" : ")[0].decode()
        lvl = int(line.split(b"level = ")[1].split(b" ,")[0])
        n = int(line.split(b"level->nnodes = ")[1])
        vals = np.frombuffer(rest[:8 * n], dtype=np.float64).copy()
        assert rest[8 * n:8 * n + 1] == b"\n"
        rest = rest[8 * n + 1:]
        blocks.append((name, lvl, vals))
    return head.decode(), params.decode(), blocks


def test_checkpoint_write_read_roundtrip(tmp_path):
    levels = _levels()
    rng = np.random.default_rng(9)
    data = {}
    for li, lev in enumerate(levels):
        en = [0, 1, 2] if li == 0 else [0, 1]       # variable 2 has storage on level 0 only
        lev.enable(en)
        for v in en:
            data[(li, v)] = rng.standard_normal(lev.npoints)
            lev.put(v, data[(li, v)])
    names = ["alpha", "betax", "ham"]
    path = tmp_path / "checkpoint.0"
    B.checkpoint_write(path, levels, [4, 8], [0.5, 0.5], [0, 1, 2], names, params={"order_RP": 4, "physics": "z4"})
    head, params, blocks = _parse(path)
    assert head.split("\n")[0] == "$BEGIN_checkpoint_header:  $lmin = 0 , $lmax = 1 , $time = 0.500"
    assert head.split("\n")[1:3] == ["level 0 : iteration = 4 , time = 0.5", "level 1 : iteration = 8 , time = 0.5"]
    assert params == "$BEGIN_parameter_database:  $number_of_parameters = 2\norder_RP = 4\nphysics = z4\n"
    assert [(n, l) for n, l, _ in blocks] == [("alpha", 0), ("betax", 0), ("ham", 0), ("alpha", 1), ("betax", 1)]
    for name, l, vals in blocks:
        np.testing.assert_array_equal(vals, data[(l, names.index(name))])
    # restart into fresh levels: listed variables restored, "ham" not asked for
    fresh = _levels()
    for lev in fresh:
        lev.enable([0, 1, 2])
        for v in range(3):
            lev.fill(v, -1.0)
    it, tm = B.checkpoint_read(path, fresh, [1, 0], ["betax", "alpha"])
    assert it == [4, 8] and tm == [0.5, 0.5]
    # with the level time steps the restart time is iteration * dt (the
    # header's %g keeps 6 digits): dt = 1/3 / 4 on level 0, half on level 1
    dt = [1.0 / 12.0, 1.0 / 24.0]
    it, tm = B.checkpoint_read(path, fresh, [1, 0], ["betax", "alpha"], dt=dt)
    assert it == [4, 8] and tm == [4 * dt[0], 8 * dt[1]]
    for li, lev in enumerate(fresh):
        np.testing.assert_array_equal(lev.get(0), data[(li, 0)])
        np.testing.assert_array_equal(lev.get(1), data[(li, 1)])
        np.testing.assert_array_equal(lev.get(2), -1.0)
    # a level of another size is refused
    other = [B.Level([B.make_box(10, 9, 7, -1.0, -1.0, -1.0, 0.25)], 4, l=0), fresh[1]]
    other[0].enable([0, 1])
    with pytest.raises(B.BamhipError):
        B.checkpoint_read(path, other, [0], ["alpha"])
    for lev in levels + fresh + other[:1]:
        lev.free()
