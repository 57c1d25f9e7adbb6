"""0d and 1d output of a device level in the reference's file formats
(bamhip_output_0d / bamhip_output_1d) against the CPU oracle of the
reductions (write_level_0d / bampi_allreduce) and of the 1d point lists
(makeorderedlist1d / bufferorderedpoints): file names, line layout, values.
max, min and max |u| are exact; the l2 norm and the integral sum in another
order than the oracle and agree to 1e-13.  The 1d files are compared as text
(the values are gathered, the coordinates computed by the same formula)."""
import numpy as np
import pytest

import baum_amd as B
import oracle as O

pytestmark = pytest.mark.gpu

H = 0.25
SHAPES = [(12, 10, 8, (-1.375, -1.125, -0.875)), (6, 10, 8, (2.125, -1.125, -0.875))]
NAMES = ["alpha", "ham"]


def _setup():
    oboxes = [O.make_box(m, n, o, *x0, H) for m, n, o, x0 in SHAPES]
    gboxes = [B.make_box(m, n, o, *x0, H) for m, n, o, x0 in SHAPES]
    rng = np.random.default_rng(4)
    fields = [[rng.standard_normal(b.m * b.n * b.o) for b in oboxes] for _ in NAMES]
    lev = B.Level(gboxes, 2, l=3)
    lev.enable(range(2))
    for v in range(2):
        lev.put(v, np.concatenate(fields[v]))
    return lev, oboxes, fields


def test_output_0d_files(tmp_path):
    lev, oboxes, fields = _setup()
    for t in (0.0, 0.125):
        for kind in B.OUTPUT_0D:
            lev.output_0d([0, 1], NAMES, kind, t, tmp_path)
    for kind in B.OUTPUT_0D:
        want = O.reduce_0d(oboxes, fields, kind)
        for v, name in enumerate(NAMES):
            text = (tmp_path / f"{name}_{kind}.l3").read_text()
            lines = text.split("\n")
            assert lines[-1] == "" and len(lines) == 3
            for t, ln in zip((0.0, 0.125), lines[:2]):
                tv, vv = (float(x) for x in ln.split())
                assert ln == f"{tv:22.15e}  {vv:22.15e}  " and tv == t
                if kind in ("max", "min", "norminf"):
                    assert vv == float(f"{want[v]:22.15e}")
                else:
                    assert abs(vv - want[v]) <= 1e-13 * abs(want[v])
    lev.free()


def test_output_1d_files(tmp_path):
    lev, oboxes, fields = _setup()
    origins = {"x": (0.0, 0.0, 0.0), "y": (0.3, 0.0, -0.2), "z": (2.5, 0.4, 0.0)}
    for line, org in origins.items():
        lev.output_1d([0, 1], NAMES, line, 0.5, tmp_path, origin=org)
    for line, org in origins.items():
        crd, bo, ix = O.line_1d(oboxes, line, origin=org)
        assert len(crd) > 0
        for v, name in enumerate(NAMES):
            rows = [f"{c:22.15e}  {(fields[v][b][i] if b >= 0 else -1.0):22.15e}\n" for c, b, i in zip(crd, bo, ix)]
            want = '"Time = 0.500000"\n' + "".join(rows) + "\n"
            got = (tmp_path / f"{name}.{line}3").read_text()
            assert got == want, (name, line)
    # the x line crosses the gap between the boxes
    crd, bo, ix = O.line_1d(oboxes, "x")
    assert (bo == -1).sum() == 2
    lev.free()
