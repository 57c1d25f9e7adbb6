"""The Z4c configurations at their stated sizes, on the device against the
oracle.

C1: 64^3 interior points on [-8, 8]^3 (h = 0.25, cell centred), one
    radiative boundary layer per face (the single-process BAM box: process
    ghosts exist only between processes; the 3-point reach of the
    lop-sided stencils is served by the reference's edge rules), Brill-
    Lindquist m = 1 at the origin, precollapsed lapse, 1+log withshift,
    Gamma driver eta = 2, KO 0.1, dt = h/4, 8 RK4 steps.  Per point and
    variable after every step, with the Hamiltonian and momentum constraint
    norms tracked against the oracle.
C2: 192^3 periodic box (+3 ghosts) of flat space plus uniform noise of
    amplitude 1e-10 on [-8, 8]^3: one right-hand side and one RK4 step at
    full size; the 100-step run and its Hamiltonian history are a property
    test (bounded, finite) reported by the benchmark too.
C3: the benchmark's 7-level hierarchy (128^3 and two boxes of 48..96^3 per
    level around the binary at x = -/+3): one RK4 step of every level's real
    boxes against the oracle box by box, the R/P operators on every real
    level pair bitwise, and after a full coarse step of the device
    hierarchy the restriction identity (coarse points under a fine box are
    the restriction of the fine data, bitwise) with finite constraint norms
    per level.  A coarse step of the whole hierarchy is 2^l RK4 steps of
    level l (3.3e8 point right-hand sides), beyond the oracle's reach in a
    test; the per-level steps and R/P cover every operation of it.

Tolerances (written per test).  After n RK4 steps the state change of each
variable agrees to 1e-11 of its magnitude (the order of floating-point
operations differs: FMA contraction, association).  For C2's noise data the
right-hand side is a difference of O(1) stencil terms whose sum is ~1e-8:
its rounding floor is eps * max|u| * sum|w| / h^2 (sum|w| = 64/12 for the
second differences), so the bar is 1e-12 of the magnitude or 100 times
that floor, whichever is larger; a plain 1e-12 of the result is below
what FP64 resolves on these data.
"""
import numpy as np
import pytest

import baum_amd as B
import bench
import oracle as O
from spacetimes import flat_z4c, grid, puncture_z4c

pytestmark = pytest.mark.gpu

U, P, R, V = (list(range(25 * k, 25 * k + 25)) for k in range(4))
EPS = np.finfo(float).eps


def _rel(a, b, floor=1e-300):
    return np.array([np.abs(a[v] - b[v]).max() / max(np.abs(b[v]).max(), floor) for v in range(a.shape[0])])


def _level(obox, bflag, nvars=100):
    bx = B.make_box(obox.m, obox.n, obox.o, obox.x0, obox.y0, obox.z0, obox.dx, skip=tuple(obox.skip), bflag=bflag)
    lev = B.Level([bx], nvars)
    lev.enable(range(nvars))
    for v in range(nvars):
        lev.fill(v, 0.0)
    return lev


def test_c1_puncture_64_eight_steps():
    h, nin = 0.25, 64
    n = nin + 2
    x0 = -8.0 - 0.5 * h                     # the boundary layer at -8 - h/2; interior -8 + (i + 1/2) h
    ob = O.make_box(n, n, n, x0, x0, x0, h, skip=(1,) * 6)
    u0 = puncture_z4c(grid(ob))
    dt = 0.25 * h
    cpu = u0.copy()
    work = np.zeros((75,) + u0.shape[1:])
    lev = _level(ob, (B.PHYBOUND,) * 6)
    lev.put_state(U, u0.reshape(25, -1))
    par, opar = B.z4c_params(), O.z4c_params()
    worst = 0.0
    for step in range(8):
        assert O.z4c_rk4_step(ob, opar, dt, cpu, bc=1, pgh=0, work=work) == 0
        lev.z4c_rk4_step(par, dt, U, P, R, V)
        gpu = lev.get_state(U).reshape(u0.shape)
        err = _rel(gpu - u0, cpu - u0)
        assert np.all(np.isfinite(gpu))
        assert err.max() <= 1e-11, (step, dict(zip(B.VAR_NAMES, err)))
        worst = max(worst, err.max())
        c_cpu = O.z4c_constraints_l2(ob, cpu)
        c_gpu = np.array(lev.z4c_constraints_l2(U))
        assert np.all(np.abs(c_gpu - c_cpu) <= 1e-10 * np.abs(c_cpu)), (step, c_gpu, c_cpu)
    lev.free()
    print("C1: 8 RK4 steps of 64^3, worst relative error of the change %.2e; |H| %.6e" % (worst, c_cpu[0]),
          flush=True)


def _c2_box(n=192, g=3):
    nst = n + 2 * g
    h = 16.0 / n
    x0 = -8.0 - (g - 0.5) * h
    ob = O.make_box(nst, nst, nst, x0, x0, x0, h, skip=(g,) * 6)
    u = bench.flat_noise(nst, g, seed=20261202)
    return ob, u, h, g


def _c2_level(ob, g):
    bx = B.make_box(ob.m, ob.n, ob.o, ob.x0, ob.y0, ob.z0, ob.dx, skip=(g,) * 6, bflag=(B.PERBOUND,) * 6)
    lev = B.Level([bx], 100)
    lev.enable(range(100))
    for v in range(100):
        lev.fill(v, 0.0)
    return lev


def test_c2_rhs_192_matches_oracle():
    ob, u, h, g = _c2_box()
    cpu, nbad = O.z4c_rhs(ob, O.z4c_params(), u)
    assert nbad == 0
    lev = _c2_level(ob, g)
    lev.put_state(U, u.reshape(25, -1))
    lev.z4c_rhs(B.z4c_params(), R, U)
    gpu = lev.get_state(R).reshape(u.shape)
    lev.free()
    inner = (slice(None),) + (slice(g, -g),) * 3
    gpu, cpu = gpu[inner], cpu[inner]
    floor = 100 * EPS * np.abs(u).max() * (64.0 / 12.0) / h ** 2
    worst_rel, worst_floor = 0.0, 0.0
    for v in range(25):
        mag = np.abs(cpu[v]).max()
        err = np.abs(gpu[v] - cpu[v]).max()
        assert err <= max(1e-12 * mag, floor), (B.VAR_NAMES[v], err, mag, floor)
        worst_rel = max(worst_rel, err / max(mag, 1e-300))
        worst_floor = max(worst_floor, err / floor)
    print("C2 RHS 192^3: max error / magnitude %.2e, / (100x rounding floor) %.2e" % (worst_rel, worst_floor),
          flush=True)


def test_c2_rk4_step_192_matches_oracle():
    ob, u, h, g = _c2_box()
    dt = 0.25 * h
    cpu = u.copy()
    assert O.z4c_rk4_step(ob, O.z4c_params(), dt, cpu, bc=2, pgh=g) == 0
    lev = _c2_level(ob, g)
    lev.put_state(U, u.reshape(25, -1))
    lev.z4c_rk4_step(B.z4c_params(), dt, U, P, R, V, pgh=g)
    gpu = lev.get_state(U).reshape(u.shape)
    lev.free()
    # the change is dt * RHS; its bar follows the right-hand side's
    floor = dt * 100 * EPS * np.abs(u).max() * (64.0 / 12.0) / h ** 2
    for v in range(25):
        d_cpu, d_gpu = cpu[v] - u[v], gpu[v] - u[v]
        err = np.abs(d_gpu - d_cpu).max()
        assert err <= max(1e-11 * np.abs(d_cpu).max(), floor), (B.VAR_NAMES[v], err, np.abs(d_cpu).max(), floor)


def test_c2_hundred_steps_hamiltonian_bounded():
    """robust stability: 100 RK4 steps of the noise data; the Hamiltonian
    constraint's L2 norm stays finite and does not grow (KO damps the
    grid-scale noise), and the state stays within 1e-6 of flat space"""
    ob, u, h, g = _c2_box()
    lev = _c2_level(ob, g)
    lev.put_state(U, u.reshape(25, -1))
    del u
    par = B.z4c_params()
    hist = [lev.z4c_constraints_l2(U)[0]]
    for step in range(1, 101):
        lev.z4c_rk4_step(par, 0.25 * h, U, P, R, V, pgh=g)
        if step % 10 == 0:
            hist.append(lev.z4c_constraints_l2(U)[0])
    flat = flat_z4c((1, 1, 1)).reshape(25)
    dev = max(np.abs(lev.get(U[v]) - flat[v]).max() for v in range(25))
    lev.free()
    hist = np.array(hist)
    print("C2 |H| every 10 steps: %s; max |u - flat| %.3e" % (" ".join("%.3e" % x for x in hist), dev), flush=True)
    assert np.all(np.isfinite(hist))
    assert hist[-1] <= hist[0] and hist.max() <= 2 * hist[0]
    assert dev < 1e-6


# ------------------------------------------------------------------ C3
FP, FR = 125, 126
NVAR = 127


def _c3_boxes(make):
    boxes, parents = bench.c3_geometry(make, B.PHYBOUND, B.REFBOUND)
    return boxes, parents


def _omake(m, n, o, x0, y0, z0, h, bflag=None):
    return O.make_box(m, n, o, x0, y0, z0, h)


def _c3_params(l, oracle=False):
    """the benchmark's parameters: chi floor 1e-5, KO 0.1 with the reference's
    level rule (0.2 x on levels above amr_move_lcube = 0)"""
    f = B.dissipation_factor(l, 0.1)
    if oracle:
        return O.z4c_params(z4_chi_div_floor=1e-5, dissipation_factor=16 * f)
    p = B.z4c_params(z4_chi_div_floor=1e-5)
    p.dissfactor = f
    return p


def _c3_data(obx):
    return [bench.puncture_fields(b.x0 + b.dx * np.arange(b.m), b.y0 + b.dx * np.arange(b.n),
                                  b.z0 + b.dx * np.arange(b.o)) for b in obx]


def _ulp_sensitivity(step, u0, seed):
    """how much the oracle's own one-step change moves when its input is
    perturbed at the rounding level (u (1 + 2 eps xi), xi = +-1): the scale
    of what a different order of floating-point operations can change,
    per variable (near the punctures the metric's change is a small
    difference of 1/chi-weighted terms, and chi crosses its floor)"""
    xi = np.random.default_rng(seed).choice([-1.0, 1.0], size=u0.shape)
    up = u0 * (1 + 2 * EPS * xi)
    a, b = u0.copy(), up.copy()
    step(a)
    step(b)
    return np.abs((b - up) - (a - u0)).reshape(25, -1).max(axis=1)


@pytest.mark.parametrize("l", range(7))
def test_c3_level_step_matches_oracle(l):
    """the change of each variable agrees to 1e-11 of its magnitude, or to
    100 times the oracle's own sensitivity to rounding-level perturbations of
    its input where that is larger"""
    boxes, _ = _c3_boxes(B.make_box)
    oboxes, _ = _c3_boxes(_omake)
    obx = oboxes[l]
    us = _c3_data(obx)
    dt = 0.25 * boxes[0][0].dx / 2 ** l
    lev = B.Level(boxes[l], 100, l=l)
    lev.enable(range(100))
    for v in range(100):
        lev.fill(v, 0.0)
    lev.put_state(U, np.concatenate([u.reshape(25, -1) for u in us], axis=1))
    lev.z4c_rk4_step(_c3_params(l), dt, U, P, R, V, algcon=False)
    got = lev.get_state(U)
    lev.free()
    off = 0
    for b, (box, u0) in enumerate(zip(obx, us)):
        def step(u):
            assert O.z4c_rk4_step(box, _c3_params(l, oracle=True), dt, u, bc=1 if l == 0 else 0, pgh=0,
                                  algcon=False) == 0
        cpu = u0.copy()
        step(cpu)
        sens = _ulp_sensitivity(step, u0, 70 + 10 * l + b)
        npb = box.m * box.n * box.o
        gpu = got[:, off:off + npb].reshape(u0.shape)
        off += npb
        assert np.all(np.isfinite(gpu))
        d_cpu = (cpu - u0).reshape(25, -1)
        err = np.abs((gpu - u0).reshape(25, -1) - d_cpu).max(axis=1)
        mag = np.abs(d_cpu).max(axis=1)
        bar = np.maximum(1e-11 * mag, 100 * sens)
        assert np.all(err <= bar), (l, b, {B.VAR_NAMES[v]: (err[v], mag[v], sens[v]) for v in range(25)
                                           if err[v] > bar[v]})
        print("C3 level %d box %d: max error / change %.2e, max error / (100 x sensitivity) %.2e"
              % (l, b, np.max(err / np.maximum(mag, 1e-300)), np.max(err / np.maximum(100 * sens, 1e-300))),
              flush=True)


@pytest.mark.parametrize("l", range(1, 7))
def test_c3_level_pair_rp_bitwise(l):
    boxes, parents = _c3_boxes(B.make_box)
    oboxes, _ = _c3_boxes(_omake)
    nvar = 4
    lc, lf = B.Level(boxes[l - 1], 2 + nvar, l=l - 1), B.Level(boxes[l], 2 + nvar, l=l)
    lc.enable(range(2 + nvar))
    lf.enable(range(2 + nvar))
    B.set_rp_flags(lc, lf, parents[l], 6, 0, 1)
    rng = np.random.default_rng(40 + l)
    cv = rng.standard_normal((nvar, lc.npoints))
    fv = rng.standard_normal((nvar, lf.npoints))
    idx = list(range(2, 2 + nvar))
    for k, v in enumerate(idx):
        lc.put(v, cv[k])
        lf.put(v, fv[k])
    B.restrict_prolong_varlist(lc, lf, idx, idx, nbuffer=3, order=4)
    fp, fr = O.rp_flags(oboxes[l], oboxes[l - 1], parents[l], 6)
    np.testing.assert_array_equal(lf.get(0), fp)
    np.testing.assert_array_equal(lc.get(1), fr)
    ec, ef = cv.copy(), fv.copy()
    assert O.restrict_level(oboxes[l], oboxes[l - 1], parents[l], fr, 3, 4, ef, ec) == 0
    assert O.prolong_level(oboxes[l], oboxes[l - 1], parents[l], fp, 3, 6, 4, ec, ef) == 0
    np.testing.assert_array_equal(np.array([lc.get(v) for v in idx]), ec)
    np.testing.assert_array_equal(np.array([lf.get(v) for v in idx]), ef)
    lc.free()
    lf.free()


def test_c3_coarse_step_restriction_identity():
    UC, UP, UPP, RR, VV = (list(range(25 * k, 25 * k + 25)) for k in range(5))
    boxes, parents = _c3_boxes(B.make_box)
    oboxes, _ = _c3_boxes(_omake)
    levels = []
    for l, bl in enumerate(boxes):
        lv = B.Level(bl, NVAR, l=l)
        lv.enable(range(NVAR))
        for v in range(NVAR):
            lv.fill(v, 0.0)
        lv.put_state(UC, np.concatenate([u.reshape(25, -1) for u in _c3_data(oboxes[l])], axis=1))
        levels.append(lv)
    for l in range(1, 7):
        B.set_rp_flags(levels[l - 1], levels[l], parents[l], 6, FP, FR)
    H = B.Hierarchy(levels, [_c3_params(l) for l in range(7)], 0.25 * boxes[0][0].dx, UC, UP, UPP, RR, VV,
                    order_RP=4, timeinterpolation=3)
    H.set_check(0)
    H.advance(0)
    assert H.nonfinite() == 0
    assert list(H.iteration) == [2 ** l for l in range(7)]
    norms = []
    for l in range(7):
        norms.append(levels[l].z4c_constraints_l2(UC))
    for l in range(1, 7):
        fine = levels[l].get_state(UC)
        coarse = levels[l - 1].get_state(UC)
        fr = levels[l - 1].get(FR)
        expect = coarse.copy()
        assert O.restrict_level(oboxes[l], oboxes[l - 1], parents[l], fr, 3, 4, fine, expect) == 0
        # coarse points under the fine boxes whose restriction stencil reads
        # no fine buffer point (the buffer is refilled by the prolongation
        # that follows the restriction), minus the coarse level's own buffer
        # (refilled from its parent afterwards); a stencil touches the buffer
        # where the restriction of a positive random field held on the
        # buffer only is nonzero
        fp = levels[l].get(FP)
        ind = np.where(fp > 0, 1.0 + np.random.default_rng(l).random(fp.size), 0.0)[None, :]
        touched = np.zeros((1, coarse.shape[1]))
        assert O.restrict_level(oboxes[l], oboxes[l - 1], parents[l], fr, 3, 4, ind, touched) == 0
        under = (fr >= 3) & (touched[0] == 0)
        if l >= 2:
            under &= levels[l - 1].get(FP) == 0
        assert under.sum() > 0
        np.testing.assert_array_equal(coarse[:, under], expect[:, under])
        assert np.all(np.isfinite(fine))
    for lv in levels:
        lv.free()
    norms = np.array(norms)
    print("C3 after one coarse step, |H| per level: %s" % " ".join("%.3e" % x for x in norms[:, 0]), flush=True)
    # the ADM constraints take chi^(1/chipsipower) without the evolution's
    # chi floor: on the finest level, whose points come within h/2 of the
    # punctures, chi can dip below zero there and the norm is not finite
    assert np.all(np.isfinite(norms[:6]))
