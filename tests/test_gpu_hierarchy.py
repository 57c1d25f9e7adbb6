"""Berger-Oliger evolution of a three-level FMR hierarchy on the device
(bamhip_advance_levelandsublevels) against the oracle composed step by step
(RK4 per box, time interpolation, algcon, restriction and prolongation in
the reference's order), and graph replay of a coarse step against eager
execution."""
import numpy as np
import pytest

import baum_amd as B
import oracle as O
from fmr_geometry import small_hierarchy
from spacetimes import grid, puncture_z4c

pytestmark = pytest.mark.gpu

U, UP, UPP, R, V = (list(range(25 * k, 25 * k + 25)) for k in range(5))
FP, FR = 125, 126
NVAR = 127
NB, ORDER, TI = 6, 4, 3
BINARY = ((0.5, (-3.0, 0.0, 0.0)), (0.5, (3.0, 0.0, 0.0)))


def _initial(boxes):
    return [np.concatenate([puncture_z4c(grid(b), BINARY).reshape(25, -1) for b in lev], axis=1) for lev in boxes]


def _gpu_hierarchy(dt0, punctures=None):
    mk0 = lambda *a: B.make_box(*a, bflag=(B.PHYBOUND,) * 6)
    mkf = lambda *a: B.make_box(*a, bflag=(B.REFBOUND,) * 6)
    boxes, parents = small_hierarchy(mk0, mkf)
    levels = []
    for l, bl in enumerate(boxes):
        lv = B.Level(bl, NVAR, l=l)
        lv.enable(range(NVAR))
        for v in range(NVAR):
            lv.fill(v, 0.0)
        levels.append(lv)
    for lv, u in zip(levels, _initial(boxes)):
        lv.put_state(U, u)
    for l in range(1, len(levels)):
        B.set_rp_flags(levels[l - 1], levels[l], parents[l], NB, FP, FR)
    par = B.z4c_params()
    H = B.Hierarchy(levels, [par] * len(levels), dt0, U, UP, UPP, R, V, order_RP=ORDER, timeinterpolation=TI,
                    punctures=punctures)
    H.set_check(0)
    return H, levels


class OracleHierarchy:
    """the same recursion with the CPU oracle (main.c:397-453, evolve.c:298-380,
    inject.c:116-132)"""

    def __init__(self, dt0, punctures=None):
        self.punc = None if punctures is None else np.array(punctures, dtype=np.float64)
        self.boxes, self.parents = small_hierarchy(O.make_box, O.make_box)
        self.u = _initial(self.boxes)
        self.upp = [np.zeros_like(u) for u in self.u]
        self.work = [[np.zeros((75, b.o, b.n, b.m)) for b in lev] for lev in self.boxes]
        self.it = [0] * len(self.boxes)
        self.dt0 = dt0
        self.par = O.z4c_params()
        self.flags = [None] + [O.rp_flags(self.boxes[l], self.boxes[l - 1], self.parents[l], NB)
                               for l in range(1, len(self.boxes))]

    def _level_step(self, l):
        lmax = len(self.boxes) - 1
        dt = self.dt0 / 2 ** l
        o = min(TI, self.it[l] + 2)
        off = 0
        for b, box in enumerate(self.boxes[l]):
            npb = box.m * box.n * box.o
            sl = slice(off, off + npb)
            ub = np.ascontiguousarray(self.u[l][:, sl].reshape(25, box.o, box.n, box.m))
            upb = self.work[l][b][0:25]
            if l < lmax:
                upp = 0.5 * ub if o == 2 else 0.75 * ub + (-0.125) * upb
            assert O.z4c_rk4_step(box, self.par, dt, ub, bc=1 if l == 0 else 0, pgh=0, algcon=False,
                                  work=self.work[l][b]) == 0
            if l < lmax:
                upp = upp + (0.5 if o == 2 else 0.375) * ub
                self.upp[l][:, sl] = upp.reshape(25, -1)
            O.z4c_algcon(box, ub)
            self.u[l][:, sl] = ub.reshape(25, -1)
            off += npb
        self.it[l] += 1

    def advance(self, l=0):
        self._level_step(l)
        if l < len(self.boxes) - 1:
            self.advance(l + 1)
            self.advance(l + 1)
        if l > 0:
            fp, fr = self.flags[l]
            aligned = self.it[l] == 2 * self.it[l - 1]
            cv = self.u[l - 1] if aligned else self.upp[l - 1]
            args = (self.boxes[l], self.boxes[l - 1], self.parents[l])
            assert O.restrict_level(*args, fr, 3, ORDER, self.u[l], cv) == 0
            assert O.prolong_level(*args, fp, 3, NB, ORDER, cv, self.u[l]) == 0
        if l == len(self.boxes) - 1 and self.punc is not None:
            self._track(l)

    def _track(self, l):
        """analyze_level on the finest level: track_puncture with the shift of
        u and of u_p (the state at the start of the step)"""
        offs = np.cumsum([0] + [b.m * b.n * b.o for b in self.boxes[l]])
        beta = [[self.u[l][19 + a, offs[b]:offs[b + 1]] for b in range(len(self.boxes[l]))] for a in range(3)]
        beta_p = [[self.work[l][b][19 + a].reshape(-1) for b in range(len(self.boxes[l]))] for a in range(3)]
        for p in range(len(self.punc)):
            self.punc[p] = O.track_puncture(self.boxes[l], beta, beta_p, self.dt0 / 2 ** l, self.punc[p], order=ORDER)


def _rel_err(a, b):
    scale = np.maximum(np.abs(b).max(axis=1), 1e-300)
    return np.abs(a - b).max(axis=1) / scale


def test_hierarchy_step_matches_oracle():
    dt0 = 0.125
    H, levels = _gpu_hierarchy(dt0)
    ref = OracleHierarchy(dt0)
    u0 = [u.copy() for u in ref.u]
    for step in range(2):
        H.advance(0)
        ref.advance(0)
        assert H.nonfinite() == 0
        assert list(H.iteration) == ref.it == [step + 1, 2 * (step + 1), 4 * (step + 1)]
        for l, lv in enumerate(levels):
            got = lv.get_state(U)
            err = _rel_err(got - u0[l], ref.u[l] - u0[l])
            assert np.all(np.isfinite(got))
            assert err.max() <= 1e-10, (step, l, dict(zip(B.VAR_NAMES, err)))
    # the fine levels actually evolved and the buffers were refilled
    assert np.abs(levels[2].get(U[18]) - u0[2][18]).max() > 1e-4


PUNC0 = np.array([[-3.55, 0.07, 0.03], [3.04, -0.05, 0.02]])


def test_hierarchy_puncture_tracking_matches_oracle():
    """the tracker runs on the finest level after each of its steps (and the
    restriction into its parent); the positions follow the oracle recursion
    to round-off of the evolved shift"""
    dt0 = 0.125
    dp = B.DeviceArray.from_host(PUNC0)
    H, levels = _gpu_hierarchy(dt0, punctures=dp)
    ref = OracleHierarchy(dt0, punctures=PUNC0)
    for step in range(2):
        H.advance(0)
        ref.advance(0)
        got = dp.get()
        np.testing.assert_allclose(got, ref.punc, rtol=0, atol=1e-12)
    assert np.abs(got - PUNC0).max() > 1e-6          # the shift moved them
    dp.free()
    for lv in levels:
        lv.free()


@pytest.mark.parametrize("fusion", [2, 1])
def test_graph_replay_matches_eager(fusion):
    """a captured coarse step replayed three times equals the eager steps
    bitwise; fusion 1 runs every level's RK4 stages inside the RHS kernels
    (the rotated lists, the skip-point fills) in the graph"""
    old = B.set_rk_fusion(fusion)
    try:
        dt0 = 0.125
        pa, pb = B.DeviceArray.from_host(PUNC0), B.DeviceArray.from_host(PUNC0)
        HA, LA = _gpu_hierarchy(dt0, punctures=pa)
        HB, LB = _gpu_hierarchy(dt0, punctures=pb)
        HA.advance(0)            # first step: time interpolation of reduced order
        HB.advance(0)
        g = HA.capture(0)        # records the second coarse step, executes nothing
        try:
            for step in range(3):
                HA.launch(g)
                HB.advance(0)
                # eager work and host synchronisation between replays
                l2a = LA[0].z4c_ham_l2(U)
                l2b = LB[0].z4c_ham_l2(U)
                assert l2a == l2b
                for la, lb in zip(LA, LB):
                    np.testing.assert_array_equal(la.get_state(U), lb.get_state(U))
                    np.testing.assert_array_equal(la.get_state(UPP), lb.get_state(UPP))
                np.testing.assert_array_equal(pa.get(), pb.get())
        finally:
            B.Hierarchy.free_graph(g)
        assert HA.nonfinite() == 0
        assert np.abs(pa.get() - PUNC0).max() > 1e-6
    finally:
        B.set_rk_fusion(old)


@pytest.mark.parametrize("iteration", [0, 1])
def test_folded_level_step_bitwise_equal_separate(iteration):
    """an RK4 level step with the time-interpolation begin / end, the u_p copy
    and algcon folded into the first and last update passes equals the
    separate vladd / stage / algcon sequence bitwise"""
    dt0 = 0.125
    res = []
    for folded in (False, True):
        H, levels = _gpu_hierarchy(dt0)
        lv = levels[1]
        par = B.z4c_params()
        for _ in range(2):      # a nontrivial u_p and u_pp
            lv.z4c_rk4_step(par, dt0 / 2, U, UP, R, V, pgh=0)
        lv.bo_interp_begin(TI, 0, UPP, U, UP)
        if folded:
            lv.z4c_rk4_step_bo(par, dt0 / 2, U, UP, R, V, UPP, TI, iteration)
        else:
            B.set_rk_fusion(0)
            lv.bo_interp_begin(TI, iteration, UPP, U, UP)
            lv.z4c_rk4_stages(par, dt0 / 2, U, UP, R, V)
            lv.bo_interp_end(TI, iteration, UPP, U)
            lv.z4c_algcon(U[0:6], U[7:13])
            B.set_rk_fusion(2)
        res.append([lv.get_state(x) for x in (U, UP, UPP, V)])
        for l in levels:
            l.free()
    for a, b in zip(*res):
        np.testing.assert_array_equal(a, b)
