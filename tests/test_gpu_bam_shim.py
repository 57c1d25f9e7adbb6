"""The BAM-shaped adapters called through function pointers of the
reference's hook types (rhs_geometry / rhs_boundary of evolve.c:22-29,
MATTER.c2p, restrict_prolong_varlist of inject.c:66, evolve_rk_rk4 of
rk.c:216) on BAM levels whose tL::v holds host arrays, against the oracle.

Tolerances as in the direct-ABI tests: 1e-12 of each variable's magnitude
for one right-hand side, 1e-13 for the radiative boundary, 1e-11 of the
change after an RK4 step; c2p and the R/P operators bitwise.
"""
import ctypes as C

import numpy as np
import pytest

import baum_amd as B
import baum_amd.bam as BB
import oracle as O
from c2p_data import sample_prims
from fmr_geometry import coarse_box, fine_box
from spacetimes import grid, puncture_z4c

pytestmark = pytest.mark.gpu

U = list(range(0, 25))
R = list(range(25, 50))
P = list(range(50, 75))
V = list(range(75, 100))


def _rel_err(g, c, floor=1e-30):
    return np.array([np.abs(g[v] - c[v]).max() / max(np.abs(c[v]).max(), floor) for v in range(g.shape[0])])


def _puncture_box(n=18, h=0.25, seed=9):
    ob = O.make_box(n, n + 2, n + 1, -(n - 1) * h / 2, -(n + 1) * h / 2, -n * h / 2, h, skip=(1,) * 6)
    u = puncture_z4c(grid(ob))
    u += 1e-3 * np.random.default_rng(seed).standard_normal(u.shape)
    return ob, u


def _bam_level(ob, l=0, dt=0.0, nvars=100):
    bx = B.make_box(ob.m, ob.n, ob.o, ob.x0, ob.y0, ob.z0, ob.dx, skip=tuple(ob.skip), bflag=(B.PHYBOUND,) * 6)
    dev = B.Level([bx], nvars, l=l)
    return dev, BB.BamLevel(dev, l=l, dt=dt)


def _rule(**kw):
    d = dict(dissipation_factor=0.1, order_dissipation=4, amr_move_lcube=0, physics_z4d=0,
             factor_level0=-1.0, factor_shells=-1.0, factor_move=-1.0)
    d.update(kw)
    return B.DissipationRule(**d)


def test_rhs_and_boundary_through_hooks():
    ob, u = _puncture_box()
    cpu, nbad = O.z4c_rhs(ob, O.z4c_params(), u)
    assert nbad == 0
    O.z4c_radiative_bc(ob, O.z4c_params(), cpu, u)
    err = BB.Errors()
    dev, lev = _bam_level(ob)
    ucur, unew = lev.varlist(U), lev.varlist(R)
    for k, v in enumerate(U):
        lev.set(v, u[k])
    ucur.put()
    unew.put()
    BB.set_z4c(B.z4c_params(), _rule(), 0)
    hooks = BB.hooks()
    # evolve_rhs: rhs_geometry(unew, upre, c, ucur), then rhs_boundary (evolve.c:51-102)
    hooks.rhs_geometry(unew.ref(), None, 0.0, ucur.ref())
    hooks.rhs_boundary(unew.ref(), None, 0.0, ucur.ref())
    err.raise_if_any()
    gpu = unew.get().reshape(u.shape)
    e = _rel_err(gpu, cpu)
    assert e.max() <= 1e-12, dict(zip(B.VAR_NAMES, e))
    lev.detach()
    dev.free()


def test_dissipation_rule_applied_per_level():
    """the adapter's KO factor follows get_dissipation_factor for the level
    it is called on: level 2 > amr_move_lcube = 1 gets 0.2 sigma / 16"""
    ob, u = _puncture_box(seed=3)
    cpu, _ = O.z4c_rhs(ob, O.z4c_params(dissipation_factor=0.02), u)
    O.z4c_radiative_bc(ob, O.z4c_params(), cpu, u)
    err = BB.Errors()
    dev, lev = _bam_level(ob, l=2)
    ucur, unew = lev.varlist(U), lev.varlist(R)
    for k, v in enumerate(U):
        lev.set(v, u[k])
    ucur.put()
    unew.put()
    BB.set_z4c(B.z4c_params(), _rule(amr_move_lcube=1), 0)
    BB.hooks().rhs_geometry(unew.ref(), None, 0.0, ucur.ref())
    BB.hooks().rhs_boundary(unew.ref(), None, 0.0, ucur.ref())
    err.raise_if_any()
    e = _rel_err(unew.get().reshape(u.shape), cpu)
    assert e.max() <= 1e-12, dict(zip(B.VAR_NAMES, e))
    lev.detach()
    dev.free()


def test_evolve_rk4_through_hook():
    h, n = 0.5, 22
    ob = O.make_box(n, n, n, -(n - 1) * h / 2, -(n - 1) * h / 2, -(n - 1) * h / 2, h, skip=(1,) * 6)
    u0 = puncture_z4c(grid(ob))
    dt = 0.25 * h
    cpu = u0.copy()
    for _ in range(2):
        assert O.z4c_rk4_step(ob, O.z4c_params(), dt, cpu, bc=1, pgh=0) == 0
    err = BB.Errors()
    dev, lev = _bam_level(ob, dt=dt)
    lists = [lev.varlist(ix) for ix in (U, P, R, V)]
    for k, v in enumerate(U):
        lev.set(v, u0[k])
    for vl in lists:
        vl.put()
    BB.set_z4c(B.z4c_params(), _rule(), 0)
    BB.set_rk_lists(*lists, algcon=True)
    for _ in range(2):
        BB.hooks().evolve_rk4(C.byref(lev.tl), 0)
    err.raise_if_any()
    gpu = lists[0].get().reshape(u0.shape)
    e = _rel_err(gpu - u0, cpu - u0)
    assert np.all(np.isfinite(gpu))
    assert e.max() <= 1e-11, dict(zip(B.VAR_NAMES, e))
    lev.detach()
    dev.free()


@pytest.mark.parametrize("method", ["hybrid", "plain"])
def test_c2p_through_hook(method):
    """MATTER.c2p(ucur, primVars) on a level: ucur holds unrelated variables
    before the conserved block (INDX_VAR_q = 2), the metric sits at its own
    variable index (INDX_VAR_gxx).  The plain p-root stops on a failed point,
    so its case keeps the points where the oracle's plain root converges."""
    n = 24 * 20 * 18
    prim5, g, _ = sample_prims(n, seed=21)
    Po = O.c2p_params(errormax=1e-10, rhoatm=1e-20)
    cons = O.p2c(Po, prim5, g)
    if method == "plain":
        _, st, _, _ = O.c2p(Po, cons, g, np.zeros(n), rhoguess=prim5[0], method="plain")
        keep = st == O.C2P_OK
        prim5, g, cons = prim5[:, keep], g[:, keep], cons[:, keep]
        n = cons.shape[1]
    cp, cs, ci, cc = O.c2p(Po, cons, g, np.zeros(n), rhoguess=prim5[0], method=method)
    bx = B.make_box(n, 1, 1, 0.0, 0.0, 0.0, 1.0)
    dev = B.Level([bx], 40)
    lev = BB.BamLevel(dev)
    UCUR = [30, 31, 0, 1, 2, 3, 4]     # two unrelated variables, then D, tau, S_x, S_y, S_z
    GXX = 10                           # adm_gxx .. adm_gzz at 10..15
    PRIM = list(range(20, 28))         # rho, epsl, vx, vy, vz, p, v2, detg
    for k in range(5):
        lev.set(k, cons[k])
    for k in range(6):
        lev.set(GXX + k, g[k])
    lev.set(PRIM[0], prim5[0])
    ucur = lev.varlist(UCUR)
    metric = lev.varlist(list(range(GXX, GXX + 6)))
    prim = lev.varlist(PRIM)
    for vl in (ucur, metric, prim):
        vl.put()
    err = BB.Errors()
    Pg = B.c2p_params(errormax=1e-10, rhoatm=1e-20)
    BB.set_grhd(Pg, {"hybrid": B.C2P_HYBRID, "plain": B.C2P_PLAIN}[method], 2, GXX)
    BB.hooks().c2p(ucur.ref(), prim.ref())
    err.raise_if_any()
    np.testing.assert_array_equal(prim.get(), cp)
    np.testing.assert_array_equal(ucur.get()[2:], cc)
    lev.detach()
    dev.free()


def test_c2p_plain_failure_reaches_errorexit():
    """grhd_c2p_proot stops on a point it cannot invert: the adapter reports
    it through the errorexit hook"""
    n = 256
    prim5, g, _ = sample_prims(n, seed=4)
    Po = O.c2p_params(errormax=1e-10, rhoatm=1e-20)
    cons = O.p2c(Po, prim5, g)
    cons[1, 17] = -5.0 * abs(cons[0, 17]) - 1.0  # tau far below zero: no pressure root
    bx = B.make_box(n, 1, 1, 0.0, 0.0, 0.0, 1.0)
    dev = B.Level([bx], 20)
    lev = BB.BamLevel(dev)
    for k in range(5):
        lev.set(k, cons[k])
    for k in range(6):
        lev.set(5 + k, g[k])
    lev.set(11, prim5[0])
    ucur, metric, prim = lev.varlist(range(5)), lev.varlist(range(5, 11)), lev.varlist(range(11, 19))
    for vl in (ucur, metric, prim):
        vl.put()
    err = BB.Errors()
    BB.set_grhd(B.c2p_params(errormax=1e-10, rhoatm=1e-20), B.C2P_PLAIN, 0, 5)
    BB.hooks().c2p(ucur.ref(), prim.ref())
    assert len(err.messages) == 1 and "con2prim failed at point" in err.messages[0]
    lev.detach()
    dev.free()


def _fmr_levels(mk):
    cb0 = coarse_box(mk, 44, 28, 26, 0.5)
    cb1 = coarse_box(mk, 24, 22, 24, 0.5, x0=(40.0, 0.25, -3.0))
    f1 = fine_box(mk, cb0, (4, 5, 4), (20, 16, 16))
    f2 = fine_box(mk, cb0, (24, 4, 5), (28, 18, 14))
    f3 = fine_box(mk, cb1, (5, 6, 4), (24, 20, 28))
    return [cb0, cb1], [f1, f2, f3], [0, 0, 1]


def test_restrict_prolong_through_hook():
    nb, nvar, order, U0 = 6, 5, 4, 2
    mk = lambda m, n, o, x0, y0, z0, dx: B.make_box(m, n, o, x0, y0, z0, dx, bflag=(B.REFBOUND,) * 6)
    cbs, fbs, parent = _fmr_levels(mk)
    dc, df = B.Level(cbs, U0 + nvar, l=0), B.Level(fbs, U0 + nvar, l=1)
    dc.enable(range(U0 + nvar))
    df.enable(range(U0 + nvar))
    B.set_rp_flags(dc, df, parent, nb, 0, 1)
    lc, lf = BB.BamLevel(dc, l=0), BB.BamLevel(df, l=1)
    rng = np.random.default_rng(17)
    cv = rng.standard_normal((nvar, dc.npoints))
    fv = rng.standard_normal((nvar, df.npoints))
    idx = list(range(U0, U0 + nvar))
    for k, v in enumerate(idx):
        lc.set(v, cv[k])
        lf.set(v, fv[k])
    uc, uf = lc.varlist(idx), lf.varlist(idx)
    uc.put()
    uf.put()
    g = BB.grid([lc, lf])
    err = BB.Errors()
    BB.set_fmr(order)
    BB.hooks().restrict_prolong_varlist(C.byref(g), 0, 1, uc.ref(), uf.ref(), 3)
    err.raise_if_any()
    ocbs, ofbs, _ = _fmr_levels(lambda m, n, o, x0, y0, z0, dx: O.make_box(m, n, o, x0, y0, z0, dx))
    fp, fr = O.rp_flags(ofbs, ocbs, parent, nb)
    ec, ef = cv.copy(), fv.copy()
    assert O.restrict_level(ofbs, ocbs, parent, fr, 3, order, ef, ec) == 0
    assert O.prolong_level(ofbs, ocbs, parent, fp, 3, nb, order, ec, ef) == 0
    np.testing.assert_array_equal(uc.get(), ec)
    np.testing.assert_array_equal(uf.get(), ef)
    for L in (lc, lf):
        L.detach()
    dc.free()
    df.free()
