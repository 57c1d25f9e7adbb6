"""GPU parity of the Z4c path (HIP kernels through the C ABI) against the
CPU oracle on identical seeded inputs.

Tolerance (written per test): for each of the 25 variables,
    max |gpu - cpu|  <=  tol * max(max |cpu|, floor)
with tol = 1e-12 for one right-hand side evaluation and 1e-11 after several
RK4 steps; the reference order of floating-point operations is not
reproduced (FMA contraction, different association), so bit equality is
not expected, only agreement to a few ulps of the magnitudes involved.
"""
import numpy as np
import pytest

import baum_amd as B
import oracle as O
from spacetimes import KerrSchild, flat_z4c, grid, puncture_z4c

pytestmark = pytest.mark.gpu

U = list(range(0, 25))
R = list(range(25, 50))
P = list(range(50, 75))
V = list(range(75, 100))


def _rel_err(g, c, floor=1e-30):
    out = []
    for v in range(g.shape[0]):
        scale = max(np.abs(c[v]).max(), floor)
        out.append(np.abs(g[v] - c[v]).max() / scale)
    return np.array(out)


def _level_for(box_o, bflag, nvars=100):
    bx = B.make_box(box_o.m, box_o.n, box_o.o, box_o.x0, box_o.y0, box_o.z0, box_o.dx, box_o.dy, box_o.dz,
                    skip=tuple(box_o.skip), bflag=bflag)
    lev = B.Level([bx], nvars)
    lev.enable(range(nvars))
    for v in range(nvars):
        lev.fill(v, 0.0)
    return lev


def _perturbed_ks(box, seed=11, amp=0.05):
    """Kerr-Schild data plus resolved waves of amplitude `amp` in every field
    and small-scale noise in B: every right-hand side is then O(0.1..1), so
    the comparison is not dominated by the rounding floor of the stencils
    (about eps * max|u| / h^2 for a second difference of O(1) data)."""
    X = grid(box)
    u = KerrSchild().z4c(X)
    rng = np.random.default_rng(seed)
    x, y, z = X
    for v in range(25):
        k = rng.uniform(2.0, 4.0, 3)
        ph = rng.uniform(0, 2 * np.pi, 3)
        u[v] += amp * np.sin(k[0] * x + ph[0]) * np.cos(k[1] * y + ph[1]) * np.sin(k[2] * z + ph[2])
    u[22:25] += 0.01 * rng.standard_normal((3,) + X.shape[1:])
    return u


@pytest.mark.parametrize("params", [
    dict(),                                           # puncture gauge, KO 0.1
    dict(z4_shift="withB gamma0 withGadv withBadv withShiftadv", z4_shiftalphapower=1.0,
         z4_kappa1=0.02, z4_kappa2=0.1),
    dict(z4_lapse="1+log2 harmonic", order_dissipation=0),
])
def test_rhs_bulk_matches_oracle(params):
    h = 0.05
    n = 38
    ob = O.make_box(n, n + 4, n - 2, 2.0 - 19 * h, 1.3 - 21 * h, -1.8 - 18 * h, h, skip=(3,) * 6)
    u = _perturbed_ks(ob)
    cpu, nbad = O.z4c_rhs(ob, O.z4c_params(**params), u)
    assert nbad == 0
    lev = _level_for(ob, (B.PHYBOUND,) * 6)
    lev.put_state(U, u.reshape(25, -1))
    lev.z4c_rhs(B.z4c_params(**params), R, U)
    gpu = lev.get_state(R).reshape(u.shape)
    err = _rel_err(gpu, cpu)
    assert err.max() <= 1e-12, dict(zip(B.VAR_NAMES, err))
    lev.free()


def test_rhs_edge_points_bam_box():
    """BAM single box: only the outermost layer is flagged; points 1 and 2
    from the edge use the reduced/shifted stencils of the reference."""
    h = 0.25
    n = 20
    ob = O.make_box(n, n, n, -(n - 1) * h / 2, -(n - 1) * h / 2 + 0.1, -(n - 1) * h / 2 - 0.05, h, skip=(1,) * 6)
    X = grid(ob)
    u = puncture_z4c(X, masses=((1.0, (0.05, -0.02, 0.03)),))
    rng = np.random.default_rng(5)
    u[19:22] = 0.1 * rng.standard_normal((3,) + X.shape[1:])     # shift of both signs
    u[13] = 0.05 * np.sin(X[0]) * np.cos(X[1])
    p = dict()
    cpu, _ = O.z4c_rhs(ob, O.z4c_params(**p), u)
    lev = _level_for(ob, (B.PHYBOUND,) * 6)
    lev.put_state(U, u.reshape(25, -1))
    lev.z4c_rhs(B.z4c_params(**p), R, U)
    gpu = lev.get_state(R).reshape(u.shape)
    err = _rel_err(gpu, cpu)
    assert err.max() <= 1e-12, dict(zip(B.VAR_NAMES, err))
    lev.free()


def test_rhs_tile_variants_bitwise_equal_on_uneven_level():
    """three uneven boxes in one level (one launch per tile layout covers
    all of them): every tile variant of the bulk and of the box-edge shell
    gives bitwise the same RHS, the tuner's choice included, and it matches
    the oracle box by box"""
    h = 0.2
    shapes = [((20, 20, 20), (-1.9, -1.8, -2.0)), ((22, 17, 26), (2.3, -1.5, -2.4)), ((8, 9, 7), (-0.4, 2.6, 0.3))]
    obs, us = [], []
    for (m, n, o), x0 in shapes:
        ob = O.make_box(m, n, o, *x0, h, skip=(1,) * 6)
        X = grid(ob)
        u = puncture_z4c(X, masses=((1.0, (0.05, -0.02, 0.03)),))
        rng = np.random.default_rng(m * n)
        u[19:22] = 0.1 * rng.standard_normal((3,) + X.shape[1:])
        u[13] = 0.05 * np.sin(X[0]) * np.cos(X[1])
        obs.append(ob)
        us.append(u)
    cpu = np.concatenate([O.z4c_rhs(ob, O.z4c_params(), u)[0].reshape(25, -1) for ob, u in zip(obs, us)], axis=1)
    boxes = [B.make_box(ob.m, ob.n, ob.o, ob.x0, ob.y0, ob.z0, ob.dx, skip=(1,) * 6) for ob in obs]
    lev = B.Level(boxes, 50)
    lev.enable(range(50))
    lev.put_state(U, np.concatenate([u.reshape(25, -1) for u in us], axis=1))
    par = B.z4c_params()
    results = {}
    try:
        for kind, other in ((B.BULK, B.SHELL), (B.SHELL, B.BULK)):
            B.force_variant(other, 0)
            for v, name in enumerate(B.variants(kind)):
                B.force_variant(kind, v)
                for r in R:
                    lev.fill(r, 0.0)
                lev.z4c_rhs(par, R, U)
                results[name] = lev.get_state(R)
            B.force_variant(kind, -1)
            B.force_variant(other, -1)
        B.set_xcd_map(False)
        lev.z4c_rhs(par, R, U)
        results["plain tile order"] = lev.get_state(R)
        B.set_xcd_map(True)
        B.tune_clear()
        B.tune_mode(1)
        lev.z4c_rhs(par, R, U)
        results["tuned"] = lev.get_state(R)
        report = B.tune_report()
    finally:
        B.force_variant(B.BULK, -1)
        B.force_variant(B.SHELL, -1)
        B.set_xcd_map(True)
    assert report.count("\n") == 2, report
    first = next(iter(results.values()))
    for name, got in results.items():
        np.testing.assert_array_equal(got, first, err_msg=name)
    err = _rel_err(first, cpu)
    assert err.max() <= 1e-12, dict(zip(B.VAR_NAMES, err))
    lev.free()


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_shell_rows_with_radial_shift_match_oracle(sign):
    """A shift pointing out of the box (sign +1) or into it (-1) on every face:
    the box-edge shell then takes the shifted advection rows on every face,
    edge and corner (order-4 rows two points from an edge, order-2 rows one
    point from it), together with the order-2 derivatives and the dissipation
    switched off one point from an edge.  The shell's per-point advection
    weights and derivative and dissipation coefficients must reproduce the
    oracle's per-point rules to 1e-12 of each field's magnitude."""
    h = 0.2
    ob = O.make_box(24, 22, 20, -2.3, -2.1, -1.9, h, skip=(1,) * 6)
    X = grid(ob)
    u = puncture_z4c(X, masses=((1.0, (0.05, -0.02, 0.03)),))
    rng = np.random.default_rng(29)
    for a in range(3):
        c = 0.5 * (X[a].min() + X[a].max())
        half = 0.5 * (X[a].max() - X[a].min())
        u[19 + a] = sign * 0.4 * (X[a] - c) / half + 0.02 * rng.standard_normal(X.shape[1:])
    u[13] = 0.05 * np.sin(X[0]) * np.cos(X[1])
    cpu, _ = O.z4c_rhs(ob, O.z4c_params(), u)
    lev = _level_for(ob, (B.PHYBOUND,) * 6, nvars=50)
    lev.put_state(U, u.reshape(25, -1))
    lev.z4c_rhs(B.z4c_params(), R, U)
    err = _rel_err(lev.get_state(R).reshape(u.shape), cpu)
    assert err.max() <= 1e-12, dict(zip(B.VAR_NAMES, err))
    lev.free()


def test_rhs_pair_variants_rejected_on_odd_rows():
    """16-byte staging needs even box rows: on a 21-wide box a pinned pair
    variant is refused, and the tuner skips it and still matches the oracle"""
    h = 0.2
    ob = O.make_box(21, 18, 16, -2.0, -1.7, -1.5, h, skip=(1,) * 6)
    X = grid(ob)
    u = puncture_z4c(X, masses=((1.0, (0.05, -0.02, 0.03)),))
    cpu, _ = O.z4c_rhs(ob, O.z4c_params(), u)
    lev = _level_for(ob, (B.PHYBOUND,) * 6, nvars=50)
    lev.put_state(U, u.reshape(25, -1))
    names = B.variants(B.BULK)
    pair = [v for v, nm in enumerate(names) if nm.endswith("x4")][0]
    try:
        B.force_variant(B.BULK, pair)
        with pytest.raises(B.BamhipError):
            lev.z4c_rhs(B.z4c_params(), R, U)
    finally:
        B.force_variant(B.BULK, -1)
    lev.z4c_rhs(B.z4c_params(), R, U)
    err = _rel_err(lev.get_state(R).reshape(u.shape), cpu)
    assert err.max() <= 1e-12, dict(zip(B.VAR_NAMES, err))
    assert "x4=-1.0000ms" in B.tune_report()
    lev.free()


@pytest.mark.parametrize("params", [dict(advection_lopsided=0), dict(order_centered=2, order_advection=2),
                                    dict(advection_lopsided=2, order_advection=4)])
def test_rhs_generic_orders(params):
    h = 0.1
    n = 16
    ob = O.make_box(n, n, n, 1.5, 1.0, -1.2, h, skip=(1,) * 6)
    u = _perturbed_ks(ob)
    cpu, _ = O.z4c_rhs(ob, O.z4c_params(**params), u)
    lev = _level_for(ob, (B.PHYBOUND,) * 6)
    lev.put_state(U, u.reshape(25, -1))
    lev.z4c_rhs(B.z4c_params(**params), R, U)
    gpu = lev.get_state(R).reshape(u.shape)
    err = _rel_err(gpu, cpu)
    assert err.max() <= 1e-12, dict(zip(B.VAR_NAMES, err))
    lev.free()


def test_rhs_addlinear():
    h = 0.05
    n = 24
    ob = O.make_box(n, n, n, 2.0, 1.0, -1.5, h, skip=(3,) * 6)
    u = _perturbed_ks(ob)
    rng = np.random.default_rng(2)
    upre = u + 1e-3 * rng.standard_normal(u.shape)
    cpu = np.zeros_like(u)
    O.z4c_rhs(ob, O.z4c_params(), u, unew=cpu, upre=upre, c=0.01)
    lev = _level_for(ob, (B.PHYBOUND,) * 6)
    lev.put_state(U, u.reshape(25, -1))
    lev.put_state(P, upre.reshape(25, -1))
    lev.z4c_rhs(B.z4c_params(), R, U, upre=P, c=0.01)
    gpu = lev.get_state(R).reshape(u.shape)
    sl = (slice(None),) + (slice(3, -3),) * 3
    err = _rel_err(gpu[sl], cpu[sl])
    assert err.max() <= 1e-13, dict(zip(B.VAR_NAMES, err))
    lev.free()


def test_radiative_boundary_matches_oracle():
    h = 0.25
    n = 18
    ob = O.make_box(n, n + 2, n + 1, -(n - 1) * h / 2, -(n + 1) * h / 2, -n * h / 2, h, skip=(1,) * 6)
    X = grid(ob)
    u = puncture_z4c(X)
    rng = np.random.default_rng(9)
    u += 1e-3 * rng.standard_normal(u.shape)
    cpu = np.zeros_like(u)
    O.z4c_radiative_bc(ob, O.z4c_params(), cpu, u)
    lev = _level_for(ob, (B.PHYBOUND,) * 6)
    lev.put_state(U, u.reshape(25, -1))
    lev.z4c_set_boundary(B.z4c_params(), R, U)
    gpu = lev.get_state(R).reshape(u.shape)
    err = _rel_err(gpu, cpu)
    assert err.max() <= 1e-13, dict(zip(B.VAR_NAMES, err))
    lev.free()


def _run_steps(ob, bflag, params, u0, dt, nsteps, pgh):
    cpu = u0.copy()
    work = np.zeros((75,) + u0.shape[1:])
    for _ in range(nsteps):
        nb = O.z4c_rk4_step(ob, O.z4c_params(**params), dt, cpu, bc=1 if bflag[0] == B.PHYBOUND else 2, pgh=pgh,
                            work=work)
        assert nb == 0
    lev = _level_for(ob, bflag)
    lev.put_state(U, u0.reshape(25, -1))
    for _ in range(nsteps):
        lev.z4c_rk4_step(B.z4c_params(**params), dt, U, P, R, V, pgh=pgh)
    gpu = lev.get_state(U).reshape(u0.shape)
    lev.free()
    return gpu, cpu


def test_rk4_puncture_box_radiative():
    """small version of the single-box puncture run: BL m=1, precollapsed
    lapse, radiative boundary, dt = h/4, 4 RK4 steps"""
    h = 0.5
    n = 22
    ob = O.make_box(n, n, n, -(n - 1) * h / 2, -(n - 1) * h / 2, -(n - 1) * h / 2, h, skip=(1,) * 6)
    u0 = puncture_z4c(grid(ob))
    gpu, cpu = _run_steps(ob, (B.PHYBOUND,) * 6, dict(), u0, 0.25 * h, 4, 0)
    err = _rel_err(gpu - u0, cpu - u0)
    assert np.all(np.isfinite(gpu))
    assert err.max() <= 1e-11, dict(zip(B.VAR_NAMES, err))


@pytest.mark.parametrize("case", ["radiative", "periodic", "two_boxes"])
def test_fused_rk4_stages_bitwise_equal_unfused(case):
    """RK4 steps with the stage update fused into the RHS and boundary kernels
    give bitwise the state of the separate RHS / boundary / update passes"""
    if case == "periodic":
        g, nin = 3, 20
        n = nin + 2 * g
        h = 1.0 / 12
        ob = O.make_box(n, n, n, -(n - 1) * h / 2, -(n - 1) * h / 2, -(n - 1) * h / 2, h, skip=(g,) * 6)
        rng = np.random.default_rng(77)
        u0 = flat_z4c((n, n, n)) + rng.uniform(-1e-6, 1e-6, (25, n, n, n))
        O.periodic_fill(ob, u0, g)
        boxes = [B.make_box(n, n, n, ob.x0, ob.y0, ob.z0, h, skip=(g,) * 6, bflag=(B.PERBOUND,) * 6)]
        state, pgh, dt = u0.reshape(25, -1), g, 0.25 * h
    else:
        h = 0.5
        specs = [(22, 20, 18, (-5.25, -4.7, -4.3))]
        if case == "two_boxes":
            specs.append((16, 18, 14, (6.0, -4.0, -3.0)))
        boxes, us = [], []
        for m, nn, o, x0 in specs:
            ob = O.make_box(m, nn, o, *x0, h, skip=(1,) * 6)
            us.append(puncture_z4c(grid(ob), masses=((1.0, (0.05, -0.02, 0.03)),)).reshape(25, -1))
            flag = B.PHYBOUND if case == "radiative" else B.REFBOUND
            boxes.append(B.make_box(m, nn, o, *x0, h, skip=(1,) * 6, bflag=(flag,) * 6))
        state, pgh, dt = np.concatenate(us, axis=1), 0, 0.25 * h
    out = []
    try:
        for fused in (0, 1):
            B.set_rk_fusion(fused)
            lev = B.Level(boxes, 100)
            lev.enable(range(100))
            for v in range(100):
                lev.fill(v, 0.0)
            lev.put_state(U, state)
            for _ in range(3):
                lev.z4c_rk4_step(B.z4c_params(), dt, U, P, R, V, pgh=pgh)
            out.append(lev.get_state(U))
            lev.free()
    finally:
        B.set_rk_fusion(2)
    assert np.abs(out[0] - state).max() > 1e-9
    np.testing.assert_array_equal(out[1], out[0])


def test_fused_then_separate_stages_on_one_level():
    """two fused RK4 steps, then one with the separate passes on the same
    level, give bitwise the state of three steps with the separate passes:
    the fused stages leave the RHS list as the separate passes expect it
    (zero at the points no kernel computes)"""
    h = 0.5
    specs = [(22, 20, 18, (-5.25, -4.7, -4.3)), (16, 18, 14, (6.0, -4.0, -3.0))]
    boxes, us = [], []
    for m, nn, o, x0 in specs:
        ob = O.make_box(m, nn, o, *x0, h, skip=(1,) * 6)
        us.append(puncture_z4c(grid(ob), masses=((1.0, (0.05, -0.02, 0.03)),)).reshape(25, -1))
        boxes.append(B.make_box(m, nn, o, *x0, h, skip=(1,) * 6, bflag=(B.REFBOUND,) * 6))
    state, dt = np.concatenate(us, axis=1), 0.25 * h
    out = []
    try:
        for modes in ((1, 1, 0), (0, 0, 0)):
            lev = B.Level(boxes, 100)
            lev.enable(range(100))
            for v in range(100):
                lev.fill(v, 0.0)
            lev.put_state(U, state)
            for mode in modes:
                B.set_rk_fusion(mode)
                lev.z4c_rk4_step(B.z4c_params(), dt, U, P, R, V, pgh=0)
            out.append(lev.get_state(U))
            lev.free()
    finally:
        B.set_rk_fusion(2)
    assert np.abs(out[1] - state).max() > 1e-9
    np.testing.assert_array_equal(out[0], out[1])


def test_constraint_norms_tracked_over_time():
    """Hamiltonian and momentum constraint norms over 6 RK4 steps of the
    small puncture box (adm.c:14-74 at every step), device against oracle:
    the states agree to ~1e-11 of their change, so the norms agree to 1e-9"""
    h = 0.5
    n = 22
    ob = O.make_box(n, n, n, -(n - 1) * h / 2, -(n - 1) * h / 2 + 0.1, -(n - 1) * h / 2 - 0.05, h, skip=(1,) * 6)
    u = puncture_z4c(grid(ob), masses=((1.0, (0.05, -0.02, 0.03)),))
    dt = 0.25 * h
    lev = _level_for(ob, (B.PHYBOUND,) * 6)
    lev.put_state(U, u.reshape(25, -1))
    work = np.zeros((75,) + u.shape[1:])
    par_o, par_g = O.z4c_params(), B.z4c_params()
    history = []
    for step in range(7):
        lc = O.z4c_constraints_l2(ob, u)
        lg = lev.z4c_constraints_l2(U)
        history.append(lc)
        np.testing.assert_allclose(lg, lc, rtol=1e-9, atol=0, err_msg="step %d" % step)
        assert O.z4c_rk4_step(ob, par_o, dt, u, bc=1, pgh=0, work=work) == 0
        lev.z4c_rk4_step(par_g, dt, U, P, R, V, pgh=0)
    history = np.array(history)
    assert np.all(np.isfinite(history)) and np.all(history[:, 0] > 0)
    # K_ij = 0 initially, so the momentum constraint starts at exactly zero and
    # is generated by the evolution
    assert np.all(history[0, 1:] == 0) and np.all(history[1:, 1:] > 1e-4)
    lev.free()


def test_rk4_periodic_noise():
    """small version of the robust-stability run: flat space + 1e-10 noise,
    periodic with 3 ghost layers, 3 RK4 steps"""
    g = 3
    nin = 20
    n = nin + 2 * g
    h = 1.0 / 12
    ob = O.make_box(n, n, n, -(n - 1) * h / 2, -(n - 1) * h / 2, -(n - 1) * h / 2, h, skip=(g,) * 6)
    rng = np.random.default_rng(1234)
    u0 = flat_z4c((n, n, n)) + rng.uniform(-1e-10, 1e-10, (25, n, n, n))
    O.periodic_fill(ob, u0, g)
    gpu, cpu = _run_steps(ob, (B.PERBOUND,) * 6, dict(), u0, 0.25 * h, 3, g)
    # the updates are O(dt * 1e-10 / h^2); the stencils' rounding floor on
    # O(1) data is ~ eps / h^2, so the bound here is absolute: 64 eps / h^2 * dt
    dt = 0.25 * h
    bound = 64 * np.finfo(float).eps / h ** 2 * dt * 3
    assert np.abs(gpu - cpu).max() <= bound, (np.abs(gpu - cpu).max(), bound)
    # and the evolution itself is far above that floor
    assert np.abs(cpu - u0).max() > 100 * bound


def test_algcon_matches_oracle():
    n = 20
    ob = O.make_box(n, n, n, 0, 0, 0, 0.1)
    rng = np.random.default_rng(4)
    u = flat_z4c((n, n, n))
    u[0:6] += 0.02 * rng.standard_normal((6, n, n, n))
    u[7:13] = 0.1 * rng.standard_normal((6, n, n, n))
    cpu = u.copy()
    O.z4c_algcon(ob, cpu)
    lev = _level_for(ob, (B.PHYBOUND,) * 6)
    lev.put_state(U, u.reshape(25, -1))
    lev.z4c_algcon(U[0:6], U[7:13])
    gpu = lev.get_state(U).reshape(u.shape)
    err = _rel_err(gpu, cpu)
    assert err.max() <= 1e-14
    lev.free()


def test_hamiltonian_norm_matches_oracle():
    h = 0.1
    n = 24
    ob = O.make_box(n, n, n, 1.8, 0.9, -1.7, h, skip=(1,) * 6)
    u = _perturbed_ks(ob)
    l2c = O.z4c_ham_l2(ob, u)
    lev = _level_for(ob, (B.PHYBOUND,) * 6)
    lev.put_state(U, u.reshape(25, -1))
    l2g = lev.z4c_ham_l2(U)
    assert abs(l2g - l2c) <= 1e-11 * l2c
    lev.free()


def test_constraint_fields_and_norms_match_oracle():
    """Hamiltonian and momentum constraint fields and their L2 norms
    (adm.c:14-74) on perturbed Kerr-Schild data, including the order-2 points
    one from the box edge"""
    h = 0.1
    ob = O.make_box(22, 24, 26, 1.8, 0.9, -1.7, h, skip=(1,) * 6)
    u = _perturbed_ks(ob)
    l2c, hc, mc = O.z4c_constraints_l2(ob, u, want_fields=True)
    lev = _level_for(ob, (B.PHYBOUND,) * 6, nvars=29)
    lev.put_state(U, u.reshape(25, -1))
    l2g = lev.z4c_constraints_l2(U, ham=25, mom=[26, 27, 28])
    np.testing.assert_allclose(l2g, l2c, rtol=1e-11, atol=0)
    sl = (slice(1, -1),) * 3
    for v, ref in zip(range(25, 29), [hc, *mc]):
        got = lev.get(v).reshape(ref.shape)
        scale = np.abs(ref[sl]).max()
        assert scale > 1e-3
        assert np.abs(got[sl] - ref[sl]).max() <= 1e-11 * scale, v
    lev.free()


def test_nonfinite_rhs_is_reported():
    """a NaN planted in the input (host side) must come back as
    BAMHIP_ENONFINITE, the reference's errorexit"""
    n = 12
    ob = O.make_box(n, n, n, 0, 0, 0, 0.1, skip=(3,) * 6)
    u = flat_z4c((n, n, n))
    u[6, 6, 6, 6] = np.nan
    lev = _level_for(ob, (B.PHYBOUND,) * 6)
    lev.put_state(U, u.reshape(25, -1))
    with pytest.raises(B.BamhipError, match="non-finite"):
        lev.z4c_rhs(B.z4c_params(), R, U)
    lev.free()
