"""Pins the CPU oracle's con2prim against what the mathematics fixes.

The reference c2p (grhd_c2p_v2.c, grhd_c2p.c) includes bam.h and cannot be
built standalone, and the reference holds no c2p golden data.  Pins:

* round trip prim -> cons (grhd_compute_q_pt restated) -> c2p recovers the
  primitives: rho and v^i to 1e-12, p and eps within the conditioning of
  the inversion (eps_mach * W^2 (tau+D)/p, the W^2 from the cancellation in
  (tau+p+D)^2 - S^2);
* the pressure root against an independent high-precision solve (mpmath,
  60 digits, bisection on the same FP64 conserved inputs);
* closed form: for a static fluid (S = 0) the ideal-gas inversion is
  explicit: rho = D, p = (Gamma-1) tau;
* the reference's control flow: countmax caps (NR status, including the
  count == countmax edge), the cold fallback that rewrites tau, the
  atmosphere that rewrites all conserved variables, D <= 0.
"""
import mpmath as mp
import numpy as np
import pytest

import oracle as O
from c2p_data import sample_prims

P12 = dict(errormax=1e-12, countmax=100, rhoatm=1e-20)


def _undens(cons, g, q):
    gm = g[:, q]
    c11 = gm[3] * gm[5] - gm[4] ** 2
    c12 = gm[2] * gm[4] - gm[1] * gm[5]
    c13 = gm[1] * gm[4] - gm[2] * gm[3]
    c22 = gm[0] * gm[5] - gm[2] ** 2
    c23 = gm[1] * gm[2] - gm[0] * gm[4]
    c33 = gm[0] * gm[3] - gm[1] ** 2
    det = gm[0] * c11 + gm[1] * c12 + gm[2] * c13
    gi = np.array([c11, c12, c13, c22, c23, c33]) / det
    oo = 1 / np.sqrt(det)
    D, T, Sx, Sy, Sz = [cons[k, q] * oo for k in range(5)]
    Six = gi[0] * Sx + gi[1] * Sy + gi[2] * Sz
    Siy = gi[1] * Sx + gi[3] * Sy + gi[4] * Sz
    Siz = gi[2] * Sx + gi[4] * Sy + gi[5] * Sz
    return D, T, Sx * Six + Sy * Siy + Sz * Siz


@pytest.fixture(scope="module")
def sample():
    prim5, g, W = sample_prims(20000, seed=3)
    P = O.c2p_params(**P12)
    cons = O.p2c(P, prim5, g)
    return P, prim5, g, W, cons


def test_bracketed_round_trip(sample):
    P, prim5, g, W, cons = sample
    prim, st, it, _ = O.c2p(P, cons, g, np.zeros(prim5.shape[1]), method="bracketed")
    assert np.all(st == O.C2P_OK)
    assert it.max() <= 8
    np.testing.assert_allclose(prim[0], prim5[0], rtol=1e-12)
    for k in (2, 3, 4):
        np.testing.assert_allclose(prim[k], prim5[k], rtol=1e-11, atol=1e-14)
    ptrue = (P.gamma - 1) * prim5[0] * prim5[1]
    kappa = W ** 2 * (cons[1] + cons[0]) / ptrue
    bound = 64 * np.finfo(float).eps * kappa
    assert np.all(np.abs(prim[5] / ptrue - 1) <= np.maximum(bound, 1e-12))
    assert np.all(np.abs(prim[1] / prim5[1] - 1) <= np.maximum(bound, 1e-12))


def test_pressure_matches_high_precision_root(sample):
    P, prim5, g, W, cons = sample
    prim, st, it, _ = O.c2p(P, cons, g, np.zeros(prim5.shape[1]), method="bracketed")
    mp.mp.dps = 60
    rng = np.random.default_rng(0)
    for q in rng.choice(prim5.shape[1], 12, replace=False):
        D, T, S2 = (mp.mpf(x) for x in _undens(cons, g, q))

        def f(p):
            t = T + p + D
            sq = mp.sqrt(t * t - S2)
            Wl = t / sq
            return p - (P.gamma - 1) * (D / Wl) * ((sq - Wl * p) / D - 1)

        lo, hi = max(mp.sqrt(S2) - D - T, mp.mpf(0)), (P.gamma - 1) * (T + D)
        for _ in range(300):
            mid = (lo + hi) / 2
            lo, hi = (mid, hi) if f(mid) < 0 else (lo, mid)
        pex = float((lo + hi) / 2)
        kappa = W[q] ** 2 * float(T + D) / pex
        assert abs(prim[5, q] / pex - 1) <= max(64 * np.finfo(float).eps * kappa, 1e-13)


def test_static_fluid_closed_form():
    n = 50
    rng = np.random.default_rng(8)
    prim5 = np.array([10 ** rng.uniform(-8, -2, n), 10 ** rng.uniform(-4, 0, n), np.zeros(n), np.zeros(n),
                      np.zeros(n)])
    g = np.zeros((6, n))
    g[0] = g[3] = g[5] = 1.0
    P = O.c2p_params(**P12)
    cons = O.p2c(P, prim5, g)
    for method in ("hybrid", "plain", "bracketed"):
        prim, st, it, _ = O.c2p(P, cons, g, np.zeros(n), rhoguess=prim5[0], method=method)
        assert np.all(st == O.C2P_OK), method
        np.testing.assert_allclose(prim[0], cons[0], rtol=1e-14)
        np.testing.assert_allclose(prim[5], (P.gamma - 1) * cons[1], rtol=1e-11)


def test_reference_nr_cap_and_edge(sample):
    """try_hot_part returns NR when the loop ran countmax+1 times, and also
    when it converged exactly on that last pass (count == countmax)."""
    P0, prim5, g, W, cons = sample
    rho = O.C.c_double()
    eps = O.C.c_double()
    p = O.C.c_double()
    it = O.C.c_int()
    P = O.c2p_params(errormax=1e-10, countmax=100, rhoatm=1e-20)
    for q in range(prim5.shape[1]):          # first point the reference solves
        D, T, S2 = _undens(cons, g, q)
        st = O.lib().ob_try_hot(O.C.byref(P), D, S2, T, O.C.byref(rho), O.C.byref(eps), O.C.byref(p), 0.0,
                                O.C.byref(it))
        if st == O.C2P_OK and it.value >= 3:
            break
    assert st == O.C2P_OK
    n_needed = it.value
    for cm, expect in ((n_needed - 1, O.C2P_NR), (n_needed - 2, O.C2P_NR), (n_needed, O.C2P_OK)):
        if cm < 1:
            continue
        P = O.c2p_params(errormax=1e-10, countmax=cm, rhoatm=1e-20)
        st = O.lib().ob_try_hot(O.C.byref(P), D, S2, T, O.C.byref(rho), O.C.byref(eps), O.C.byref(p), 0.0,
                                O.C.byref(it))
        assert st == expect, (cm, n_needed, st)


def test_hybrid_cold_fallback_and_atmosphere():
    P = O.c2p_params(eos="ideal", gamma=2.0, K=100.0, rhoatm=1e-12, errormax=1e-10)
    g = np.zeros((6, 4))
    g[0] = g[3] = g[5] = 1.0
    # 0: S^2 > (tau + D + pmin)^2 impossible... force tempSQRTarg <= 0 with tau < 0
    # 1: D <= 0 -> cold part returns ATM
    # 2: density below the atmosphere threshold
    # 3: a regular state
    cons = np.array([[1e-3, -1e-3, 1e-15, 1e-3],
                     [-2e-3, 1e-4, 1e-16, 1e-4],
                     [0.0, 0.0, 0.0, 1e-5],
                     [0.0, 0.0, 0.0, 0.0],
                     [0.0, 0.0, 0.0, 0.0]])
    prim, st, it, cons_out = O.c2p(P, cons, g, np.zeros(4), rhoguess=np.full(4, 1e-3), method="hybrid")
    assert st[3] == O.C2P_OK
    # atmosphere points carry the atmosphere state in prims and conserved variables
    for q in (1, 2):
        assert st[q] in (O.C2P_ATM, O.C2P_COLD, O.C2P_FAIL, O.C2P_NR)
        assert prim[0, q] == P.rhoatm and prim[5, q] == P.patm
        assert cons_out[0, q] == P.rhoatm and cons_out[2, q] == 0.0
        np.testing.assert_allclose(cons_out[1, q], P.rhoatm * P.epslatm)
    # the cold branch rewrites tau from the cold EOS
    assert st[0] in (O.C2P_COLD, O.C2P_OK, O.C2P_ATM)
    if st[0] == O.C2P_OK:
        h = 1 + prim[1, 0] + prim[5, 0] / prim[0, 0]
        Wl = cons_out[0, 0] / prim[0, 0]
        np.testing.assert_allclose(cons_out[1, 0], Wl * Wl * prim[0, 0] * h - prim[5, 0] - cons_out[0, 0], rtol=1e-10)


def test_atmosphere_reset_of_upre():
    """grhd_atm_set_u on u_p (grhd_atm.c:217-262).  Pinned by the static
    atmosphere itself: the reset conserved state is what prim -> cons
    (grhd_compute_q_pt, restated independently in ob_p2c) gives for
    rho = rho_atm, eps = eps_atm, v = 0 on the point's metric, i.e.
    D = sqrt(detg) rho_atm and tau / D = eps_atm, S_i = 0; points at or above
    f_atm rho_atm are untouched bitwise; detg is floored at DETGMIN = 1e-5
    (grhd.h), a non-finite sqrt(detg) becomes 1, and a point left non-finite
    is zeroed and counted (the reference's errorexit)."""
    rhoatm = 1.28e-13
    P = O.c2p_params(eos="ideal", gamma=2.0, K=100.0, rhoatm=rhoatm, errormax=1e-12)
    prim5, g, _ = sample_prims(4000, seed=11)
    n = prim5.shape[1]
    rng = np.random.default_rng(3)
    thr = P.fatm * rhoatm
    rho = np.where(rng.uniform(size=n) < 0.5, thr * 10.0 ** rng.uniform(-4, 0, n), thr * 10.0 ** rng.uniform(0, 4, n))
    rho[:4] = [thr, np.nextafter(thr, 0), 0.0, -1e-30]             # boundary of the test, zero, negative
    cons = O.p2c(P, prim5, g)
    prim = np.zeros((8, n))
    prim[0] = rho
    c11 = g[3] * g[5] - g[4] ** 2
    c12 = g[2] * g[4] - g[1] * g[5]
    c13 = g[1] * g[4] - g[2] * g[3]
    prim[7] = g[0] * c11 + g[1] * c12 + g[2] * c13
    prim[7, 4:8] = [1e-7, 0.0, np.inf, np.nan]                    # floored / non-finite detg
    rho[4:8] = 0.0
    prim[0] = rho
    before, detg0 = cons.copy(), prim[7].copy()
    bad = O.grhd_atm_set_u(P, cons, prim)
    assert bad == 1                                                   # the inf detg point (below)
    atm = rho < thr
    assert atm[1] and not atm[0] and atm[2] and atm[3]
    np.testing.assert_array_equal(cons[:, ~atm], before[:, ~atm])
    np.testing.assert_array_equal(prim[7, ~atm], detg0[~atm])
    # independent prim -> cons of the atmosphere state on the same metrics
    ga = g[:, atm]
    pa = np.zeros((5, ga.shape[1]))
    pa[0], pa[1] = rhoatm, P.epslatm
    ca = O.p2c(P, pa, ga)
    finite_det = np.isfinite(detg0[atm]) & (detg0[atm] > 1e-5)
    np.testing.assert_allclose(cons[0, atm][finite_det], ca[0][finite_det], rtol=2e-16, atol=0)
    ok = np.arange(n)[atm] != 6
    np.testing.assert_allclose(cons[1, atm][ok] / cons[0, atm][ok], P.epslatm, rtol=4e-16)
    np.testing.assert_array_equal(cons[2:, atm], 0.0)
    # the floor (a NaN detg fails DMAX's comparison and takes the floor too)
    assert prim[7, 4] == 1e-5 and prim[7, 5] == 1e-5 and prim[7, 7] == 1e-5
    assert cons[0, 4] == np.sqrt(1e-5) * rhoatm
    # detg = inf: sqrt(inf) -> 1, but detg itself stays non-finite: zeroed, detg = 1, counted
    assert prim[7, 6] == 1.0 and np.all(cons[:, 6] == 0.0)
    # a non-finite conserved value at a point above the threshold is not cured
    cons2, prim2 = before.copy(), np.zeros((8, n))
    prim2[0] = 1.0
    prim2[7] = 1.0
    cons2[2, 10] = np.nan
    assert O.grhd_atm_set_u(P, cons2, prim2) == 1
    assert np.all(cons2[:, 10] == 0.0) and prim2[7, 10] == 1.0


# ------------------------------------------------ analytic recovery (vanal)
def _vanal_kappa(coef, v):
    """conditioning of rho = sqrt(1 - v^2) D through the quartic's root:
    (sum_k |a_k v^k|) / (v |q'(v)|) * v^2 / (1 - v^2), plus 1"""
    a3, a2, a1, a0 = coef
    q1 = 4 * v ** 3 + 3 * a3 * v ** 2 + 2 * a2 * v + a1
    sc = v ** 4 + np.abs(a3) * v ** 3 + np.abs(a2) * v ** 2 + np.abs(a1) * v + np.abs(a0)
    return sc / np.abs(q1) * v / (1 - v * v) + 1


def test_vanal_quartic_vanishes_at_the_true_speed():
    """flat metric: the quartic built from (D, tau, S) has the original |v|
    as a root, to rounding of its terms (a dropped term or a wrong sign in
    a3..a0 leaves an O(1) residual)"""
    prim5, g, _ = sample_prims(20000, seed=3, metric_noise=0.0)
    P = O.c2p_params(**P12)
    cons = O.p2c(P, prim5, g)
    _, st, coef = O.c2p_vanal(P, cons, g)
    assert np.all(st == O.C2P_OK)
    v = np.sqrt((prim5[2:5] ** 2).sum(axis=0))
    a3, a2, a1, a0 = coef
    res = v ** 4 + a3 * v ** 3 + a2 * v ** 2 + a1 * v + a0
    sc = v ** 4 + np.abs(a3) * v ** 3 + np.abs(a2) * v ** 2 + np.abs(a1) * v + np.abs(a0)
    assert np.max(np.abs(res) / sc) < 1e-14


def test_vanal_root_against_high_precision():
    """the Ferrari closed form (resolvent cubic, quadratic factor) picks the
    physical root: rho against sqrt(1 - v^2) D with v from a 50-digit
    polynomial root of the same FP64 coefficients, within 32 eps times the
    root's conditioning"""
    mp.mp.dps = 50
    prim5, g, _ = sample_prims(300, seed=5, metric_noise=0.0)
    P = O.c2p_params(**P12)
    cons = O.p2c(P, prim5, g)
    prim, st, coef = O.c2p_vanal(P, cons, g)
    assert np.all(st == O.C2P_OK)
    vtrue = np.sqrt((prim5[2:5] ** 2).sum(axis=0))
    for q in range(prim5.shape[1]):
        a = [mp.mpf(float(c)) for c in coef[:, q]]
        roots = mp.polyroots([1] + a, maxsteps=200, extraprec=200)
        real = [r.real for r in roots if abs(r.imag) < mp.mpf(10) ** -30 and 0 <= r.real < 1]
        v = min(real, key=lambda r: abs(r - vtrue[q]))
        rho = float(mp.sqrt(1 - v * v) * mp.mpf(float(cons[0, q])))
        kap = _vanal_kappa(coef[:, q], float(v))
        assert abs(prim[0, q] - rho) <= 32 * np.finfo(float).eps * kap * rho, q


def test_vanal_static_fluid_closed_form():
    """S = 0: v = 0, rho = D, p = (Gamma - 1) tau, eps = tau / D, v^i = 0;
    v^2 (prim[6]) is left as it was"""
    n = 64
    rng = np.random.default_rng(4)
    cons = np.zeros((5, n))
    cons[0] = 10.0 ** rng.uniform(-8, -2, n)
    cons[1] = cons[0] * 10.0 ** rng.uniform(-4, 0, n)
    g = np.zeros((6, n))
    g[0] = g[3] = g[5] = 1.0
    prim0 = np.full((8, n), 7.0)
    P = O.c2p_params(**P12)
    prim, st, coef = O.c2p_vanal(P, cons, g, prim0=prim0)
    assert np.all(st == O.C2P_OK) and np.all(coef == 0.0)
    np.testing.assert_array_equal(prim[0], cons[0])
    # p = (Gamma - 1)((tau + D) - D): the rounding of tau + D is relative to D
    tol = 4 * np.finfo(float).eps * (cons[0] + cons[1]) / cons[1]
    assert np.all(np.abs(prim[5] - (P.gamma - 1) * cons[1]) <= tol * (P.gamma - 1) * cons[1])
    assert np.all(np.abs(prim[1] - cons[1] / cons[0]) <= 2 * tol * cons[1] / cons[0])
    assert np.all(prim[2:5] == 0.0) and np.all(prim[6] == 7.0) and np.all(prim[7] == 1.0)


def test_vanal_conformally_flat_pressure_term():
    """on gamma_ij = a delta_ij the reference forms S_i v_i with the lowered
    S_i and v_i = S_i |v| / |S|, i.e. a |S| |v| instead of |S| |v|: its
    pressure is the true one minus (Gamma - 1)(a - 1)|S||v| (rho and |v|,
    from invariants, are exact)"""
    a = 1.01
    prim5, g, _ = sample_prims(2000, seed=8, metric_noise=0.0)
    g[0] = g[3] = g[5] = a
    # contravariant v^i with a |v^i|^2 = 1 - 1/W^2
    prim5[2:5] /= np.sqrt(a)
    P = O.c2p_params(**P12)
    cons = O.p2c(P, prim5, g)
    prim, st, _ = O.c2p_vanal(P, cons, g)
    ok = st == O.C2P_OK
    assert ok.sum() > 100
    S = np.sqrt((cons[2:5] ** 2).sum(axis=0) / a) / a ** 1.5     # |S| = sqrt(S_i S^i), undensitised
    v = np.sqrt(a * (prim5[2:5] ** 2).sum(axis=0))
    ptrue = (P.gamma - 1) * prim5[0] * prim5[1]
    expect = ptrue - (P.gamma - 1) * (a - 1) * S * v
    big = ok & (np.abs(expect) > 1e-3 * (cons[1] + cons[0]) / a ** 1.5)
    np.testing.assert_allclose(prim[5][big], expect[big], rtol=1e-6)
    np.testing.assert_allclose(prim[0][ok], prim5[0][ok], rtol=1e-8)
    # where the term drives the pressure negative the reference stops (eps < 0)
    scale = 1e-10 * (cons[0] + cons[1]) / a ** 1.5
    assert np.all(st[expect < -scale] == O.C2P_FAIL) and np.all(st[expect > scale] == O.C2P_OK)


def test_vanal_agrees_with_the_newton_solvers_flat():
    """an independent method on the same inputs: rho from the analytic root
    and from the hybrid p-root (tolerance 1e-12) agree within the two
    methods' conditioning"""
    prim5, g, _ = sample_prims(20000, seed=9, metric_noise=0.0)
    P = O.c2p_params(**P12)
    cons = O.p2c(P, prim5, g)
    pa, sa, coef = O.c2p_vanal(P, cons, g)
    ph, sh, _, _ = O.c2p(P, cons, g, np.zeros(cons.shape[1]), rhoguess=prim5[0], method="bracketed")
    ok = (sa == O.C2P_OK) & (sh == O.C2P_OK)
    assert ok.mean() > 0.99
    v = np.sqrt((prim5[2:5] ** 2).sum(axis=0))
    kap = _vanal_kappa(coef, v)
    rel = np.abs(pa[0] - ph[0]) / ph[0]
    assert np.all(rel[ok] <= 1e-12 + 64 * np.finfo(float).eps * kap[ok])
