"""GPU parity of con2prim (HIP kernels through the C ABI) with the CPU oracle.

Both sides evaluate the same operations in the same order with no FMA
contraction (x86-64 SSE2 on the host, `fp contract(off)` on the device,
IEEE-correct division and square root on both), so status codes, iteration
counts and every output double must agree exactly.
"""
import numpy as np
import pytest

import baum_amd as B
import oracle as O
from c2p_data import sample_prims

pytestmark = pytest.mark.gpu

METHODS = {"hybrid": B.C2P_HYBRID, "plain": B.C2P_PLAIN, "bracketed": B.C2P_BRACKETED}


def _run_gpu(Pg, method, cons, g, prim0):
    n = cons.shape[1]
    dc = B.DeviceArray.from_host(cons)
    dg = B.DeviceArray.from_host(g)
    dp = B.DeviceArray.from_host(prim0)
    ds = B.DeviceArray((n,), np.int32)
    di = B.DeviceArray((n,), np.int32)
    B.c2p_batch(Pg, METHODS[method], dc, dg, dp, ds, di)
    out = dc.get(), dp.get(), ds.get(), di.get()
    for a in (dc, dg, dp, ds, di):
        a.free()
    return out


def _edge_states(n=64, seed=5):
    """unphysical and atmosphere-level states mixed with regular ones"""
    prim5, g, _ = sample_prims(n, seed=seed)
    P = O.c2p_params(errormax=1e-10)
    cons = O.p2c(P, prim5, g)
    cons[1, 0:8] *= -3.0            # negative tau: cold branch
    cons[0, 8:12] = 0.0             # D = 0
    cons[0, 12:16] *= 1e-14         # below the atmosphere threshold
    cons[2, 16:20] *= 1e3           # S^2 far above (tau + D)^2
    return cons, g


@pytest.mark.parametrize("method", ["hybrid", "plain", "bracketed"])
@pytest.mark.parametrize("tol", [1e-10, 1e-12])
def test_c2p_matches_oracle_exactly(method, tol):
    n = 100000
    prim5, g, _ = sample_prims(n, seed=11)
    P = O.c2p_params(errormax=tol, countmax=100, rhoatm=1e-20)
    Pg = B.c2p_params(errormax=tol, countmax=100, rhoatm=1e-20)
    cons = O.p2c(P, prim5, g)
    guess = np.zeros(n)
    prim0 = np.zeros((8, n))
    prim0[0] = prim5[0]
    prim0[5] = guess
    cp, cs, ci, cc = O.c2p(P, cons, g, guess, rhoguess=prim5[0], method=method)
    gc, gp, gs, gi = _run_gpu(Pg, method, cons, g, prim0)
    np.testing.assert_array_equal(gs, cs)
    np.testing.assert_array_equal(gi, ci)
    ok = cs == O.C2P_OK
    np.testing.assert_array_equal(gp[:, ok], cp[:, ok])
    np.testing.assert_array_equal(gc, cc)


@pytest.mark.parametrize("method", ["hybrid", "plain", "bracketed"])
def test_c2p_edge_states_match_oracle(method):
    cons, g = _edge_states()
    n = cons.shape[1]
    P = O.c2p_params(eos="ideal", rhoatm=1e-12, errormax=1e-10)
    Pg = B.c2p_params(eos="ideal", rhoatm=1e-12, errormax=1e-10)
    prim0 = np.zeros((8, n))
    prim0[0] = 1e-4
    cp, cs, ci, cc = O.c2p(P, cons, g, np.zeros(n), rhoguess=prim0[0], method=method)
    gc, gp, gs, gi = _run_gpu(Pg, method, cons, g, prim0)
    np.testing.assert_array_equal(gs, cs)
    np.testing.assert_array_equal(gi, ci)
    fin = np.isfinite(cp)
    np.testing.assert_array_equal(np.isfinite(gp), fin)
    np.testing.assert_array_equal(gp[fin], cp[fin])
    np.testing.assert_array_equal(gc, cc)


def test_c2p_polytrope_hybrid():
    """polytropic EOS: pow() is not correctly rounded on either side (host
    libm, device OCML), so values are compared to 1e-11 instead of bitwise
    and 99.9% of the status codes must agree"""
    n = 20000
    rng = np.random.default_rng(4)
    rho = 10 ** rng.uniform(-6, -2.5, n)
    K, G = 100.0, 2.0
    eps = K * rho ** (G - 1) / (G - 1)
    prim5 = np.array([rho, eps, 0.3 * rng.uniform(-1, 1, n), 0.3 * rng.uniform(-1, 1, n), 0.2 * rng.uniform(-1, 1, n)])
    g = np.zeros((6, n))
    g[0] = g[3] = g[5] = 1.0
    P = O.c2p_params(eos="poly", K=K, gamma=G, rhoatm=1e-12, errormax=1e-10)
    Pg = B.c2p_params(eos="poly", K=K, gamma=G, rhoatm=1e-12, errormax=1e-10)
    cons = O.p2c(P, prim5, g)
    prim0 = np.zeros((8, n))
    prim0[0] = rho
    cp, cs, ci, cc = O.c2p(P, cons, g, np.zeros(n), rhoguess=rho, method="hybrid")
    gc, gp, gs, gi = _run_gpu(Pg, "hybrid", cons, g, prim0)
    assert np.mean(gs == cs) > 0.999
    same = gs == cs
    np.testing.assert_allclose(gp[:, same], cp[:, same], rtol=1e-11, atol=1e-300)


def test_p2c_matches_oracle_exactly():
    """grhd_compute_q_pt on the device (the generator of the c2p benchmark
    inputs) against the oracle's restatement"""
    n = 50000
    prim5, g, _ = sample_prims(n, seed=21)
    P = O.c2p_params(errormax=1e-10)
    Pg = B.c2p_params(errormax=1e-10)
    dp, dg, dc = B.DeviceArray.from_host(prim5), B.DeviceArray.from_host(g), B.DeviceArray((5, n))
    B.p2c_batch(Pg, dp, dg, dc)
    got = dc.get()
    for a in (dp, dg, dc):
        a.free()
    np.testing.assert_array_equal(got, O.p2c(P, prim5, g))


def _vanal_bounds(cons, g, coef, prim):
    """per point: the root conditioning kappa_v = sum|a_k v^k| / (v |q'(v)|) + 1
    (v from the oracle's rho: v = sqrt(1 - (rho/D)^2)) and the magnitude of
    the terms of p / (Gamma - 1) = E - S_i v_i - rho"""
    det = (g[0] * (g[3] * g[5] - g[4] ** 2) - g[1] * (g[1] * g[5] - g[2] * g[4]) + g[2] * (g[1] * g[4] - g[2] * g[3]))
    oo = 1 / np.sqrt(det)
    D, E = cons[0] * oo, (cons[0] + cons[1]) * oo
    x = np.clip(prim[0] / D, 0.0, 1.0)
    v = np.sqrt(np.maximum(1 - x * x, 1e-300))
    a3, a2, a1, a0 = coef
    q1 = 4 * v ** 3 + 3 * a3 * v ** 2 + 2 * a2 * v + a1
    sc = v ** 4 + np.abs(a3) * v ** 3 + np.abs(a2) * v ** 2 + np.abs(a1) * v + np.abs(a0)
    kv = sc / np.maximum(np.abs(q1) * v, 1e-300) + 1
    Sabs = np.sqrt((cons[2:5] ** 2).sum(axis=0)) * oo
    return kv, v, E + Sabs * v + prim[0]


@pytest.mark.parametrize("metric_noise", [0.0, 1e-3])
def test_c2p_vanal_matches_oracle(metric_noise):
    """grhd_c2p_vanal on the device against the oracle.  pow, cos and atan2
    of the device library and of glibc differ in the last bits, so values
    agree within the conditioning of the quartic's root: rho to 64 eps
    kappa_v v^2/(1-v^2), p to 64 eps kappa_v times the magnitude of its
    terms; statuses agree except where the oracle's pressure is within that
    bar of zero (the eps < 0 stop)"""
    n = 100000
    prim5, g, _ = sample_prims(n, seed=31, metric_noise=metric_noise)
    P = O.c2p_params(errormax=1e-12, rhoatm=1e-20)
    Pg = B.c2p_params(errormax=1e-12, rhoatm=1e-20)
    cons = O.p2c(P, prim5, g)
    prim0 = np.zeros((8, n))
    prim0[6] = 0.5                       # v^2: not written by the analytic recovery
    cp, cs, coef = O.c2p_vanal(P, cons, g, prim0=prim0)
    dc, dg, dp = B.DeviceArray.from_host(cons), B.DeviceArray.from_host(g), B.DeviceArray.from_host(prim0)
    ds, di = B.DeviceArray((n,), np.int32), B.DeviceArray((n,), np.int32)
    B.c2p_batch(Pg, B.C2P_VANAL, dc, dg, dp, ds, di)
    gp, gs, gi, gc = dp.get(), ds.get(), di.get(), dc.get()
    for a in (dc, dg, dp, ds, di):
        a.free()
    np.testing.assert_array_equal(gc, cons)          # the analytic recovery never rewrites
    assert np.all(gi == 0) and np.all(gp[6] == 0.5)
    np.testing.assert_array_equal(gp[7], cp[7])      # detg: the same cofactor sums
    eps = np.finfo(float).eps
    kv, v, pterms = _vanal_bounds(cons, g, coef, cp)
    pbar = 64 * eps * kv * (P.gamma - 1) * pterms
    flip = gs != cs
    assert np.all(np.abs(cp[5][flip]) <= pbar[flip] + 64 * eps * np.abs(cp[5][flip])), flip.sum()
    ok = (gs == O.C2P_OK) & (cs == O.C2P_OK)
    assert ok.mean() > (0.99 if metric_noise == 0.0 else 0.3)
    rbar = 64 * eps * (1 + kv * v * v / (1 - v * v)) * cp[0]
    assert np.all(np.abs(gp[0] - cp[0])[ok] <= rbar[ok])
    assert np.all(np.abs(gp[5] - cp[5])[ok] <= pbar[ok] + 64 * eps * np.abs(cp[5][ok]))
    print("vanal: %d of %d points OK on both sides, %d status flips" % (ok.sum(), n, flip.sum()), flush=True)


def test_level_bracketed_follows_hybrid_control_flow():
    """MATTER.c2p with the bracketed solver on a level: the reference hybrid's
    cold branch, atmosphere and conserved rewrites (bitwise where the state
    sits below the atmosphere density or has negative tau), the same root as
    the reference where its Newton iteration converges, and a converged root
    where the reference stalls at its rounding floor"""
    n = 1 << 15
    prim5, g, W = sample_prims(n, seed=53)
    Po = O.c2p_params(errormax=1e-12, countmax=100, rhoatm=1e-12)
    cons = O.p2c(Po, prim5, g)
    cons[:, :256] *= 1e-6            # atmosphere-level densities
    cons[1, 256:512] *= -2.0         # negative tau: the cold branch
    pprev = (Po.gamma - 1) * prim5[0] * prim5[1] * 1.3
    cp, cs, ci, cc = O.c2p(Po, cons, g, pprev, rhoguess=prim5[0], method="hybrid")
    lev = B.Level([B.make_box(n, 1, 1, 0.0, 0.0, 0.0, 1.0)], 20)
    lev.enable(range(20))
    for k in range(5):
        lev.put(k, cons[k])
    for k in range(6):
        lev.put(5 + k, g[k])
    prim0 = np.zeros((8, n))
    prim0[0], prim0[5] = prim5[0], pprev
    for k in range(8):
        lev.put(11 + k, prim0[k])
    ds, di = B.DeviceArray((n,), np.int32), B.DeviceArray((n,), np.int32)
    Pg = B.c2p_params(errormax=1e-12, countmax=100, rhoatm=1e-12)
    lev.grhd_c2p(Pg, B.C2P_BRACKETED, list(range(5)), list(range(5, 11)), list(range(11, 19)), ds, di)
    st = ds.get()
    gp = np.array([lev.get(11 + k) for k in range(8)])
    gc = np.array([lev.get(k) for k in range(5)])
    lev.free()
    ds.free()
    di.free()
    low = (cons[0] / np.sqrt(np.abs(g[0] * (g[3] * g[5] - g[4] ** 2) - g[1] * (g[1] * g[5] - g[2] * g[4])
                                    + g[2] * (g[1] * g[4] - g[2] * g[3])))) < Po.fatm * Po.rhoatm
    # below the atmosphere density: exactly the reference hybrid
    np.testing.assert_array_equal(st[low], cs[low])
    np.testing.assert_array_equal(gp[:, low], cp[:, low])
    np.testing.assert_array_equal(gc[:, low], cc[:, low])
    # Newton failures only where the reference fails too (a handful); where
    # the reference converged, the same status
    assert np.all(cs[st == B.C2P_NR] == O.C2P_NR) and (st == B.C2P_NR).sum() <= 1e-3 * n
    okc = cs == O.C2P_OK
    assert np.all(st[okc] == O.C2P_OK)
    kappa = W ** 2 * (cons[1] + cons[0]) / np.maximum(cp[5], 1e-300)
    bound = np.maximum(64 * np.finfo(float).eps * kappa, 1e-12)
    assert np.all(np.abs(gp[5] / cp[5] - 1)[okc] <= bound[okc])
    # where the reference stalled (NR -> atmosphere), the bracketed root is a root
    stalled = cs == O.C2P_NR
    assert stalled.sum() > 0 and np.mean(st[stalled] == O.C2P_OK) > 0.99
