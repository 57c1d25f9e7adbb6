"""Pins the CPU oracle of BAM's GRHD flux divergence (matter_fluxes with
grhd_flx1d_fvpr) against what the mathematics fixes:
- WENOZ / WENO5 reproduce linear data at the half points, LAG4 cubics;
- on a uniform state the interface flux is the physical flux, whose closed
  form (conserved variables, alpha v - beta transport, pressure terms) is
  written here independently;
- in flat space with motion along the line the LLF / HLL dissipation uses
  the relativistic characteristic speeds (v +- cs) / (1 +- v cs);
- the divergence converges at second order to -d_x F(u) for smooth data, in
  each of the three directions, with the momentum component of that
  direction (the permutation tables, the sign, 1/h and the index offset);
- mirror symmetry, the c != 0 form, the cold EOS (no tau flux) and the
  zeroing of non-finite interface fluxes."""
import numpy as np
import pytest

import oracle as O


def test_reconstructions_exact_on_polynomials():
    x = np.arange(12, dtype=np.float64)
    lin = 0.3 - 1.7 * x
    cub = 0.3 - 1.7 * x + 0.2 * x ** 2 - 0.05 * x ** 3
    for i in range(3, 8):
        for kind in ("WENOZ", "WENO5"):
            assert abs(O.hrsc_rec(kind, 0, lin, i) - (0.3 - 1.7 * (i + 0.5))) < 1e-13
            assert abs(O.hrsc_rec(kind, 1, lin, i) - (0.3 - 1.7 * (i - 0.5))) < 1e-13
        xp, xm = i + 0.5, i - 0.5
        assert abs(O.hrsc_rec("LAG4", 0, cub, i) - (0.3 - 1.7 * xp + 0.2 * xp ** 2 - 0.05 * xp ** 3)) < 1e-12
        assert abs(O.hrsc_rec("LAG4", 1, cub, i) - (0.3 - 1.7 * xm + 0.2 * xm ** 2 - 0.05 * xm ** 3)) < 1e-12
    # a step: the WENO values stay within the data (no new extrema)
    step = np.where(x < 6, 0.0, 1.0)
    for i in range(3, 9):
        for kind in ("WENOZ", "WENO5"):
            for side in (0, 1):
                v = O.hrsc_rec(kind, side, step, i)
                assert -1e-6 <= v <= 1 + 1e-6


def _metric(seed=1):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-0.1, 0.1, (3, 3))
    g = np.eye(3) * 1.2 + 0.5 * (a + a.T)
    alpha, beta = 0.8, np.array([0.05, -0.03, 0.02])
    return alpha, beta, g


def _physical_flux(alpha, beta, g, rho, eps, v, gamma=2.0):
    """closed form: p = (G-1) rho eps, W = 1/sqrt(1 - v_i v^i),
    D = sqrt(g) W rho, tau = sqrt(g) (W^2 h rho - p - W rho), S_i = sqrt(g) W^2 h rho v_i;
    F = (alpha v^x - beta^x) q + alpha sqrt(g) p (0, v^x, delta_ix)"""
    p = (gamma - 1) * rho * eps
    vl = g @ v
    W = 1 / np.sqrt(1 - vl @ v)
    sg = np.sqrt(np.linalg.det(g))
    hrho = rho * (1 + eps) + p
    q = np.array([sg * W * rho, sg * (W * W * hrho - p - W * rho), *(sg * W * W * hrho * vl)])
    tr = alpha * v[0] - beta[0]
    return tr * q + alpha * sg * p * np.array([0, v[0], 1, 0, 0]), q


def test_uniform_state_gives_the_physical_flux():
    alpha, beta, g = _metric()
    gv = [alpha, beta[0], g[0, 0], g[0, 1], g[0, 2], g[1, 1], g[1, 2], g[2, 2]]
    w = [1.3e-3, 0.21, 0.31, -0.22, 0.12]
    for riemann in ("LLF", "HLL"):
        F = O.grhd_interface_flux(O.hrsc_params(riemann=riemann), gv, w, w)
        want, _ = _physical_flux(alpha, beta, g, w[0], w[1], np.array(w[2:]))
        np.testing.assert_allclose(F, want, rtol=1e-13, atol=1e-20)
    # a cold EOS carries no tau flux
    Fp = O.grhd_interface_flux(O.hrsc_params(eos=1), gv, w, w)
    assert Fp[1] == 0.0


def test_riemann_dissipation_uses_relativistic_characteristic_speeds():
    """flat space, motion along the line: lambda = alpha (v +- cs)/(1 +- v cs) - beta"""
    alpha, bx = 0.9, 0.04
    gv = [alpha, bx, 1, 0, 0, 1, 0, 1]
    wl = [1.0e-3, 0.3, 0.25, 0.0, 0.0]
    wr = [0.6e-3, 0.1, -0.1, 0.0, 0.0]
    lam = []
    fs, qs = [], []
    for w in (wl, wr):
        rho, eps, v = w[0], w[1], w[2]
        p = rho * eps                                   # ideal gas, Gamma = 2
        cs = np.sqrt(2.0 * p / (rho * (1 + eps + p / rho)))   # cs^2 = Gamma p / (rho h)
        lam += [alpha * (v + cs) / (1 + v * cs) - bx, alpha * (v - cs) / (1 - v * cs) - bx, alpha * v - bx]
        f, q = _physical_flux(alpha, np.array([bx, 0, 0]), np.eye(3), rho, eps, np.array(w[2:]))
        fs.append(f)
        qs.append(q)
    ap, am = abs(max(0.0, max(lam))), abs(min(0.0, min(lam)))
    F = O.grhd_interface_flux(O.hrsc_params(riemann="LLF"), gv, wl, wr)
    want = 0.5 * (fs[1] + fs[0] - max(ap, am) * (qs[1] - qs[0]))
    np.testing.assert_allclose(F, want, rtol=1e-12, atol=1e-22)
    F = O.grhd_interface_flux(O.hrsc_params(riemann="HLL"), gv, wl, wr)
    want = (ap * fs[0] + am * fs[1] - ap * am * (qs[1] - qs[0])) / (ap + am)
    np.testing.assert_allclose(F, want, rtol=1e-12, atol=1e-22)


def _box(m, n, o, h):
    return O.make_box(m, n, o, 0.0, 0.0, 0.0, h)


def _smooth_state(box, axis, h):
    """density wave along `axis`, uniform velocity along it, flat space"""
    shape = (box.o, box.n, box.m)
    s = (np.arange(shape[2 - axis]) * h).reshape([-1 if d == 2 - axis else 1 for d in range(3)])
    rho = 1e-3 * (1 + 0.2 * np.sin(2 * np.pi * s / 2.0)) * np.ones(shape)
    eps = 0.2 + 0.05 * np.cos(2 * np.pi * s / 2.0) * np.ones(shape)
    metric = np.zeros((10,) + shape)
    metric[0] = 1.0
    metric[4] = metric[7] = metric[9] = 1.0
    prim = np.zeros((5,) + shape)
    prim[0], prim[1] = rho, eps
    prim[2 + axis] = 0.3
    return metric, prim, s


def _exact_rhs(s, axis, gamma=2.0):
    """-d/ds of the physical flux of the wave above"""
    v = 0.3
    W = 1 / np.sqrt(1 - v * v)
    k = 2 * np.pi / 2.0
    rho = 1e-3 * (1 + 0.2 * np.sin(k * s))
    drho = 1e-3 * 0.2 * k * np.cos(k * s)
    eps = 0.2 + 0.05 * np.cos(k * s)
    deps = -0.05 * k * np.sin(k * s)
    p = (gamma - 1) * rho * eps
    dp = (gamma - 1) * (drho * eps + rho * deps)
    hrho = rho + rho * eps + p
    dhrho = drho + drho * eps + rho * deps + dp
    out = np.zeros((5,) + s.shape)
    out[0] = -W * drho * v
    out[1] = -((W * W * dhrho - dp - W * drho) * v + dp * v)
    out[2 + axis] = -(W * W * dhrho * v * v + dp)
    return out


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_divergence_converges_at_second_order(axis):
    errs = []
    for npts in (40, 80):
        h = 2.0 / npts
        dims = [8, 8, 8]
        dims[axis] = npts
        box = _box(*dims, h)
        metric, prim, s = _smooth_state(box, axis, h)
        rhs = np.zeros((5, box.o, box.n, box.m))
        O.grhd_fluxes(box, O.hrsc_params(), metric, prim, rhs)
        want = _exact_rhs(s, axis) * np.ones_like(rhs)
        inner = [slice(None)] * 4
        inner[3 - axis] = slice(4, npts - 4)
        err = np.abs(rhs - want)[tuple(inner)]
        scale = np.abs(want).max(axis=(1, 2, 3))
        for comp in (0, 1, 2 + axis):
            assert scale[comp] > 0
        others = [c for c in (2, 3, 4) if c != 2 + axis]
        assert np.abs(rhs[others]).max() == 0.0          # no momentum across the wave
        errs.append(np.array([err[c].max() / scale[c] for c in (0, 1, 2 + axis)]))
    ratio = errs[0] / errs[1]
    assert np.all(errs[1] < 2e-2), errs
    assert np.all(ratio > 3.5), ratio


def _random_level_state(box, seed):
    rng = np.random.default_rng(seed)
    shape = (box.o, box.n, box.m)
    metric = np.zeros((10,) + shape)
    metric[0] = 0.8 + 0.1 * rng.random(shape)
    metric[1:4] = 0.05 * rng.standard_normal((3,) + shape)
    a = 0.03 * rng.standard_normal((6,) + shape)
    metric[4:10] = a + np.array([1.1, 0, 0, 1.1, 0, 1.1])[:, None, None, None]
    prim = np.zeros((5,) + shape)
    prim[0] = 1e-3 * (1 + 0.5 * rng.random(shape))
    prim[1] = 0.1 + 0.1 * rng.random(shape)
    prim[2:5] = 0.2 * (rng.random((3,) + shape) - 0.5)
    return metric, prim


def test_mirror_symmetry_along_x():
    box = _box(14, 9, 7, 0.1)
    metric, prim = _random_level_state(box, 5)
    P = O.hrsc_params()
    rhs = O.grhd_fluxes(box, P, metric, prim, np.zeros((5, box.o, box.n, box.m)))
    mm, pm = metric[..., ::-1].copy(), prim[..., ::-1].copy()
    mm[1] *= -1          # beta^x
    mm[5] *= -1          # g_xy
    mm[6] *= -1          # g_xz
    pm[2] *= -1          # v^x
    rm = O.grhd_fluxes(box, P, mm, pm, np.zeros((5, box.o, box.n, box.m)))[..., ::-1]
    rm[2] *= -1
    scale = np.abs(rhs).max(axis=(1, 2, 3))[:, None, None, None]
    assert np.abs(rm - rhs).max() / scale.max() < 1e-12
    assert (np.abs(rm - rhs) / scale).max() < 1e-11


def test_c_form_and_cold_eos():
    box = _box(10, 9, 8, 0.1)
    metric, prim = _random_level_state(box, 6)
    P = O.hrsc_params()
    base = np.random.default_rng(7).standard_normal((5, box.o, box.n, box.m))
    div = O.grhd_fluxes(box, P, metric, prim, np.zeros_like(base))
    got = O.grhd_fluxes(box, P, metric, prim, base.copy(), c=0.37)
    np.testing.assert_allclose(got, base + 0.37 * div, rtol=1e-13, atol=1e-15)
    cold = O.grhd_fluxes(box, O.hrsc_params(eos=1, K=100.0), metric, prim, base.copy())
    np.testing.assert_array_equal(cold[1], base[1])     # no tau flux for a cold EOS
    assert np.abs(cold[0] - base[0]).max() > 0


def test_non_finite_interfaces_are_zeroed():
    box = _box(16, 6, 6, 0.1)
    metric, prim = _random_level_state(box, 8)
    prim[0, 3, 3, 8] = np.nan
    rhs = O.grhd_fluxes(box, O.hrsc_params(), metric, prim, np.zeros((5, box.o, box.n, box.m)))
    assert np.all(np.isfinite(rhs))
    # v >= vmax: W hits grhd_Wlor_max and the interface fluxes are zeroed
    prim2 = prim.copy()
    prim2[0, 3, 3, 8] = 1e-3
    prim2[2, :, :, :] = 0.999999
    rhs2 = O.grhd_fluxes(box, O.hrsc_params(vmax=0.9), metric, prim2, np.zeros((5, box.o, box.n, box.m)))
    assert np.all(np.isfinite(rhs2))
