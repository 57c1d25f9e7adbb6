"""Pins the numpy oracle of the learned pressure guess: its analytic
gradient against central finite differences (float64, no bf16 rounding),
and its bf16 rounding against known round-to-nearest-even cases."""
import numpy as np

from oracle import mlp_oracle as M


def _params(seed):
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((64, 3)) * 0.8
    b1 = rng.standard_normal(64) * 0.1
    W2 = rng.standard_normal((64, 64)) * 0.18
    b2 = rng.standard_normal(64) * 0.1
    W3 = rng.standard_normal(64) * 0.12
    return M.pack(W1, b1, W2, b2, W3, 0.05)


def test_bf16_round_to_nearest_even():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)                  # bf16 spacing at 1
    cases = {
        one + ulp / 2: one,                      # tie -> even (1.0)
        one + 3 * ulp / 2: one + 2 * ulp,        # tie -> even (1 + 2 ulp)
        one + ulp / 2 + ulp / 64: one + ulp,     # above the tie
        np.float32(-3.0): np.float32(-3.0),
        np.float32(1e-30): np.float32(1e-30),    # same exponent range as float32
    }
    for x, want in cases.items():
        got = M.bf16(np.array([x], np.float32))[0]
        if x == np.float32(1e-30):
            assert abs(got / want - 1) < 2 ** -8
        else:
            assert got == want, (x, got, want)


def test_gradient_matches_finite_differences():
    p = _params(1)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((200, 3))
    t = rng.standard_normal(200)
    loss, g = M.loss_and_grad(p, x, t, emulate=False)
    h = 1e-6
    idx = np.concatenate([rng.choice(np.arange(M.OFF_W1, M.OFF_B1), 10, replace=False),
                          rng.choice(np.arange(M.OFF_B1, M.OFF_W2), 10, replace=False),
                          rng.choice(np.arange(M.OFF_W2, M.OFF_B2), 20, replace=False),
                          rng.choice(np.arange(M.OFF_B2, M.OFF_W3), 10, replace=False),
                          rng.choice(np.arange(M.OFF_W3, M.OFF_B3), 10, replace=False), [M.OFF_B3]])
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grad(pp, x, t, emulate=False)[0] - M.loss_and_grad(pm, x, t, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-8, (i, fd, g[i])


def test_emulated_forward_is_close_to_exact():
    """bf16 rounding perturbs the output by O(2^-8) of its scale, not more"""
    p = _params(3)
    x = np.random.default_rng(4).standard_normal((500, 3))
    ye, _ = M.forward(p, x, emulate=True)
    yx, _ = M.forward(p, x, emulate=False)
    assert np.max(np.abs(ye - yx)) < 0.05 * np.max(np.abs(yx))
    assert np.max(np.abs(ye - yx)) > 0
