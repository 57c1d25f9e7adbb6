"""Synthetic GRHD states for the con2prim tests and benchmark.

Primitive distribution of the con2prim configuration: rho log-uniform in
[1e-10, 1e-2], eps log-uniform in [1e-6, 1], Lorentz factor W uniform in
[1, 10] along a direction uniform on the sphere, spatial metric flat plus a
seeded symmetric perturbation of 1e-3; the ideal-gas EOS with Gamma = 2.
"""
import numpy as np


def sample_prims(n, seed=0, metric_noise=1e-3):
    rng = np.random.default_rng(seed)
    rho = 10.0 ** rng.uniform(-10, -2, n)
    eps = 10.0 ** rng.uniform(-6, 0, n)
    W = rng.uniform(1.0, 10.0, n)
    d = rng.standard_normal((3, n))
    d /= np.linalg.norm(d, axis=0)
    g = np.zeros((6, n))
    g[0] = g[3] = g[5] = 1.0
    g += metric_noise * rng.uniform(-1, 1, (6, n))
    # contravariant v^i = s d^i with gamma_ij v^i v^j = 1 - 1/W^2
    gdd = (g[0] * d[0] ** 2 + g[3] * d[1] ** 2 + g[5] * d[2] ** 2
           + 2 * (g[1] * d[0] * d[1] + g[2] * d[0] * d[2] + g[4] * d[1] * d[2]))
    s = np.sqrt((1.0 - 1.0 / W ** 2) / gdd)
    v = s * d
    prim5 = np.array([rho, eps, v[0], v[1], v[2]])
    return prim5, g, W
