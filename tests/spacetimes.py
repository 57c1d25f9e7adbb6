"""Analytic initial data for the tests (numpy, host side).

* Kerr-Schild Kerr black hole (stationary vacuum solution), optionally
  rotated and displaced so that every metric component is non-trivial.
  With its own lapse and shift the data are time independent, so the Z4c
  right-hand side of the non-gauge variables must vanish up to the
  truncation error of the finite differences (4th order).
* Brill-Lindquist punctures (time-symmetric, conformally flat).
* A generic smooth non-solution ("warped" data) for constraint checks.

Derivatives of the analytic expressions are taken by 8th-order centred
differences with a small step (1e-3), accurate to ~1e-12, far below the
truncation errors the tests measure.
"""
import numpy as np

SY = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
_D8 = np.array([1 / 280, -4 / 105, 1 / 5, -4 / 5, 0, 4 / 5, -1 / 5, 4 / 105, -1 / 280])


def grid(box):
    """coordinates (3, o, n, m) of a box given by an oracle.Box"""
    x = box.x0 + box.dx * np.arange(box.m)
    y = box.y0 + box.dy * np.arange(box.n)
    z = box.z0 + box.dz * np.arange(box.o)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return np.array([X, Y, Z])


def fd_grad(f, X, eps=1e-3):
    """d f / d x^k for a vectorised function f(X) -> array(..., *X.shape[1:])"""
    out = []
    for k in range(3):
        acc = 0
        for s, w in zip(range(-4, 5), _D8):
            if w == 0:
                continue
            Xs = X.copy()
            Xs[k] = Xs[k] + s * eps
            acc = acc + w * f(Xs)
        out.append(acc / eps)
    return np.array(out)


def rotation(axis_angles):
    ax, ay, az = axis_angles
    Rx = np.array([[1, 0, 0], [0, np.cos(ax), -np.sin(ax)], [0, np.sin(ax), np.cos(ax)]])
    Ry = np.array([[np.cos(ay), 0, np.sin(ay)], [0, 1, 0], [-np.sin(ay), 0, np.cos(ay)]])
    Rz = np.array([[np.cos(az), -np.sin(az), 0], [np.sin(az), np.cos(az), 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


class KerrSchild:
    def __init__(self, M=1.0, a=0.6, center=(0.3, -0.2, 0.1), angles=(0.4, -0.7, 0.25)):
        self.M, self.a = M, a
        self.c = np.array(center).reshape(3, 1, 1, 1)
        self.R = rotation(angles)

    def H_l(self, X):
        Xp = np.einsum("ai,i...->a...", self.R, X - self.c)
        x, y, z = Xp
        a, M = self.a, self.M
        rho2 = x * x + y * y + z * z
        r2 = 0.5 * (rho2 - a * a) + np.sqrt(0.25 * (rho2 - a * a) ** 2 + a * a * z * z)
        r = np.sqrt(r2)
        H = M * r ** 3 / (r2 * r2 + a * a * z * z)
        lp = np.array([(r * x + a * y) / (r2 + a * a), (r * y - a * x) / (r2 + a * a), z / r])
        l = np.einsum("ai,a...->i...", self.R, lp)
        return H, l

    def gamma(self, X):
        H, l = self.H_l(X)
        g = np.zeros((3, 3) + X.shape[1:])
        for i in range(3):
            for j in range(3):
                g[i, j] = (i == j) + 2 * H * l[i] * l[j]
        return g

    def beta_low(self, X):
        H, l = self.H_l(X)
        return 2 * H * l

    def alpha(self, X):
        H, _ = self.H_l(X)
        return 1.0 / np.sqrt(1 + 2 * H)

    def adm(self, X):
        """gamma_ij, K_ij, alpha, beta^i of the stationary slice"""
        g = self.gamma(X)
        gi = np.linalg.inv(np.moveaxis(g, (0, 1), (-2, -1)))
        gi = np.moveaxis(gi, (-2, -1), (0, 1))
        bl = self.beta_low(X)
        alp = self.alpha(X)
        dg = fd_grad(self.gamma, X)             # dg[k, i, j]
        db = fd_grad(self.beta_low, X)          # db[k, j] = d_k beta_j
        # gdo[k,i,j] = Gamma_kij = 1/2(d_i g_kj + d_j g_ki - d_k g_ij)
        gdo = 0.5 * (np.einsum("ikj...->kij...", dg) + np.einsum("jki...->kij...", dg) - dg)
        gam = np.einsum("kl...,lij...->kij...", gi, gdo)
        betau = np.einsum("ij...,j...->i...", gi, bl)
        Dbeta = db - np.einsum("kij...,k...->ij...", gam, bl)   # D_i beta_j
        K = (Dbeta + np.swapaxes(Dbeta, 0, 1)) / (2 * alp)
        return g, K, alp, betau

    def z4c(self, X, chipsipower=-4.0):
        """the 25 Z4c fields (list order) plus exact gauge data"""
        g, K, alp, betau = self.adm(X)
        gi = np.moveaxis(np.linalg.inv(np.moveaxis(g, (0, 1), (-2, -1))), (-2, -1), (0, 1))
        detg = np.linalg.det(np.moveaxis(g, (0, 1), (-2, -1)))
        chi = detg ** (-1.0 / 3.0)
        trK = np.einsum("ij...,ij...->...", gi, K)
        u = np.zeros((25,) + X.shape[1:])
        for c, (i, j) in enumerate(SY):
            u[c] = chi * g[i, j]
            u[7 + c] = chi * (K[i, j] - g[i, j] * trK / 3)
        u[6] = detg ** (chipsipower / 12.0)
        u[13] = trK

        def gtinv(Xs):
            gg = self.gamma(Xs)
            ggi = np.moveaxis(np.linalg.inv(np.moveaxis(gg, (0, 1), (-2, -1))), (-2, -1), (0, 1))
            dd = np.linalg.det(np.moveaxis(gg, (0, 1), (-2, -1)))
            return ggi * dd ** (1.0 / 3.0)

        dgti = fd_grad(gtinv, X)          # [k, i, j]
        Gam = -np.einsum("jij...->i...", dgti)
        u[14:17] = Gam
        u[17] = 0.0
        u[18] = alp
        u[19:22] = betau
        return u


def brill_lindquist(X, masses=((1.0, (0.0, 0.0, 0.0)),)):
    psi = 1.0
    for m, c in masses:
        r = np.sqrt((X[0] - c[0]) ** 2 + (X[1] - c[1]) ** 2 + (X[2] - c[2]) ** 2)
        psi = psi + m / (2 * r)
    return psi


def puncture_z4c(X, masses=((1.0, (0.0, 0.0, 0.0)),)):
    """BL data, chi = psi^-4, precollapsed lapse psi^-2, zero shift
    (z4.c:291-299, ConformalFactor.c:16-37)"""
    psi = brill_lindquist(X, masses)
    u = np.zeros((25,) + X.shape[1:])
    for c, (i, j) in enumerate(SY):
        u[c] = 1.0 if i == j else 0.0
    u[6] = psi ** -4
    u[18] = psi ** -2
    return u


def flat_z4c(shape):
    u = np.zeros((25,) + tuple(shape))
    for c, (i, j) in enumerate(SY):
        u[c] = 1.0 if i == j else 0.0
    u[6] = 1.0
    u[18] = 1.0
    return u
