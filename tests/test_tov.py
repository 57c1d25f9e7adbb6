"""The TOV solver in the log enthalpy (bamhip_tov_h, tov_h of src/projects/
BBIDdataReader/tov.c:104-236) on the host, against an independent
integration of the TOV equations in the Schwarzschild radius (scipy, tight
tolerances) for the reference's default star (hc = 0.227869, K = 100,
Gamma = 2, rho_c = 1.28e-3: M = 1.400, R = 9.59)."""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

import baum_amd as B

K, G, HC = 100.0, 2.0, 0.227869


@pytest.fixture(scope="module")
def star():
    return B.tov_h(HC, K, G, 100000)


def scipy_star(rhoc):
    def rhs(r, y):
        m, p, nu, lnrb = y
        rho = (max(p, 0.0) / K) ** (1 / G)
        e = rho + p / (G - 1)
        a = (m + 4 * np.pi * r ** 3 * p) / (r * (r - 2 * m))
        return [4 * np.pi * r * r * e, -(e + p) * a, a, 1 / (r * np.sqrt(1 - 2 * m / r))]

    pc = K * rhoc ** G
    r0 = 1e-6
    y0 = [4 / 3 * np.pi * r0 ** 3 * (rhoc + pc / (G - 1)), pc, 0.0, np.log(r0)]
    surf = lambda r, y: y[1] - 1e-16
    surf.terminal = True
    s = solve_ivp(rhs, (r0, 20.0), y0, rtol=1e-12, atol=1e-20, events=surf, dense_output=True)
    R = s.t_events[0][0]
    return s, R


def test_default_star_mass_and_radius(star):
    rhoc = ((G - 1) / (G * K) * np.expm1(HC)) ** (1 / (G - 1))
    rho0 = ((G - 1) / (G * K) * np.expm1(HC - HC / 100000)) ** (1 / (G - 1))   # one step below the centre
    assert abs(star["rho"][0] - rho0) < 1e-13 * rho0 and abs(rho0 - 1.28e-3) < 1e-6
    s, R = scipy_star(rhoc)
    M = s.sol(R)[0]
    assert abs(star["M"] - M) < 2e-5 * M and abs(star["R"] - R) < 2e-4 * R
    assert abs(star["M"] - 1.400) < 2e-3 and abs(star["R"] - 9.59) < 1e-2
    # isotropic surface radius: the exterior Schwarzschild relation
    Rb = 0.5 * (R - M + np.sqrt(R * (R - 2 * M)))
    assert abs(star["Riso"] - Rb) < 2e-4 * Rb
    # the potential at the centre and the isotropic radius inside, from the same integration
    y = s.sol(R)
    nu_c = 0.5 * np.log(1 - 2 * M / R) - y[2]
    assert abs(star["phi"][0] - nu_c) < 1e-5
    rm = 0.5 * R
    j = np.searchsorted(star["r"], rm)
    rb_mid = Rb * np.exp(s.sol(rm)[3] - y[3])
    assert abs(np.interp(rm, star["r"], star["riso"]) - rb_mid) < 1e-5 * rb_mid, j
    # baryonic mass: 4 pi int r^2 rho / sqrt(1 - 2m/r) dr (first-order sum in the reference)
    rr = np.linspace(1e-6, R, 20001)
    yy = s.sol(rr)
    rho = (np.maximum(yy[1], 0) / K) ** (1 / G)
    Mb = 4 * np.pi * np.trapezoid(rr ** 2 * rho / np.sqrt(1 - 2 * yy[0] / rr), rr)
    assert abs(star["Mb"] - Mb) < 1e-3 * Mb


def test_profiles_are_monotone_and_consistent(star):
    assert np.all(np.diff(star["h"]) < 0) and np.all(np.diff(star["r"]) > 0) and np.all(np.diff(star["riso"]) > 0)
    np.testing.assert_allclose(star["pre"], K * star["rho"] ** G, rtol=1e-12, atol=1e-30)
    assert abs(star["phi"][-1] - 0.5 * np.log(1 - 2 * star["M"] / star["R"])) < 1e-14
    assert star["h"][-1] < 1e-5 and star["rho"][-1] < 1e-7


def test_bad_arguments():
    with pytest.raises(Exception):
        B.tov_h(-0.1, K, G, 1000)
    with pytest.raises(Exception):
        B.tov_h(HC, K, 1.0, 1000)
    with pytest.raises(Exception):
        B.tov_h(HC, K, G, 3)
