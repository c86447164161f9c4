"""Pins of the analytic references (oracle/analytic.py).

The exact Riemann solver is checked against the published star states of
Toro's test 1 (Sod; Toro, table 4.3) and against the Rankine-Hugoniot
jump conditions; the Sedov-Taylor constant against the classical values of
xi0 for gamma = 1.4 and 5/3, and its profiles against the strong-shock
state and the conserved total energy.
"""
import numpy as np
import pytest

from oracle import analytic as A


def test_sod_star_state_toro_table_4_3():
    ps, us = A.star_state((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4)
    assert abs(ps - 0.30313) < 5e-6
    assert abs(us - 0.92745) < 5e-6
    w = A.exact_riemann((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4, [0.5 * us, 1.2 * us])
    assert abs(w[0, 0] - 0.42632) < 5e-6   # rho*L
    assert abs(w[1, 0] - 0.26557) < 5e-6   # rho*R


@pytest.mark.parametrize("case", [((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01)),      # Toro test 3
                                  ((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950))])  # test 5
def test_shock_jump_conditions(case):
    g = 1.4
    L, R = case
    ps, us = A.star_state(L, R, g)
    # right-moving shock: mass and momentum fluxes equal in its frame
    rr, ur, pr = R
    if ps > pr:
        s = ur + np.sqrt(g * pr / rr) * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
        rs = A.exact_riemann(L, R, g, [0.5 * (us + s)])[0, 0]
        m1, m2 = rr * (ur - s), rs * (us - s)
        assert abs(m1 - m2) < 1e-9 * abs(m1)
        assert abs((pr + rr * (ur - s) ** 2) - (ps + rs * (us - s) ** 2)) < 1e-9 * ps


def test_rarefaction_is_isentropic():
    g = 1.4
    L, R = (1.0, 0.0, 1.0), (0.125, 0.0, 0.1)
    w = A.exact_riemann(L, R, g, np.linspace(-1.18, -0.08, 50))
    np.testing.assert_allclose(w[:, 2] / w[:, 0] ** g, 1.0, rtol=1e-12)


@pytest.mark.parametrize("gamma,xi0", [(1.4, 1.033), (5.0 / 3.0, 1.152)])
def test_sedov_constant(gamma, xi0):
    assert abs(A.sedov_xi0(gamma) - xi0) < 1e-3


def test_sedov_profiles_behind_the_shock():
    g = 5.0 / 3.0
    xi, U, W, P = A.sedov_profiles(g)
    assert xi[0] == 1.0 and abs(W[0] - 4.0) < 1e-12 and abs(U[0] - 0.75) < 1e-12
    # density falls to zero at the centre, pressure tends to a plateau
    assert abs(W[-1]) < 1e-3 and 0.2 < P[-1] / P[0] < 0.6
    assert np.all(np.diff(W) < 1e-12)
    # homologous core: u = r / (gamma t), i.e. U / xi -> 1 / gamma
    assert abs(U[-1] / xi[-1] - 1.0 / g) < 1e-3
