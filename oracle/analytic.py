"""ORACLE — test infrastructure, not product code.

Analytic references for the moving-mesh solver (the reference code has no
moving mesh, so these pin the physics instead of a reference output):

* exact_riemann: the exact solution of the 1D Riemann problem for the
  ideal-gas Euler equations (Toro, "Riemann Solvers and Numerical Methods
  for Fluid Dynamics", ch. 4: pressure function, Newton iteration for p*,
  sampling of the wave pattern);
* sedov_xi0: the constant of the Sedov-Taylor point blast,
  R(t) = xi0 (E t^2 / rho0)^(1/5) in 3D, from this module's own integration
  of the self-similar ODEs and the energy integral.
"""
from __future__ import annotations

import numpy as np


# ------------------------------------------------------------------ Riemann
def _f_and_df(p, rho, u, pk, gamma):
    """Toro (4.6)-(4.7): pressure function of one side and its derivative."""
    c = np.sqrt(gamma * pk / rho)
    if p > pk:  # shock
        A = 2.0 / ((gamma + 1.0) * rho)
        B = (gamma - 1.0) / (gamma + 1.0) * pk
        q = np.sqrt(A / (p + B))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (B + p))
    # rarefaction
    e = (gamma - 1.0) / (2.0 * gamma)
    return 2.0 * c / (gamma - 1.0) * ((p / pk) ** e - 1.0), (p / pk) ** (-(gamma + 1.0) / (2.0 * gamma)) / (rho * c)


def star_state(L, R, gamma):
    """(p*, u*) of the Riemann problem with primitive states L, R = (rho, u, p)."""
    rl, ul, pl = L
    rr, ur, pr = R
    cl, cr = np.sqrt(gamma * pl / rl), np.sqrt(gamma * pr / rr)
    if 2.0 * (cl + cr) / (gamma - 1.0) <= ur - ul:
        raise ValueError("vacuum is generated")
    # two-rarefaction guess (Toro 4.46)
    e = (gamma - 1.0) / (2.0 * gamma)
    p = ((cl + cr - 0.5 * (gamma - 1.0) * (ur - ul)) / (cl / pl ** e + cr / pr ** e)) ** (1.0 / e)
    p = max(p, 1e-12)
    for _ in range(100):
        fl, dl = _f_and_df(p, rl, ul, pl, gamma)
        fr, dr = _f_and_df(p, rr, ur, pr, gamma)
        pn = p - (fl + fr + ur - ul) / (dl + dr)
        pn = max(pn, 1e-12)
        if abs(pn - p) <= 1e-14 * 0.5 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _f_and_df(p, rl, ul, pl, gamma)
    fr, _ = _f_and_df(p, rr, ur, pr, gamma)
    return p, 0.5 * (ul + ur) + 0.5 * (fr - fl)


def exact_riemann(L, R, gamma, s):
    """Primitive state (rho, u, p) at the similarity coordinates s = (x - x0)/t
    (Toro's sampling procedure, section 4.5)."""
    rl, ul, pl = L
    rr, ur, pr = R
    ps, us = star_state(L, R, gamma)
    cl, cr = np.sqrt(gamma * pl / rl), np.sqrt(gamma * pr / rr)
    g1, g2 = (gamma - 1.0) / (gamma + 1.0), 2.0 / (gamma + 1.0)
    s = np.atleast_1d(np.asarray(s, dtype=np.float64))
    out = np.empty((s.size, 3))
    for i, si in enumerate(s):
        if si <= us:  # left of the contact
            if ps > pl:  # left shock
                sl = ul - cl * np.sqrt((gamma + 1) / (2 * gamma) * ps / pl + (gamma - 1) / (2 * gamma))
                if si <= sl:
                    out[i] = (rl, ul, pl)
                else:
                    out[i] = (rl * (ps / pl + g1) / (g1 * ps / pl + 1.0), us, ps)
            else:  # left rarefaction
                shl = ul - cl
                csl = cl * (ps / pl) ** ((gamma - 1) / (2 * gamma))
                stl = us - csl
                if si <= shl:
                    out[i] = (rl, ul, pl)
                elif si >= stl:
                    out[i] = (rl * (ps / pl) ** (1.0 / gamma), us, ps)
                else:
                    c = g2 * (cl + 0.5 * (gamma - 1) * (ul - si))
                    u = g2 * (cl + 0.5 * (gamma - 1) * ul + si)
                    r = rl * (c / cl) ** (2.0 / (gamma - 1))
                    out[i] = (r, u, pl * (c / cl) ** (2 * gamma / (gamma - 1)))
        else:  # right of the contact
            if ps > pr:  # right shock
                sr = ur + cr * np.sqrt((gamma + 1) / (2 * gamma) * ps / pr + (gamma - 1) / (2 * gamma))
                if si >= sr:
                    out[i] = (rr, ur, pr)
                else:
                    out[i] = (rr * (ps / pr + g1) / (g1 * ps / pr + 1.0), us, ps)
            else:  # right rarefaction
                shr = ur + cr
                csr = cr * (ps / pr) ** ((gamma - 1) / (2 * gamma))
                st = us + csr
                if si >= shr:
                    out[i] = (rr, ur, pr)
                elif si <= st:
                    out[i] = (rr * (ps / pr) ** (1.0 / gamma), us, ps)
                else:
                    c = g2 * (cr - 0.5 * (gamma - 1) * (ur - si))
                    u = g2 * (-cr + 0.5 * (gamma - 1) * ur + si)
                    r = rr * (c / cr) ** (2.0 / (gamma - 1))
                    out[i] = (r, u, pr * (c / cr) ** (2 * gamma / (gamma - 1)))
    return out


# -------------------------------------------------------------------- Sedov
def sedov_profiles(gamma, xi_min=1e-3, n=4000):
    """Self-similar profiles of the 3D strong point blast behind the shock.

    With R = A t^(2/5), xi = r/R, u = R' U(xi), rho = rho0 W(xi),
    p = rho0 R'^2 P(xi), mass, momentum and entropy conservation become
      (U - xi) W' + W U' = -2 W U / xi
      (U - xi) U' + P' / W = 3/2 U
      (U - xi) (P'/P - gamma W'/W) = 3
    integrated inward from the strong-shock state at xi = 1:
    W = (g+1)/(g-1), U = P = 2/(g+1).  Returns xi (descending), U, W, P.
    Below xi ~ 1e-3 the density (~ xi^(3/(g-1))) underflows the tolerance
    and the core is the homologous limit U/xi = 1/g, P = const."""
    from scipy.integrate import solve_ivp

    g = gamma

    def rhs(xi, y):
        U, W, P = y
        D = U - xi
        num = 1.5 * U * D - 3.0 * P / W + 2.0 * g * P * U / (W * xi)
        dU = num / (D * D - g * P / W)
        dW = -(W * dU + 2.0 * W * U / xi) / D
        dP = P * (3.0 / D + g * dW / W)
        return [dU, dW, dP]

    y0 = [2.0 / (g + 1.0), (g + 1.0) / (g - 1.0), 2.0 / (g + 1.0)]
    xs = np.geomspace(1.0, xi_min, n)
    sol = solve_ivp(rhs, (1.0, xi_min), y0, t_eval=xs, rtol=1e-11, atol=1e-14, method="DOP853")
    if not sol.success:
        raise RuntimeError(sol.message)
    return sol.t, sol.y[0], sol.y[1], sol.y[2]


def sedov_xi0(gamma):
    """xi0 with E = (16 pi / 25) rho0 R^5 / t^2 * I, I = int_0^1 (W U^2/2 +
    P/(gamma-1)) xi^2 dxi, so that R = xi0 (E t^2 / rho0)^(1/5)."""
    xi, U, W, P = sedov_profiles(gamma)
    f = (0.5 * W * U * U + P / (gamma - 1.0)) * xi * xi
    # ascending order for the trapezoid rule, plus the core below xi_min
    # (constant pressure, vanishing density)
    I = np.trapezoid(f[::-1], xi[::-1]) + P[-1] / (gamma - 1.0) * xi[-1] ** 3 / 3.0
    return (25.0 / (16.0 * np.pi * I)) ** 0.2
