"""Exact solution of the Riemann problem for the 1D Euler equations of an
ideal gas (Toro 2009, "Riemann Solvers and Numerical Methods for Fluid
Dynamics", chapter 4: pressure function eqs. 4.6-4.8, Newton iteration for p*
with the PVRS guess eq. 4.47, sampling section 4.5).  Independent of the
oracle; used to check it."""
import math

import numpy as np


def _f(p, rho, pk, ak, g):
    if p > pk:    # shock (eq. 4.6)
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pk
        q = math.sqrt(A / (B + p))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (B + p))
    # rarefaction (eq. 4.7)
    pr = p / pk
    f = 2.0 * ak / (g - 1.0) * (pr ** ((g - 1.0) / (2.0 * g)) - 1.0)
    df = 1.0 / (rho * ak) * pr ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(L, R, g):
    rL, uL, pL = L
    rR, uR, pR = R
    aL, aR = math.sqrt(g * pL / rL), math.sqrt(g * pR / rR)
    p = max(1e-10, 0.5 * (pL + pR) - 0.125 * (uR - uL) * (rL + rR) * (aL + aR))
    for _ in range(100):
        fL, dL = _f(p, rL, pL, aL, g)
        fR, dR = _f(p, rR, pR, aR, g)
        pn = p - (fL + fR + uR - uL) / (dL + dR)
        pn = max(pn, 1e-10)
        if abs(pn - p) / (0.5 * (pn + p)) < 1e-14:
            p = pn
            break
        p = pn
    fL, _ = _f(p, rL, pL, aL, g)
    fR, _ = _f(p, rR, pR, aR, g)
    u = 0.5 * (uL + uR) + 0.5 * (fR - fL)
    return p, u


def sample(L, R, g, xi):
    """(rho, u, p) at x/t = xi for each xi."""
    rL, uL, pL = L
    rR, uR, pR = R
    aL, aR = math.sqrt(g * pL / rL), math.sqrt(g * pR / rR)
    ps, us = star_state(L, R, g)
    out = np.zeros((len(xi), 3))
    gm = (g - 1.0) / (g + 1.0)
    for n, s in enumerate(xi):
        if s <= us:   # left of contact
            if ps > pL:
                SL = uL - aL * math.sqrt((g + 1) / (2 * g) * ps / pL + (g - 1) / (2 * g))
                out[n] = (rL, uL, pL) if s <= SL else (rL * (ps / pL + gm) / (gm * ps / pL + 1), us, ps)
            else:
                aLs = aL * (ps / pL) ** ((g - 1) / (2 * g))
                SHL, STL = uL - aL, us - aLs
                if s <= SHL:
                    out[n] = (rL, uL, pL)
                elif s >= STL:
                    out[n] = (rL * (ps / pL) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) + (g - 1) / ((g + 1) * aL) * (uL - s)
                    out[n] = (rL * c ** (2 / (g - 1)), 2 / (g + 1) * (aL + (g - 1) / 2 * uL + s),
                              pL * c ** (2 * g / (g - 1)))
        else:
            if ps > pR:
                SR = uR + aR * math.sqrt((g + 1) / (2 * g) * ps / pR + (g - 1) / (2 * g))
                out[n] = (rR, uR, pR) if s >= SR else (rR * (ps / pR + gm) / (gm * ps / pR + 1), us, ps)
            else:
                aRs = aR * (ps / pR) ** ((g - 1) / (2 * g))
                SHR, STR = uR + aR, us + aRs
                if s >= SHR:
                    out[n] = (rR, uR, pR)
                elif s <= STR:
                    out[n] = (rR * (ps / pR) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) - (g - 1) / ((g + 1) * aR) * (uR - s)
                    out[n] = (rR * c ** (2 / (g - 1)), 2 / (g + 1) * (-aR + (g - 1) / 2 * uR + s),
                              pR * c ** (2 * g / (g - 1)))
    return out
