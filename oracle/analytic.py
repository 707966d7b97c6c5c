"""Closed-form solutions the reference tests compare against (test
infrastructure only).

* ``sod_exact``  restates ``_sod_exact`` (test/test_sod3d.jl:13-75): the
  exact Riemann solution of the Sod tube by Newton iteration on p* (Toro 4.2).
* ``sedov_radius`` is the self-similar shock radius used at
  test/test_sedov.jl:29-30: R = (E / (alpha rho0))^(1/5) t^(2/5), alpha = 0.4942
  for gamma = 5/3 in 3-D.
"""
from __future__ import annotations

import math

import numpy as np


def sod_exact(x, t, gamma=1.4, left=(1.0, 0.0, 1.0), right=(0.125, 0.0, 0.1)):
    g = gamma
    rL, vL, pL = left
    rR, vR, pR = right
    csL = math.sqrt(g * pL / rL)
    csR = math.sqrt(g * pR / rR)

    def fK(p, rK, pK, csK):
        if p > pK:
            AK = 2 / ((g + 1) * rK)
            BK = (g - 1) / (g + 1) * pK
            sq = math.sqrt(AK / (p + BK))
            return (p - pK) * sq, sq * (1 - (p - pK) / (2 * (p + BK)))
        fv = 2 * csK / (g - 1) * ((p / pK) ** ((g - 1) / (2 * g)) - 1)
        dv = (p / pK) ** (-(g + 1) / (2 * g)) / (rK * csK)
        return fv, dv

    p = 0.5 * (pL + pR)
    for _ in range(100):
        fL, dfL = fK(p, rL, pL, csL)
        fR, dfR = fK(p, rR, pR, csR)
        dp = (fL + fR + vR - vL) / (dfL + dfR)
        p = max(1e-12, p - dp)
        if abs(dp) / p < 1e-12:
            break
    fL, _ = fK(p, rL, pL, csL)
    fR, _ = fK(p, rR, pR, csR)
    v_star = 0.5 * (vL + vR) + 0.5 * (fR - fL)
    r_starL = (rL * (p / pL + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pL + 1)
               if p > pL else rL * (p / pL) ** (1 / g))
    r_starR = (rR * (p / pR + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pR + 1)
               if p > pR else rR * (p / pR) ** (1 / g))
    if p > pL:
        SL_shock = vL - csL * math.sqrt((g + 1) / (2 * g) * p / pL + (g - 1) / (2 * g))
        SL_head = SL_tail = SL_shock
    else:
        SL_head = vL - csL
        SL_tail = v_star - math.sqrt(g * p / r_starL)
        SL_shock = SL_head
    SR = (vR + csR * math.sqrt((g + 1) / (2 * g) * p / pR + (g - 1) / (2 * g))
          if p > pR else vR + csR)

    x = np.asarray(x, dtype=np.float64)
    rho = np.empty_like(x)
    vel = np.empty_like(x)
    prs = np.empty_like(x)
    for k, xk in enumerate(x):
        xi = xk / t if t > 0 else math.copysign(math.inf, xk)
        if p > pL:
            if xi < SL_shock:
                rho[k], vel[k], prs[k] = rL, vL, pL
            elif xi < v_star:
                rho[k], vel[k], prs[k] = r_starL, v_star, p
            elif xi < SR:
                rho[k], vel[k], prs[k] = r_starR, v_star, p
            else:
                rho[k], vel[k], prs[k] = rR, vR, pR
        else:
            if xi < SL_head:
                rho[k], vel[k], prs[k] = rL, vL, pL
            elif xi < SL_tail:
                cs_fan = (2 * csL + (g - 1) * (vL - xi)) / (g + 1)
                vel[k] = xi + cs_fan
                rho[k] = rL * (cs_fan / csL) ** (2 / (g - 1))
                prs[k] = pL * (rho[k] / rL) ** g
            elif xi < v_star:
                rho[k], vel[k], prs[k] = r_starL, v_star, p
            elif xi < SR:
                rho[k], vel[k], prs[k] = r_starR, v_star, p
            else:
                rho[k], vel[k], prs[k] = rR, vR, pR
    return rho, vel, prs


SEDOV_ALPHA_53 = 0.4942


def sedov_radius(E, rho0, t, alpha=SEDOV_ALPHA_53):
    return (E / (alpha * rho0)) ** 0.2 * t ** 0.4
