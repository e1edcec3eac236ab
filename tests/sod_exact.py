"""Exact solution of the Sod shock tube (left rarefaction, contact, right
shock) at time t, for the multi-step tests of the kick / drift / timestep
loop.  The star state (p*, u*) comes from the oracle's exact Riemann solver,
which tests/test_oracle_elements.py holds to Toro's Table 4.3; the wave
structure around it is the self-similar solution of Toro, "Riemann Solvers
and Numerical Methods for Fluid Dynamics", section 4.4-4.5 (eqs. 4.53-4.59:
shock density and speed from the Rankine-Hugoniot conditions, isentropic
rarefaction fan)."""
import numpy as np

import oracle as O


def star_state(rho_l, p_l, rho_r, p_r, gamma):
    out = O.riemann((rho_l, p_l, 0, 0, 0), (rho_r, p_r, 0, 0, 0), n=(1, 0, 0), press_tot_limiter=0.0, gamma=gamma)
    assert out["mask"] & 4
    return out["P_M"], out["S_M"]


def sod_profile(xi, rho_l, p_l, rho_r, p_r, gamma):
    """rho, u, p at self-similar coordinates xi = (x - x0)/t (left state at
    xi -> -inf, both states at rest)."""
    ps, us = star_state(rho_l, p_l, rho_r, p_r, gamma)
    assert p_r < ps < p_l  # left rarefaction, right shock
    g = gamma
    cl = np.sqrt(g * p_l / rho_l)
    cr = np.sqrt(g * p_r / rho_r)
    # right shock (Toro 4.57, 4.59)
    pr_ratio = ps / p_r
    rho_sr = rho_r * (pr_ratio + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * pr_ratio + 1)
    s_shock = cr * np.sqrt((g + 1) / (2 * g) * pr_ratio + (g - 1) / (2 * g))
    # left rarefaction (Toro 4.53-4.56)
    rho_sl = rho_l * (ps / p_l) ** (1 / g)
    c_sl = cl * (ps / p_l) ** ((g - 1) / (2 * g))
    head, tail = -cl, us - c_sl
    xi = np.asarray(xi, dtype=np.float64)
    rho = np.empty_like(xi)
    u = np.empty_like(xi)
    p = np.empty_like(xi)
    a = xi < head
    rho[a], u[a], p[a] = rho_l, 0.0, p_l
    f = (xi >= head) & (xi < tail)
    fan = 2 / (g + 1) + (g - 1) / ((g + 1) * cl) * (0.0 - xi[f])
    rho[f] = rho_l * fan ** (2 / (g - 1))
    u[f] = 2 / (g + 1) * (cl + (g - 1) / 2 * 0.0 + xi[f])
    p[f] = p_l * fan ** (2 * g / (g - 1))
    b = (xi >= tail) & (xi < us)
    rho[b], u[b], p[b] = rho_sl, us, ps
    c = (xi >= us) & (xi < s_shock)
    rho[c], u[c], p[c] = rho_sr, us, ps
    d = xi >= s_shock
    rho[d], u[d], p[d] = rho_r, 0.0, p_r
    return rho, u, p
