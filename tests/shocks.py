"""Exact normal shocks (Rankine-Hugoniot) for pinning Riemann fluxes.

A Roe linearisation maps the jump across a single shock onto one of its
eigenvectors with the shock speed as eigenvalue, so Roe's flux at a lone
shock is the upwind side's physical flux (Roe 1981, "approximate Riemann
solvers, parameter vectors and difference schemes": exact resolution of
isolated discontinuities). The states here follow from the textbook shock
relations alone, not from any flux code.
"""
import math

import numpy as np


def shock_pair(g, M, rho1, p1, direction, n, w=0.0, vt=0.0):
    """A normal shock of Mach M running into gas (rho1, p1), which moves
    with normal velocity w; both sides carry tangential velocity vt.
    direction +1: the shock runs along the normal n (the left side is the
    shocked gas); -1: against it (the right side is). Returns left and right
    as (rho, (u_x, u_y), p) and the shock's speed along n."""
    n = np.asarray(n, dtype=float)
    t = np.array([-n[1], n[0]])
    c1 = math.sqrt(g * p1 / rho1)
    p2 = p1 * (1.0 + 2.0 * g / (g + 1.0) * (M * M - 1.0))
    rho2 = rho1 * (g + 1.0) * M * M / ((g - 1.0) * M * M + 2.0)
    s = M * c1
    u2 = s * (1.0 - rho1 / rho2)  # the shocked gas follows the shock
    pre = (rho1, tuple(w * n + vt * t), p1)
    if direction > 0:
        post = (rho2, tuple((w + u2) * n + vt * t), p2)
        return post, pre, w + s
    post = (rho2, tuple((w - u2) * n + vt * t), p2)
    return pre, post, w - s


def physical_flux(n, rho, u, p, g):
    un = u[0] * n[0] + u[1] * n[1]
    E = p / (g - 1.0) / rho + 0.5 * (u[0] ** 2 + u[1] ** 2)
    return np.array([rho * un, rho * u[0] * un + p * n[0], rho * u[1] * un + p * n[1], (rho * E + p) * un])


def enthalpy(g, rho, p):
    return p / ((g - 1.0) * rho) + p / rho


def cases(g=1.4):
    """(normal, left, right, expected flux): shocks of Mach 1.5, 2 and 4 in
    both directions, at rest and moving, on three normals, with a
    tangential velocity; the expected flux is the upwind side's."""
    out = []
    for th in (0.0, 0.5 * math.pi, 0.7):
        n = np.array([math.cos(th), math.sin(th)])
        for M in (1.5, 2.0, 4.0):
            for d in (1, -1):
                for w in (0.0, 0.3 * d):
                    L, R, s = shock_pair(g, M, 1.0, 1.0, d, n, w=w, vt=0.2)
                    up = L if s > 0 else R
                    out.append((n, L, R, physical_flux(n, up[0], up[1], up[2], g)))
    return out
