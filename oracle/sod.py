"""Exact solution of Sod's shock tube (TEST INFRASTRUCTURE ONLY).

Restates model_problems/Euler1D/sod_shock_tube/analytic_sod.go: left state
(rho, p, u) = (1, 1, 0), right state (0.125, 0.1, 0), gamma 1.4, diaphragm at
x0 = 0.5 on [0, 1]. The post-shock pressure is the root of the reference's
pressure function `sod_func` (analytic_sod.go:136-156); the reference finds
it with a secant iteration to |f| < 1e-7 (`fzero`, analytic_sod.go:119-134),
this module brackets it and bisects to round-off, which moves nothing at the
1e-3 bar of the reference's own test. Wave positions follow
`calc` (analytic_sod.go:36-63), point values `Getx` (analytic_sod.go:65-101)
and the sample points of `Get` (analytic_sod.go:103-117).
"""
from __future__ import annotations

import math

import numpy as np

GAMMA = 1.4
LEFT = (1.0, 1.0, 0.0)     # rho, p, u
RIGHT = (0.125, 0.1, 0.0)


def _sod_func(P: float) -> float:
    """analytic_sod.go:136-156 (the corrected form the reference uses)."""
    g = GAMMA
    rl, pl = LEFT[0], LEFT[1]
    rr, pr = RIGHT[0], RIGHT[1]
    mu2 = (g - 1) / (g + 1)
    return ((P - pr) * math.sqrt((1 - mu2) / (rr * (P + mu2 * pr)))
            - (pl ** ((g - 1) / (2 * g)) - P ** ((g - 1) / (2 * g)))
            * math.sqrt(((1 - mu2 * mu2) * pl ** (1 / g)) / (mu2 * mu2 * rl)))


def post_shock_pressure() -> float:
    """Root of _sod_func in (p_right, p_left), by bisection to round-off."""
    lo, hi = RIGHT[1], LEFT[1]
    flo = _sod_func(lo)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        fm = _sod_func(mid)
        if fm == 0 or hi - lo < 1e-16:
            return mid
        if (fm < 0) == (flo < 0):
            lo, flo = mid, fm
        else:
            hi = mid
    return 0.5 * (lo + hi)


class SodExact:
    """Wave positions x1 (rarefaction head), x2 (tail), x3 (contact),
    x4 (shock) at time t (SOD_Exact.calc, analytic_sod.go:36-63)."""

    def __init__(self, t: float):
        g = GAMMA
        mu2 = (g - 1) / (g + 1)
        rl, pl, ul = LEFT
        rr, pr, ur = RIGHT
        self.t = t
        self.x0 = 0.5
        P = post_shock_pressure()
        self.p_post = P
        self.rho_middle = rl * (P / pl) ** (1 / g)
        self.rho_post = rr * ((P / pr) + mu2) / (1 + mu2 * (P / pr))
        self.u_post = ur + (P - pr) / math.sqrt(0.5 * rr * ((g + 1) * P + (g - 1) * pr))
        self.c_left = math.sqrt(g * pl / rl)
        ratio = self.rho_post / rr
        self.v_shock = self.u_post * ratio / (ratio - 1)
        self.x1 = self.x0 - self.c_left * t
        self.x3 = self.x0 + self.u_post * t
        self.x4 = self.x0 + self.v_shock * t
        c2 = self.c_left - 0.5 * (g - 1) * self.u_post
        self.x2 = self.x0 + t * (self.u_post - c2)

    def at(self, x) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(rho, p, u) at the points x (Getx, analytic_sod.go:65-101)."""
        g = GAMMA
        mu2 = (g - 1) / (g + 1)
        x = np.asarray(x, dtype=np.float64)
        rl, pl, ul = LEFT
        rr, pr, ur = RIGHT
        rho = np.empty_like(x)
        p = np.empty_like(x)
        u = np.empty_like(x)
        fan = (x >= self.x1) & (x <= self.x2)
        mid = (x > self.x2) & (x <= self.x3)
        post = (x > self.x3) & (x <= self.x4)
        left, right = x < self.x1, x > self.x4
        rho[left], p[left], u[left] = rl, pl, ul
        rho[right], p[right], u[right] = rr, pr, ur
        if self.t > 0 and fan.any():
            c = mu2 * ((self.x0 - x[fan]) / self.t) + (1 - mu2) * self.c_left
            rho[fan] = rl * (c / self.c_left) ** (2 / (g - 1))
            p[fan] = pl * (rho[fan] / rl) ** g
            u[fan] = (1 - mu2) * (-(self.x0 - x[fan]) / self.t + self.c_left)
        rho[mid], p[mid], u[mid] = self.rho_middle, self.p_post, self.u_post
        rho[post], p[post], u[post] = self.rho_post, self.p_post, self.u_post
        return rho, p, u

    def conserved(self, x) -> np.ndarray:
        """[rho, rho u, 0, E] at the points x."""
        rho, p, u = self.at(x)
        return np.stack([rho, rho * u, 0 * rho, p / (GAMMA - 1) + 0.5 * rho * u * u])

    def sample_points(self) -> np.ndarray:
        """The 20 plot points of SOD_Exact.Get (analytic_sod.go:103-117)."""
        tol = 1e-4
        mid = (self.x2 - self.x1) / 10
        pts = [0.0, self.x1 - tol, self.x1 + tol]
        pts += [self.x1 + i * mid for i in range(1, 10)]
        pts += [self.x1 + 10 * mid - 2 * tol, self.x2 - tol, self.x2 + tol,
                self.x3 - tol, self.x3 + tol, self.x4 - tol, self.x4 + tol, 1.0]
        return np.array(pts)
