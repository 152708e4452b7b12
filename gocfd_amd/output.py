"""Output files of a finished run, written from the device state by the host
(Euler.OutputFinal, model_problems/Euler2D/euler.go:222-318):

* ``plotfile.dat`` (freestream cases with PlotFields): one line per point of
  every wall edge, on the owner's face, with the RT edge point coordinates
  and each plot field interpolated there (euler.go:251-315);
* ``shocktube.dat`` (InitType shocktube): density, x momentum and energy
  sampled along the channel centre line (euler.go:222-238,
  sod_shock_tube/shock_tube.go);
* ``shocktube_analytic.dat``: the exact Sod solution at FinalTime at the
  reference's sample points (euler.go:239-249,
  Euler1D/sod_shock_tube/analytic_sod.go).

Where the reference's result depends on Go map iteration order (the wall
edges of plotfile.dat come from ranging over the edge map; a PlotFields name
that ties two flow functions), the order here is fixed: edges by key
(min vertex + max vertex << 32), ties by the flow-function list order of
fluids.go:112-118.
"""
from __future__ import annotations

import math
from typing import Dict, List, Sequence

import numpy as np

# FlowFunction (fluids.go:202-228) and their display names (fluids.go:162-200)
FLOW_FUNCTIONS = [
    (0, "Density"), (1, "XMomentum"), (2, "YMomentum"), (3, "Energy"), (4, "Mach"),
    (5, "Static Pressure"), (6, "Dynamic Pressure"), (7, "Pressure Coefficient"), (8, "Sound Speed"),
    (9, "Velocity"), (10, "XVelocity"), (11, "YVelocity"), (12, "Enthalpy"), (13, "Entropy"),
    (100, "ShockFunction"), (101, "Artificial Dissipation Epsilon"), (102, "Artificial Dissipation Epsilon C0"),
    (200, "R Direction Gradient of Density"), (201, "R Direction Gradient of R Momentum"),
    (202, "R Direction Gradient of S Momentum"), (203, "R Direction Gradient of Energy"),
    (300, "S Direction Gradient of Density"), (301, "S Direction Gradient of R Momentum"),
    (302, "S Direction Gradient of S Momentum"), (303, "S Direction Gradient of Energy"),
]
BC_WALL = 5  # types.BC_Wall (types/cfd.go:13-26)


def _tokenize(s: str) -> List[str]:
    """fluids.go:128-134"""
    return s.lower().replace("-", " ").replace("_", " ").split()


def best_match_flow_function(name: str):
    """BestMatchFlowFunction (fluids.go:137-160): the flow function whose name
    tokens contain the most input tokens; (code, matched)."""
    inp = _tokenize(name)
    best, score_best = -1, -1
    for code, disp in FLOW_FUNCTIONS:
        toks = _tokenize(disp)
        score = sum(1 for t in inp if any(t in tok for tok in toks))
        if score > score_best:
            best, score_best = code, score
    return best, score_best > 0


def flow_function(Q: np.ndarray, code: int, gamma: float, pinf: float, qqinf: float) -> np.ndarray:
    """GetFlowFunctionBase (fluids.go:289-336) at every point of Q [4][...]."""
    rho, rhoU, rhoV, E = Q[0], Q[1], Q[2], Q[3]
    if code <= 3:
        return Q[code].copy()
    oorho = 1.0 / rho
    if code == 10:
        return rhoU * oorho
    if code == 11:
        return rhoV * oorho
    u, v = rhoU * oorho, rhoV * oorho
    U2 = u * u + v * v
    q = 0.5 * rho * U2
    p = (gamma - 1.0) * (E - q)
    if code == 9:
        return np.sqrt(U2)
    if code == 6:
        return q
    if code == 5:
        return p
    if code == 7:
        return (p - pinf) / qqinf
    if code == 8:
        return np.sqrt(np.abs(gamma * p * oorho))
    if code == 12:
        return (E + p) / rho
    if code == 13:
        return np.log(p) - gamma * np.log(rho)
    if code == 4:
        return np.sqrt(U2) / np.sqrt(np.abs(gamma * p * oorho))
    raise ValueError(f"flow function {code} is not a point function of the state; the host writer "
                     "supports Density..Entropy (fluids.go:202-215)")


def far_field(ip) -> tuple:
    """(gamma, Pinf, QQinf) of FSFar: NewFreeStream (fluids.go:245-262) or,
    for IVortex, Qinf = {1, 1, 0, 3} at gamma 1.4 (euler.go:773)."""
    if ip.InitType.lower() == "ivortex":
        g, q = 1.4, (1.0, 1.0, 0.0, 3.0)
    else:
        g, M, a = ip.Gamma, ip.Minf, math.radians(ip.Alpha)
        q = (1.0, M * math.cos(a), M * math.sin(a), 1.0 / (g * (g - 1.0)) + 0.5 * M * M)
    u, v = q[1] / q[0], q[2] / q[0]
    qq = 0.5 * q[0] * (u * u + v * v)
    return g, (g - 1.0) * (q[3] - qq), qq


def plotfile_lines(solver, Q: np.ndarray, field_names: Sequence[str]) -> List[str]:
    """The lines of plotfile.dat (euler.go:251-315) for state Q [4][NpInt][K]:
    per wall edge, on its owner's face e, for each edge point i (only i = 0
    at N = 0): "x, y," at RT point 2 NpInt + e NpEdge + i, then " f," per
    field, the field at the solution points interpolated to that point
    (GetPlotField plot.go:14-85: GraphInterp rows of the edge points are the
    FluxEdgeInterp rows; the vertex averaging touches only the vertices)."""
    codes = []
    for name in field_names:
        code, ok = best_match_flow_function(name)
        if not ok:
            raise ValueError("Unable to find matching flow function named: " + name)
        codes.append(code)
    g, pinf, qqinf = far_field(solver.ip)
    ops, xy = solver.operators(), solver.coords()
    interp = ops["Interp"]
    NpInt, NpEdge, N = solver.NpInt, solver.NpEdge, solver.N
    fields = [flow_function(Q, c, g, pinf, qqinf) for c in codes]  # [NpInt][K] each
    lines = []
    for E in solver.edges():  # sorted by edge key
        if E[9] != BC_WALL:
            continue
        k, e = int(E[3]), int(E[5])
        for i in range(NpEdge):
            if N == 0 and i > 0:
                continue
            row = 2 * NpInt + e * NpEdge + i
            s = "%.5e, %.5e," % (xy["FluxX"][row, k], xy["FluxY"][row, k])
            for f in fields:
                s += " %.5e," % float(interp[e * NpEdge + i] @ f[:, k])
            lines.append(s)
    return lines


def write_plotfile(solver, path: str = "plotfile.dat", Q: np.ndarray | None = None) -> int:
    """Write plotfile.dat; returns the number of wall edges."""
    names = list(getattr(solver.ip, "PlotFields", []) or [])
    if not names:
        return 0
    lines = plotfile_lines(solver, solver.Q if Q is None else Q, names)
    with open(path, "w") as f:
        for ln in lines:
            f.write(ln + "\n")
    return sum(1 for E in solver.edges() if E[9] == BC_WALL)


# ---- shock tube ----------------------------------------------------------

def shocktube_x_locations(npts: int) -> np.ndarray:
    """NewSODShockTube (shock_tube.go:27-46): npts equal steps over [0, 1],
    the end points pulled in by 1e-5."""
    x = np.arange(npts, dtype=np.float64) * (1.0 / float(npts - 1))
    x[0] += 0.00001
    x[-1] -= 0.00001
    return x


def _uv_coords(VX, VY, EToV, x, y):
    """getUVCoords (shock_tube.go:81-125): the first element (k ascending)
    whose coordinates r along C-A and s along B-A satisfy r, s >= 0,
    r + s <= 1, with those (r, s)."""
    A = np.stack([VX[EToV[:, 0]], VY[EToV[:, 0]]], 1)
    B = np.stack([VX[EToV[:, 1]], VY[EToV[:, 1]]], 1)
    Cc = np.stack([VX[EToV[:, 2]], VY[EToV[:, 2]]], 1)
    v0, v1 = Cc - A, B - A
    v2 = np.array([x, y])[None, :] - A
    d00 = (v0 * v0).sum(1)
    d01 = (v0 * v1).sum(1)
    d02 = (v0 * v2).sum(1)
    d11 = (v1 * v1).sum(1)
    d12 = (v1 * v2).sum(1)
    inv = 1.0 / (d00 * d11 - d01 * d01)
    r = (d11 * d02 - d01 * d12) * inv
    s = (d00 * d12 - d01 * d02) * inv
    hit = np.nonzero((r >= 0) & (s >= 0) & ((r + s) <= 1.0))[0]
    if hit.size == 0:
        raise RuntimeError(f"unable to find point within elements: [{x:5.3f},{y:5.3f}]")
    k = int(hit[0])
    return k, float(r[k]), float(s[k])


def shocktube_samples(solver, Q: np.ndarray):
    """SODShockTube sampling (shock_tube.go:27-142) of Q [4][NpInt][K]:
    X, Rho, RhoU, E at 4K/5 centre-line points (euler.go:771). The (r, s) of
    getUVCoords lie in [0, 1] and go to the solution basis as they are
    (calculateInterpolation :60-79), as in the reference."""
    VX, VY, EToV = solver.mesh_arrays()
    npts = 4 * solver.K // 5
    X = shocktube_x_locations(npts)
    ymid = 0.5 * (VY.max() - VY.min()) + VY.min()
    ks, rs, ss = [], [], []
    for x in X:
        k, r, s = _uv_coords(VX, VY, EToV, x, ymid)
        ks.append(k)
        rs.append(r)
        ss.append(s)
    M = solver.interp_matrix(np.array(rs), np.array(ss))  # [npts][NpInt]
    ks = np.array(ks)
    rho = np.einsum("pi,ip->p", M, Q[0][:, ks])
    rhou = np.einsum("pi,ip->p", M, Q[1][:, ks])
    ener = np.einsum("pi,ip->p", M, Q[3][:, ks])
    return X, rho, rhou, ener


def write_shocktube(solver, path: str = "shocktube.dat", Q: np.ndarray | None = None) -> None:
    """shocktube.dat (euler.go:228-238)"""
    X, rho, rhou, ener = shocktube_samples(solver, solver.Q if Q is None else Q)
    with open(path, "w") as f:
        f.write(f"Meshfile: {solver.MeshFile}\n")
        f.write("X\tRho\tRhoU\tE\n")
        for i in range(len(X)):
            f.write("%.8f\t%.8f\t%.8f\t%.8f\n" % (X[i], rho[i], rhou[i], ener[i]))


class SodExact:
    """SOD_Exact (Euler1D/sod_shock_tube/analytic_sod.go): left (rho, p, u) =
    (1, 1, 0), right (0.125, 0.1, 0), gamma 1.4, diaphragm at 0.5 on [0, 1].
    The post-shock pressure is the root of sod_func found by the reference's
    secant iteration fzero from pi to |f| < 1e-7 (analytic_sod.go:119-156),
    so the file matches the reference's to the digits it prints."""

    def __init__(self, t: float):
        g = self.gamma = 1.4
        self.l_s = (1.0, 1.0, 0.0)   # rho, p, u
        self.r_s = (0.125, 0.1, 0.0)
        self.t = t
        mu2 = (g - 1.0) / (g + 1.0)
        P = self._fzero(math.pi)
        rl, pl, ul = self.l_s
        rr, pr, ur = self.r_s
        cl = math.sqrt(g * pl / rl)
        self.rho_middle = rl * (P / pl) ** (1.0 / g)
        self.x0 = 0.5
        post_rho = rr * ((P / pr) + mu2) / (1 + mu2 * (P / pr))
        post_u = ur + (P - pr) / math.sqrt(0.5 * rr * ((g + 1) * P + (g - 1) * pr))
        self.post = (post_rho, P, post_u)
        v_shock = post_u * (post_rho / rr) / ((post_rho / rr) - 1.0)
        self.x1 = self.x0 - cl * t
        self.x3 = self.x0 + post_u * t
        self.x4 = self.x0 + v_shock * t
        c2 = cl - 0.5 * (g - 1.0) * post_u
        self.x2 = self.x0 + t * (post_u - c2)

    def _sod_func(self, P):
        """analytic_sod.go:136-156"""
        g = self.gamma
        rl, pl, _ = self.l_s
        rr, pr, _ = self.r_s
        mu2 = (g - 1) / (g + 1)
        return ((P - pr) * math.sqrt((1 - mu2) / (rr * (P + mu2 * pr)))
                - (pl ** ((g - 1) / (2 * g)) - P ** ((g - 1) / (2 * g)))
                * math.sqrt(((1 - mu2 * mu2) * pl ** (1 / g)) / (mu2 * mu2 * rl)))

    def _fzero(self, start):
        """fzero (analytic_sod.go:119-134)"""
        tol = 0.0000001
        start_old = start / 2
        res = self._sod_func(start_old)
        while abs(res) > tol:
            res_new = self._sod_func(start)
            deriv = (start - start_old) / (res_new - res)
            start_new = abs(start - res_new * deriv)
            start_old = start
            start = start_new
            res = res_new
        return start

    def getx(self, x):
        """Getx (analytic_sod.go:65-101): rho, p, u, e, rhou"""
        g = self.gamma
        mu2 = (g - 1) / (g + 1)
        rl, pl, ul = self.l_s
        rr, pr, ur = self.r_s
        cl = math.sqrt(g * pl / rl)
        if x < self.x1:
            rho, p, u = rl, pl, ul
        elif self.x1 <= x <= self.x2:
            c = mu2 * ((self.x0 - x) / self.t) + (1.0 - mu2) * cl
            rho = rl * (c / cl) ** (2 / (g - 1))
            p = pl * (rho / rl) ** g
            u = (1.0 - mu2) * ((-(self.x0 - x) / self.t) + cl)
        elif self.x2 <= x <= self.x3:
            rho, p, u = self.rho_middle, self.post[1], self.post[2]
        elif self.x3 <= x <= self.x4:
            rho, p, u = self.post
        else:
            rho, p, u = rr, pr, ur
        e = p / (g - 1.0) + 0.5 * u * u * rho
        return rho, p, u, e, rho * u

    def get(self):
        """Get (analytic_sod.go:103-117): X, Rho, P, RhoU, E at the reference's
        20 sample points around the waves"""
        tol = 0.0001
        mid = (self.x2 - self.x1) / 10.0
        X = [0.0, self.x1 - tol, self.x1 + tol] + [self.x1 + j * mid for j in range(1, 10)] + [
            self.x1 + 10 * mid - 2.0 * tol, self.x2 - tol, self.x2 + tol, self.x3 - tol, self.x3 + tol,
            self.x4 - tol, self.x4 + tol, 1.0]
        vals = [self.getx(x) for x in X]
        return (np.array(X), np.array([v[0] for v in vals]), np.array([v[1] for v in vals]),
                np.array([v[4] for v in vals]), np.array([v[3] for v in vals]))


def write_shocktube_analytic(t: float, path: str = "shocktube_analytic.dat") -> None:
    """shocktube_analytic.dat (euler.go:239-249)"""
    X, rho, _, rhou, ener = SodExact(t).get()
    with open(path, "w") as f:
        f.write("X\tRho\tRhoU\tE\n")
        for i in range(len(X)):
            f.write("%.8f\t%.8f\t%.8f\t%.8f\n" % (X[i], rho[i], rhou[i], ener[i]))


def output_final(solver, directory: str = ".") -> List[str]:
    """OutputFinal (euler.go:222-318): the files of the case; returns their paths."""
    import os
    init = solver.ip.InitType.lower()
    out = []
    if init == "shocktube":
        p1, p2 = os.path.join(directory, "shocktube.dat"), os.path.join(directory, "shocktube_analytic.dat")
        write_shocktube(solver, p1)
        write_shocktube_analytic(solver.FinalTime, p2)
        out = [p1, p2]
    elif init == "freestream" and getattr(solver.ip, "PlotFields", None):
        p = os.path.join(directory, "plotfile.dat")
        write_plotfile(solver, p)
        out = [p]
    return out
