"""CPU checks of Sod's shock tube (InitType shocktube, euler.go:738-771).

The exact solution (oracle/sod.py) is pinned against the reference's own test
vectors (analytic_sod_test.go, kept in tests/golden/sod_exact.json) and
against the jump conditions that fix it independently of the formulas it
restates. The oracle's solver on the channel mesh is then held to the exact
solution: the In/Out boundary states (edges.go:425-431) and the split
initial state are what a plausible mistake would break.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from gocfd_amd import meshgen
from oracle.sod import GAMMA, LEFT, RIGHT, SodExact

G = os.path.join(os.path.dirname(__file__), "golden")


def test_exact_sod_matches_reference_vectors():
    g = json.load(open(os.path.join(G, "sod_exact.json")))
    s = SodExact(g["t"])
    x = s.sample_points()
    rho, _, _ = s.at(x)
    assert np.abs(x - np.array(g["x"])).max() < g["tol"]
    assert np.abs(rho - np.array(g["rho"])).max() < g["tol"]
    for t, x4 in g["x4"].items():
        assert abs(SodExact(float(t)).x4 - x4) < g["x4_tol"]


def test_exact_sod_satisfies_jump_conditions():
    """Rankine-Hugoniot across the shock (mass, momentum, energy in the
    shock frame), isentropy and the Riemann invariant u + 2c/(g-1) through
    the rarefaction, and pressure/velocity continuity at the contact."""
    s = SodExact(0.1)
    g = GAMMA
    rr, pr, _ = RIGHT
    rl, pl, _ = LEFT
    P, u, V = s.p_post, s.u_post, s.v_shock
    r2 = s.rho_post
    # shock frame: w = u - V
    m1, m2 = rr * (0 - V), r2 * (u - V)
    assert abs(m1 - m2) < 1e-12
    assert abs((pr + rr * V * V) - (P + r2 * (u - V) ** 2)) < 1e-12
    h1 = g / (g - 1) * pr / rr + 0.5 * V * V
    h2 = g / (g - 1) * P / r2 + 0.5 * (u - V) ** 2
    assert abs(h1 - h2) < 1e-12
    # rarefaction: isentropic, J+ = u + 2c/(g-1) constant
    rm = s.rho_middle
    assert abs(P / rm ** g - pl / rl ** g) < 1e-12
    cm = math.sqrt(g * P / rm)
    assert abs((u + 2 * cm / (g - 1)) - (2 * s.c_left / (g - 1))) < 1e-12
    x = np.linspace(s.x1, s.x2, 11)
    rho, p, uu = s.at(x)
    c = np.sqrt(g * p / rho)
    assert np.abs(uu + 2 * c / (g - 1) - 2 * s.c_left / (g - 1)).max() < 1e-12
    # fan head/tail are characteristics: x/t = u - c there
    assert abs((s.x2 - s.x0) / s.t - (u - cm)) < 1e-12
    # textbook values (Toro, Riemann Solvers, Table 4.2 test 1)
    assert abs(P - 0.30313) < 1e-5 and abs(u - 0.92745) < 1e-5
    assert abs(rm - 0.42632) < 1e-5 and abs(r2 - 0.26557) < 1e-5


@pytest.fixture(scope="module")
def tube(tmp_path_factory):
    return meshgen.shocktube_mesh(str(tmp_path_factory.mktemp("st") / "st100.su2"), 100, 2, height=0.02)


def _run(mesh, N, limiter, t_end=0.1, **kw):
    o = oracle.OracleSolver(mesh, N, flux="roe", init="shocktube", CFL=kw.get("CFL", 0.5), FinalTime=t_end,
                            Minf=0.0, limiter=limiter, Kappa=kw.get("Kappa", 3.0))
    for _ in range(10000):
        if o.time >= t_end:
            break
        o.step()
    return o


def test_shocktube_initial_state(tube):
    o = oracle.OracleSolver(tube, 1, flux="roe", init="shocktube")
    X = o.coords()["SolX"]
    q = o.state()
    left = X < 0.5
    assert np.array_equal(q[0], np.where(left, 1.0, 0.125))
    assert np.array_equal(q[3], np.where(left, 1 / (1.4 - 1), 0.1 / (1.4 - 1)))
    assert not q[1].any() and not q[2].any()
    # the In/Out states are the initial states: nothing moves near x=0, x=1
    # (round-off of a divergence over elements of size 0.01 is ~1e-10)
    r0 = o.rhs(q, 0.0)
    far = (X < 0.4) | (X > 0.6)
    assert np.abs(r0[:, far]).max() < 1e-9
    assert np.abs(r0[:, ~far]).max() > 1.0


@pytest.mark.parametrize("N,limiter,l1", [(0, "", 0.015), (1, "PerssonC0", 0.006)])
def test_oracle_shocktube_converges_to_exact(tube, N, limiter, l1):
    """Density at the solution points at t=0.1 within an L1 bound of the
    exact solution (100 x 2 quads, 400 triangles); the undisturbed states at
    both ends stay exact to round-off (In/Out swapped or a wrong state there
    would start a wave at the ends)."""
    o = _run(tube, N, limiter)
    assert abs(o.time - 0.1) < 1e-14
    q = o.state()
    X = o.coords()["SolX"]
    ex = SodExact(o.time).conserved(X)
    err = np.abs(q[0] - ex[0]).mean()
    assert err < l1, err
    quiet = (X < 0.15) | (X > 0.85)
    assert np.abs(q[:, quiet] - ex[:, quiet]).max() < 1e-10
