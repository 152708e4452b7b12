"""GPU checks of Sod's shock tube (InitType shocktube): the In/Out far-field
states (FSIn/FSOut, edges.go:425-431) and the split initial state
(euler.go:738-771), through the C ABI, against the CPU oracle and against
the exact solution.

Bars: FP64 against the oracle within 1e-10 relative L2 in density after
the steps (1e-8 over the 500 steps to t = 0.1 with the shock sensor active,
whose log10/sin and vertex maxima amplify last-bit differences); FP32 within
1e-5 of the FP64 oracle; both within the oracle's L1 bound of the exact
solution at t = 0.1.
"""
import numpy as np
import pytest

import oracle
from gocfd_amd import Ensemble, Euler, InputParameters2D, meshgen
from oracle.sod import SodExact

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tube(tmp_path_factory):
    return meshgen.shocktube_mesh(str(tmp_path_factory.mktemp("st") / "st100.su2"), 100, 2, height=0.02)


def _ip(N, limiter, final=0.1):
    return InputParameters2D(CFL=0.5, FluxType="Roe", InitType="shocktube", PolynomialOrder=N, FinalTime=final,
                             Minf=0.0, MaxIterations=100000, Limiter=limiter, Kappa=3.0)


def _oracle(mesh, N, limiter, final=0.1):
    return oracle.OracleSolver(mesh, N, flux="roe", init="shocktube", CFL=0.5, FinalTime=final, Minf=0.0,
                               limiter=limiter, Kappa=3.0)


def _run_gpu(g, t_end):
    for _ in range(100):
        t, _, _ = g.time_info()
        if t >= t_end:
            break
        g.step(100)
    return g.time_info()


def _run_oracle(o, t_end):
    for _ in range(100000):
        if o.time >= t_end:
            break
        o.step()


def _l2(a, b):
    return np.linalg.norm(a[0] - b[0]) / np.linalg.norm(b[0])


@pytest.mark.parametrize("N,limiter,nsteps", [(0, "", 40), (1, "PerssonC0", 40), (2, "PerssonC0", 40),
                                              (3, "", 10)])
def test_shocktube_step_parity(tube, N, limiter, nsteps):
    g = Euler(_ip(N, limiter), tube)
    o = _oracle(tube, N, limiter)
    assert np.array_equal(g.Q, o.state())  # the split initial state, bit-exact
    g.step(nsteps)
    for _ in range(nsteps):
        o.step()
    t, dt, steps = g.time_info()
    assert steps == nsteps
    assert abs(t - o.time) <= 1e-10 * o.time and abs(dt - o.dt) <= 1e-10 * o.dt
    assert _l2(g.Q, o.state()) < 1e-10


@pytest.mark.parametrize("N,limiter,l1", [(0, "", 0.015), (1, "PerssonC0", 0.006)])
def test_shocktube_to_exact_solution_fp64(tube, N, limiter, l1):
    g = Euler(_ip(N, limiter), tube)
    t, _, steps = _run_gpu(g, 0.1)
    assert t == pytest.approx(0.1, rel=1e-14)
    o = _oracle(tube, N, limiter)
    _run_oracle(o, 0.1)
    assert o.step_count == steps
    q = g.Q
    assert _l2(q, o.state()) < 1e-8
    X = o.coords()["SolX"]
    ex = SodExact(t).conserved(X)
    assert np.abs(q[0] - ex[0]).mean() < l1


@pytest.mark.parametrize("N,limiter,l1", [(0, "", 0.015), (2, "PerssonC0", 0.015)])
def test_shocktube_fp32(tube, N, limiter, l1):
    g = Euler(_ip(N, limiter), tube, precision="fp32")
    t, _, _ = _run_gpu(g, 0.1)
    o = _oracle(tube, N, limiter)
    _run_oracle(o, 0.1)
    assert abs(t - o.time) < 1e-6
    assert _l2(g.Q, o.state()) < 1e-5
    ex = SodExact(t).conserved(o.coords()["SolX"])
    assert np.abs(g.Q[0] - ex[0]).mean() < l1


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_shocktube_in_an_ensemble(tube, tmp_path, prec):
    """A shock tube beside a vortex case: the per-case In/Out states travel
    with the case; each case is bit-identical to its single-case run."""
    vortex = meshgen.vortex_mesh(str(tmp_path / "v8.su2"), 8)
    cases = [(_ip(0, ""), tube),
             (InputParameters2D(CFL=1.0, FluxType="Roe", InitType="ivortex", PolynomialOrder=2,
                                FinalTime=1e9, MaxIterations=1000), vortex),
             (_ip(2, ""), tube)]
    ens = Ensemble(cases, precision=prec)
    ens.step(15)
    for c, (ip, path) in enumerate(cases):
        single = Euler(ip, path, precision=prec)
        single.step(15)
        assert np.array_equal(ens.state(c), single.Q), c
        assert ens.time_info(c) == single.time_info(), c
