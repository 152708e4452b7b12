"""GPU checks of the tangent-linear (jvp) and discrete-adjoint (vjp) step
through an active PerssonC0 shock sensor (global time stepping): the sensor
Se and its sigma ramp, eps = eps0 hK sigma, the vertex max-merges, the
vertex-mean sigma's modal filters of the RHS and of the stage-2 state, eps
at the RT points, the RT gradient times eps (AddDissipation), the viscous
edge flux with its penalty, and the viscous time step.

* The tangent-linear run's primal state equals the solver's own steps to
  1e-12 while the sensor flags elements.
* Dot-product test <s, J v> = <J^T s, v> to 1e-10 relative on the shock tube
  (N = 1 Lax with Kappa 5, N = 2 Roe with Kappa 10) and on a Mach 2 wedge
  (N = 2, 3, Kappa 10), each from a state with sigma > 0 on some elements.
* Central differences of the CPU oracle (the reference's algorithm
  restated) along a random state direction, on the wedge (a jittered
  Delaunay mesh, so the vertex maxima and the dt's arg-max have no ties):
  the map is smooth there but steep (Se = log10 of a ratio whose numerator
  is the small top-mode energy, so the sensor's higher derivatives are
  large), and central differences converge to the adjoint as h^2 (measured
  at N = 3: 7.6e-2, 1.4e-4, 1.4e-6 relative at h = 1e-5, 1e-6, 1e-7; at N = 2
  3.9e-3, 4.4e-6, 3.7e-8; the oracle's and the solver's differences agree to
  1e-10). The test takes h = 5e-7 and 2.5e-7, checks that the error shrinks,
  and compares the Richardson extrapolation (4 D(h/2) - D(h)) / 3 with the
  adjoint within 1e-6 relative.
"""
import numpy as np
import pytest

import oracle
from gocfd_amd import Euler, InputParameters2D, meshgen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def meshes(tmp_path_factory):
    d = tmp_path_factory.mktemp("adjshock")
    return dict(
        tube=meshgen.shocktube_mesh(str(d / "st.su2"), 30, 2, height=0.05),
        wedge=meshgen.wedge_mesh(str(d / "wedge.su2"), 600, seed=3),
    )


def _ip_tube(N, flux="Roe", kappa=5.0):
    return InputParameters2D(CFL=0.5, FluxType=flux, InitType="Shocktube", PolynomialOrder=N, Minf=0.0,
                             FinalTime=1e9, MaxIterations=100000, Limiter="PerssonC0", Kappa=kappa)


def _ip_wedge(N, flux="Roe", kappa=10.0):
    return InputParameters2D(CFL=1.0, FluxType=flux, InitType="Freestream", PolynomialOrder=N, Minf=2.0,
                             FinalTime=1e9, MaxIterations=100000, Limiter="PerssonC0", Kappa=kappa)


def _active(mesh, ip, steps):
    """A solver `steps` steps into the run, with the sensor flagging elements."""
    e = Euler(ip, mesh, precision="fp64")
    e.step(steps)
    _, sigma, _ = e.shock()
    n = int((sigma > 0).sum())
    print(f"{n} of {e.K} elements flagged after {steps} steps", flush=True)
    assert n > 0, "the sensor flags no element"
    return e


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("case", [("tube", 2, "Roe", 10.0), ("wedge", 2, "Roe", 10.0), ("wedge", 3, "Lax", 10.0)])
def test_tangent_linear_primal_matches_solver_with_active_sensor(meshes, case):
    name, N, flux, kappa = case
    ip = _ip_tube(N, flux, kappa) if name == "tube" else _ip_wedge(N, flux, kappa)
    steps = 20 if name == "tube" else 6
    e = _active(meshes[name], ip, steps)
    qn, _ = e.jvp(4, np.zeros_like(e.Q))
    ref = Euler(ip, meshes[name], precision="fp64")
    ref.step(steps + 4)
    assert _rel(qn, ref.Q) < 1e-12


@pytest.mark.parametrize("case", [("tube", 1, "Lax", 5.0), ("tube", 2, "Roe", 10.0), ("wedge", 2, "Roe", 10.0),
                                  ("wedge", 3, "Roe", 10.0)])
def test_dot_product_with_active_sensor(meshes, case):
    name, N, flux, kappa = case
    ip = _ip_tube(N, flux, kappa) if name == "tube" else _ip_wedge(N, flux, kappa)
    e = _active(meshes[name], ip, 20 if name == "tube" else 6)
    q = e.Q
    rng = np.random.default_rng(31 + N + len(name))
    v = rng.standard_normal(q.shape) * np.abs(q)
    w = rng.standard_normal(4)
    s = rng.standard_normal(q.shape)
    _, jv = e.jvp(3, v, w)
    d0, dfar = e.vjp(3, s)
    lhs = float(np.sum(s * jv))
    rhs = float(np.sum(d0 * v) + dfar @ w)
    print(f"<s, J v> = {lhs!r}, <J^T s, v> = {rhs!r}", flush=True)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs), (lhs, rhs)
    assert np.array_equal(e.Q, q)


@pytest.mark.parametrize("N", [2, 3])
def test_active_sensor_against_oracle_differences(meshes, N):
    """d(s . q_n)/dq_0 along v: the adjoint against central differences of
    the oracle's steps from a state with flagged elements"""
    mesh, n, h = meshes["wedge"], 3, 5e-7
    ip = _ip_wedge(N)
    e = _active(mesh, ip, 6)
    q0 = e.Q
    rng = np.random.default_rng(77 + N)
    v = rng.standard_normal(q0.shape) * np.abs(q0)
    s = rng.standard_normal(q0.shape)
    d0, _ = e.vjp(n, s)
    adj = float(np.sum(d0 * v))

    def J(q):
        o = oracle.OracleSolver(mesh, N, flux="roe", init="freestream", CFL=1.0, Minf=2.0, limiter="PerssonC0",
                                Kappa=10.0)
        o.set_state(q)
        for _ in range(n):
            o.step()
        return float(np.sum(s * o.state()))

    fd1 = (J(q0 + h * v) - J(q0 - h * v)) / (2 * h)
    fd2 = (J(q0 + 0.5 * h * v) - J(q0 - 0.5 * h * v)) / h
    rich = (4 * fd2 - fd1) / 3
    print(f"adjoint {adj!r}; oracle central differences {fd1!r} (h), {fd2!r} (h/2), extrapolated {rich!r}, "
          f"rel {abs(adj - rich) / abs(rich):.3e}", flush=True)
    assert abs(fd2 - adj) < abs(fd1 - adj)
    assert adj == pytest.approx(rich, rel=1e-6)
