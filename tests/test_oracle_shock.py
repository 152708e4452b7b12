"""Pins the oracle's shock-capturing path (PerssonC0: sensor, vertex merge,
RT gradient, viscous fluxes, modal filter) and local time stepping against
the reference's fixtures and closed forms. CPU only.
"""
import math
import os

import numpy as np
import pytest

import oracle
from gocfd_amd import meshgen

G = os.path.join(os.path.dirname(__file__), "golden")


def _solver(path, N, **kw):
    kw.setdefault("flux", "roe")
    kw.setdefault("init", "freestream")
    kw.setdefault("Minf", 2.0)
    kw.setdefault("limiter", "PerssonC0")
    return oracle.OracleSolver(path, N, **kw)


def test_eps_vertex_fixture():
    """TestDissipation / TestDissipation2 (model_problems/Euler2D/euler_test.go:
    480-515,528-545): element value k merged to the vertices by max gives
    [1,3,4,5,7,9,9,6,8,9] on test_tris_9."""
    s = _solver(os.path.join(G, "test_tris_9.neu"), 1)
    vert = np.zeros(10)
    oracle.lib().or_vertex_max(s._h, oracle._p(np.arange(s.K, dtype=np.float64)), oracle._p(vert))
    assert vert.tolist() == [1, 3, 4, 5, 7, 9, 9, 6, 8, 9]


@pytest.mark.parametrize("mesh", ["test_tris_5.neu", "test_tris_9.neu"])
def test_rt_gradient_exact_on_polynomials(mesh):
    """TestDissipation2 gradient check (euler_test.go:548-668): at N=4 the RT
    gradient of x+y, x^2+y^2, x^3+y^3, x^4+y^4 sampled at the solution points
    is exact at every RT point (reference tolerance 1e-6)."""
    s = _solver(os.path.join(G, mesh), 4, flux="average")
    c = s.coords()
    X, Y = c["SolX"], c["SolY"]
    q = np.stack([X + Y, X ** 2 + Y ** 2, X ** 3 + Y ** 3, X ** 4 + Y ** 4])
    dx, dy = s.gradient(q)
    FX, FY = c["FluxX"], c["FluxY"]
    for n in range(4):
        p = n + 1
        ex, ey = p * FX ** (p - 1), p * FY ** (p - 1)
        sc = max(np.abs(ex).max(), np.abs(ey).max(), 1.0)
        assert np.abs(dx[n] - ex).max() < 1e-11 * sc, (n, np.abs(dx[n] - ex).max() / sc)
        assert np.abs(dy[n] - ey).max() < 1e-11 * sc


def test_rt_gradient_linear_on_periodic_mesh(tmp_path):
    """A linear field in x on a mesh with periodic partners: the gradient is
    exact where the field is continuous; it checks the non-owner's reversed
    read of the owner trace on interior faces of a jittered mesh."""
    path = meshgen.vortex_mesh(str(tmp_path / "v6.su2"), 6)
    s = _solver(path, 2, init="ivortex")
    c = s.coords()
    X, Y = c["SolX"], c["SolY"]
    q = np.stack([3 * X - 2 * Y, X * Y, Y, np.ones_like(X)])
    dx, dy = s.gradient(q)
    etoe, _ = s.connectivity()
    # elements away from the periodic seam (x in [2,8]) see a continuous field
    cx = X.mean(axis=0)
    inner = (cx > 2) & (cx < 8)
    assert np.abs(dx[0][:, inner] - 3).max() < 1e-10
    assert np.abs(dy[0][:, inner] + 2).max() < 1e-10
    assert np.abs(dx[1][:, inner] - c["FluxY"][:, inner]).max() < 1e-10
    assert np.abs(dy[2][:, inner] - 1).max() < 1e-10
    assert np.abs(dx[3]).max() < 1e-10 and np.abs(dy[3]).max() < 1e-10


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_sensor_closed_forms(N):
    """UpdateSeMoment (DG2D/dfr_shock_capturing.go:142-166): a field in the
    span of modes of degree < N has no top-mode energy (Se -> -inf); a field
    made of degree-N modes only has Se = log10(1) = 0."""
    s = _solver(os.path.join(G, "test_tris_9.neu"), N)
    se_el = s.solution_element()
    V = se_el["V"]
    modes = [(i, j) for i in range(N + 1) for j in range(N + 1 - i)]
    low = np.array([i + j < N for (i, j) in modes])
    rng = np.random.default_rng(N)
    c_low = rng.standard_normal((s.NpInt, s.K)) * low[:, None]
    c_low[0] += 3.0
    c_top = rng.standard_normal((s.NpInt, s.K)) * (~low)[:, None]
    se = s.sensor(V @ c_low)
    assert (se < -12).all(), se
    se = s.sensor(V @ c_top)
    assert np.abs(se).max() < 1e-12


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_modal_filter_closed_forms(N):
    """limitAndFilterSolution (dissipation.go:606-622) with ModalFilter
    (dfr_shock_capturing.go:43-69): sigma = 1 keeps only mode 0, so a
    polynomial of degree <= N becomes its mean over the element (r -> -1/3 on
    the reference triangle); every sigma keeps the element mean; sigma = 0 at
    N = 1 is the identity."""
    s = _solver(os.path.join(G, "test_tris_9.neu"), N)
    se_el = s.solution_element()
    R, S, W = se_el["R"], se_el["S"], se_el["W"]
    u = np.repeat(R[:, None], s.K, axis=1)
    f = s.filter(u, 1.0)
    assert np.abs(f + 1.0 / 3.0).max() < 1e-12
    rng = np.random.default_rng(7)
    u = rng.standard_normal((4, s.NpInt, s.K))
    sig = rng.uniform(0, 1, s.K)
    f = s.filter(u, sig)
    assert np.abs(np.einsum("i,nik->nk", W, f - u)).max() < 1e-12
    if N == 1:
        assert np.abs(s.filter(u, 0.0) - u).max() < 1e-13
    # the top mode is damped by mf * (1 - sin(pi sigma / 2)) < 1
    top = se_el["V"][:, -1]
    ft = s.filter(np.repeat(top[:, None], s.K, axis=1), 0.5)
    assert np.abs(ft).max() < np.abs(top).max()


@pytest.mark.parametrize("N", [1, 2, 3])
@pytest.mark.parametrize("flux", ["lax", "roe"])
def test_shock_rhs_conserves_on_periodic_mesh(tmp_path, N, flux):
    """On a mesh periodic in x and y the Euler and viscous edge fluxes are
    single valued per edge and the modal filter keeps mode 0, so the
    integral of the shock-path RHS over the domain vanishes:
    sum_k Jdet_k sum_i W_i RHS_ik = 0 for each variable."""
    path = meshgen.periodic_mesh(str(tmp_path / "p6.su2"), 6)
    s = _solver(path, N, flux=flux, Minf=0.5)
    q = s.state()
    rng = np.random.default_rng(3)
    q = q * (1.0 + 0.05 * rng.standard_normal(q.shape))
    r = s.rhs(q)
    assert np.isfinite(r).all()
    se, sigma, eps = s.shock()
    assert (sigma > 0).any() and (eps > 0).any()
    d = s.dissipation()
    assert np.abs(d["DissX"]).max() > 1e-3
    W = s.solution_element()["W"]
    J = s.geometry()["Jdet"]
    tot = np.einsum("i,nik,k->n", W, r, J)
    scale = np.einsum("i,nik,k->n", W, np.abs(r), J)
    assert (np.abs(tot) < 1e-12 * scale).all(), tot / scale


def test_shock_rhs_reduces_to_euler_when_smooth(tmp_path):
    """With a smooth state the sensor is far below S0 - kappa, sigma = eps = 0
    and the filter at sigma 0 only damps the modes by mf; at N=1 (mf = 1) the
    shock-path RHS equals the plain Euler RHS."""
    path = meshgen.vortex_mesh(str(tmp_path / "v6.su2"), 6)
    s = _solver(path, 1, init="ivortex", Kappa=0.1)
    p = oracle.OracleSolver(path, 1, flux="roe", init="ivortex")
    q = p.state()
    r_s, r_p = s.rhs(q), p.rhs(q)
    se, sigma, eps = s.shock()
    assert (eps == 0).all()
    assert np.abs(r_s - r_p).max() < 1e-12 * np.abs(r_p).max()


def test_local_time_stepping_dt(tmp_path):
    """CalculateLocalDT (model_problems/Euler2D/euler.go:973-1002): with local
    time stepping each element's DT is CFL / (max edge wave speed of its
    three edges); the global Time does not advance (GlobalDT keeps 0)."""
    path = meshgen.vortex_mesh(str(tmp_path / "v6.su2"), 6)
    s = oracle.OracleSolver(path, 2, flux="lax", init="ivortex", CFL=0.8, LocalTimeStepping=True)
    g = oracle.OracleSolver(path, 2, flux="lax", init="ivortex", CFL=0.8)
    s.step()
    g.step()
    dt = s.element_dt()
    assert s.time == 0.0
    assert (dt > 0).all() and dt.max() > dt.min()
    # every local dt is at least the global one (the global dt takes the
    # minimum over elements of the same quantity at stage 4)
    assert dt.min() >= 0.5 * g.dt
