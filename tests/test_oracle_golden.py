"""Pins the CPU oracle against the reference's own known-answer tests and
against closed forms. CPU only.

Golden vectors come from the reference tests named at each test; the mesh
fixtures under tests/golden/ are the reference's DG2D/test_data files.
"""
import math
import os

import numpy as np
import pytest

import oracle

G = os.path.join(os.path.dirname(__file__), "golden")


def _solver(name, N, **kw):
    kw.setdefault("flux", "roe")
    kw.setdefault("init", "freestream")
    kw.setdefault("Minf", 2.0)
    return oracle.OracleSolver(os.path.join(G, name), N, **kw)


def test_edge_keys_tris9():
    """TestEdges (model_problems/Euler2D/edges_test.go:71-96)."""
    s = _solver("test_tris_9.neu", 1)
    e = s.edges()
    assert len(e) == 19
    assert sorted(e[:, 1].tolist()) == [1, 2, 3, 4, 4, 4, 5, 5, 5, 6, 6, 7, 7, 8, 8, 8, 9, 9, 9]
    assert sorted(e[:, 0].tolist()) == [0, 0, 0, 1, 1, 1, 2, 2, 3, 3, 4, 4, 4, 5, 5, 5, 6, 7, 8]


def test_edge_structure_tris5():
    """TestDFR2D edge construction (DG2D/dfr_startup_test.go:134-155)."""
    s = _solver("test_tris_5.neu", 1)
    e = {(int(r[0]), int(r[1])): r for r in s.edges()}
    r = e[(0, 2)]
    assert r[2] == 2 and r[3] == 0 and r[5] == 2 and r[4] == 1 and r[6] == 0
    assert e[(0, 1)][9] == 1 and e[(2, 3)][9] == 1  # BC_In
    etoe, _ = s.connectivity()
    assert etoe[0].tolist() == [-1, -1, 1]
    assert etoe[1].tolist() == [0, -1, -1]


def test_vertex_to_element_area_tris9():
    """TestDissipation: element area = 2 * Jdet = 0.25 (euler_test.go:442-447)."""
    s = _solver("test_tris_9.neu", 1)
    assert np.allclose(2 * s.geometry()["Jdet"], 0.25, atol=1e-12)


@pytest.mark.parametrize("name", ["test_tris_5.neu", "test_tris_9.neu", "vortex-new.su2"])
def test_transformed_normals(name):
    """TestDFR2D Piola check (dfr_startup_test.go:157-203): the scaled normal of
    a whole edge maps to [0,-2], [2,2], [-2,0] in the unit triangle, and the
    ||n|| factor makes it unit length there."""
    s = _solver(name, 1, init="ivortex" if name.endswith("su2") else "freestream")
    g = s.geometry()
    K = s.K
    VX, VY, EToV = _mesh_arrays(os.path.join(G, name))
    expect = {0: (0, -2), 1: (2, 2), 2: (-2, 0)}
    for k in range(K):
        v = EToV[k]
        xr, yr = 0.5 * (VX[v[1]] - VX[v[0]]), 0.5 * (VY[v[1]] - VY[v[0]])
        xs, ys = 0.5 * (VX[v[2]] - VX[v[0]]), 0.5 * (VY[v[2]] - VY[v[0]])
        det = xr * ys - xs * yr
        for f in range(3):
            a, b = v[f], v[(f + 1) % 3]
            length = math.hypot(VX[b] - VX[a], VY[b] - VY[a])
            n = np.array([g["FaceNx"][f, k], g["FaceNy"][f, k]]) * length
            nT = np.array([xr * n[0] + yr * n[1], xs * n[0] + ys * n[1]]) / det
            assert np.allclose(nT, expect[f], atol=1e-9), (k, f, nT)
            scaled = np.array([g["FaceNx"][f, k], g["FaceNy"][f, k]]) * g["IInII"][f, k]
            nTu = np.array([xr * scaled[0] + yr * scaled[1], xs * scaled[0] + ys * scaled[1]]) / det
            assert abs(np.hypot(*nTu) - 1.0) < 1e-9


def _mesh_arrays(path):
    s = oracle.OracleSolver(path, 0, flux="lax", init="freestream", Minf=0.5)
    return _read_vertices(path, s.K)


def _read_vertices(path, K):
    if path.endswith(".su2"):
        lines = open(path).read().splitlines()
        i = next(j for j, l in enumerate(lines) if l.startswith("NELEM"))
        E = [list(map(int, lines[i + 1 + k].split()[1:4])) for k in range(K)]
        i = next(j for j, l in enumerate(lines) if l.startswith("NPOIN"))
        nv = int(lines[i].split("=")[1])
        V = [list(map(float, lines[i + 1 + j].split()[:2])) for j in range(nv)]
        V = np.array(V)
        return V[:, 0], V[:, 1], np.array(E)
    lines = open(path).read().splitlines()
    nv, ke = map(int, lines[6].split()[:2])
    V = np.zeros((nv, 2))
    for j in range(nv):
        t = lines[9 + j].split()
        V[int(t[0]) - 1] = [float(t[1]), float(t[2])]
    E = np.zeros((ke, 3), dtype=int)
    base = 9 + nv + 2
    for k in range(ke):
        t = list(map(int, lines[base + k].split()))
        E[t[0] - 1] = [t[3] - 1, t[4] - 1, t[5] - 1]
    return V[:, 0], V[:, 1], E


def test_fluid_functions():
    """TestFluidFunctions (euler_test.go:32-45): M=2 freestream."""
    q, p, c = oracle.freestream(2.0, 1.4, 0.0)
    mach = math.hypot(q[1] / q[0], q[2] / q[0]) / c
    assert np.allclose([q[0], q[1], q[2], q[3], mach, p], [1, 2, 0, 3.78571, 2, 0.71429], atol=1e-5)


def test_ivortex_known_answer():
    """TestIVortex (isentropic_vortex/analytic_vortex_test.go:11-24): state and flux at the centre."""
    q = oracle.ivortex_state(0.0, 5.0, 0.0)
    assert np.allclose(q, [0.361673, 0.361673, 0.0, 0.782817], atol=1e-5)
    Fx = oracle.numerical_flux("average", 1.4, q, q, [1.0, 0.0])
    Fy = oracle.numerical_flux("average", 1.4, q, q, [0.0, 1.0])
    assert np.allclose(Fx, [0.361673, 0.602465, 0.0, 1.023609], atol=1e-5)
    assert np.allclose(Fy, [0.0, 0.0, 0.240792, 0.0], atol=1e-5)


def test_ivortex_far_field_is_freestream():
    """Closed form: far from the centre the vortex is the (1, 1, 0, 1/0.4 + 0.5) freestream."""
    q = oracle.ivortex_state(0.0, 5.0 + 50.0, 0.0)
    assert np.allclose(q, [1, 1, 0, 1 / 0.4 + 0.5], atol=1e-14)


@pytest.mark.parametrize("flux", ["average", "lax", "roe", "roe-er"])
def test_flux_consistency(flux):
    """Closed form: every numerical flux of equal states is the physical flux
    F(Q).n (fluxes.go:135-503)."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        rho = rng.uniform(0.5, 2.0); u, v = rng.uniform(-1, 1, 2); p = rng.uniform(0.5, 2.0)
        q = np.array([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])
        th = rng.uniform(0, 2 * np.pi)
        n = np.array([np.cos(th), np.sin(th)])
        Fn = n[0] * np.array([q[1], q[1] * u + p, q[1] * v, u * (q[3] + p)]) + \
            n[1] * np.array([q[2], q[2] * u, q[2] * v + p, v * (q[3] + p)])
        F = oracle.numerical_flux(flux, 1.4, q, q, n)
        assert np.allclose(F, Fn, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("flux", ["lax", "roe"])
def test_flux_conservation_symmetry(flux):
    """Closed form: F*(L,R,n) = -F*(R,L,-n) for Lax-Friedrichs and Roe."""
    rng = np.random.default_rng(5)
    for _ in range(20):
        qs = []
        for _ in range(2):
            rho = rng.uniform(0.5, 2.0); u, v = rng.uniform(-1, 1, 2); p = rng.uniform(0.5, 2.0)
            qs.append(np.array([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)]))
        th = rng.uniform(0, 2 * np.pi)
        n = np.array([np.cos(th), np.sin(th)])
        a = oracle.numerical_flux(flux, 1.4, qs[0], qs[1], n)
        b = oracle.numerical_flux(flux, 1.4, qs[1], qs[0], -n)
        assert np.allclose(a, -b, rtol=1e-12, atol=1e-12)


def test_riemann_bc_invariants():
    """Closed forms of RiemannBC (bcs.go:70-133): a state equal to the far
    field is returned unchanged (subsonic); supersonic inflow copies Qinf and
    supersonic outflow copies the interior."""
    q, p, c = oracle.freestream(0.5, 1.4, 10.0)
    for th in np.linspace(0, 2 * np.pi, 9):
        n = [np.cos(th), np.sin(th)]
        qb = oracle.riemann_bc(1.4, 0.5, p, c, q, q, n)
        assert np.allclose(qb, q, rtol=1e-13, atol=1e-13)
    q2, p2, c2 = oracle.freestream(2.0, 1.4, 0.0)
    qi = q2 * 1.1
    assert np.allclose(oracle.riemann_bc(1.4, 2.0, p2, c2, qi, q2, [-1.0, 0.0]), q2)  # inflow
    assert np.allclose(oracle.riemann_bc(1.4, 2.0, p2, c2, qi, q2, [1.0, 0.0]), qi)   # outflow


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_interpolation_exactness(N):
    """TestDFR2D interpolation accuracy (dfr_startup_test.go:38-63): r^N + s^N
    interpolates exactly from the solution points to every RT point."""
    s = _solver("test_tris_5.neu", N)
    o = s.operators()
    R, S = o["R"], o["S"]
    NpInt = s.NpInt
    f = R ** N + S ** N
    # interior points: the first NpInt RT points coincide with the solution points
    se = s.solution_element()
    assert np.allclose(se["R"], R[:NpInt], atol=0) and np.allclose(se["S"], S[:NpInt], atol=0)
    edge = o["Interp"] @ f[:NpInt]
    assert np.allclose(edge, f[2 * NpInt:], atol=1e-11)


def _rt_divergence_check(s, order):
    """DivergenceCheck (dfr_startup_test.go:293-360): project F=(x^o, y^o) on
    the RT DOFs (ProjectFluxOntoRTSpace, dfr_startup.go:316-357) and compare
    Div.Fp/Jdet with o(x^(o-1) + y^(o-1))."""
    o = s.operators()
    g = s.geometry()
    c = s.coords()
    NpInt, NpEdge, NpFlux = s.NpInt, s.NpEdge, s.NpFlux
    worst = 0.0
    for k in range(s.K):
        x, y = c["FluxX"][:, k], c["FluxY"][:, k]
        Fx, Fy = x ** order, y ** order
        J, Ji = g["Jdet"][k], g["Jinv"][k]
        fT0 = J * (Ji[0] * Fx + Ji[1] * Fy)
        fT1 = J * (Ji[2] * Fx + Ji[3] * Fy)
        Fp = np.empty(NpFlux)
        Fp[:NpInt] = fT0[:NpInt]
        Fp[NpInt:2 * NpInt] = fT1[NpInt:2 * NpInt]
        e1 = slice(2 * NpInt, 2 * NpInt + NpEdge)
        e2 = slice(2 * NpInt + NpEdge, 2 * NpInt + 2 * NpEdge)
        e3 = slice(2 * NpInt + 2 * NpEdge, NpFlux)
        Fp[e1] = -fT1[e1]
        Fp[e2] = (fT0[e2] + fT1[e2]) / math.sqrt(2)
        Fp[e3] = -fT0[e3]
        div = o["Div"] @ Fp / J
        exact = order * (x ** (order - 1) + y ** (order - 1))
        tol = 1e-9 * max(np.abs(Fx).max(), 1.0)
        worst = max(worst, np.abs(div - exact).max() / tol)
    return worst


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_rt_divergence_exact(N):
    """TestDivergence (dfr_startup_test.go:362-368) on test_tris_5: the RT
    divergence is exact for polynomial fluxes up to order N."""
    s = _solver("test_tris_5.neu", N)
    for order in range(1, N + 1):
        assert _rt_divergence_check(s, order) <= 1.0


@pytest.mark.parametrize("N", [0, 1, 2, 3, 4])
def test_freestream_preservation(N):
    """TestEuler part 2 (euler_test.go:176-226): a freestream has zero divergence."""
    s = _solver("test_tris_6_nowall.neu", N, flux="average", Minf=2.0)
    r = s.rhs(s.state(), 0.0)
    assert np.abs(r).max() < 1e-10


@pytest.mark.parametrize("N", [0, 1, 2, 3, 4])
def test_edge_interpolation_of_constants(N):
    """TestEuler part 1 (euler_test.go:55-91): constants interpolate exactly to edges."""
    s = _solver("test_tris_5.neu", N)
    o = s.operators()
    assert np.allclose(o["Interp"].sum(axis=1), 1.0, atol=1e-12)


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_mass_matrix_orthonormal_basis(N):
    """M = V^T W V (lagrange_element.go:79-82). The WSJ weights sum to 1 and the
    constant mode of the orthonormal basis is 1/sqrt(2), so M[0,0] = 1/2 for
    every N; for N <= 2 the rule integrates degree 2N exactly and M = I/2
    (CheckBasisOrthogonality, lagrange_element.go:94-121)."""
    s = _solver("test_tris_5.neu", N)
    se = s.solution_element()
    assert abs(se["W"].sum() - 1.0) < 1e-13
    M = se["Mass"]
    assert abs(M[0, 0] - 0.5) < 1e-13
    assert np.allclose(M, M.T, atol=1e-14)
    if N <= 2:
        assert np.allclose(M, 0.5 * np.eye(s.NpInt), atol=1e-12)


def test_shock_finder_flags():
    """TestShockFinder (DG2D/dfr_shock_capturing_test.go:11-49): shocked
    elements [F,F,T,T,F,F,F,T,T,F] for the Mach-2 normal shock field, N=2..4,
    kappa=3, threshold 0.0075."""
    for N in (2, 3, 4):
        s = _solver("test_10tris_centered.neu", N, limiter="PerssonC0", Kappa=3.0)
        c = s.coords()
        rho = _normal_shock_density(c["SolX"], c["SolY"])
        sigma = s.shock_indicator(rho, 3.0)
        flags = (sigma > 0.0075).tolist()
        assert flags == [False, False, True, True, False, False, False, True, True, False], (N, sigma)


def _normal_shock_density(X, Y):
    """NORMALSHOCKTESTM2 density (DG2D/test_functions2.go:62-69,133-238): the
    pre-shock state rho=1, p=1, u=2 for x < 0 and the Rankine-Hugoniot
    post-shock density for x >= 0, with normal Mach u/sqrt(gamma p/rho)."""
    g = 1.4
    M1n = 2.0 / math.sqrt(g)
    ratio = (g + 1) * M1n * M1n / ((g - 1) * M1n * M1n + 2)
    return np.where(X < 0, 1.0, ratio)
