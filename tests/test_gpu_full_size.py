"""GPU checks at the full benchmark sizes, through properties that do not
depend on the size, and short parity runs against the oracle (which runs
with OpenMP over the host cores).

* 80,000 triangles (200x200 jittered quads), P=4, Roe: the freestream is
  preserved to round-off over 20 steps (1e-11 FP64; 1e-3 FP32, whose
  divergence rounding is ~1e-7 x sum|DivInt| / Jdet per stage); 5 FP64 steps of the isentropic vortex match the
  oracle to 1e-12 relative L2 in density, time and dt to 1e-12; 50 steps
  stay within 1e-6 of the analytic convected vortex; FP32 bulk stays within
  1e-5 of FP64.
* The 125k-point wedge channel (about 250k triangles), P=2, Roe, PerssonC0
  (Kappa 10): 10 FP64 steps from the freestream match the oracle to 1e-10,
  with identical shock-cell flags; after 10, 500, 1000 and 2000 steps the
  flags are the oracle's bit for bit (a committed fixture: the oracle needs
  about 90 minutes for 2000 steps).
"""
import math
import os

import numpy as np
import pytest

import oracle
from gocfd_amd import Euler, InputParameters2D, meshgen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vortex200(tmp_path_factory):
    return meshgen.vortex_mesh(str(tmp_path_factory.mktemp("vortex200") / "vortex200.su2"), 200)


def _vortex_density(t, X, Y):
    g, beta = 1.4, 5.0
    r2 = (X - t - 5.0) ** 2 + Y ** 2
    tv1 = 1.0 - (g - 1) * beta * beta * np.exp(2.0 * (1.0 - r2)) / (16 * g * math.pi ** 2)
    return tv1 ** (1.0 / (g - 1))


def _l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_freestream_preserved_full_size(vortex200, prec):
    ip = InputParameters2D(CFL=1.0, FluxType="Roe", InitType="Freestream", PolynomialOrder=4, Minf=0.5,
                           Alpha=2.0, FinalTime=1e9)
    g = Euler(ip, vortex200, precision=prec)
    assert g.K == 80000
    q0 = g.Q
    g.step(20)
    bar = 1e-11 if prec == "fp64" else 1e-3
    assert np.abs(g.Q - q0).max() < bar * np.abs(q0).max()


def test_vortex_full_size_parity_and_analytic(vortex200):
    ip = InputParameters2D(CFL=1.0, FluxType="Roe", InitType="IVortex", PolynomialOrder=4, FinalTime=1e9)
    g = Euler(ip, vortex200, precision="fp64")
    o = oracle.OracleSolver(vortex200, 4, flux="roe", init="ivortex", CFL=1.0)
    g.step(5)
    for _ in range(5):
        o.step()
    t, dt, steps = g.time_info()
    assert steps == 5
    assert abs(t - o.time) <= 1e-12 * o.time and abs(dt - o.dt) <= 1e-12 * o.dt
    assert _l2(g.Q[0], o.state()[0]) < 1e-12
    g.step(45)
    t, _, _ = g.time_info()
    c = g.coords()
    assert _l2(g.Q[0], _vortex_density(t, c["SolX"], c["SolY"])) < 1e-6
    g32 = Euler(ip, vortex200, precision="fp32")
    g32.step(50)
    assert _l2(g32.Q[0], g.Q[0]) < 1e-5


def test_wedge_full_size_shock_parity(tmp_path):
    mesh = meshgen.wedge_mesh(str(tmp_path / "wedge.su2"), 125_000, seed=1)
    ip = InputParameters2D(CFL=1.0, FluxType="Roe", InitType="Freestream", PolynomialOrder=2, Minf=2.0,
                           FinalTime=1e9, Limiter="PerssonC0", Kappa=10.0)
    g = Euler(ip, mesh, precision="fp64")
    assert g.K > 240_000
    o = oracle.OracleSolver(mesh, 2, flux="roe", init="freestream", CFL=1.0, Minf=2.0, limiter="PerssonC0",
                            Kappa=10.0)
    g.step(10)
    for _ in range(10):
        o.step()
    t, dt, steps = g.time_info()
    assert abs(t - o.time) <= 1e-10 * o.time
    assert _l2(g.Q[0], o.state()[0]) < 1e-10
    _, sg_g, _ = g.shock()
    _, sg_o, _ = o.shock()
    print(f"wedge: {(sg_o > 0).sum()} of {o.K} elements flagged after 10 steps", flush=True)
    assert (sg_o > 0).sum() > 0  # the comparison below is not vacuous
    assert np.array_equal(sg_g > 0, sg_o > 0)


def test_wedge_flags_2000_steps_match_oracle(tmp_path):
    """The sensor flags (sigma > 0 per element) of the FP64 wedge after 10,
    500, 1000 and 2000 steps are the oracle's, bit for bit
    (tests/golden/wedge_flags_2000.npz, made by make_wedge_flags.py beside
    it); time and dt agree to 1e-12 relative, and the density's sum and max
    at 2000 steps to 1e-12."""
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "wedge_flags_2000.npz"))
    mesh = meshgen.wedge_mesh(str(tmp_path / "wedge.su2"), 125_000, seed=1)
    ip = InputParameters2D(CFL=1.0, FluxType="Roe", InitType="Freestream", PolynomialOrder=2, Minf=2.0,
                           FinalTime=1e9, Limiter="PerssonC0", Kappa=10.0)
    g = Euler(ip, mesh, precision="fp64")
    K = int(gold["K"])
    assert g.K == K
    done = 0
    for n in gold["checkpoints"]:
        g.step(int(n) - done)
        done = int(n)
        t, dt, steps = g.time_info()
        _, sg, _ = g.shock()
        want = np.unpackbits(gold[f"flags_{n}"])[:K].astype(bool)
        print(f"wedge step {n}: {int((sg > 0).sum())} flagged (oracle {int(want.sum())})", flush=True)
        assert want.sum() > 0
        assert np.array_equal(sg > 0, want), f"step {n}: {int(np.sum((sg > 0) != want))} flags differ"
        assert abs(t - float(gold[f"time_{n}"])) <= 1e-12 * t
        assert abs(dt - float(gold[f"dt_{n}"])) <= 1e-12 * dt
    rho = g.Q[0]
    assert abs(rho.sum() - float(gold["rho_sum_2000"])) <= 1e-12 * abs(float(gold["rho_sum_2000"]))
    assert abs(rho.max() - float(gold["rho_max_2000"])) <= 1e-12 * float(gold["rho_max_2000"])
