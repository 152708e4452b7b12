"""GPU checks of batched ensembles: each case of an FP64 or FP32 ensemble is
bit-identical to its own single-case run in the same precision (same
tiling, same arithmetic; 25 steps cover the eager steps and four replays of
the captured 6-step graph); FP32 cases stay within 1e-5 relative L2 of the
FP64 oracle; cases stop at their own FinalTime."""
import numpy as np
import pytest

import oracle
from gocfd_amd import Ensemble, Euler, InputParameters2D, Mesh, meshgen
from gocfd_amd._lib import GocfdError

pytestmark = pytest.mark.gpu


def _mesh(n, seed):
    VX, VY, EToV, tags = meshgen.jittered_quads(n, n, seed=seed)
    return Mesh.from_arrays(VX, VY, EToV, tags)


def _cases(tmp_path):
    spec = [  # (quads per side, order, init, Minf, Alpha, flux CFL, FinalTime)
        (10, 2, "IVortex", 0.5, 0.0, 1.0, 1e9),
        (7, 2, "Freestream", 0.4, 3.0, 0.8, 1e9),
        (12, 3, "IVortex", 0.5, 0.0, 1.0, 1e9),
        (9, 1, "Freestream", 0.7, -4.0, 1.0, 1e9),
        (6, 4, "Freestream", 0.3, 1.0, 0.9, 1e9),
    ]
    out = []
    for i, (n, N, init, M, a, cfl, ft) in enumerate(spec):
        path = str(tmp_path / f"case{i}.su2")
        VX, VY, EToV, tags = meshgen.jittered_quads(n, n, seed=100 + i)
        meshgen.write_su2(path, VX, VY, EToV, tags)
        ip = InputParameters2D(CFL=cfl, FluxType="Roe", InitType=init, PolynomialOrder=N, Minf=M, Alpha=a,
                               FinalTime=ft, MaxIterations=1000)
        out.append((ip, path))
    return out


def test_ensemble_fp64_is_bitwise_the_single_case_run(tmp_path):
    cases = _cases(tmp_path)
    ens = Ensemble(cases, precision="fp64")
    ens.step(25)
    for c, (ip, path) in enumerate(cases):
        single = Euler(ip, path, precision="fp64")
        single.step(25)
        assert np.array_equal(ens.state(c), single.Q), c
        assert ens.time_info(c) == single.time_info(), c


def test_ensemble_fp32_against_single_case_and_oracle(tmp_path):
    """FP32 ensemble: bit-identical to each case's single-case FP32 run, and
    within 1e-5 relative L2 of the FP64 oracle. (At N<=1 the FP32 path keeps
    FP32 stage states but computes in FP64: its RT edge points sit 0.056
    apart and |DivInt| reaches 1.2e3.)"""
    cases = _cases(tmp_path)
    ens = Ensemble(cases, precision="fp32")
    ens.step(20)
    for c, (ip, path) in enumerate(cases):
        single = Euler(ip, path, precision="fp32")
        single.step(20)
        assert np.array_equal(ens.state(c), single.Q), c
        o = oracle.OracleSolver(path, ip.PolynomialOrder, flux="roe", init=ip.InitType.lower(), CFL=ip.CFL,
                                FinalTime=ip.FinalTime, Minf=ip.Minf, Alpha=ip.Alpha)
        for _ in range(20):
            o.step()
        qo = o.state()
        l2 = np.linalg.norm(ens.state(c)[0] - qo[0]) / np.linalg.norm(qo[0])
        assert l2 < 1e-5, (c, l2)
        t, dt, steps = ens.time_info(c)
        assert steps == 20 and abs(t - o.time) < 1e-6 * o.time


def test_ensemble_cases_stop_at_their_final_time(tmp_path):
    cases = _cases(tmp_path)[:3]
    cases[1][0].FinalTime = 0.02
    ens = Ensemble(cases, precision="fp64")
    ens.step(40)
    t1, _, s1 = ens.time_info(1)
    assert t1 == pytest.approx(0.02, rel=1e-14) and s1 < 40
    t0, _, s0 = ens.time_info(0)
    assert s0 == 40 and t0 > 0.02
    single = Euler(cases[1][0], cases[1][1], precision="fp64")
    single.step(40)
    assert np.array_equal(ens.state(1), single.Q)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_launch_configurations_are_bitwise_identical(tmp_path, prec):
    """Order groups as parallel graph branches or (FP64) in one multi-order
    launch per stage, 24-step graphs, eager launches and the plain tile
    order change the schedule, not the arithmetic. The FP32 multi-order
    kernel is switched off: it faulted the GPU in its first run, cause not
    found."""
    cases = _cases(tmp_path)
    ref = Ensemble(cases, precision=prec)
    assert ref.launch() == {"parallel": 0, "graph_steps": 6, "xcd_tiles": True}
    ref.step(31)
    configs = [(1, 6, True), (1, 24, False), (0, 12, False)]
    if prec == "fp64":  # the multi-order launch is FP64 only (its FP32 variant is switched off)
        configs += [(2, 6, True), (2, 24, False), (2, 0, True)]
    else:
        with pytest.raises(GocfdError, match="no multi-order launch"):
            ref.set_launch(2, 6, True)
        assert ref.launch() == {"parallel": 0, "graph_steps": 6, "xcd_tiles": True}
    for par, gs, xcd in configs:
        ens = Ensemble(cases, precision=prec)
        ens.set_launch(par, gs, xcd)
        ens.step(31)
        for c in range(len(cases)):
            assert np.array_equal(ens.state(c), ref.state(c)), (par, gs, xcd, c)
            assert ens.time_info(c) == ref.time_info(c)


@pytest.mark.skip(reason="the FP32 multi-order stage kernel faulted the GPU in its first run; reading the kernel "
                         "and its launch found no cause, so the launch mode is switched off (set_launch refuses it)")
def test_fp32_multi_order_launch_is_bitwise_identical(tmp_path):
    """The FP32 ensemble in one multi-order launch per stage against its
    groups launched in sequence, bit for bit."""
    cases = _cases(tmp_path)
    ref = Ensemble(cases, precision="fp32")
    ref.step(31)
    ens = Ensemble(cases, precision="fp32")
    ens.set_launch(2, 6, True)
    ens.step(31)
    for c in range(len(cases)):
        assert np.array_equal(ens.state(c), ref.state(c)), c


@pytest.mark.parametrize("par", [0, 1])
def test_interrupted_capture_leaves_the_ensemble_usable(tmp_path, monkeypatch, par):
    """An error inside a graph capture (forced from the host by the
    GOCFD_TEST_FAIL_CAPTURE hook after the first captured step; no kernel is
    affected) is reported, ends the capture (joining the forked branch
    streams of the parallel schedule) and leaves the handle usable: the
    steps taken afterwards give the states of an uninterrupted run of the
    same total length, bit for bit."""
    cases = _cases(tmp_path)
    ens = Ensemble(cases, precision="fp64")
    ens.set_launch(par, 6, True)
    monkeypatch.setenv("GOCFD_TEST_FAIL_CAPTURE", "1")
    with pytest.raises(GocfdError, match="graph capture interrupted"):
        ens.step(12)
    monkeypatch.delenv("GOCFD_TEST_FAIL_CAPTURE")
    done = ens.time_info(0)[2]  # eager steps before the capture started
    ens.step(13)
    ref = Ensemble(cases, precision="fp64")
    ref.set_launch(par, 6, True)
    ref.step(done + 13)
    for c in range(len(cases)):
        assert np.array_equal(ens.state(c), ref.state(c)), (par, c)
        assert ens.time_info(c) == ref.time_info(c)


def test_tune_restores_the_initial_state(tmp_path):
    cases = _cases(tmp_path)
    ens = Ensemble(cases, precision="fp32")
    choice = ens.tune(1)
    assert choice["graph_steps"] in (6, 24) and choice["ms_per_step"] > 0
    for c in range(len(cases)):
        t, _, steps = ens.time_info(c)
        assert t == 0.0 and steps == 0
    ens.step(30)
    ref = Ensemble(cases, precision="fp32")
    ref.step(30)
    for c in range(len(cases)):
        assert np.array_equal(ens.state(c), ref.state(c)), c
    with pytest.raises(GocfdError):
        ens.tune(1)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_shock_and_lts_cases_in_an_ensemble(tmp_path, prec):
    """PerssonC0 cases (two Kappa values, so two groups of one order),
    LocalTimeStepping cases (with and without the limiter) and a plain case
    step together; each is bit-identical to its single-case run over the
    eager steps and replays of the captured graph, also with the groups
    without PerssonC0 (a LocalTimeStepping one among them) in one
    multi-order launch per stage (FP64)."""
    tube = meshgen.shocktube_mesh(str(tmp_path / "st.su2"), 60, 2, height=0.02)
    wedge = meshgen.wedge_mesh(str(tmp_path / "wedge.su2"), 900, seed=3)

    def ip(N, init, lim="", kappa=0.0, lts=False, minf=0.0, cfl=0.5):
        return InputParameters2D(CFL=cfl, FluxType="Roe", InitType=init, PolynomialOrder=N, Minf=minf,
                                 FinalTime=1e9, MaxIterations=1000, Limiter=lim, Kappa=kappa,
                                 LocalTimeStepping=lts)

    cases = [(ip(2, "shocktube", "PerssonC0", 3.0), tube),
             (ip(2, "Freestream", "PerssonC0", 10.0, minf=2.0, cfl=1.0), wedge),
             (ip(2, "shocktube", "PerssonC0", 3.0), tube),
             (ip(1, "Freestream", "", lts=True, minf=2.0, cfl=1.0), wedge),
             (ip(2, "Freestream", "PerssonC0", 10.0, lts=True, minf=2.0, cfl=1.0), wedge),
             (ip(3, "shocktube", ""), tube)]
    ens = Ensemble(cases, precision=prec)
    ens.step(25)
    multi = None
    if prec == "fp64":  # the two groups without PerssonC0 in one launch (FP64 only)
        multi = Ensemble(cases, precision=prec)
        multi.set_launch(2, 6, True)
        multi.step(25)
    for c, (p, path) in enumerate(cases):
        single = Euler(p, path, precision=prec)
        single.step(25)
        assert np.array_equal(ens.state(c), single.Q), c
        assert ens.time_info(c) == single.time_info(), c
        if multi is not None:
            assert np.array_equal(multi.state(c), single.Q), ("multi-order launch", c)
