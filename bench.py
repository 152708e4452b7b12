#!/usr/bin/env python3
"""Benchmark of the 2D Euler DFR SSP54 step on one MI355X.

Workload: isentropic vortex on [0,10]x[-5,5], 200x200 jittered quads split into
80,000 triangles, P=4 (15 solution points per element, 1.2M DOF x 4
variables), Roe flux, SSP54, CFL 1. One "step" is one SSP54 step (5 stage
launches) over the whole mesh with the state resident in HBM.

`value` is the FP32-bulk step's DOF-stage/s (DOF = K*NpInt; 5 stages per
step; FP32 stage states and flux/divergence arithmetic, FP64 accumulated
state q0 and residual R); the FP64 drop-in is measured in the same run and
reported beside it,
with the density L2/Linf error of the FP32 run against the FP64 run and of
both against the analytic convected vortex.

The same run measures the shock-capturing path (PerssonC0) on a Mach 2
wedge-channel mesh of about 250k triangles at P=2 (FP64, 2000 steps) under
"shock_path", and a batched ensemble of 64 independent cases (2k..40k
triangles, orders 1..4, FP32, 1000 steps) under "ensemble", and the adjoint
workload (a 100k-triangle airfoil O-mesh at P=3, FP64) under "airfoil": the
forward step, and the lift gradients of one discrete-adjoint sweep against
central finite differences of the solver.

python bench.py [--gpus 1] [--steps 500] [--warmup 10] [--no-cpu] [--no-shock]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "DOF-stage/s at P=4 (FP32 vs FP64 GPU, L2 err vs FP64 oracle); % HBM peak"


def alg_bytes_per_elem_stage(np_int: int, state_bytes: int) -> float:
    """Algorithmic HBM bytes per element per SSP54 stage: 19 state-array
    accesses of 4*NpInt values per step -- 6 of q0 and 3 of R, which are FP64
    in every precision, and 10 of the stage states q1..q4 (state_bytes each)
    -- divided over the 5 stages, plus 144 B of per-element geometry,
    connectivity and face codes. FP64: 30.4 B per value and stage; FP32
    stage states: 22.4 B."""
    per_value = (6 * 8 + 3 * 8 + 10 * state_bytes) / 5.0
    return 4 * np_int * per_value + 144.0


def analytic_density(t, X, Y):
    """Isentropic vortex density (beta=5, centre (5,0), gamma 1.4, u=1)."""
    g, beta = 1.4, 5.0
    r2 = (X - t - 5.0) ** 2 + Y ** 2
    tv1 = 1.0 - (g - 1) * beta * beta * np.exp(2.0 * (1.0 - r2)) / (16 * g * math.pi ** 2)
    return tv1 ** (1.0 / (g - 1))


def run_gpu(mesh, precision, steps, warmup):
    from gocfd_amd import Euler, InputParameters2D
    ip = InputParameters2D(Title="bench", CFL=1.0, FluxType="Roe", InitType="IVortex",
                           PolynomialOrder=4, FinalTime=1e9, MaxIterations=warmup + steps)
    e = Euler(ip, mesh, precision=precision)
    e.step(warmup)
    # timed region: inputs already resident in HBM; synchronised on both sides
    e.sync()
    t0 = time.perf_counter()
    e.step(steps)  # enqueues 5*steps stage kernels and waits for them
    elapsed = time.perf_counter() - t0
    stage_ms, launches = e.kernel_time()
    t, dt, nst = e.time_info()
    q = e.Q
    c = e.coords()
    return dict(e=e, elapsed=elapsed, stage_ms=stage_ms, launches=launches, time=t, dt=dt,
                steps_done=nst, rho=q[0], X=c["SolX"], Y=c["SolY"], K=e.K, NpInt=e.NpInt,
                tiling=e.tiling())


def shock_alg_bytes_per_elem_stage(np_int: int, np_edge: int, state_bytes: int = 8, diss_bytes: int = 8,
                                   active: float = 1.0) -> float:
    """Algorithmic HBM bytes per element per stage of the shock path: the
    plain stage (alg_bytes_per_elem_stage) plus the sensor (density of the
    stage input in, Se/eps out, 3 vertex maxima x 2), the element fields (3
    vertex pairs in, eps at the NR used RT points and sigma out), the
    gradient pass (stage input, eps, 144 B geometry in; eps*grad at
    2 x 4 x NR points out, diss_bytes each) and the stage kernel's reads of
    eps and eps*grad; NR = NpInt + 3 NpEdge. The stage input is q0 (FP64) at
    one stage of five and a stage state (state_bytes) at the other four.
    eps at the RT points and eps*grad are written and read only for active
    elements (a vertex carries eps > 0; elsewhere both are identically zero):
    those terms scale with the active fraction."""
    nr = np_int + 3 * np_edge
    nv = 4 * np_int
    sin = (8 + 4 * state_bytes) / 5.0
    plain = alg_bytes_per_elem_stage(np_int, state_bytes)
    sensor = np_int * sin + 16 + 48
    prep = 48 + nr * 8 * active + 8
    grad = nv * sin + nr * 8 + 144 + 8 * nr * diss_bytes * active
    stage_extra = (nr * 8 + 8 * nr * diss_bytes) * active + 8
    return plain + sensor + prep + grad + stage_extra


SHOCK_KAPPA = 10.0


def run_gpu_shock(mesh, steps, warmup, precision):
    """The shock-capturing path: Mach 2 wedge channel, P=2, Roe, PerssonC0,
    uniform freestream start. Kappa 10: with the reference wedge's 6 the
    sensor flags no element in these steps (the RHS modal filter, on with
    PerssonC0 whatever sigma is, keeps the top-mode energy ratio near 1e-7.5
    while S0 - Kappa = -5.95), so the dissipation would stay zero."""
    from gocfd_amd import Euler, InputParameters2D, Mesh
    ip = InputParameters2D(Title="wedge", CFL=1.0, FluxType="Roe", InitType="Freestream",
                           PolynomialOrder=2, Minf=2.0, FinalTime=1e9, MaxIterations=warmup + steps,
                           Limiter="PerssonC0", Kappa=SHOCK_KAPPA)
    e = Euler(ip, mesh, precision=precision)
    EToV = Mesh.read(mesh).arrays()[2]

    def active_fraction():
        # elements with a vertex whose eps maximum is positive
        _, _, eps = e.shock()
        vmax = np.zeros(int(EToV.max()) + 1)
        np.maximum.at(vmax, EToV.ravel(), np.repeat(eps, 3))
        return float((vmax[EToV] > 0).any(axis=1).mean())

    e.step(warmup)
    # two timed halves; the active fraction is sampled between them (outside
    # the timed region) and at the end
    elapsed, ms_sum, act = 0.0, 0.0, []
    for n in (steps // 2, steps - steps // 2):
        e.sync()
        t0 = time.perf_counter()
        e.step(n)
        elapsed += time.perf_counter() - t0
        ms, launches = e.kernel_time()
        ms_sum += ms * launches
        act.append(active_fraction())
    stage_ms = ms_sum / (5 * steps)
    t, dt, nst = e.time_info()
    se, sigma, eps = e.shock()
    return dict(elapsed=elapsed, stage_ms=stage_ms, time=t, dt=dt, K=e.K, NpInt=e.NpInt,
                NpEdge=e.NpEdge, flagged=int((sigma > 0).sum()), sigma_max=float(sigma.max()),
                rho=e.Q[0], flags=sigma > 0, active=float(np.mean(act)))


def wedge_flags_vs_oracle(mesh):
    """The FP64 wedge from the freestream for 2000 steps: the sensor flags
    after 10, 500, 1000 and 2000 steps against the oracle's, bit for bit
    (tests/golden/wedge_flags_2000.npz; the oracle takes about 90 minutes)."""
    from gocfd_amd import Euler, InputParameters2D
    gold = np.load(os.path.join(ROOT, "tests", "golden", "wedge_flags_2000.npz"))
    ip = InputParameters2D(Title="wedge", CFL=1.0, FluxType="Roe", InitType="Freestream", PolynomialOrder=2,
                           Minf=2.0, FinalTime=1e9, Limiter="PerssonC0", Kappa=SHOCK_KAPPA)
    e = Euler(ip, mesh, precision="fp64")
    K = int(gold["K"])
    out = {"checkpoints": [], "flagged": [], "flags_identical": [], "time_rel": [], "dt_rel": []}
    if e.K != K:
        return {"error": f"mesh has {e.K} triangles, the oracle fixture {K}"}
    done = 0
    for n in (int(x) for x in gold["checkpoints"]):
        e.step(n - done)
        done = n
        t, dt, _ = e.time_info()
        _, sg, _ = e.shock()
        want = np.unpackbits(gold[f"flags_{n}"])[:K].astype(bool)
        out["checkpoints"].append(n)
        out["flagged"].append(int((sg > 0).sum()))
        out["flags_identical"].append(bool(np.array_equal(sg > 0, want)))
        out["time_rel"].append(abs(t - float(gold[f"time_{n}"])) / t)
        out["dt_rel"].append(abs(dt - float(gold[f"dt_{n}"])) / dt)
    rho = e.Q[0]
    out["rho_sum_rel_2000"] = abs(rho.sum() - float(gold["rho_sum_2000"])) / abs(float(gold["rho_sum_2000"]))
    return out


def ensemble_cases(ncases: int, seed: int = 7):
    """The batched workload: ncases independent cases, K log-uniform in
    [2k, 40k] triangles (jittered quads, n = sqrt(K/2) per side), order drawn
    from {1,2,3,4}, Mach from U[0.3,0.8], angle from U[-5,5] deg; even cases
    start from the isentropic vortex, odd ones from the freestream. Roe flux,
    CFL 1."""
    from gocfd_amd import InputParameters2D, Mesh, meshgen
    u = meshgen.splitmix64(seed, 4 * ncases).reshape(ncases, 4)
    cases = []
    for c in range(ncases):
        K = 2000.0 * (20.0 ** u[c, 0])
        n = max(4, int(round(math.sqrt(K / 2))))
        N = 1 + int(u[c, 1] * 4) % 4
        M = 0.3 + 0.5 * u[c, 2]
        a = -5.0 + 10.0 * u[c, 3]
        VX, VY, EToV, tags = meshgen.jittered_quads(n, n, seed=1000 + c)
        ip = InputParameters2D(Title=f"case{c}", CFL=1.0, FluxType="Roe",
                               InitType="IVortex" if c % 2 == 0 else "Freestream",
                               PolynomialOrder=N, Minf=M, Alpha=a, FinalTime=1e9)
        cases.append((ip, Mesh.from_arrays(VX, VY, EToV, tags)))
    return cases


def run_ensemble(ncases, steps, warmup):
    """The ensemble in FP32 (one launch per order group and stage, graph
    replays), the same cases run one after another as single solvers, and
    every case of the FP32 ensemble against its own single-case FP64 run."""
    from gocfd_amd import Ensemble, Euler
    cases = ensemble_cases(ncases)
    ens = Ensemble(cases, precision="fp32")
    # autotuned launch configuration (results do not depend on it)
    t_tune = time.perf_counter()
    tuned = ens.tune(2)
    t_tune = time.perf_counter() - t_tune
    ens.step(warmup)
    t0 = time.perf_counter()
    ens.step(steps)
    elapsed = time.perf_counter() - t0
    ms, _ = ens.kernel_time()
    dof = sum(K * Np for K, Np in ens.dims)
    # sequential single-case runs of the same cases (a bounded sample of steps)
    seq_steps = max(6, steps // 20)
    seq = 0.0
    for ip, m in cases:
        e = Euler(ip, m, precision="fp32")
        e.step(1)
        e.sync()
        t1 = time.perf_counter()
        e.step(seq_steps)
        seq += time.perf_counter() - t1
    seq_rate = dof * 5 * seq_steps / seq
    # every case against its own single-case FP64 run of the same steps
    per_case = []
    for c, (ip, m) in enumerate(cases):
        e = Euler(ip, m, precision="fp64")
        e.step(warmup + steps)
        q64 = e.Q
        q32 = ens.state(c)
        d = q32[0] - q64[0]
        per_case.append({"case": c, "N": ip.PolynomialOrder, "K": ens.dims[c][0],
                         "rho_l2": float(np.linalg.norm(d) / np.linalg.norm(q64[0])),
                         "rho_linf": float(np.abs(d).max() / np.abs(q64[0]).max())})
        del e
    errs = {"cases_checked": len(per_case), "rho_l2_max": max(r["rho_l2"] for r in per_case),
            "rho_linf_max": max(r["rho_linf"] for r in per_case),
            "rho_l2_max_by_order": {n: max(r["rho_l2"] for r in per_case if r["N"] == n)
                                    for n in sorted({r["N"] for r in per_case})},
            "per_case": per_case}
    orders = [ip.PolynomialOrder for ip, _ in cases]
    return dict(value=dof * 5 * steps / elapsed, elapsed=elapsed, ms=ms, dof=dof,
                K_total=sum(K for K, _ in ens.dims), K_min=min(K for K, _ in ens.dims),
                K_max=max(K for K, _ in ens.dims), orders={n: orders.count(n) for n in sorted(set(orders))},
                seq_rate=seq_rate, errs=errs, tuned=tuned, tune_seconds=t_tune)


def run_airfoil(mesh, steps, warmup, h=1e-6, ndir=16, seed=3, developed=20, window=200, every=20):
    """The adjoint workload: 4-digit airfoil in an O-domain (Delaunay mesh of
    50k wall-clustered points), P=3, M=0.5, AoA 2 deg, Roe, FP64. Times the
    forward step (warmup + steps); computes the lift functional's gradients
    over a `window`-step run by one discrete-adjoint sweep (Euler.vjp,
    checkpoint every `every` steps) with respect to Mach, angle of attack
    (degrees) and the initial state, and checks them against central finite
    differences (step h) of the solver: in Mach and angle from the freestream
    start, and in `ndir` seeded random state directions from a state
    `developed` steps into the run (at the uniform start every wave-speed
    candidate of the dt maximum ties, so the state derivative is one-sided
    there). Also the tangent-linear run's dot product against the adjoint."""
    from gocfd_amd import Euler, InputParameters2D, adjoint

    def make(M, a):
        ip = InputParameters2D(Title="airfoil", CFL=1.0, FluxType="Roe", InitType="Freestream",
                               PolynomialOrder=3, Minf=M, Alpha=a, FinalTime=1e9,
                               MaxIterations=warmup + steps + developed + window)
        return Euler(ip, mesh, precision="fp64")

    M0, a0 = 0.5, 2.0
    e = make(M0, a0)
    e.step(warmup)
    e.sync()
    t0 = time.perf_counter()
    e.step(steps)
    elapsed = time.perf_counter() - t0
    stage_ms, _ = e.kernel_time()
    t_end, dt, _ = e.time_info()
    n = window

    def J_of(M, a):
        g = make(M, a)
        g.step(n)
        return adjoint.lift(g, g.Q, a)

    # adjoint from the freestream start: Mach and angle
    e0 = make(M0, a0)
    ta = time.perf_counter()
    ga = adjoint.lift_gradients(e0, n, every)
    grad_seconds = time.perf_counter() - ta
    seed_n = adjoint.lift_state_gradient(e0, e0.Q, a0)
    tv = time.perf_counter()
    e0.vjp(n, seed_n, every)
    vjp_seconds = time.perf_counter() - tv
    t1 = time.perf_counter()
    fdM = (J_of(M0 + h, a0) - J_of(M0 - h, a0)) / (2 * h)
    fda = (J_of(M0, a0 + h) - J_of(M0, a0 - h)) / (2 * h)
    # state directions from a developed state; one solver re-run from each
    # perturbed state (set_state / set_time restart the step sequence)
    ed = make(M0, a0)
    ed.step(developed)
    qd = ed.Q
    ed.set_time(0.0)
    gd = adjoint.lift_gradients(ed, n, every)
    rng = np.random.default_rng(seed)
    fd_run = make(M0, a0)

    def J_from(q0):
        fd_run.Q = q0
        fd_run.set_time(0.0)
        fd_run.step(n)
        return adjoint.lift(fd_run, fd_run.Q, a0)

    dirs = []
    for _ in range(ndir):
        v = rng.standard_normal(qd.shape) * np.abs(qd)
        fd = (J_from(qd + h * v) - J_from(qd - h * v)) / (2 * h)
        ad = float(np.sum(gd["dJ_dQ0"] * v))
        dirs.append({"adjoint": ad, "fd": fd, "rel_err": abs(ad - fd) / abs(fd)})
    fd_seconds = time.perf_counter() - t1
    # tangent-linear dot product at full size: <dJ/dQn, J v> = <J^T dJ/dQn, v>
    v = rng.standard_normal(qd.shape) * np.abs(qd)
    tj = time.perf_counter()
    qn, jv = ed.jvp(n, v)
    jvp_seconds = time.perf_counter() - tj
    tlm = float(np.sum(adjoint.lift_state_gradient(ed, qn, a0) * jv))
    adj = float(np.sum(gd["dJ_dQ0"] * v))
    errs = [d["rel_err"] for d in dirs]
    return dict(elapsed=elapsed, stage_ms=stage_ms, K=e.K, NpInt=e.NpInt, J=ga["J"], time=t_end, dt=dt, steps=n,
                adjoint={"window_steps": n, "checkpoint_every": every, "h": h,
                         "dJ_dMinf": ga["dJ_dMinf"], "fd_dJ_dMinf": fdM,
                         "rel_err_Minf": abs(ga["dJ_dMinf"] - fdM) / abs(fdM),
                         "dJ_dalpha_deg": ga["dJ_dalpha_deg"], "fd_dJ_dalpha_deg": fda,
                         "rel_err_alpha": abs(ga["dJ_dalpha_deg"] - fda) / abs(fda),
                         "state_dirs": len(dirs), "state_dir_rel_err_max": max(errs),
                         "state_dir_rel_err_median": float(np.median(errs)),
                         "state_dirs_from_developed_state": dirs,
                         "tangent_linear_dot_rel": abs(tlm - adj) / abs(tlm),
                         "vjp_seconds": vjp_seconds, "vjp_ms_per_step": vjp_seconds / n * 1e3,
                         "vjp_over_forward_step": (vjp_seconds / n) / (elapsed / steps),
                         "jvp_ms_per_step": jvp_seconds / n * 1e3,
                         "gradient_seconds_with_primal": grad_seconds, "fd_seconds": fd_seconds})


def host_cpu():
    """Model name of the host CPU (/proc/cpuinfo, what lscpu prints) and its
    logical CPU count."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(mesh, target_seconds=15.0, nan_scan=2, max_steps=50):
    """The CPU oracle (plain C, FP64, OpenMP over the host cores) on the same mesh and order:
    whole SSP54 steps until about `target_seconds` of CPU work. nan_scan 2 scans
    the reference's eleven arrays for NaN before every stage (euler.go:471-484,
    on in the reference), 0 none."""
    import oracle
    s = oracle.OracleSolver(mesh, 4, flux="roe", init="ivortex", CFL=1.0, FinalTime=1e9)
    s.set_nan_scan(nan_scan)
    n, t0 = 0, time.perf_counter()
    while True:
        s.step()
        n += 1
        el = time.perf_counter() - t0
        if el >= target_seconds or n >= max_steps:
            break
    return dict(value=s.K * s.NpInt * 5 * n / el, steps=n, seconds=el, K=s.K, threads=oracle.threads(),
                rho=s.state()[0])


def gpu_vs_oracle(mesh, steps, rho_oracle):
    """FP64 and FP32 GPU runs of the same number of steps from the same
    initial state, against the oracle's density (relative L2 and Linf)."""
    from gocfd_amd import Euler, InputParameters2D
    ip = InputParameters2D(Title="bench", CFL=1.0, FluxType="Roe", InitType="IVortex",
                           PolynomialOrder=4, FinalTime=1e9, MaxIterations=steps)
    out = {"steps": steps}
    for prec in ("fp64", "fp32"):
        e = Euler(ip, mesh, precision=prec)
        e.step(steps)
        d = e.Q[0] - rho_oracle
        out[prec + "_rho_l2"] = float(np.linalg.norm(d) / np.linalg.norm(rho_oracle))
        out[prec + "_rho_linf"] = float(np.abs(d).max() / np.abs(rho_oracle).max())
    return out


def copy_probe(mb: float) -> None:
    """A plain device-to-device copy of `mb` MB with a torch elementwise copy
    kernel, run 5 times: the reference point that scales the HBM byte counters."""
    import torch
    n = int(mb * 1e6 / 4)
    a = torch.rand(n, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=200, help="quads per side (2n^2 triangles)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--only", choices=["fp32", "fp64", "shock", "ensemble", "airfoil"], default=None,
                    help="run one leg only (for profiler runs)")
    ap.add_argument("--shock-points", type=int, default=125_000,
                    help="interior points of the wedge mesh (about twice as many triangles)")
    ap.add_argument("--shock-steps", type=int, default=2000)
    ap.add_argument("--no-shock", action="store_true", help="skip the shock-capturing leg")
    ap.add_argument("--ensemble-cases", type=int, default=64)
    ap.add_argument("--ensemble-steps", type=int, default=1000)
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--no-airfoil", action="store_true", help="skip the airfoil (adjoint-workload forward) leg")
    ap.add_argument("--airfoil-steps", type=int, default=200)
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="MB",
                    help="run a plain device copy of MB megabytes (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported; the step "
              "would shard by element partition with a per-stage face-trace halo exchange and "
              "an 8-byte max all-reduce for dt, which this one-GPU build does not implement",
              file=sys.stderr)
        sys.exit(2)

    from gocfd_amd import meshgen
    tmpdir = tempfile.mkdtemp(prefix="gocfd_bench_")
    mesh = meshgen.vortex_mesh(os.path.join(tmpdir, f"vortex{args.n}.su2"), args.n, seed=1)

    wedge = None
    if args.only == "shock" or (args.only is None and not args.no_shock):
        wedge = meshgen.wedge_mesh(os.path.join(tmpdir, "wedge.su2"), args.shock_points, seed=1)
    if args.only == "shock":
        out = {"leg": "shock"}
        for prec in ("fp64", "fp32"):
            r = run_gpu_shock(wedge, args.shock_steps, args.warmup, prec)
            out[prec] = {"value": r["K"] * r["NpInt"] * 5 * args.shock_steps / r["elapsed"],
                         "stage_ms": r["stage_ms"], "ms_per_step": r["elapsed"] / args.shock_steps * 1e3,
                         "K": r["K"], "sensor_flagged_elements": r["flagged"]}
        print(json.dumps(out))
        return
    airfoil_mesh = None
    if args.only == "airfoil" or (args.only is None and not args.no_airfoil):
        airfoil_mesh = meshgen.airfoil_delaunay_mesh(os.path.join(tmpdir, "airfoil.su2"), seed=1)
    if args.only == "airfoil":
        r = run_airfoil(airfoil_mesh, args.airfoil_steps, args.warmup)
        r["value"] = r["K"] * r["NpInt"] * 5 * args.airfoil_steps / r["elapsed"]
        print(json.dumps({"leg": "airfoil", **r}))
        return
    if args.only == "ensemble":
        r = run_ensemble(args.ensemble_cases, args.ensemble_steps, args.warmup)
        print(json.dumps({"leg": "ensemble", "value": r["value"], "ms_per_step": r["elapsed"] / args.ensemble_steps * 1e3,
                          "dof": r["dof"], "seq_rate": r["seq_rate"], "errs": r["errs"], "tuned": r["tuned"]}))
        return
    if args.only:
        r = run_gpu(mesh, args.only, args.steps, args.warmup)
        v = r["K"] * r["NpInt"] * 5 * args.steps / r["elapsed"]
        print(json.dumps({"precision": args.only, "value": v, "stage_ms": r["stage_ms"],
                          "ms_per_step": r["elapsed"] / args.steps * 1e3, "K": r["K"]}))
        return
    r64 = run_gpu(mesh, "fp64", args.steps, args.warmup)
    r32 = run_gpu(mesh, "fp32", args.steps, args.warmup)
    K, NpInt = r32["K"], r32["NpInt"]
    dof_stages = K * NpInt * 5 * args.steps
    v32 = dof_stages / r32["elapsed"]
    v64 = dof_stages / r64["elapsed"]

    # accuracy: FP32 vs FP64 GPU (FP64 GPU is parity-tested against the CPU
    # oracle), both vs the analytic convected vortex
    rho32, rho64 = r32["rho"], r64["rho"]
    l2_32_64 = float(np.linalg.norm(rho32 - rho64) / np.linalg.norm(rho64))
    linf_32_64 = float(np.abs(rho32 - rho64).max())
    exact = analytic_density(r64["time"], r64["X"], r64["Y"])
    l2_64_an = float(np.linalg.norm(rho64 - exact) / np.linalg.norm(exact))
    l2_32_an = float(np.linalg.norm(rho32 - exact) / np.linalg.norm(exact))

    b32 = alg_bytes_per_elem_stage(NpInt, 4) * K
    b64 = alg_bytes_per_elem_stage(NpInt, 8) * K
    ach32 = b32 / (r32["stage_ms"] * 1e-3) / 1e9
    ach64 = b64 / (r64["stage_ms"] * 1e-3) / 1e9

    # HBM bytes per launch of the stage kernel at this workload, from the
    # committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (calibrated by a copy
    # of known size); None when the summary is absent
    traffic = traffic64 = None
    traffic_src = "profiles/stage_fp32_p4_k80000_traffic.json"
    tfile = os.path.join(ROOT, traffic_src)
    if os.path.exists(tfile) and args.n == 200:
        with open(tfile) as f:
            tj = json.load(f)
        traffic, traffic64 = tj.get("bytes_per_launch"), tj.get("fp64_bytes_per_launch")
        traffic_src += f" (from {tj.get('source', '?')})"
    else:
        traffic_src = None

    shock = None
    if wedge is not None:
        legs = {}
        for prec, sb in (("fp64", 8), ("fp32", 4)):
            rs = run_gpu_shock(wedge, args.shock_steps, args.warmup, prec)
            bs = shock_alg_bytes_per_elem_stage(rs["NpInt"], rs["NpEdge"], sb, sb, rs["active"]) * rs["K"]
            legs[prec] = rs
            legs[prec + "_out"] = {
                "value": rs["K"] * rs["NpInt"] * 5 * args.shock_steps / rs["elapsed"], "unit": "DOF-stage/s",
                "ms_per_step": rs["elapsed"] / args.shock_steps * 1e3,
                "stage_ms": rs["stage_ms"], "alg_bytes_per_stage": bs,
                "achieved_GBs": bs / (rs["stage_ms"] * 1e-3) / 1e9,
                "frac_hbm_peak": bs / (rs["stage_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "time": rs["time"], "dt": rs["dt"], "sensor_flagged_elements": rs["flagged"],
                "sigma_max": rs["sigma_max"], "active_element_fraction": rs["active"]}
        r64s, r32s = legs["fp64"], legs["fp32"]
        d = r32s["rho"] - r64s["rho"]
        shock = {"workload": f"Mach 2 wedge channel 3x1, 15 deg ramp, {r64s['K']} triangles (Delaunay of a "
                             f"jittered lattice), P=2, Roe, PerssonC0 Kappa {SHOCK_KAPPA:g}, "
                             f"{args.shock_steps} steps, FP64 and FP32 bulk",
                 "fp64": legs["fp64_out"], "fp32": legs["fp32_out"],
                 "fp32_over_fp64": r64s["elapsed"] / r32s["elapsed"],
                 "rho_l2_fp32_vs_fp64": float(np.linalg.norm(d) / np.linalg.norm(r64s["rho"])),
                 "rho_linf_fp32_vs_fp64": float(np.abs(d).max() / np.abs(r64s["rho"]).max()),
                 "flags_agree_fp32_vs_fp64": float(np.mean(r32s["flags"] == r64s["flags"]))}
        if SHOCK_KAPPA == 10.0 and args.shock_points == 125_000:
            shock["fp64_flags_vs_oracle"] = wedge_flags_vs_oracle(wedge)

    ensemble = None
    if not args.no_ensemble:
        re_ = run_ensemble(args.ensemble_cases, args.ensemble_steps, args.warmup)
        ensemble = {"workload": f"{args.ensemble_cases} independent cases, K {re_['K_min']}..{re_['K_max']} "
                                f"(total {re_['K_total']}), orders {re_['orders']}, Mach U[0.3,0.8], "
                                f"AoA U[-5,5] deg, Roe, FP32 bulk, {args.ensemble_steps} steps",
                    "value": re_["value"], "unit": "DOF-stage/s",
                    "ms_per_step": re_["elapsed"] / args.ensemble_steps * 1e3,
                    "sequential_single_case_value": re_["seq_rate"],
                    "speedup_vs_sequential": re_["value"] / re_["seq_rate"],
                    "launch_autotuned": re_["tuned"], "autotune_seconds": re_["tune_seconds"],
                    "fp32_vs_own_single_case_fp64": re_["errs"]}

    airfoil = None
    if airfoil_mesh is not None:
        ra = run_airfoil(airfoil_mesh, args.airfoil_steps, args.warmup)
        airfoil = {"workload": f"symmetric 4-digit airfoil (12 %, chord 1) in an O-domain of radius 20, {ra['K']} "
                               f"triangles (Delaunay of 50,400 seeded wall-clustered points), P=3, M=0.5, AoA 2 deg, "
                               f"Roe, FP64, {args.airfoil_steps} timed steps; gradients over 200 steps, "
                               f"checkpoints every 20, 16 random state directions",
                   "value": ra["K"] * ra["NpInt"] * 5 * args.airfoil_steps / ra["elapsed"], "unit": "DOF-stage/s",
                   "ms_per_step": ra["elapsed"] / args.airfoil_steps * 1e3, "stage_ms": ra["stage_ms"],
                   "time": ra["time"], "dt": ra["dt"], "lift_functional": ra["J"],
                   "adjoint_vs_central_fd_h1e-6": ra["adjoint"]}

    cpu = None
    vs_oracle = None
    if not args.no_cpu:
        model, ncpu = host_cpu()
        c = cpu_baseline(mesh, args.cpu_seconds, nan_scan=2)
        vs_oracle = gpu_vs_oracle(mesh, c["steps"], c["rho"])
        c0 = cpu_baseline(mesh, args.cpu_seconds, nan_scan=0, max_steps=c["steps"])
        cpu = {"value": c["value"], "unit": "DOF-stage/s", "cores": c["threads"], "kind": "port",
               "sample": f"CPU oracle (plain C, FP64, OpenMP {c['threads']} threads), same {K}-triangle P=4 Roe "
                         f"mesh, {c['steps']} SSP54 steps in {c['seconds']:.1f} s, with the reference's "
                         f"per-stage NaN scan of eleven arrays (euler.go:471-484)",
               "host_cpu": model, "host_logical_cpus": ncpu,
               "no_nan_scan": {"value": c0["value"], "steps": c0["steps"], "seconds": c0["seconds"],
                               "cores": c0["threads"]}}

    line = {
        "metric": METRIC,
        "value": v32,
        "unit": "DOF-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["elapsed"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "dtype_detail": "f32 stage states q1..q4 and bulk flux/divergence arithmetic; f64 geometry, dt "
                        "reduction, SSP sums and accumulated state q0/R",
        "data": "synthetic: deterministic jittered-quad isentropic-vortex mesh",
        "config": {"workload": f"isentropic vortex, {K} triangles ({args.n}x{args.n} jittered quads), "
                               f"P=4, Roe, SSP54, CFL 1", "K": K, "NpInt": NpInt,
                   "dof": K * NpInt, "precision": "fp32-bulk", "tile": r32["tiling"][0]},
        "fp64": {"value": v64, "ms_per_step": r64["elapsed"] / args.steps * 1e3,
                 "stage_ms": r64["stage_ms"],
                 "roofline": {"bound": "hbm", "achieved": ach64, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": ach64 / HBM_PEAK_GBS, "traffic": traffic64,
                              "kernel": "k_stage<4,double,double>", "avg_launch_ms": r64["stage_ms"],
                              "alg_bytes_per_launch": b64}},
        "fp32_over_fp64": v32 / v64,
        "accuracy": {"rho_l2_fp32_vs_fp64": l2_32_64, "rho_linf_fp32_vs_fp64": linf_32_64,
                     "rho_l2_fp64_vs_analytic": l2_64_an, "rho_l2_fp32_vs_analytic": l2_32_an,
                     "time": r64["time"], "vs_oracle": vs_oracle},
        "roofline": {"bound": "hbm", "achieved": ach32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach32 / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "k_stage<4,float,float>", "avg_launch_ms": r32["stage_ms"],
                     "alg_bytes_per_launch": b32},
        "cpu_baseline": cpu,
        "shock_path": shock,
        "airfoil": airfoil,
        "ensemble": ensemble,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
