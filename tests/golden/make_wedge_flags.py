"""Oracle sensor flags of the Mach 2 wedge channel (P=2, Roe, PerssonC0
Kappa 10, FP64, uniform freestream start) after 10, 500, 1000 and 2000 steps:
wedge_flags_2000.npz holds, per checkpoint, the packed bits of sigma > 0 per
element (reference element order), the time and the last dt, and the
density's sum and max at 2000 steps.

The mesh is meshgen.wedge_mesh(125_000 interior points, seed 1), about 249k
triangles; the oracle runs the 2000 steps in about 90 minutes on 8 threads.

    python tests/golden/make_wedge_flags.py
"""
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

CHECKPOINTS = (10, 500, 1000, 2000)
POINTS, SEED, KAPPA, MINF = 125_000, 1, 10.0, 2.0


def pack(results):
    """results: {steps: (sigma, time, dt)} and 'rho' (density at the last
    checkpoint) -> the fixture's arrays"""
    out = {"checkpoints": np.array(CHECKPOINTS), "K": np.int64(len(results[CHECKPOINTS[0]][0]))}
    for n in CHECKPOINTS:
        sg, t, dt = results[n]
        out[f"flags_{n}"] = np.packbits(np.asarray(sg) > 0)
        out[f"time_{n}"] = np.float64(t)
        out[f"dt_{n}"] = np.float64(dt)
    rho = np.asarray(results["rho"])
    out["rho_sum_2000"] = np.float64(rho.sum())
    out["rho_max_2000"] = np.float64(rho.max())
    return out


def main():
    import oracle
    from gocfd_amd import meshgen
    d = tempfile.mkdtemp()
    mesh = meshgen.wedge_mesh(os.path.join(d, "wedge.su2"), POINTS, seed=SEED)
    o = oracle.OracleSolver(mesh, 2, flux="roe", init="freestream", CFL=1.0, Minf=MINF, FinalTime=1e9,
                            limiter="PerssonC0", Kappa=KAPPA)
    res, t0 = {}, time.perf_counter()
    for i in range(1, CHECKPOINTS[-1] + 1):
        o.step()
        if i in CHECKPOINTS:
            _, sg, _ = o.shock()
            res[i] = (sg, o.time, o.dt)
            print(i, o.time, int((sg > 0).sum()), f"{time.perf_counter() - t0:.0f} s", flush=True)
    res["rho"] = o.state()[0]
    np.savez_compressed(os.path.join(HERE, "wedge_flags_2000.npz"), **pack(res))


if __name__ == "__main__":
    main()
