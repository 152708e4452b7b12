"""Host-side mirror of the reference Euler2D solver API over the C ABI.

Reference: model_problems/Euler2D/euler.go — NewEuler (:61), (*Euler).Solve
(:138), (*RungeKutta5SSP).Step (:408), CheckIfFinished (:796), PrintUpdate
(:814) and PrintFinal (:851). All arithmetic runs in libgocfd_euler2d.so on
the GPU; this module only moves parameters, meshes and state across the
boundary and keeps the reference's control flow.
"""
from __future__ import annotations

import ctypes as C
import time as _time
from typing import Optional

import numpy as np

from . import _lib
from .input_parameters import InputParameters2D

FLUX_TYPES = {"average": 0, "lax": 1, "roe": 2, "roe-er": 3}   # fluxes.go:25-33
INIT_TYPES = {"freestream": 0, "ivortex": 1, "shocktube": 2}    # initialization.go:24-31
LIMITERS = {"": 0, "perssonc0": 1, "persson c0": 1}              # filter.go:15-23
PRECISIONS = {"fp64": 0, "fp32": 1}


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i32(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Mesh:
    """A triangle mesh as returned by readfiles.ReadSU2 / ReadGambit2d."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def read(cls, path: str) -> "Mesh":
        L = _lib.load()
        h = C.c_void_p()
        _lib.check(L.gocfd_mesh_read(path.encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, VX, VY, EToV, tags: dict) -> "Mesh":
        """tags: {"far": [(v0, v1), ...], "periodic-1": [...], ...} in file order."""
        L = _lib.load()
        VX = np.ascontiguousarray(VX, dtype=np.float64)
        VY = np.ascontiguousarray(VY, dtype=np.float64)
        EToV = np.ascontiguousarray(EToV, dtype=np.int32)
        names = list(tags.keys())
        counts = np.array([len(tags[n]) for n in names], dtype=np.int32)
        offsets = np.zeros(len(names), dtype=np.int32)
        if len(names):
            offsets[1:] = np.cumsum(counts)[:-1]
        verts = np.ascontiguousarray(
            np.concatenate([np.asarray(tags[n], dtype=np.int32).reshape(-1, 2) for n in names])
            if names else np.zeros((0, 2), np.int32), dtype=np.int32)
        arr = (C.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
        h = C.c_void_p()
        _lib.check(L.gocfd_mesh_from_arrays(EToV.shape[0], VX.shape[0], _dp(VX), _dp(VY), _i32(EToV),
                                            len(names), arr, _i32(offsets), _i32(counts), _i32(verts),
                                            C.byref(h)))
        return cls(h)

    def dims(self):
        K, Nv, nt = C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(_lib.load().gocfd_mesh_dims(self._h, C.byref(K), C.byref(Nv), C.byref(nt)))
        return K.value, Nv.value, nt.value

    def arrays(self):
        K, Nv, _ = self.dims()
        VX, VY = np.empty(Nv), np.empty(Nv)
        EToV = np.empty((K, 3), dtype=np.int32)
        _lib.check(_lib.load().gocfd_mesh_arrays(self._h, _dp(VX), _dp(VY), _i32(EToV)))
        return VX, VY, EToV

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.load().gocfd_mesh_free(self._h)
            self._h = None


class RungeKutta5SSP:
    """Time-stepper view (euler.go:320-341): Step, Time, GlobalDT, StepCount, Residual."""

    def __init__(self, c: "Euler"):
        self._c = c

    def Step(self, c: Optional["Euler"] = None) -> None:
        """One SSP54 step (euler.go:408-418) plus Time += GlobalDT, StepCount++."""
        self._c.step(1)

    @property
    def Time(self) -> float:
        return self._c.time_info()[0]

    @property
    def GlobalDT(self) -> float:
        return self._c.time_info()[1]

    @property
    def StepCount(self) -> int:
        return self._c.time_info()[2]

    @property
    def Residual(self) -> np.ndarray:
        return self._c.residual()


def params_of(ip: InputParameters2D, precision: str = "fp64", device: int = 0) -> "_lib.Params":
    """The C-ABI parameter block of an InputParameters2D; unknown names raise
    as the reference panics (fluxes.go:18-33, initialization.go:18-31,
    filter.go:15-23)."""
    ft = ip.FluxType.lower()
    if ft not in FLUX_TYPES:
        raise ValueError(f"unable to use flux named {ft}")
    it = ip.InitType.lower()
    if not it:
        raise ValueError(f"empty init type, must be one of {list(INIT_TYPES)}")
    if it not in INIT_TYPES:
        raise ValueError(f"unable to use init type named {it}")
    lim = ip.Limiter.strip().lower()
    if lim not in LIMITERS:
        raise ValueError(f"unable to use limiter named [{lim}]")
    return _lib.Params(ip.PolynomialOrder, FLUX_TYPES[ft], INIT_TYPES[it], ip.MaxIterations,
                       ip.CFL, ip.FinalTime, ip.Minf, ip.Gamma, ip.Alpha,
                       int(ip.LocalTimeStepping), LIMITERS[lim], ip.Kappa,
                       PRECISIONS[precision], device)


class Euler:
    """The device-resident 2D Euler DFR solver (reference: euler.go:31-136)."""

    def __init__(self, ip: InputParameters2D, meshFile: str | Mesh, ProcLimit: int = 0,
                 verbose: bool = False, profile: bool = False, precision: str = "fp64",
                 device: int = 0):
        L = _lib.load()
        p = params_of(ip, precision, device)
        it = ip.InitType.lower()
        self.ip = ip
        self.MeshFile = meshFile if isinstance(meshFile, str) else "<arrays>"
        self.CFL, self.FinalTime, self.MaxIterations = ip.CFL, ip.FinalTime, ip.MaxIterations
        self.precision = precision
        mesh = Mesh.read(meshFile) if isinstance(meshFile, str) else meshFile
        self._mesh_xy = mesh.arrays()  # VX, VY, EToV for the output writers
        h = C.c_void_p()
        _lib.check(L.gocfd_euler2d_create(mesh._h, C.byref(p), C.byref(h)))
        self._h = h
        K, NpInt, NpEdge, NpFlux, ne = (C.c_int32() for _ in range(5))
        _lib.check(L.gocfd_euler2d_dims(h, C.byref(K), C.byref(NpInt), C.byref(NpEdge),
                                        C.byref(NpFlux), C.byref(ne)))
        self.K, self.NpInt, self.NpEdge, self.NpFlux, self.nedges = (
            K.value, NpInt.value, NpEdge.value, NpFlux.value, ne.value)
        self.N = ip.PolynomialOrder
        self.RK = RungeKutta5SSP(self)
        self.verbose = verbose
        if verbose:
            print("Euler Equations in 2 Dimensions")
            print(f"Solving {it}")
            print(f"CFL = {ip.CFL:8.4f}, Polynomial Degree N = {self.N} (1 is linear), "
                  f"Num Elements K = {self.K}\n\n")

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.load().gocfd_euler2d_destroy(self._h)
            self._h = None

    # ---- state -----------------------------------------------------------
    @property
    def Q(self) -> np.ndarray:
        q = np.empty((4, self.NpInt, self.K))
        _lib.check(_lib.load().gocfd_euler2d_get_state(self._h, _dp(q)))
        return q

    @Q.setter
    def Q(self, q) -> None:
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(4, self.NpInt, self.K)
        _lib.check(_lib.load().gocfd_euler2d_set_state(self._h, _dp(q)))

    def residual(self) -> np.ndarray:
        r = np.empty((4, self.NpInt, self.K))
        _lib.check(_lib.load().gocfd_euler2d_get_residual(self._h, _dp(r)))
        return r

    def rhs(self, q, t: float = 0.0) -> np.ndarray:
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(4, self.NpInt, self.K)
        r = np.empty_like(q)
        _lib.check(_lib.load().gocfd_euler2d_rhs(self._h, _dp(q), t, _dp(r)))
        return r

    # ---- stepping --------------------------------------------------------
    def step(self, nsteps: int = 1) -> None:
        rc = _lib.load().gocfd_euler2d_step(self._h, nsteps)
        if rc == _lib.ERR_NAN:
            raise FloatingPointError("NAN found")
        _lib.check(rc)

    def step_async(self, nsteps: int) -> None:
        _lib.check(_lib.load().gocfd_euler2d_step_async(self._h, nsteps))

    def sync(self) -> None:
        rc = _lib.load().gocfd_euler2d_sync(self._h)
        if rc == _lib.ERR_NAN:
            raise FloatingPointError("NAN found")
        _lib.check(rc)

    def time_info(self):
        t, dt, s = C.c_double(), C.c_double(), C.c_int32()
        _lib.check(_lib.load().gocfd_euler2d_time(self._h, C.byref(t), C.byref(dt), C.byref(s)))
        return t.value, dt.value, s.value

    def set_time(self, t: float) -> None:
        _lib.check(_lib.load().gocfd_euler2d_set_time(self._h, t))

    def shock(self):
        """Se, sigma, eps per element of the last stage (PerssonC0 only)."""
        se, sg, ep = np.empty(self.K), np.empty(self.K), np.empty(self.K)
        _lib.check(_lib.load().gocfd_euler2d_get_shock(self._h, _dp(se), _dp(sg), _dp(ep)))
        return se, sg, ep

    def element_dt(self):
        """Element time steps of the last stage (LocalTimeStepping only)."""
        d = np.empty(self.K)
        _lib.check(_lib.load().gocfd_euler2d_get_element_dt(self._h, _dp(d)))
        return d

    # ---- derivatives (FP64, global time stepping, no limiter) -------------
    def vjp(self, nsteps: int, dQn_bar, checkpoint_every: int = 0):
        """Discrete adjoint of `nsteps` steps from the current state:
        returns (dQ0, dQinf) = the transposed Jacobians of the state after
        the steps, with respect to the current state and the far-field state
        FSFar.Qinf, applied to dQn_bar. The solver is left unchanged."""
        lam = np.ascontiguousarray(dQn_bar, dtype=np.float64).reshape(4, self.NpInt, self.K)
        d0 = np.empty_like(lam)
        dfar = np.empty(4)
        _lib.check(_lib.load().gocfd_euler2d_vjp(self._h, nsteps, checkpoint_every, _dp(lam), _dp(d0),
                                                 _dp(dfar)))
        return d0, dfar

    def jvp(self, nsteps: int, dQ0, dQinf=None):
        """Tangent-linear run of `nsteps` steps from the current state:
        returns (Qn, dQn), the state after the steps and its directional
        derivative along (dQ0, dQinf). The solver is left unchanged."""
        v = np.ascontiguousarray(dQ0, dtype=np.float64).reshape(4, self.NpInt, self.K)
        qn, dqn = np.empty_like(v), np.empty_like(v)
        w = None if dQinf is None else np.ascontiguousarray(dQinf, dtype=np.float64).reshape(4)
        _lib.check(_lib.load().gocfd_euler2d_jvp(self._h, nsteps, _dp(v), None if w is None else _dp(w),
                                                 _dp(qn), _dp(dqn)))
        return qn, dqn

    def kernel_time(self):
        ms, n = C.c_double(), C.c_int32()
        _lib.check(_lib.load().gocfd_euler2d_kernel_time(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def tiling(self):
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(_lib.load().gocfd_euler2d_tiling(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def CheckIfFinished(self, Time: float, steps: int) -> bool:
        """euler.go:796-801"""
        return Time >= self.FinalTime or steps >= self.MaxIterations

    def Solve(self, output_dir: Optional[str] = ".") -> float:
        """euler.go:138-220: step until FinalTime or MaxIterations; returns the
        step-loop wall time in seconds (the scope the reference times). Then
        OutputFinal (euler.go:222-318) writes the case's files into
        output_dir (None: no files)."""
        self.PrintInitialization()
        steps, elapsed, finished = 0, 0.0, False
        while not finished:
            # run up to the next print point without host round trips; the
            # device stops by itself once Time >= FinalTime
            batch = 1 if steps == 0 else min(100 - steps % 100, max(1, self.MaxIterations - steps))
            t0 = _time.perf_counter()
            self.step(batch)
            elapsed += _time.perf_counter() - t0
            Time, dt, _ = self.time_info()
            steps += batch
            finished = self.CheckIfFinished(Time, steps)
            if finished or steps == 1 or steps % 100 == 0:
                self.PrintUpdate(Time, dt, steps)
        self.PrintFinal(elapsed, steps)
        if output_dir is not None:
            self.OutputFinal(output_dir)
        return elapsed

    def OutputFinal(self, output_dir: str = ".") -> list:
        """euler.go:222-318: shocktube.dat and shocktube_analytic.dat for the
        shock tube, plotfile.dat for freestream cases with PlotFields."""
        from . import output
        return output.output_final(self, output_dir)

    def PrintInitialization(self) -> None:
        print(f"Using mesh from file: [{self.MeshFile}]")
        print(f"Solving until finaltime = {self.FinalTime:8.5f}")
        print("    iter    time      dt       Res0       Res1       Res2       Res3         L1         L2")

    def PrintUpdate(self, Time: float, dt: float, steps: int) -> None:
        """euler.go:814-849"""
        R = self.residual()
        maxR = [float(R[n].max()) for n in range(4)]
        l1 = max(0.0, *maxR)
        l2 = float(np.sqrt(sum(m * m for m in maxR)) / 4.0)
        print(f"{steps:8d}{Time:8.5f}{dt:8.5f}" + "".join(f"{m:11.4e}" for m in maxR) + f"{l1:11.4e}{l2:11.4e}")

    def PrintFinal(self, elapsed: float, steps: int) -> None:
        rate = elapsed * 1e6 / (self.K * steps)
        print(f"\nRate of execution = {rate:8.5f} us/(element*iteration) over {steps} iterations")

    # ---- setup introspection --------------------------------------------
    def mesh_arrays(self):
        """VX [Nv], VY [Nv], EToV [K][3] of the mesh the solver was built on"""
        VX, VY, EToV = self._mesh_xy
        return VX, VY, np.asarray(EToV).reshape(-1, 3)

    def interp_matrix(self, r, s) -> np.ndarray:
        """The solution basis at reference points (r, s): [npts][NpInt]
        (GetInterpMatrix, DG2D/basis_polynomials.go:49-53)"""
        r = np.ascontiguousarray(r, dtype=np.float64)
        s = np.ascontiguousarray(s, dtype=np.float64)
        out = np.empty((r.size, self.NpInt))
        _lib.check(_lib.load().gocfd_euler2d_interp_matrix(self._h, r.size, _dp(r), _dp(s), _dp(out)))
        return out

    def edges(self) -> np.ndarray:
        e = np.empty((self.nedges, 10), dtype=np.int64)
        _lib.check(_lib.load().gocfd_euler2d_get_edges(self._h, e.ctypes.data_as(C.POINTER(C.c_int64))))
        return e

    def connectivity(self):
        etoe = np.empty((self.K, 3), dtype=np.int32)
        fe = np.empty((3, self.K), dtype=np.int32)
        _lib.check(_lib.load().gocfd_euler2d_get_connectivity(self._h, _i32(etoe), _i32(fe)))
        return etoe, fe

    def operators(self):
        a = np.empty((self.NpInt, self.NpFlux)); d = np.empty((self.NpFlux, self.NpFlux))
        it = np.empty((3 * self.NpEdge, self.NpInt)); r = np.empty(self.NpFlux); s = np.empty(self.NpFlux)
        _lib.check(_lib.load().gocfd_euler2d_get_operators(self._h, _dp(a), _dp(d), _dp(it), _dp(r), _dp(s)))
        return dict(DivInt=a, Div=d, Interp=it, R=r, S=s)

    def geometry(self):
        K = self.K
        out = dict(Jdet=np.empty(K), Jinv=np.empty((K, 4)), FaceNx=np.empty((3, K)),
                   FaceNy=np.empty((3, K)), IInII=np.empty((3, K)), EdgeLenMax=np.empty(K))
        _lib.check(_lib.load().gocfd_euler2d_get_geometry(
            self._h, *[_dp(out[k]) for k in ("Jdet", "Jinv", "FaceNx", "FaceNy", "IInII", "EdgeLenMax")]))
        return out

    def coords(self):
        K = self.K
        sx, sy = np.empty((self.NpInt, K)), np.empty((self.NpInt, K))
        fx, fy = np.empty((self.NpFlux, K)), np.empty((self.NpFlux, K))
        _lib.check(_lib.load().gocfd_euler2d_get_coords(self._h, _dp(sx), _dp(sy), _dp(fx), _dp(fy)))
        return dict(SolX=sx, SolY=sy, FluxX=fx, FluxY=fy)


def NewEuler(ip: InputParameters2D, meshFile: str, ProcLimit: int = 0, verbose: bool = False,
             profile: bool = False, **kw) -> Euler:
    """Euler2D.NewEuler (euler.go:61)."""
    return Euler(ip, meshFile, ProcLimit, verbose, profile, **kw)
