"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: a plain-C FP64 restatement of the reference Go
solver (model_problems/Euler2D, DG2D). Imported by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
The product package (gocfd_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

FLUX = {"average": 0, "lax": 1, "roe": 2, "roe-er": 3}
INIT = {"freestream": 0, "ivortex": 1, "shocktube": 2}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrParams(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("flux", C.c_int), ("init", C.c_int),
        ("CFL", C.c_double), ("FinalTime", C.c_double), ("MaxIterations", C.c_int),
        ("Minf", C.c_double), ("Gamma", C.c_double), ("Alpha", C.c_double),
        ("LocalTimeStepping", C.c_int), ("limiter", C.c_int), ("Kappa", C.c_double),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.c_char_p, C.POINTER(OrParams), C.c_char_p, C.c_int]
        L.or_destroy.argtypes = [C.c_void_p]
        for f in ("or_K", "or_NpInt", "or_NpEdge", "or_NpFlux", "or_num_edges", "or_step_count"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = C.c_int
        for f in ("or_time", "or_dt"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = C.c_double
        L.or_get_state.argtypes = [C.c_void_p, _dp]
        L.or_set_state.argtypes = [C.c_void_p, _dp]
        L.or_get_residual.argtypes = [C.c_void_p, _dp]
        L.or_rhs.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
        L.or_rhs.restype = C.c_int
        L.or_get_flux_dofs.argtypes = [C.c_void_p, _dp]
        L.or_step.argtypes = [C.c_void_p]
        L.or_threads.restype = C.c_int
        L.or_step.restype = C.c_int
        L.or_get_operators.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.or_get_solution_element.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
        L.or_get_geometry.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
        L.or_get_coords.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.or_get_edges.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
        L.or_get_connectivity.argtypes = [C.c_void_p, _ip, _ip]
        L.or_get_shock.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.or_shock_indicator.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
        L.or_vertex_max.argtypes = [C.c_void_p, _dp, _dp]
        L.or_gradient.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.or_sensor.argtypes = [C.c_void_p, _dp, _dp]
        L.or_sensor.restype = C.c_int
        L.or_filter.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.or_filter.restype = C.c_int
        L.or_get_dissipation.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.or_get_dissipation.restype = C.c_int
        L.or_get_dt.argtypes = [C.c_void_p, _dp]
        L.or_ivortex_state.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        L.or_freestream.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp, _dp]
        L.or_numerical_flux.argtypes = [C.c_int, C.c_double, _dp, _dp, _dp, _dp]
        L.or_riemann_bc.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        L.or_set_nan_scan.argtypes = [C.c_void_p, C.c_int]
        L.or_sigma_ramp.argtypes = [C.c_double, C.c_double, C.c_int]
        L.or_sigma_ramp.restype = C.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


class OracleSolver:
    """The reference Euler2D solver restated on the CPU (NewEuler, euler.go:61)."""

    def __init__(self, meshfile: str, N: int, flux: str = "lax", init: str = "ivortex",
                 CFL: float = 1.0, FinalTime: float = 1e9, MaxIterations: int = 0,
                 Minf: float = 0.1, Gamma: float = 1.4, Alpha: float = 0.0,
                 LocalTimeStepping: bool = False, limiter: str = "", Kappa: float = 0.0):
        L = lib()
        p = OrParams(N, FLUX[flux.lower()], INIT[init.lower()], CFL, FinalTime, MaxIterations,
                     Minf, Gamma, Alpha, int(LocalTimeStepping),
                     1 if limiter.lower().replace(" ", "") == "perssonc0" else 0, Kappa)
        err = C.create_string_buffer(512)
        h = L.or_create(meshfile.encode(), C.byref(p), err, 512)
        if not h:
            raise RuntimeError(err.value.decode())
        self._h = h
        self.K = L.or_K(h)
        self.NpInt = L.or_NpInt(h)
        self.NpEdge = L.or_NpEdge(h)
        self.NpFlux = L.or_NpFlux(h)
        self.N = N

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_destroy(self._h)
            self._h = None

    def state(self):
        q = np.empty((4, self.NpInt, self.K))
        lib().or_get_state(self._h, _p(q))
        return q

    def set_state(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        lib().or_set_state(self._h, _p(q))

    def rhs(self, q, t=0.0):
        q = np.ascontiguousarray(q, dtype=np.float64)
        r = np.empty_like(q)
        lib().or_rhs(self._h, _p(q), t, _p(r))
        return r

    def flux_dofs(self):
        F = np.empty((4, self.NpFlux, self.K))
        lib().or_get_flux_dofs(self._h, _p(F))
        return F

    def rhs_term_scale(self):
        """sum_j |DivInt_ij| |F_RT_j| / |Jdet| of the last RHS evaluation: the
        magnitude of the terms the divergence sums, which bounds its rounding."""
        F = self.flux_dofs()
        A = np.abs(self.operators()["DivInt"])
        J = np.abs(self.geometry()["Jdet"])
        return np.einsum("ij,njk->nik", A, np.abs(F)) / J[None, None, :]

    def set_nan_scan(self, mode: int):
        """NaN scan before each stage: 0 none, 1 stage input, q0 and DT (the
        default), 2 the reference's eleven arrays (euler.go:471-484)"""
        lib().or_set_nan_scan(self._h, int(mode))

    def step(self):
        rc = lib().or_step(self._h)
        if rc:
            raise FloatingPointError("NAN found")

    @property
    def time(self):
        return lib().or_time(self._h)

    @property
    def dt(self):
        return lib().or_dt(self._h)

    @property
    def step_count(self):
        return lib().or_step_count(self._h)

    def residual(self):
        r = np.empty((4, self.NpInt, self.K))
        lib().or_get_residual(self._h, _p(r))
        return r

    def operators(self):
        a = np.empty((self.NpInt, self.NpFlux)); d = np.empty((self.NpFlux, self.NpFlux))
        it = np.empty((3 * self.NpEdge, self.NpInt)); r = np.empty(self.NpFlux); s = np.empty(self.NpFlux)
        lib().or_get_operators(self._h, _p(a), _p(d), _p(it), _p(r), _p(s))
        return dict(DivInt=a, Div=d, Interp=it, R=r, S=s)

    def solution_element(self):
        n = self.NpInt
        V, Vi, M = np.empty((n, n)), np.empty((n, n)), np.empty((n, n))
        r, s, w = np.empty(n), np.empty(n), np.empty(n)
        lib().or_get_solution_element(self._h, _p(V), _p(Vi), _p(M), _p(r), _p(s), _p(w))
        return dict(V=V, Vinv=Vi, Mass=M, R=r, S=s, W=w)

    def geometry(self):
        K = self.K
        out = dict(Jdet=np.empty(K), Jinv=np.empty((K, 4)), FaceNx=np.empty((3, K)),
                   FaceNy=np.empty((3, K)), IInII=np.empty((3, K)), EdgeLenMax=np.empty(K))
        lib().or_get_geometry(self._h, *[_p(out[k]) for k in
                                        ("Jdet", "Jinv", "FaceNx", "FaceNy", "IInII", "EdgeLenMax")])
        return out

    def coords(self):
        K = self.K
        sx, sy = np.empty((self.NpInt, K)), np.empty((self.NpInt, K))
        fx, fy = np.empty((self.NpFlux, K)), np.empty((self.NpFlux, K))
        lib().or_get_coords(self._h, _p(sx), _p(sy), _p(fx), _p(fy))
        return dict(SolX=sx, SolY=sy, FluxX=fx, FluxY=fy)

    def edges(self):
        ne = lib().or_num_edges(self._h)
        e = np.empty((ne, 10), dtype=np.int64)
        lib().or_get_edges(self._h, e.ctypes.data_as(C.POINTER(C.c_longlong)))
        return e

    def connectivity(self):
        etoe = np.empty((self.K, 3), dtype=np.int32)
        fe = np.empty((3, self.K), dtype=np.int32)
        lib().or_get_connectivity(self._h, etoe.ctypes.data_as(_ip), fe.ctypes.data_as(_ip))
        return etoe, fe

    def shock(self):
        K = self.K
        se, sg, ep = np.zeros(K), np.zeros(K), np.zeros(K)
        lib().or_get_shock(self._h, _p(se), _p(sg), _p(ep))
        return se, sg, ep

    def shock_indicator(self, rho, kappa):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        out = np.empty(self.K)
        lib().or_shock_indicator(self._h, _p(rho), kappa, _p(out))
        return out


    # ---- shock-capturing pieces (PerssonC0) ----
    def gradient(self, q):
        """RT gradient of the 4 variables with eps = 1: DX, DY [4][NpFlux][K]."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        dx = np.empty((4, self.NpFlux, self.K)); dy = np.empty_like(dx)
        lib().or_gradient(self._h, _p(q), _p(dx), _p(dy))
        return dx, dy

    def sensor(self, rho):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        se = np.empty(self.K)
        if lib().or_sensor(self._h, _p(rho), _p(se)):
            raise RuntimeError("solver made without the PerssonC0 limiter")
        return se

    def filter(self, u, sigma):
        u = np.array(u, dtype=np.float64, order="C")
        nv = u.size // (self.NpInt * self.K)
        sig = np.ascontiguousarray(np.broadcast_to(sigma, (self.K,)), dtype=np.float64)
        if lib().or_filter(self._h, _p(u), _p(sig), nv):
            raise RuntimeError("solver made without the PerssonC0 limiter")
        return u

    def dissipation(self):
        nf = (self.NpFlux, self.K)
        eps = np.empty(nf); dx = np.empty((4,) + nf); dy = np.empty((4,) + nf)
        ne = lib().or_num_edges(self._h)
        ev = np.empty((ne, self.NpEdge, 4))
        if lib().or_get_dissipation(self._h, _p(eps), _p(dx), _p(dy), _p(ev)):
            raise RuntimeError("solver made without the PerssonC0 limiter")
        return dict(EpsF=eps, DissX=dx, DissY=dy, EdgeVisc=ev)

    def element_dt(self):
        d = np.empty(self.K)
        lib().or_get_dt(self._h, _p(d))
        return d


def threads() -> int:
    """OpenMP threads of the oracle's parallel loops."""
    return lib().or_threads()


def ivortex_state(t, x, y):
    q = np.empty(4)
    lib().or_ivortex_state(t, x, y, _p(q))
    return q


def freestream(Minf, Gamma=1.4, Alpha=0.0):
    q = np.empty(4); pinf = C.c_double(); cinf = C.c_double()
    lib().or_freestream(Minf, Gamma, Alpha, _p(q), C.byref(pinf), C.byref(cinf))
    return q, pinf.value, cinf.value


def numerical_flux(flux, gamma, QL, QR, n):
    QL = np.ascontiguousarray(QL, dtype=np.float64); QR = np.ascontiguousarray(QR, dtype=np.float64)
    n = np.ascontiguousarray(n, dtype=np.float64); F = np.empty(4)
    lib().or_numerical_flux(FLUX[flux], gamma, _p(QL), _p(QR), _p(n), _p(F))
    return F


def riemann_bc(gamma, minf, pinf, cinf, Qint, Qinf, n):
    Qint = np.ascontiguousarray(Qint, dtype=np.float64); Qinf = np.ascontiguousarray(Qinf, dtype=np.float64)
    n = np.ascontiguousarray(n, dtype=np.float64); Q = np.empty(4)
    lib().or_riemann_bc(gamma, minf, pinf, cinf, _p(Qint), _p(Qinf), _p(n), _p(Q))
    return Q


def sigma_ramp(se, kappa, N):
    """The stepping path's sensor ramp sigma(Se) (dissipation.go:491-518)."""
    return lib().or_sigma_ramp(float(se), float(kappa), int(N))
