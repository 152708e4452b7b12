"""Batched ensembles of independent cases on one GPU (no counterpart in the
reference, which runs one case per process): each case has its own mesh,
order, initial state, Mach number, angle, CFL, FinalTime, limiter and
time-stepping mode; cases of one order, limiter (PerssonC0 with its Kappa)
and time-stepping mode share the kernel launches of a stage, and runs of 6
steps replay a captured HIP graph."""
from __future__ import annotations

import ctypes as C
from typing import Sequence, Tuple

import numpy as np

from . import _lib
from .euler2d import Mesh, params_of, _dp
from .input_parameters import InputParameters2D


class Ensemble:
    def __init__(self, cases: Sequence[Tuple[InputParameters2D, "str | Mesh"]], precision: str = "fp32",
                 device: int = 0):
        L = _lib.load()
        self._meshes = [m if isinstance(m, Mesh) else Mesh.read(m) for _, m in cases]
        n = len(cases)
        params = (_lib.Params * n)(*[params_of(ip, precision, device) for ip, _ in cases])
        mh = (C.c_void_p * n)(*[m._h for m in self._meshes])
        h = C.c_void_p()
        _lib.check(L.gocfd_ensemble_create(n, mh, params, C.byref(h)))
        self._h = h
        self.ncases = n
        self.dims = []
        for c in range(n):
            K, Np = C.c_int32(), C.c_int32()
            _lib.check(L.gocfd_ensemble_case_dims(h, c, C.byref(K), C.byref(Np)))
            self.dims.append((K.value, Np.value))

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.load().gocfd_ensemble_destroy(self._h)
            self._h = None

    def step(self, nsteps: int = 1) -> None:
        rc = _lib.load().gocfd_ensemble_step(self._h, nsteps)
        if rc == _lib.ERR_NAN:
            raise FloatingPointError(_lib.load().gocfd_last_error().decode())
        _lib.check(rc)

    def state(self, c: int) -> np.ndarray:
        K, Np = self.dims[c]
        q = np.empty((4, Np, K))
        _lib.check(_lib.load().gocfd_ensemble_get_state(self._h, c, _dp(q)))
        return q

    def time_info(self, c: int):
        t, dt, s = C.c_double(), C.c_double(), C.c_int32()
        _lib.check(_lib.load().gocfd_ensemble_time(self._h, c, C.byref(t), C.byref(dt), C.byref(s)))
        return t.value, dt.value, s.value

    def kernel_time(self):
        ms, n = C.c_double(), C.c_int32()
        _lib.check(_lib.load().gocfd_ensemble_kernel_time(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_launch(self, parallel: int, graph_steps: int = 6, xcd_tiles: bool = True) -> None:
        """Launch configuration: parallel 0 = order groups one after another,
        1 (or True) = as parallel graph branches, 2 = the groups without shock
        capturing in one multi-order launch per stage; steps per captured
        graph (a multiple of 6, or 0 for eager launches); XCD-contiguous tile
        order."""
        _lib.check(_lib.load().gocfd_ensemble_set_launch(self._h, int(parallel), int(graph_steps), int(xcd_tiles)))

    def launch(self) -> dict:
        p, g, x = C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(_lib.load().gocfd_ensemble_get_launch(self._h, C.byref(p), C.byref(g), C.byref(x)))
        return {"parallel": p.value, "graph_steps": g.value, "xcd_tiles": bool(x.value)}

    def tune(self, trial_blocks: int = 2) -> dict:
        """Autotune the launch configuration from the initial state (before
        the first step); returns the choice and its ms per step."""
        _lib.check(_lib.load().gocfd_ensemble_tune(self._h, int(trial_blocks)))
        out = self.launch()
        out["ms_per_step"] = self.kernel_time()[0]
        return out
