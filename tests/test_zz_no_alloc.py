"""The no-allocation RHS entry point gocfd_euler2d_rhs_noalloc, through the C
ABI, against gocfd_euler2d_rhs on the same state: bit for bit.

Two handles on one device share the entry point's fixed scratch array; the
calls run on the larger mesh, then on a smaller one (the array keeps the
larger call's contents), then on the larger one again. Device memory for the
inputs and outputs comes from the HIP runtime the solver library links.
"""
import ctypes as C

import numpy as np
import pytest

from gocfd_amd import Euler, InputParameters2D, _lib, meshgen

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    _lib.load()
    return C.CDLL("libamdhip64.so.7")  # the runtime libgocfd_euler2d.so is linked against


class _DevArray:
    def __init__(self, hip, host: np.ndarray):
        self.hip, self.nbytes = hip, host.nbytes
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0
        assert hip.hipMemcpy(self.ptr, host.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), H2D) == 0

    def get(self, shape):
        out = np.empty(shape)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes), D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _noalloc_rhs(hip, gpu, q, t, stream=None):
    q = np.ascontiguousarray(q, dtype=np.float64)
    dq = _DevArray(hip, q)
    dr = _DevArray(hip, np.full_like(q, np.nan))
    print(f"rhs_noalloc K={gpu.K} N={gpu.N} stream={'handle' if stream is None else 'caller'}", flush=True)
    rc = _lib.load().gocfd_euler2d_rhs_noalloc(gpu._h, dq.ptr, C.c_double(t), dr.ptr, stream)
    _lib.check(rc)
    if stream is not None:
        assert hip.hipStreamSynchronize(stream) == 0
    r = dr.get(q.shape)
    dq.free()
    dr.free()
    return r


def _perturbed(q, seed, amp=1e-2):
    rng = np.random.default_rng(seed)
    return q * (1.0 + amp * rng.standard_normal(q.shape))


def test_rhs_noalloc_matches_rhs(tmp_path):
    hip = _hip()
    big = Euler(InputParameters2D(CFL=1.0, FluxType="Roe", InitType="IVortex", PolynomialOrder=3,
                                  FinalTime=1e9, MaxIterations=10),
                meshgen.vortex_mesh(str(tmp_path / "v20.su2"), 20), precision="fp64")
    small = Euler(InputParameters2D(CFL=1.0, FluxType="Lax", InitType="IVortex", PolynomialOrder=2,
                                    FinalTime=1e9, MaxIterations=10, Limiter="PerssonC0", Kappa=10.0),
                  meshgen.vortex_mesh(str(tmp_path / "v4.su2"), 4), precision="fp32")
    assert small.K < big.K
    qb, qs = _perturbed(big.Q, 1), _perturbed(small.Q, 2)
    # first call: the larger mesh on the handle's stream
    r1 = _noalloc_rhs(hip, big, qb, 0.3)
    assert np.array_equal(r1, big.rhs(qb, 0.3))
    # second call: a smaller mesh, FP32 shock-capturing handle, on a caller stream
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    r2 = _noalloc_rhs(hip, small, qs, 0.1, stream)
    assert np.array_equal(r2, small.rhs(qs, 0.1))
    # the larger mesh again, after the smaller call left its contents behind
    r3 = _noalloc_rhs(hip, big, qb * 1.01, 0.7, stream)
    assert np.array_equal(r3, big.rhs(qb * 1.01, 0.7))
    assert hip.hipStreamDestroy(stream) == 0
    assert np.isfinite(r1).all() and np.isfinite(r2).all() and np.abs(r1).max() > 0


def test_rhs_noalloc_rejects_oversize(tmp_path):
    hip = _hip()
    # K = 2 * 135^2 = 36,450 triangles at N=4 need 35 MB of scratch (max 32 MiB)
    gpu = Euler(InputParameters2D(CFL=1.0, FluxType="Roe", InitType="IVortex", PolynomialOrder=4,
                                  FinalTime=1e9, MaxIterations=10),
                meshgen.vortex_mesh(str(tmp_path / "v135.su2"), 135), precision="fp32")
    q = np.ascontiguousarray(gpu.Q)
    dq = _DevArray(hip, q)
    dr = _DevArray(hip, q)
    rc = _lib.load().gocfd_euler2d_rhs_noalloc(gpu._h, dq.ptr, C.c_double(0.0), dr.ptr, None)
    msg = _lib.load().gocfd_last_error().decode()
    dq.free()
    dr.free()
    assert rc == _lib.ERR_ARG and "scratch" in msg, (rc, msg)
