"""The C ABI on the CPU: the library loads and exports what the header
declares, host-only handles (device -1) reproduce the oracle's setup
(edges, connectivity, geometry bit-exact; operators to rounding), the
parameter checks fail the way the reference fails, and the YAML parameters
bind with the reference's rules. No device work is done here.
"""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import oracle
from gocfd_amd import Euler, InputParameters2D, Mesh, _lib
from gocfd_amd._lib import GocfdError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = os.path.join(ROOT, "tests", "golden")


def _declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        for m in re.finditer(r"\b(gocfd_\w+)\s*\(", src):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = _declared_symbols()
    assert len(names) >= 25
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def _host(mesh, N, flux="Lax", init="ivortex", **kw):
    ip = InputParameters2D(CFL=1.0, FluxType=flux, InitType=init, PolynomialOrder=N, FinalTime=1e9,
                           Minf=kw.get("Minf", 0.5), Alpha=kw.get("Alpha", 0.0))
    return Euler(ip, mesh, device=-1)


CASES = [
    ("vortex20", 2, "ivortex"),
    ("vortex4", 4, "ivortex"),
    ("test_tris_5.neu", 1, "freestream"),
    ("test_tris_6.neu", 2, "ivortex"),
    ("test_tris_9.neu", 3, "freestream"),
    ("fstepA001.neu", 2, "freestream"),
    ("vortex-new.su2", 0, "ivortex"),
]


def _mesh_path(name, vortex_meshes):
    if name.startswith("vortex") and not name.endswith(".su2"):
        return vortex_meshes[int(name[6:])]
    return os.path.join(G, name)


@pytest.mark.parametrize("name,N,init", CASES)
def test_host_setup_matches_oracle(vortex_meshes, name, N, init):
    path = _mesh_path(name, vortex_meshes)
    h = _host(path, N, init=init)
    o = oracle.OracleSolver(path, N, flux="lax", init=init, Minf=0.5)
    assert (h.K, h.NpInt, h.NpEdge, h.NpFlux) == (o.K, o.NpInt, o.NpEdge, o.NpFlux)
    # integer structure: bit-exact
    assert np.array_equal(h.edges(), o.edges())
    e1, f1 = h.connectivity()
    e2, f2 = o.connectivity()
    assert np.array_equal(e1, e2) and np.array_equal(f1, f2)
    # geometry: same closed-form arithmetic, bit-exact
    g1, g2 = h.geometry(), o.geometry()
    for k in g1:
        assert np.array_equal(g1[k], g2[k]), k
    c1, c2 = h.coords(), o.coords()
    for k in c1:
        assert np.allclose(c1[k], c2[k], rtol=0, atol=1e-13), k
    # operators: two independent inversions, agreement to rounding
    p1, p2 = h.operators(), o.operators()
    for k in ("R", "S"):
        assert np.abs(p1[k] - p2[k]).max() < 1e-14, k
    for k in ("DivInt", "Div", "Interp"):
        scale = max(np.abs(p2[k]).max(), 1.0)
        assert np.abs(p1[k] - p2[k]).max() / scale < 1e-11, k


def test_tiling_covers_mesh(vortex_meshes):
    h = _host(vortex_meshes[20], 2)
    tile, ntiles, hmax = h.tiling()
    assert tile == 64 and ntiles * tile >= h.K > (ntiles - 1) * tile
    assert 0 < hmax <= 3 * tile


def test_interp_reproduces_polynomials_host(vortex_meshes):
    """FluxEdgeInterp of the host operators is exact on polynomials of degree
    <= N sampled at the solution points (dfr_startup_test.go:38-63)."""
    for N in range(0, 5):
        h = _host(vortex_meshes[4], N)
        p = h.operators()
        R, S = p["R"], p["S"]
        ni, ne = h.NpInt, h.NpEdge
        rs, ss = R[:ni], S[:ni]
        re_, se_ = R[2 * ni:], S[2 * ni:]
        for a in range(N + 1):
            for b in range(N + 1 - a):
                f = rs ** a * ss ** b
                g = p["Interp"] @ f
                assert np.abs(g - re_ ** a * se_ ** b).max() < 1e-12, (N, a, b)
        assert len(re_) == 3 * ne


def test_host_handle_rejects_device_calls(vortex_meshes):
    h = _host(vortex_meshes[4], 1)
    with pytest.raises(GocfdError) as ei:
        _ = h.Q
    assert ei.value.code == _lib.ERR_DEVICE
    with pytest.raises(GocfdError):
        h.step(1)


@pytest.mark.parametrize("kw,msg,prec", [
    (dict(PolynomialOrder=5), "polynomial order", "fp64"),
])
def test_parameter_errors(vortex_meshes, kw, msg, prec):
    base = dict(CFL=1.0, FluxType="Lax", InitType="ivortex", PolynomialOrder=2, FinalTime=1.0)
    base.update(kw)
    with pytest.raises(GocfdError) as ei:
        Euler(InputParameters2D(**base), vortex_meshes[4], precision=prec, device=-1)
    assert ei.value.code == _lib.ERR_ARG and msg.lower() in str(ei.value).lower()


def test_shocktube_host_setup(tmp_path):
    """InitType shocktube builds on the host: In/Out edges of the channel
    mesh map to the In/Out boundary kinds like the oracle's."""
    from gocfd_amd import meshgen
    mesh = meshgen.shocktube_mesh(str(tmp_path / "st.su2"), 20, 2, height=0.02)
    ip = InputParameters2D(CFL=0.5, FluxType="Roe", InitType="shocktube", PolynomialOrder=1, FinalTime=0.1,
                           Limiter="PerssonC0", Kappa=3.0)
    h = Euler(ip, mesh, device=-1)
    o = oracle.OracleSolver(mesh, 1, flux="roe", init="shocktube")
    assert np.array_equal(h.edges(), o.edges())


def test_shock_and_lts_handles_build_on_host(vortex_meshes):
    """PerssonC0 and LocalTimeStepping are accepted (FP64 and FP32); host-only handles
    refuse the device-only introspection."""
    ip = InputParameters2D(CFL=1.0, FluxType="Roe", InitType="ivortex", PolynomialOrder=2, FinalTime=1.0,
                           Limiter="PerssonC0", Kappa=3.0, LocalTimeStepping=True)
    h = Euler(ip, vortex_meshes[4], device=-1)
    Euler(ip, vortex_meshes[4], precision="fp32", device=-1)
    with pytest.raises(GocfdError):
        h.shock()
    with pytest.raises(GocfdError):
        h.element_dt()


def test_unknown_names_raise_like_the_reference(vortex_meshes):
    with pytest.raises(ValueError, match="unable to use flux"):
        Euler(InputParameters2D(FluxType="hllc", InitType="ivortex", PolynomialOrder=1), vortex_meshes[4], device=-1)
    with pytest.raises(ValueError, match="unable to use init type"):
        Euler(InputParameters2D(FluxType="lax", InitType="bogus", PolynomialOrder=1), vortex_meshes[4], device=-1)


def test_mesh_read_errors(tmp_path):
    with pytest.raises(GocfdError) as ei:
        Mesh.read(str(tmp_path / "missing.su2"))
    assert ei.value.code == _lib.ERR_MESH
    bad = tmp_path / "bad.su2"
    bad.write_text("NDIME= 3\n")
    with pytest.raises(GocfdError):
        Mesh.read(str(bad))


def test_mesh_from_arrays_round_trip(vortex_meshes):
    m = Mesh.read(vortex_meshes[4])
    K, Nv, nt = m.dims()
    VX, VY, EToV = m.arrays()
    assert EToV.shape == (K, 3) and VX.shape == (Nv,)
    # same mesh without BC tags: all boundary edges open
    m2 = Mesh.from_arrays(VX, VY, EToV, {})
    assert m2.dims()[:2] == (K, Nv)
    ip = InputParameters2D(CFL=1.0, FluxType="Lax", InitType="freestream", PolynomialOrder=1, FinalTime=1.0)
    h = Euler(ip, m2, device=-1)
    e = h.edges()
    assert e.shape[1] == 10 and (e[:, 9] == 0).all()


def test_input_parameters_binding():
    """ghodss/yaml binds keys to Go field names case-insensitively and ignores
    unknown keys (InputParameters.go:11-32); CLI defaults Gamma 1.4, Minf 0.1."""
    ip = InputParameters2D().parse(
        "title: t\ncfl: 0.5\nfluxtype: Roe\nInitType: IVortex\npolynomialorder: 3\n"
        "FinalTime: 2\nLocalTimeStep: true\nMaxIterations: 7\nKappa: 6\nunknown: 1\n")
    assert (ip.Title, ip.CFL, ip.FluxType, ip.InitType, ip.PolynomialOrder) == ("t", 0.5, "Roe", "IVortex", 3)
    assert ip.FinalTime == 2.0 and ip.MaxIterations == 7 and ip.Kappa == 6.0
    assert ip.LocalTimeStepping is False  # "LocalTimeStep" is not the field name
    assert ip.Gamma == 1.4 and ip.Minf == 0.1
