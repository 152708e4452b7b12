"""CPU checks of the host side of the lift gradient: the state gradient and
the angle partial of the lift functional against central differences of
the functional itself, the freestream derivatives against differences of
NewFreeStream's formula (fluids.go:245-262), and the refusal of the
derivative entry points on a host-only handle."""
import math

import numpy as np
import pytest

from gocfd_amd import Euler, InputParameters2D, adjoint, meshgen
from gocfd_amd._lib import GocfdError


@pytest.fixture(scope="module")
def foil(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("foil") / "foil.su2")
    meshgen.airfoil_mesh(path, n_around=40, n_radial=6, radius=5.0, seed=2)
    ip = InputParameters2D(CFL=1.0, FluxType="Roe", InitType="Freestream", PolynomialOrder=2, Minf=0.5,
                           Alpha=2.0, FinalTime=1e9)
    return Euler(ip, path, device=-1)


def _state(e, seed):
    rng = np.random.default_rng(seed)
    q = np.empty((4, e.NpInt, e.K))
    q[0] = 1.0 + 0.1 * rng.random((e.NpInt, e.K))
    q[1] = 0.5 + 0.1 * rng.standard_normal((e.NpInt, e.K))
    q[2] = 0.1 * rng.standard_normal((e.NpInt, e.K))
    q[3] = 2.0 + 0.1 * rng.random((e.NpInt, e.K))
    return q


def test_lift_state_gradient_matches_differences(foil):
    q = _state(foil, 1)
    G = adjoint.lift_state_gradient(foil, q, 2.0)
    assert np.count_nonzero(G) > 0
    rng = np.random.default_rng(2)
    h = 1e-6
    for _ in range(3):
        v = rng.standard_normal(q.shape)
        fd = (adjoint.lift(foil, q + h * v, 2.0) - adjoint.lift(foil, q - h * v, 2.0)) / (2 * h)
        assert abs(np.sum(G * v) - fd) <= 1e-7 * abs(fd), (np.sum(G * v), fd)


def test_lift_alpha_partial_matches_differences(foil):
    # a pressure field with a net force in both directions
    q = _state(foil, 3)
    c = foil.coords()
    q[3] += 0.3 * c["SolX"] - 0.5 * c["SolY"]
    h = 1e-6
    fd = (adjoint.lift(foil, q, 2.0 + h) - adjoint.lift(foil, q, 2.0 - h)) / (2 * h)
    assert abs(adjoint.lift_alpha_partial(foil, q, 2.0) - fd) <= 1e-7 * abs(fd)


def test_far_state_derivatives_match_differences():
    g = 1.4

    def qinf(M, a):
        return np.array([1.0, M * math.cos(math.radians(a)), M * math.sin(math.radians(a)),
                         1.0 / (g * (g - 1.0)) + 0.5 * M * M])

    dM, da = adjoint.far_state_derivatives(0.5, 2.0)
    h = 1e-6
    assert np.allclose(dM, (qinf(0.5 + h, 2.0) - qinf(0.5 - h, 2.0)) / (2 * h), rtol=1e-8, atol=1e-10)
    assert np.allclose(da, (qinf(0.5, 2.0 + h) - qinf(0.5, 2.0 - h)) / (2 * h), rtol=1e-8, atol=1e-10)


def test_derivative_entry_points_need_a_device(foil):
    z = np.zeros((4, foil.NpInt, foil.K))
    with pytest.raises(GocfdError):
        foil.vjp(2, z)
    with pytest.raises(GocfdError):
        foil.jvp(2, z)
