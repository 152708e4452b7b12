"""CPU checks of the adjoint-workload harness: the airfoil O-mesh and the lift
functional (gocfd_amd.adjoint.lift), evaluated on the oracle (the functional reads only
the state, the operators, the geometry and the edge list, which the oracle
and the GPU handle expose alike).

* The mesh: 2 n_around n_radial counter-clockwise triangles, wall and far
  edges n_around each.
* The functional: zero for a uniform pressure (the wall polygon is closed);
  for a fluid at rest with pressure b*y + c, exactly -b times the area the
  wall polygon encloses (pressure is linear along each edge and the RT edge
  points are symmetric, so the edge mean is the midpoint value), at any
  angle: the force is (0, -b A) and the lift direction (-sin a, cos a).
"""
import math

import numpy as np
import pytest

import oracle
from gocfd_amd import Mesh, adjoint, meshgen


class _OracleView:
    def __init__(self, o):
        self.o, self.NpEdge, self.NpInt = o, o.NpEdge, o.NpInt

    def operators(self):
        return self.o.operators()

    def geometry(self):
        return self.o.geometry()

    def edges(self):
        return self.o.edges()


@pytest.fixture(scope="module")
def small(tmp_path_factory):
    return meshgen.airfoil_mesh(str(tmp_path_factory.mktemp("af") / "af64.su2"), 64, 20)


def test_airfoil_mesh(small):
    VX, VY, E = Mesh.read(small).arrays()
    assert E.shape == (2 * 64 * 20, 3) and len(VX) == 64 * 21
    p = [np.stack([VX[E[:, k]], VY[E[:, k]]], 1) for k in range(3)]
    cross = (p[1][:, 0] - p[0][:, 0]) * (p[2][:, 1] - p[0][:, 1]) - (p[1][:, 1] - p[0][:, 1]) * (p[2][:, 0] - p[0][:, 0])
    assert (cross > 0).all()
    o = oracle.OracleSolver(small, 1, flux="roe", init="freestream", Minf=0.5)
    bc = o.edges()[:, 9]
    assert (bc == adjoint.BC_WALL).sum() == 64 and (bc == 4).sum() == 64


@pytest.mark.parametrize("N", [1, 3])
def test_lift_functional_closure_and_linear_pressure(small, N):
    o = oracle.OracleSolver(small, N, flux="roe", init="freestream", Minf=0.5, Alpha=2.0)
    assert abs(adjoint.lift(_OracleView(o), o.state(), 2.0)) < 1e-14
    Y = o.coords()["SolY"]
    b, c = 0.3, 2.0
    q = np.stack([np.ones_like(Y), 0 * Y, 0 * Y, (b * Y + c) / 0.4])
    # area enclosed by the wall polygon (the airfoil's surface points)
    VX, VY, _ = Mesh.read(small).arrays()
    x, y = VX[:64], VY[:64]
    area = 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))
    for a in (0.0, 2.0, 30.0):
        J = adjoint.lift(_OracleView(o), q, a, gamma=1.4)
        assert abs(J - (-b * area * math.cos(math.radians(a)))) < 1e-12 * max(1.0, abs(J)), (a, J)
