"""Lift gradients through the discrete adjoint of the FP64 step.

The lift functional of the wall edges (a functional the reference does not
define; it has no derivative code) and its gradients with respect to the
initial state, the freestream Mach number and the angle of attack, from one
adjoint sweep (`Euler.vjp`) of the steps:

* J = sum over wall edges of |e| * mean_i p_i * (n . l), with p_i the
  pressure of the owner's trace at the edge's RT points (FluxEdgeInterp of
  the element state), n the element's outward normal (into the body, so
  p n is the pressure force on it) and l = (-sin a, cos a) the lift
  direction;
* the freestream enters the far-field state FSFar.Qinf = (1, M cos a,
  M sin a, 1/(g(g-1)) + M^2/2) (NewFreeStream, model_problems/Euler2D/
  fluids.go:245-262) and, with InitType Freestream, every value of the
  initial state (InitializeFS, initialization.go:50-59);
* the angle also enters J through l.
"""
from __future__ import annotations

import math

import numpy as np

BC_WALL = 5  # types.BCFLAG Wall (types/cfd.go:13-26)


def _wall_faces(e):
    ed = e.edges()
    w = ed[ed[:, 9] == BC_WALL]
    k, f = w[:, 3], w[:, 5]
    geo = e.geometry()
    length = geo["IInII"][f, k] * np.where(f == 1, 2.0 * math.sqrt(2.0), 2.0)
    return k, f, length, geo["FaceNx"][f, k], geo["FaceNy"][f, k]


def _interp(e):
    return e.operators()["Interp"].reshape(3, e.NpEdge, e.NpInt)


def lift(e, Q, alpha_deg: float, gamma: float = 1.4) -> float:
    """J of the state Q [4][NpInt][K] on the solver's mesh."""
    k, f, length, nx, ny = _wall_faces(e)
    Qf = np.einsum("epi,nie->nep", _interp(e)[f], Q[:, :, k])        # [4][nwall][NpEdge]
    rho, ru, rv, E = Qf
    p = (gamma - 1.0) * (E - 0.5 * (ru * ru + rv * rv) / rho)
    a = math.radians(alpha_deg)
    proj = -nx * math.sin(a) + ny * math.cos(a)
    return float(np.sum(length * p.mean(axis=1) * proj))


def lift_state_gradient(e, Q, alpha_deg: float, gamma: float = 1.4) -> np.ndarray:
    """dJ/dQ [4][NpInt][K]: the pressure derivatives at the wall RT points,
    weighted by |e| (n . l) / NpEdge, through the transposed FluxEdgeInterp
    rows of each wall face."""
    k, f, length, nx, ny = _wall_faces(e)
    I = _interp(e)[f]                                                  # [nwall][NpEdge][NpInt]
    Qf = np.einsum("epi,nie->nep", I, Q[:, :, k])
    rho, ru, rv, _ = Qf
    a = math.radians(alpha_deg)
    wgt = (length * (-nx * math.sin(a) + ny * math.cos(a)) / e.NpEdge)[:, None]
    gm1 = gamma - 1.0
    dp = np.stack([gm1 * 0.5 * (ru * ru + rv * rv) / (rho * rho), -gm1 * ru / rho, -gm1 * rv / rho,
                   np.full_like(rho, gm1)])                            # [4][nwall][NpEdge]
    g_face = dp * wgt                                                  # dJ/dQf
    G = np.zeros_like(Q)
    contrib = np.einsum("epi,nep->nie", I, g_face)                     # [4][NpInt][nwall]
    for n in range(4):
        np.add.at(G[n].T, k, contrib[n].T)
    return G


def lift_alpha_partial(e, Q, alpha_deg: float, gamma: float = 1.4) -> float:
    """dJ/d(angle in degrees) at fixed state: through the lift direction."""
    k, f, length, nx, ny = _wall_faces(e)
    Qf = np.einsum("epi,nie->nep", _interp(e)[f], Q[:, :, k])
    rho, ru, rv, E = Qf
    p = (gamma - 1.0) * (E - 0.5 * (ru * ru + rv * rv) / rho)
    a = math.radians(alpha_deg)
    dproj = (-nx * math.cos(a) - ny * math.sin(a)) * (math.pi / 180.0)
    return float(np.sum(length * p.mean(axis=1) * dproj))


def far_state_derivatives(minf: float, alpha_deg: float):
    """d Qinf / d Minf and d Qinf / d(angle in degrees) of NewFreeStream."""
    a = math.radians(alpha_deg)
    dM = np.array([0.0, math.cos(a), math.sin(a), minf])
    da = np.array([0.0, -minf * math.sin(a), minf * math.cos(a), 0.0]) * (math.pi / 180.0)
    return dM, da


def lift_gradients(e, nsteps: int, checkpoint_every: int = 0, gamma: float = 1.4) -> dict:
    """J after `nsteps` steps from the solver's current state and its
    gradients: dJ/dQ0 [4][NpInt][K] (current state), dJ/dQinf [4], and the
    total dJ/dMinf and dJ/dalpha (degrees) for a run that starts from the
    freestream (the initial state is Qinf at every point)."""
    ip = e.ip
    q0 = e.Q
    qn, _ = e.jvp(nsteps, np.zeros_like(q0))
    J = lift(e, qn, ip.Alpha, gamma)
    seed = lift_state_gradient(e, qn, ip.Alpha, gamma)
    dq0, dfar = e.vjp(nsteps, seed, checkpoint_every)
    dM, da = far_state_derivatives(ip.Minf, ip.Alpha)
    through_far = dfar + dq0.sum(axis=(1, 2))
    return dict(J=J, dJ_dQ0=dq0, dJ_dQinf=dfar,
                dJ_dMinf=float(through_far @ dM),
                dJ_dalpha_deg=float(through_far @ da) + lift_alpha_partial(e, qn, ip.Alpha, gamma))
