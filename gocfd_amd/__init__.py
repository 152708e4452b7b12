"""gocfd_amd: MI355X-native 2D Euler DFR right-hand side and SSP54 step.

A drop-in for the time step of Notargets/gocfd's Euler2D solver: the same
.su2/.neu ingest, DFR2D / Raviart-Thomas setup and YAML parameters, with the
step running in hand-written gfx950 HIP kernels behind a C ABI
(include/gocfd_euler2d.h).
"""
from .input_parameters import InputParameters2D
from .euler2d import Euler, Mesh, NewEuler, RungeKutta5SSP
from .ensemble import Ensemble

__all__ = ["InputParameters2D", "Euler", "Ensemble", "Mesh", "NewEuler", "RungeKutta5SSP"]
