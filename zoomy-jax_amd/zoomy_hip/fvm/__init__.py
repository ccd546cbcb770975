"""Finite-volume time step (host API over the HIP kernels)."""
from .solver import (DeviceMesh, HyperbolicSolver, IMEXSourceSolver,
                     ShallowWater)

__all__ = ["DeviceMesh", "HyperbolicSolver", "IMEXSourceSolver",
           "ShallowWater"]
