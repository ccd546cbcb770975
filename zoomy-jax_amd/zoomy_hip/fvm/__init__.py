"""Finite-volume time step and Chorin projection (host API over the HIP
kernels)."""
from .chorin import ChorinSplitVAMSolver, EllipticSolveError, elliptic_accept
from .solver import (DeviceMesh, HyperbolicSolver, IMEXSourceSolver,
                     ShallowWater)

__all__ = ["ChorinSplitVAMSolver", "DeviceMesh", "EllipticSolveError",
           "HyperbolicSolver", "IMEXSourceSolver", "ShallowWater",
           "elliptic_accept"]
