"""Finite-volume time step and Chorin projection (host API over the HIP
kernels)."""
from .chorin import ChorinSplitVAMSolver, EllipticSolveError, elliptic_accept
from .diffusion import DiffusionOperator
from .solver import (DeviceMesh, HyperbolicSolver, IMEXSourceSolver,
                     Reconstruction, ShallowWater)
from . import blocks

__all__ = ["ChorinSplitVAMSolver", "DeviceMesh", "DiffusionOperator",
           "EllipticSolveError",
           "HyperbolicSolver", "IMEXSourceSolver", "Reconstruction",
           "ShallowWater", "blocks",
           "elliptic_accept"]
