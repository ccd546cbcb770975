"""zoomy_hip: MI355X-native finite-volume time step of zoomy_jax.

``zoomy_hip.mesh`` builds meshes on the host, ``zoomy_hip.fvm`` is the
solver API (a drop-in for ``zoomy_jax.fvm.solver_jax``), and the arithmetic
runs in ``libzoomy_hip.so`` (HIP kernels for gfx950 behind a C ABI).
"""
from . import mesh  # noqa: F401

__version__ = "0.1.0"
