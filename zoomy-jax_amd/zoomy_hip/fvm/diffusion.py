"""Implicit (Crank-Nicolson) scalar diffusion on the GPU
(``zoomy_jax.fvm.reconstruction_jax.DiffusionOperatorJAX``,
``reconstruction_jax.py:1423-1520``).

``L`` is the two-point nu-Laplacian over interior faces (boundary faces carry
no diffusive flux); ``explicit(u) = L u`` and ``implicit_solve(u*, dt)``
solves ``(I - dt/2 L) u = (I + dt/2 L) u*``.  The reference assembles ``L``
densely and solves it with ``jnp.linalg.solve``; here the face coefficients
sit in the mesh's slot tables on the device and the volume-weighted (SPD)
system runs as Jacobi-preconditioned CG (``zj_diffusion_*``) to a relative
residual ``tol`` (1e-12 by default).
"""
import ctypes

import torch

from .. import _lib as L
from ..mesh import Mesh
from .solver import _device, _p, _stream, mesh_desc


class DiffusionOperator:
    """``DiffusionOperator(mesh, dim, nu)``: the reference's constructor
    signature; ``u`` arguments are FP64 device tensors of ``n_cells``."""

    def __init__(self, mesh: Mesh, dim: int = None, nu: float = 1.0,
                 tol: float = 1e-12, maxiter: int = 2000):
        if dim is not None and dim != mesh.dim:
            raise ValueError(f"dim {dim} does not match the mesh ({mesh.dim})")
        if not nu >= 0.0:
            raise ValueError("nu must be >= 0")
        self.mesh = mesh
        self.nu = float(nu)
        self.tol = float(tol)
        self.maxiter = int(maxiter)
        self.nc = mesh.n_cells
        self.device = _device()
        d, keep = mesh_desc(mesh)
        h = ctypes.c_void_p()
        L.check(L.lib().zj_diffusion_create(ctypes.byref(d), ctypes.byref(h)),
                "zj_diffusion_create")
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                L.lib().zj_diffusion_destroy(h)
            except Exception:
                pass
            self.handle = None

    def _vec(self, u):
        u = torch.as_tensor(u, dtype=torch.float64, device=self.device)
        if u.shape[-1] < self.nc or u.dim() != 1:
            raise ValueError(f"expected a vector of {self.nc} cells, "
                             f"got shape {tuple(u.shape)}")
        return u[:self.nc].contiguous()

    def explicit(self, u):
        """``L u`` over the cells (shape ``(n_cells,)``)."""
        u = self._vec(u)
        out = torch.empty_like(u)
        L.check(L.lib().zj_diffusion_explicit(self.handle, self.nu, _p(u),
                                              _p(out), _stream()),
                "zj_diffusion_explicit")
        return out

    def implicit_solve(self, u_star, dt):
        """One Crank-Nicolson step from ``u_star``; raises
        :class:`zoomy_hip._lib.ZoomyHipError` when CG misses ``tol``."""
        u = self._vec(u_star)
        out = torch.empty_like(u)
        L.check(L.lib().zj_diffusion_cn(self.handle, self.nu, float(dt),
                                        _p(u), _p(out), self.tol,
                                        self.maxiter, _stream()),
                "zj_diffusion_cn")
        return out

    def stats(self):
        it = ctypes.c_int32(0)
        rr = ctypes.c_double(0.0)
        L.check(L.lib().zj_diffusion_stats(self.handle, ctypes.byref(it),
                                           ctypes.byref(rr)),
                "zj_diffusion_stats")
        return {"iterations": it.value, "relres": rr.value}
