"""Chorin projection for the non-hydrostatic shallow-water system on MI355X.

Host mirror of ``ChorinSplitVAMSolverJax``
(``zoomy_jax/fvm/solver_chorin_vam_jax.py``):

* ``chorin_stage(dt, time, Q)``: hyperbolic Euler predictor
  (``_hyperbolic_stage`` :496-522) -> pressure solve
  (``_step_pressure_pure`` :673-883) -> corrector
  (``_step_corrector_pure`` :885-907)                               (:524-572)
* ``chorin_cycle(dt, time, Q)``: ``time_order`` 2 = two stages and the
  average with the step-start state                                 (:574-615)
* ``step(dt, time, Q)``                                             (:617-638)
* ``run_jit_steps(dt, n_steps, Q, t_start)`` -> ``(Q, t)``          (:640-671)
* :class:`EllipticSolveError` when a solve misses the accept gate
  ``max(1e-6, 10 tol)``                                             (:72-102)

The state is a ``(5, N)`` float64 tensor with rows ``(b, h, hu, hv, hw)`` on
an ``nx x ny`` wall-bounded ``Mesh.quad`` grid (cell id ``i*ny + j``).  The
pressure equation ``S p = -D(Q*)/dt`` has the symmetric positive definite
5-point operator of the depth-averaged non-hydrostatic model (linear vertical
pressure profile, bottom pressure 2p).  Preconditioners of the FP64 CG:
``"mg"`` (FP64 geometric multigrid V-cycle), ``"mixed"`` (the FP32 V-cycle,
flexible CG) and ``"learned"`` (the FP32 V-cycle plus one correction by the
BF16 message-passing network, flexible CG); a low-precision solve whose fresh
FP64 residual misses ``tol`` is redone with ``"mg"``.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np
import torch

from .. import _lib as L
from ..mesh import Mesh
from .solver import DeviceMesh, ShallowWater, _Runtime, _device, _p, _stream


class EllipticSolveError(RuntimeError):
    """The pressure solve missed its accept gate (raised, never warned)."""


def elliptic_accept(tol: float) -> float:
    """Residual above which a pressure solve is declared failed."""
    return max(1e-6, 10.0 * float(tol))


class ChorinSplitVAMSolver:
    """Chorin projection driver (see the module docstring).

    ``model``: the hyperbolic numerics of the predictor (a
    :class:`ShallowWater` of dimension 2); ``tol``: relative residual target
    of the pressure solve (``c_elliptic_tol``)."""

    def __init__(self, model: Optional[ShallowWater] = None, *,
                 tol: float = 1e-8, time_order: int = 2,
                 precond: str = "mg", maxiter: int = 200, nu: int = 1,
                 coarse_sweeps: int = 2, min_coarse: int = 4,
                 flux_dtype: torch.dtype = torch.float64):
        if precond not in L.PRECOND:
            raise ValueError(f"precond must be one of {sorted(L.PRECOND)}")
        if time_order not in (1, 2):
            raise ValueError("time_order must be 1 or 2")
        self.model = model or ShallowWater(dimension=2, order=2,
                                           limiter="minmod", riemann="hll")
        if self.model.dimension != 2:
            raise ValueError("the Chorin projection needs a 2-D model")
        self.tol = float(tol)
        self.time_order = int(time_order)
        self.precond = precond
        self._prm = L.ChorinParams(float(tol),
                                   float(self.model.wet_dry_eps), int(maxiter),
                                   int(nu), int(coarse_sweeps),
                                   int(min_coarse), L.PRECOND[precond],
                                   int(time_order))
        if flux_dtype not in (torch.float64, torch.float32):
            raise ValueError("flux_dtype must be torch.float64 or float32")
        self.flux_dtype = flux_dtype
        self.handle = None
        self.net = None

    # ── setup ─────────────────────────────────────────────────────────
    def setup_simulation(self, mesh: Mesh, nx: int, ny: int, Q0=None,
                         fv_order: str = "morton"):
        """Bind to an ``nx x ny`` ``Mesh.quad`` grid (walls); returns the
        state tensor (zeros, or ``Q0`` as a (4|5, N) array).  The
        hyperbolic predictor runs on the cells in ``fv_order`` ("morton":
        along the Z-order curve; "grid": the grid order)."""
        if fv_order not in ("morton", "grid"):
            raise ValueError("fv_order must be 'morton' or 'grid'")
        if mesh.dim != 2 or mesh.n_cells != nx * ny or mesh.max_faces != 4:
            raise ValueError("the Chorin projection needs an nx x ny quad "
                             "mesh in Mesh.quad order")
        if not mesh.bc:
            mesh.set_bc("wall")
        if any(k != 1 for k in mesh.bc.values()):
            raise ValueError("the Chorin projection is built for walls on "
                             "every boundary tag")
        lo = mesh.verts.min(axis=0)
        hi = mesh.verts.max(axis=0)
        dx = float(hi[0] - lo[0]) / nx
        dy = float(hi[1] - lo[1]) / ny
        self.mesh, self.nx, self.ny, self.dx, self.dy = mesh, nx, ny, dx, dy
        self.N = nx * ny
        self.device = _device()
        order = None
        fv_mesh = mesh
        if fv_order == "morton":
            order = np.ascontiguousarray(mesh.morton_permutation(), np.int32)
            fv_mesh = mesh.reorder(order).set_bc("wall")
        self._dmesh = DeviceMesh(fv_mesh)
        self._rt = _Runtime(self._dmesh, self.model, 0.9, torch.float64,
                            "conservative")
        h = ctypes.c_void_p()
        L.check(L.lib().zj_chorin_create(self._rt.handle, nx, ny, dx, dy,
                                         ctypes.byref(self._prm),
                                         ctypes.byref(h)), "zj_chorin_create")
        self.handle = h
        if order is not None:
            L.check(L.lib().zj_chorin_set_fv_order(h, order.ctypes.data),
                    "zj_chorin_set_fv_order")
        if self.flux_dtype == torch.float32:
            # mixed precision: FP32 fluxes on the same (reordered) cells
            self._rt32 = _Runtime(self._dmesh, self.model, 0.9, torch.float32,
                                  "conservative")
            L.check(L.lib().zj_chorin_set_fv32(h, self._rt32.handle),
                    "zj_chorin_set_fv32")
        Q = torch.zeros((5, self.N), dtype=torch.float64, device=self.device)
        if Q0 is not None:
            Q0 = torch.as_tensor(np.asarray(Q0), dtype=torch.float64)
            Q[:Q0.shape[0]] = Q0.to(self.device)
        return Q

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and L is not None and L._lib is not None:
            L._lib.zj_chorin_destroy(h)
            self.handle = None

    def set_mg_fused(self, on: bool = True):
        """Fused multigrid kernels (bit-identical to the separate ones)."""
        L.check(L.lib().zj_chorin_set_mg_fused(self.handle, int(bool(on))),
                "zj_chorin_set_mg_fused")

    def set_preconditioner(self, precond: str, net=None):
        """``"mg"`` (FP64 V-cycle), ``"mixed"`` (FP32 V-cycle inside FP64
        flexible CG), ``"learned"`` (FP32 V-cycle plus the BF16 network
        correction of the :class:`~zoomy_hip.learned.MPPreconditioner`
        ``net``, inside FP64 flexible CG) or ``"learned_guess"`` (that
        learned V-cycle applied once to the right-hand side as the initial
        guess of the FP64 V-cycle CG, retried from zero when it misses
        ``tol``).  A low-precision solve whose fresh FP64 residual misses
        ``tol`` is redone with ``"mg"``."""
        if precond not in L.PRECOND:
            raise ValueError(f"precond must be one of {sorted(L.PRECOND)}")
        nh = None
        if precond in ("learned", "learned_guess"):
            if net is None:
                raise ValueError("the learned preconditioner needs a network")
            nh = net.handle
        L.check(L.lib().zj_chorin_set_precond(self.handle, L.PRECOND[precond],
                                              nh), "zj_chorin_set_precond")
        self.precond = precond
        self.net = net

    # ── helpers ───────────────────────────────────────────────────────
    def _q(self, Q):
        if Q.dtype != torch.float64 or Q.device != self.device \
                or tuple(Q.shape) != (5, self.N) or not Q.is_contiguous():
            raise ValueError(f"Q must be a contiguous (5, {self.N}) float64 "
                             "tensor on the GPU")
        return Q

    def _v(self, x, rows=1):
        if x.dtype != torch.float64 or x.device != self.device \
                or x.numel() != rows * self.N or not x.is_contiguous():
            raise ValueError("expected a contiguous float64 device tensor of "
                             f"{rows} x {self.N}")
        return x

    def _check(self, rc, what):
        if rc == L.ZJ_ESOLVE:
            st = self.stats()
            raise EllipticSolveError(
                f"{what}: pressure solve rel. residual {st['relres']:.3e} "
                f"above accept gate {elliptic_accept(self.tol):.1e} after "
                f"{st['iterations']} iterations")
        L.check(rc, what)

    def stats(self):
        s = L.ChorinStats()
        L.check(L.lib().zj_chorin_get_stats(self.handle, ctypes.byref(s)),
                "zj_chorin_get_stats")
        return {k: getattr(s, k) for k, _ in L.ChorinStats._fields_}

    def levels(self):
        n = ctypes.c_int32()
        dims = (ctypes.c_int32 * 64)()
        L.check(L.lib().zj_chorin_levels(self.handle, ctypes.byref(n), dims,
                                         32), "zj_chorin_levels")
        return [(dims[2 * i], dims[2 * i + 1]) for i in range(n.value)]

    # ── the Chorin cycle ──────────────────────────────────────────────
    def _hyperbolic_stage(self, dt, time, Q):
        out = torch.empty_like(self._q(Q))
        self._check(L.lib().zj_chorin_predict(self.handle, float(dt), _p(Q),
                                              _p(out), _stream()),
                    "zj_chorin_predict")
        return out

    def pressure_rhs(self, Qs, dt):
        rhs = torch.empty(self.N, dtype=torch.float64, device=self.device)
        self._check(L.lib().zj_chorin_rhs(self.handle, float(dt),
                                          _p(self._q(Qs)), _p(rhs),
                                          _stream()), "zj_chorin_rhs")
        return rhs

    def _step_pressure_pure(self, Q, dt, time=0.0):
        """Pressure of a predicted state: builds S(h, b), solves
        S p = -D(Q)/dt; returns p (N,)."""
        Q = self._q(Q)
        self.setup_operator(Q[0], Q[1])
        return self.solve(self.pressure_rhs(Q, dt))

    def _step_corrector_pure(self, Q, p, dt, Q0=None):
        out = torch.empty_like(self._q(Q))
        self._check(L.lib().zj_chorin_correct(
            self.handle, float(dt), _p(Q), _p(self._v(p)),
            _p(Q0) if Q0 is not None else None, _p(out), _stream()),
            "zj_chorin_correct")
        return out

    def chorin_stage(self, dt, time, Q, Q0=None):
        out = torch.empty_like(self._q(Q))
        self._check(L.lib().zj_chorin_stage(
            self.handle, float(dt), _p(Q), _p(out),
            _p(self._q(Q0)) if Q0 is not None else None, _stream()),
            "zj_chorin_stage")
        return out

    def chorin_cycle(self, dt, time, Q):
        out = torch.empty_like(self._q(Q))
        self._check(L.lib().zj_chorin_step(self.handle, float(dt), _p(Q),
                                           _p(out), _stream()),
                    "zj_chorin_step")
        return out

    def step(self, dt, time, Q, Qaux=None):
        return self.chorin_cycle(dt, time, Q)

    def run_jit_steps(self, dt, n_steps, Q, t_start=0.0):
        """``n_steps`` fixed-dt cycles; returns ``(Q, t)``."""
        t = float(t_start)
        for _ in range(int(n_steps)):
            Q = self.chorin_cycle(dt, t, Q)
            t += float(dt)
        return Q, t

    # ── the pressure operator ─────────────────────────────────────────
    def setup_operator(self, b, h):
        self._check(L.lib().zj_chorin_setup_operator(
            self.handle, _p(self._v(b.contiguous())),
            _p(self._v(h.contiguous())), _stream()),
            "zj_chorin_setup_operator")

    def apply(self, x):
        y = torch.empty_like(self._v(x))
        self._check(L.lib().zj_chorin_apply(self.handle, _p(x), _p(y),
                                            _stream()), "zj_chorin_apply")
        return y

    def vcycle(self, r):
        if r.dtype not in (torch.float64, torch.float32) or \
                r.numel() != self.N or not r.is_contiguous():
            raise ValueError("r must be a contiguous float32/64 vector")
        z = torch.empty_like(r)
        self._check(L.lib().zj_chorin_vcycle(
            self.handle, _p(r), _p(z),
            L.DTYPES["float64" if r.dtype == torch.float64 else "float32"],
            _stream()), "zj_chorin_vcycle")
        return z

    def learned_vcycle(self, r, train=False):
        """z = M r with the learned V-cycle, z = V r + N(r - S V r) in
        FP32 (the learned preconditioner must be set)."""
        if r.dtype != torch.float32 or r.numel() != self.N or \
                not r.is_contiguous():
            raise ValueError("r must be a contiguous float32 vector")
        z = torch.empty_like(r)
        self._check(L.lib().zj_chorin_learned_vcycle(
            self.handle, _p(r), _p(z), int(bool(train)), _stream()),
            "zj_chorin_learned_vcycle")
        return z

    def learned_vcycle_backward(self, g):
        if g.dtype != torch.float32 or g.numel() != self.N or \
                not g.is_contiguous():
            raise ValueError("g must be a contiguous float32 vector")
        self._check(L.lib().zj_chorin_learned_vcycle_backward(
            self.handle, _p(g), _stream()),
            "zj_chorin_learned_vcycle_backward")

    def solve(self, rhs):
        p = torch.empty_like(self._v(rhs))
        self._check(L.lib().zj_chorin_solve(self.handle, _p(rhs), _p(p),
                                            _stream()), "zj_chorin_solve")
        return p


def fixed_dt(dx: float, h_scale: float, cfl: float = 0.45, g: float = 9.81):
    """Fixed Chorin step dt = CFL dx / (sqrt(g h_scale) + 1), the rule of the
    reference's Chorin march (``tests/cases.py:545``)."""
    return cfl * dx / (np.sqrt(g * h_scale) + 1.0)
