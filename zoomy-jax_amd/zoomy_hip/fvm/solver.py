"""Drop-in host API of the finite-volume time step on MI355X.

Mirrors the reference's operator/solver surface (``zoomy_jax/fvm/
solver_jax.py``) with the arithmetic done by the HIP kernels behind the C ABI:

* ``HyperbolicSolver.setup_simulation(mesh, model) -> (Q, Qaux)``  (:901-988)
* ``get_flux_operator(mesh, runtime) -> flux_operator(dt, time, Q, Qaux,
  parameters, dQ, force_o1=None)``                                  (:563-897)
* ``compute_timestep(Q, Qaux, p=None) -> dt``                        (:1222-1236)
* ``step(dt, time, Q, Qaux, p=None) -> Q``                           (:1095-1131)
* ``post_step(time, dt, Q, Qold, Qaux, p=None) -> (Q, Qaux)``        (:1133-1172)
* ``run_simulation`` / ``solve``                                     (:1238-1368)
* ``IMEXSourceSolver`` with the local Newton source stage
  (``zoomy_jax/fvm/solver_imex_jax.py:268-342``, loop :589-646)

``model`` is a :class:`ShallowWater` description: the reference's
``NumericalSystemModel`` of ``SME(level=0)`` (state ``[b, h, q_0(, q_1)]``)
with its ``ReconstructionSpec(order, limiter)`` and Riemann class.  Arrays are
torch tensors on the GPU, ``(n_var, n_cells)`` float64 (or float32).  dt stays
on the device; the only host synchronisation is the dt guard
(``assert_dt_admissible``, ``zoomy_jax/fvm/solver2/blocks.py:98-124``), which
raises ``FloatingPointError`` exactly like the reference.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np
import torch

from .. import _lib as L
from ..mesh import Mesh


@dataclass
class ShallowWater:
    """SWE numerics description (the reference's NSM for ``SME(level=0)``).

    ``riemann``: ``"rusanov_nc"`` (NonconservativeRusanov: s_M dissipation on
    every row), ``"rusanov"`` (the positive/well-balanced Rusanov: eta jump on
    h, b untouched), ``"hll"`` (HLL with hydrostatic reconstruction).
    ``initial_conditions``: callable ``x (2,) -> state (n_var,)`` or an array.
    """
    dimension: int = 2
    order: int = 1
    limiter: str = "minmod"
    riemann: str = "rusanov_nc"
    g: float = 9.81
    manning_n: float = 0.0
    wet_dry_eps: float = 1e-8
    dt_max: float = 0.0
    initial_conditions: Optional[object] = None
    boundary_conditions: object = "extrapolation"
    # {state row (index or name "h", "hu", "hv"): nu}: Crank-Nicolson
    # diffusion of those rows in IMEXSourceSolver (the model's scalar
    # diffusion operators, solver_imex_jax.py:396-460)
    diffusion: Optional[dict] = None
    # alpha of the momentum source -alpha g h dh/dx_d, which reads the
    # derivative aux rows: a source with derivative specs, which the IMEX
    # solver's "auto" source mode integrates with the global Newton-GMRES
    # stage (solver_imex_jax.py:182-188, :344-396); IMEX only
    depth_gradient_source: float = 0.0

    @property
    def n_state(self):
        return 2 + self.dimension


def _device():
    if not torch.cuda.is_available():
        raise L.ZoomyHipError("the HIP time step needs a GPU (torch.cuda)")
    return torch.device("cuda", torch.cuda.current_device())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def mesh_desc(mesh: Mesh, patch_cells: int = 0, patch_order: str = "auto"):
    """``zj_mesh_desc`` of a host mesh; returns (desc, arrays kept alive).
    ``patch_cells`` / ``patch_order`` ("auto", "consecutive", "clusters")
    steer the stage kernel's patch decomposition (0 / "auto": chosen by the
    library)."""
    N = mesh.n_cells
    ex = mesh.extra_stencil()
    dmask, dval = mesh.bface_dirichlet()
    keep = dict(
        cv=np.ascontiguousarray(mesh.cell_vertices.T, np.int32),
        nbr=np.ascontiguousarray(mesh.cell_nbr.T, np.int32),
        ex=np.ascontiguousarray(ex.T, np.int32),
        verts=np.ascontiguousarray(mesh.verts, np.float64),
        cen=np.ascontiguousarray(mesh.cell_centers().T, np.float64),
        rin=np.ascontiguousarray(mesh.cell_inradius(), np.float64),
        bcell=np.ascontiguousarray(mesh.bface_cell, np.int32),
        bslot=np.ascontiguousarray(mesh.bface_slot, np.int32),
        bkind=np.ascontiguousarray(mesh.bface_kind(), np.int32),
        bpart=np.ascontiguousarray(
            mesh.periodic_partner if mesh.periodic_partner is not None
            else -np.ones(mesh.n_bfaces), np.int32),
        dmask=np.ascontiguousarray(dmask, np.int32),
        dval=np.ascontiguousarray(dval, np.float64),
    )
    d = L.MeshDesc(mesh.dim, N, len(mesh.verts), mesh.max_faces,
                   keep["ex"].shape[0], mesh.n_bfaces,
                   *[keep[n].ctypes.data for n in (
                       "cv", "nbr", "ex", "verts", "cen", "rin", "bcell",
                       "bslot", "bkind", "bpart", "dmask", "dval")],
                   int(patch_cells), L.PATCH_ORDER[patch_order])
    return d, keep


def patch_plan(mesh: Mesh, patch_cells: int = 0, patch_order: str = "auto"):
    """Host-only: the patch decomposition ``zj_mesh_create`` picks."""
    d, keep = mesh_desc(mesh, patch_cells, patch_order)
    st = (ctypes.c_int32 * 7)()
    rc = L.lib().zj_patch_plan(ctypes.byref(d),
                               ctypes.cast(st, ctypes.c_void_p))
    return rc, dict(zip(("P", "npatch", "ntemplates", "max_load",
                         "max_slope", "max_vert", "clustered"), list(st)))


def patch_stats(mesh: Mesh, P: int = 256):
    """Host-only build + validation of the LDS patch decomposition."""
    d, keep = mesh_desc(mesh)
    st = (ctypes.c_int32 * 6)()
    rc = L.lib().zj_patch_check(ctypes.byref(d), int(P),
                                ctypes.cast(st, ctypes.c_void_p))
    return rc, dict(zip(("P", "npatch", "ntemplates", "max_load",
                         "max_slope", "max_vert"), list(st)))


class DeviceMesh:
    """Mesh uploaded to the device (``zj_mesh_create``)."""

    def __init__(self, mesh: Mesh, patch_cells: int = 0,
                 patch_order: str = "auto"):
        self.mesh = mesh
        lib = L.lib()
        d, self._keep = mesh_desc(mesh, patch_cells, patch_order)
        h = ctypes.c_void_p()
        L.check(lib.zj_mesh_create(ctypes.byref(d), ctypes.byref(h)),
                "zj_mesh_create")
        self.handle = h
        self.cell_volumes = mesh.cell_volumes()
        self.cell_centers = mesh.cell_centers()

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and L is not None and L._lib is not None:
            L._lib.zj_mesh_destroy(h)
            self.handle = None


class _Runtime:
    """Solver context (``zj_solver_create``): mesh + numerics + scratch."""

    def __init__(self, dmesh: DeviceMesh, model: ShallowWater, cfl: float,
                 dtype: torch.dtype, recon: str):
        lib = L.lib()
        self.dmesh = dmesh
        self.model = model
        self.dtype = dtype
        self.params = L.Params(
            float(model.g), float(cfl), float(model.wet_dry_eps),
            float(model.dt_max), float(model.manning_n), int(model.order),
            L.LIMITERS[model.limiter], L.RIEMANN[model.riemann],
            L.RECON[recon],
            L.DTYPES["float64" if dtype == torch.float64 else "float32"])
        h = ctypes.c_void_p()
        L.check(lib.zj_solver_create(dmesh.handle, ctypes.byref(self.params),
                                     ctypes.byref(h)), "zj_solver_create")
        self.handle = h

    def _face_args(self, qL, qR, qauxL, qauxR, normal):
        dev = _device()
        n = qL.shape[1]
        dim = self.dmesh.mesh.dim

        def rows4(q):
            q = torch.as_tensor(q, dtype=self.dtype, device=dev)
            if q.shape[0] == 4:
                return q.contiguous()
            out = torch.zeros((4, n), dtype=self.dtype, device=dev)
            out[:q.shape[0]] = q
            return out

        def aux(a):
            if a is None:
                return None
            a = torch.as_tensor(a, dtype=self.dtype, device=dev).contiguous()
            if a.shape != (3 * dim + 1, n):
                raise ValueError(f"qaux must be ({3 * dim + 1}, {n}), got "
                                 f"{tuple(a.shape)}")
            return a
        a, b = rows4(qL), rows4(qR)
        if a.shape != (4, n) or b.shape != (4, n):
            raise ValueError("qL and qR must be (n_var, n) face states")
        nrm = torch.as_tensor(normal, dtype=torch.float64,
                              device=dev)[:dim].contiguous()
        if nrm.shape != (dim, n):
            raise ValueError(f"normal must be ({dim}, {n})")
        return n, a, b, aux(qauxL), aux(qauxR), nrm

    def numerical_flux(self, qL, qR, qauxL, qauxR, parameters, normal):
        """The runtime slot ``numerical_flux(qL, qR, qauxL, qauxR, p, n)``
        (``jax_runtime.py:789-795``): the conservative face flux (n_var, n)
        of the solver's Riemann solver on given face states."""
        n, a, b, xa, xb, nrm = self._face_args(qL, qR, qauxL, qauxR, normal)
        F = torch.empty_like(a)
        L.check(L.lib().zj_face_numerics(
            self.handle, n, _p(a), _p(b), _p(xa), _p(xb), _p(nrm), _p(F),
            None, None, _stream()), "zj_face_numerics")
        return F[:self.model.n_state]

    def numerical_fluctuations(self, qL, qR, qauxL, qauxR, parameters,
                               normal):
        """The runtime slot ``numerical_fluctuations`` (``jax_runtime.py:
        797-803``): the fluctuations (Dp, Dm) stacked as (2, n_var, n)."""
        n, a, b, xa, xb, nrm = self._face_args(qL, qR, qauxL, qauxR, normal)
        D = torch.empty((2, 4, n), dtype=self.dtype, device=a.device)
        L.check(L.lib().zj_face_numerics(
            self.handle, n, _p(a), _p(b), _p(xa), _p(xb), _p(nrm), None,
            _p(D[0]), _p(D[1]), _stream()), "zj_face_numerics")
        return D[:, :self.model.n_state]

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and L is not None and L._lib is not None:
            L._lib.zj_solver_destroy(h)
            self.handle = None


class HyperbolicSolver:
    """Explicit FV march on the GPU (``zoomy_jax.fvm.solver_jax.
    HyperbolicSolver``).  ``cfl`` is the reference's ``CFL`` of
    ``timestepping.adaptive(CFL, dimension)`` (the 1/d factor is applied
    inside, ``tests/conftest.py:80-93``)."""

    def __init__(self, time_end: float = 1.0, cfl: float = 0.9,
                 dtype: torch.dtype = torch.float64,
                 reconstruction_variables: str = "conservative",
                 compute_dt: Optional[Callable] = None, mood: bool = False,
                 positivity_method: str = "",
                 front_theta_tol: Optional[float] = None,
                 stage_aux: str = "refresh", patch_cells: int = 0,
                 patch_order: str = "auto", output_directory: str = "output",
                 output_filename: str = "simulation", snapshots: int = 10):
        """``reconstruction_variables``: ``"conservative"``, ``"eta"``
        (well-balanced wet/dry recipe on (b, eta, u)), ``"eta_hu"`` (the
        same on (b, eta, hu) with the 1e-3 depth threshold of the class it
        replaced) or ``"xz"`` (conservative with the Xing-Zhang cell-mean
        positivity scaling).  ``positivity_method`` (``""`` or
        ``"zhang_shu"``) and ``front_theta_tol`` (None or a value in (0, 1])
        act on the eta reconstructions only, as in the reference
        (``solver_jax.py:196-226``).  ``stage_aux``: ``"refresh"`` (the
        second SSP-RK2 stage's fluxes read that stage's 1/h,
        ``solver_jax.py:1026-1041``) or ``"frozen"`` (the step-start aux,
        with which the reference's order-2 goldens were recorded).
        ``patch_cells`` / ``patch_order`` steer the device layout only (the
        stage kernel's patches; results do not depend on them).
        ``output_directory``, ``output_filename`` and ``snapshots`` are the
        reference's ``settings.output`` fields read by
        ``run_simulation(write_output=True)`` (``solver_jax.py:1264-1292``)."""
        if stage_aux not in L.STAGE_AUX:
            raise ValueError(
                f"stage_aux must be one of {sorted(L.STAGE_AUX)}")
        self.stage_aux = stage_aux
        if patch_order not in L.PATCH_ORDER:
            raise ValueError(
                f"patch_order must be one of {sorted(L.PATCH_ORDER)}")
        self.patch_cells = int(patch_cells)
        self.patch_order = patch_order
        if reconstruction_variables not in L.RECON:
            raise ValueError(
                f"reconstruction_variables must be one of {sorted(L.RECON)}")
        if positivity_method not in L.POSITIVITY:
            raise ValueError(
                f"positivity_method must be one of {sorted(L.POSITIVITY)}")
        if front_theta_tol is not None and not 0.0 < front_theta_tol <= 1.0:
            raise ValueError("front_theta_tol must be None or in (0, 1]")
        self.positivity_method = positivity_method
        self.front_theta_tol = front_theta_tol
        self.time_end = float(time_end)
        self.cfl = float(cfl)
        self.dtype = dtype
        self.reconstruction_variables = reconstruction_variables
        self.compute_dt = compute_dt
        # a-posteriori MOOD (solver_jax.py:1055-1071): redo a step with the
        # troubled cells (h < 0 or non-finite) at first order
        self.mood = bool(mood)
        if int(snapshots) < 1:
            raise ValueError("snapshots must be >= 1")
        self.output_directory = output_directory
        self.output_filename = output_filename
        self.snapshots = int(snapshots)

    # ── setup ─────────────────────────────────────────────────────────
    def initialize(self, mesh: Mesh, model: ShallowWater):
        N = mesh.n_cells
        ic = model.initial_conditions
        if ic is None:
            Q = np.zeros((4, N))
        elif callable(ic):
            xc = mesh.cell_centers()
            rows = np.array([np.asarray(ic(xc[i]), np.float64)
                             for i in range(N)]).T
            Q = np.zeros((4, N))
            Q[:rows.shape[0]] = rows
        else:
            ic = np.asarray(ic, np.float64)
            Q = np.zeros((4, N))
            Q[:ic.shape[0]] = ic
        return Q

    def setup_simulation(self, mesh: Mesh, model: ShallowWater):
        if model.depth_gradient_source and not isinstance(
                self, IMEXSourceSolver):
            raise NotImplementedError(
                "the depth-gradient source is integrated implicitly by "
                "IMEXSourceSolver only")
        if model.dimension != mesh.dim:
            raise ValueError(f"model dimension {model.dimension} does not "
                             f"match mesh dimension {mesh.dim}")
        if not mesh.bc:
            mesh.set_bc(model.boundary_conditions)
        self.mesh = mesh
        self.model = model
        self.n_state = model.n_state
        self._dmesh = DeviceMesh(mesh, self.patch_cells, self.patch_order)
        self._rt = self._make_runtime(model)
        dev = _device()
        self.device = dev
        self._dt = torch.zeros((), dtype=torch.float64, device=dev)
        self._t = torch.zeros((), dtype=torch.float64, device=dev)
        self._troubled = torch.zeros((2,), dtype=torch.int32, device=dev)
        self._mass = torch.zeros((), dtype=torch.float64, device=dev)
        Q = torch.as_tensor(self.initialize(mesh, model), dtype=self.dtype,
                            device=dev).contiguous()
        Qaux = torch.zeros((3 * mesh.dim + 1, mesh.n_cells),
                           dtype=self.dtype, device=dev)
        Qaux = self.update_qaux(Q, Qaux)
        return self._view(Q), Qaux

    def _make_runtime(self, model):
        """A solver context with this solver's numerics options (MOOD,
        stage aux, positivity guards) applied."""
        rt = _Runtime(self._dmesh, model, self.cfl, self.dtype,
                      self.reconstruction_variables)
        lib = L.lib()
        if self.mood:
            L.check(lib.zj_solver_set_mood(rt.handle, 1),
                    "zj_solver_set_mood")
        L.check(lib.zj_solver_set_stage_aux(
            rt.handle, L.STAGE_AUX[self.stage_aux]),
            "zj_solver_set_stage_aux")
        if self.positivity_method or self.front_theta_tol is not None:
            L.check(lib.zj_solver_set_positivity(
                rt.handle, L.POSITIVITY[self.positivity_method],
                -1.0 if self.front_theta_tol is None
                else float(self.front_theta_tol)), "zj_solver_set_positivity")
        return rt

    @property
    def runtime(self):
        return self._rt

    # ── state helpers ─────────────────────────────────────────────────
    def _full(self, Q):
        """(4, N) contiguous device tensor for the kernels."""
        if Q.dtype != self.dtype or Q.device != self.device:
            Q = Q.to(device=self.device, dtype=self.dtype)
        if Q.shape[0] == 4 and Q.is_contiguous():
            return Q
        base = getattr(Q, "_base", None)
        if Q.shape[0] == 3 and base is not None and base.shape[0] == 4 \
                and base.is_contiguous() and Q.data_ptr() == base.data_ptr():
            return base
        out = torch.zeros((4, Q.shape[1]), dtype=self.dtype,
                          device=self.device)
        out[:Q.shape[0]] = Q
        return out

    def _view(self, Q4):
        return Q4[:self.n_state] if self.n_state == 3 else Q4

    def _dt_dev(self, dt):
        if torch.is_tensor(dt) and dt.device == self.device and \
                dt.dtype == torch.float64 and dt.numel() == 1:
            return dt.reshape(())
        self._dt.fill_(float(dt))
        return self._dt

    # ── operators ─────────────────────────────────────────────────────
    @property
    def supports_force_o1(self) -> bool:
        """Whether the reconstruction honours a per-cell first-order mask:
        the eta reconstructions do, the others ignore it, as in the
        reference (``solver_jax.py:580-582``, ``reconstruction_jax.py:
        1058-1062``)."""
        return self.reconstruction_variables in ("eta", "eta_hu")

    @property
    def runtime(self):
        """The solver's runtime: its ``numerical_flux`` /
        ``numerical_fluctuations`` slots evaluate the Riemann solver on
        given face states (``transformation/jax_runtime.py:789-806``)."""
        return self._rt

    def get_flux_operator(self, mesh=None, runtime=None):
        lib = L.lib()

        def flux_operator(dt, time, Q, Qaux, parameters=None, dQ=None,
                          force_o1=None):
            Q4 = self._full(Q)
            out = torch.empty_like(Q4)
            if force_o1 is not None and self.supports_force_o1:
                m = torch.as_tensor(force_o1, device=self.device)
                if m.numel() != Q4.shape[1]:
                    raise ValueError("force_o1 must have one entry per cell")
                m = m.reshape(-1).to(torch.uint8).contiguous()
                L.check(lib.zj_flux_operator_o1(self._rt.handle, _p(Q4),
                                                _p(m), _p(out), _stream()),
                        "zj_flux_operator_o1")
            else:
                L.check(lib.zj_flux_operator(self._rt.handle, _p(Q4),
                                             _p(out), _stream()),
                        "zj_flux_operator")
            return self._view(out)

        return flux_operator

    def _faces(self):
        """The reference-layout face arrays on the device (uploaded once)."""
        if getattr(self, "_meshrt", None) is None:
            from .blocks import prepare_mesh
            self._meshrt = prepare_mesh(self.mesh, self)
        return self._meshrt

    def get_compute_max_abs_eigenvalue(self, mesh=None, runtime=None):
        """Per-face max |eigenvalue| of both cell states, ``(n_faces,)`` in
        the order of ``Mesh.faces()`` (``solver_jax.py:434-457``)."""
        from .blocks import dt_pass

        def compute_max_abs_eigenvalue(Q, Qaux, parameters=None):
            lo, hi = dt_pass(Q, Qaux, parameters, self._faces())
            return torch.maximum(lo.abs(), hi.abs())

        return compute_max_abs_eigenvalue

    def get_reconstruction(self):
        """The reconstruction object of this solver's numerics
        (``reconstruction_jax.py:398-429``, ``:1307-1366``)."""
        return Reconstruction(self)

    def compute_timestep(self, Q, Qaux=None, parameters=None,
                         t_remaining: Optional[float] = None):
        Q4 = self._full(Q)
        dt = torch.empty((), dtype=torch.float64, device=self.device)
        L.check(L.lib().zj_compute_dt(
            self._rt.handle, _p(Q4),
            -1.0 if t_remaining is None else float(t_remaining), _p(dt),
            _stream()), "zj_compute_dt")
        return dt

    def step(self, dt, time, Q, Qaux=None, parameters=None):
        Q4 = self._full(Q)
        out = torch.empty_like(Q4)
        if self.n_state == 3:
            out[3].zero_()
        L.check(L.lib().zj_step(self._rt.handle, _p(self._dt_dev(dt)),
                                _p(Q4), _p(out), _p(self._troubled),
                                _stream()), "zj_step")
        return self._view(out)

    def mood_stats(self):
        """(steps redone, cells demoted) since setup (MOOD on)."""
        out = np.zeros(2, np.int32)
        L.check(L.lib().zj_solver_mood_stats(self._rt.handle,
                                             out.ctypes.data),
                "zj_solver_mood_stats")
        return int(out[0]), int(out[1])

    def troubled_count(self) -> int:
        """Cells of the last step with h < 0 or a non-finite component
        (``zoomy_jax/fvm/solver2/blocks.py:290-301``); synchronises."""
        return int(self._troubled[0].item())

    def _check_qaux(self, Qaux, n_cells):
        """The kernels write (3 dim + 1, N) words of the solver's dtype
        through Qaux's pointer: anything else is refused."""
        shape = (3 * self.mesh.dim + 1, n_cells)
        if not torch.is_tensor(Qaux):
            raise ValueError("Qaux must be a torch tensor on the solver's "
                             "device")
        if tuple(Qaux.shape) != shape or Qaux.dtype != self.dtype or \
                Qaux.device != self.device or not Qaux.is_contiguous():
            raise ValueError(
                f"Qaux must be a contiguous {self.dtype} tensor of shape "
                f"{shape} on {self.device}; got {Qaux.dtype} "
                f"{tuple(Qaux.shape)} on {Qaux.device}"
                f"{'' if Qaux.is_contiguous() else ' (non-contiguous)'}")
        return Qaux

    def update_qaux(self, Q, Qaux):
        Q4 = self._full(Q)
        if Qaux is None:
            Qaux = torch.zeros((3 * self.mesh.dim + 1, Q4.shape[1]),
                               dtype=self.dtype, device=self.device)
        self._check_qaux(Qaux, Q4.shape[1])
        L.check(L.lib().zj_update_aux(self._rt.handle, _p(Q4), _p(Qaux),
                                      _stream()), "zj_update_aux")
        return Qaux

    def post_step(self, time, dt, Q, Qold, Qaux, parameters=None):
        return Q, self.update_qaux(Q, Qaux)

    def assert_dt_admissible(self, dt, time=0.0, iteration=0):
        d = float(dt)
        if np.isfinite(d) and d > 0.0:
            return d
        raise FloatingPointError(
            f"reduce_dt returned dt = {d!r} at t = {float(time)} (iteration "
            f"{int(iteration)}) -- FATAL: a non-positive or non-finite dt "
            f"aborts the march")

    def time_kernels(self, Q, dt, reps: int = 20):
        """Average ms per launch of {stage1, stage2, dt, aux, stage1_aux}
        measured with HIP events on the current stream."""
        Q4 = self._full(Q)
        ms = (ctypes.c_float * 5)()
        L.check(L.lib().zj_time_kernels(self._rt.handle, _p(Q4),
                                        _p(self._dt_dev(dt)), int(reps),
                                        ctypes.cast(ms, ctypes.c_void_p),
                                        _stream()), "zj_time_kernels")
        return dict(zip(("stage1", "stage2", "dt", "aux", "stage1_aux"),
                        list(ms)))

    def compute_mass(self, Q):
        """Compensated FP64 sum of h |cell| (device scalar)."""
        Q4 = self._full(Q)
        L.check(L.lib().zj_mass(self._rt.handle, _p(Q4), _p(self._mass),
                                _stream()), "zj_mass")
        return self._mass.clone()

    # ── loops ─────────────────────────────────────────────────────────
    def march_device(self, Q, n_steps: int, t0: float = 0.0,
                     time_end: Optional[float] = None, Qaux=None):
        """``n_steps`` adaptive steps entirely on the device, dt capped by
        ``time_end - t`` (``run_simulation``'s loop body, with the post_step
        aux refresh into ``Qaux`` when given).  Q is updated in place;
        returns ``(Q, t_device, dt_device)``."""
        Q4 = self._full(Q)
        if Qaux is not None:
            self._check_qaux(Qaux, Q4.shape[1])
        self._t.fill_(float(t0))
        te = self.time_end if time_end is None else time_end
        L.check(L.lib().zj_march(self._rt.handle, _p(Q4), _p(Qaux),
                                 int(n_steps),
                                 float(te) if np.isfinite(te) else -1.0,
                                 _p(self._t), _p(self._dt),
                                 _p(self._troubled), _stream()), "zj_march")
        return self._view(Q4), self._t, self._dt

    def run_simulation(self, Q, Qaux, write_output=False, parameters=None,
                       max_steps: int = 10 ** 9):
        """Adaptive march to ``time_end``; the dt guard reads dt each step
        (the solver2 march's host guard).  With ``write_output``, HDF5
        snapshots go to ``<output_directory>/<output_filename>.h5`` at the
        reference's schedule (``solver_jax.py:1264-1322``): the initial
        state, then after every step that reaches the next multiple of
        ``time_end / (snapshots - 1)``."""
        save = None
        if write_output:
            import os
            from ..io import get_save_fields
            os.makedirs(self.output_directory, exist_ok=True)
            save = get_save_fields(os.path.join(
                self.output_directory, f"{self.output_filename}.h5"))
            dt_snapshot = self.time_end / max(self.snapshots - 1, 1)
            i_snap = save(0.0, 0.0, 0.0, Q, Qaux)
        t = 0.0
        it = 0
        try:
            while t < self.time_end and it < max_steps:
                dt = self.compute_timestep(Q, Qaux,
                                           t_remaining=self.time_end - t)
                d = self.assert_dt_admissible(dt, t, it)
                Qn = self.step(dt, t, Q, Qaux)
                Q, Qaux = self.post_step(t + d, d, Qn, Q, Qaux)
                t += d
                it += 1
                if save is not None:
                    i_snap = save(t, i_snap * dt_snapshot, i_snap, Q, Qaux)
        finally:
            if save is not None:
                save.writer.close()
        self.n_steps_taken = it
        return Q, Qaux

    def solve(self, mesh: Mesh, model: ShallowWater, write_output=False):
        Q, Qaux = self.setup_simulation(mesh, model)
        return self.run_simulation(Q, Qaux, write_output=write_output)


class Reconstruction:
    """Reconstruction object over the HIP reconstruction kernel:
    ``__call__(Q, bf_face_values, force_o1=None, Qaux=None, parameters=None,
    bf_aux=None) -> (Q_L, Q_R)`` and ``reconstruct_with_grad(...) -> (Q_L,
    Q_R, lim_grad)`` with face arrays ``(n_var, n_faces)`` in the order of
    ``Mesh.faces()`` and ``lim_grad (n_var, dim, n_cells)`` in conservative
    variables (``LSQMUSCLReconstructionJAX``, ``reconstruction_jax.py:
    398-429``; the eta reconstruction ``EmittedWLSQMUSCLJAX``, :1307-1366).
    The boundary values are computed on the device from Q (the BC ghosts of
    the cell states, what every reference caller passes); Q_R of a boundary
    face is that ghost."""

    supports_grad_recon = True

    def __init__(self, solver: HyperbolicSolver):
        self.solver = solver
        self.supports_force_o1 = solver.supports_force_o1

    def _run(self, Q, force_o1, with_grad):
        from .blocks import build_operators, reconstruct
        s = self.solver
        meshrt = s._faces()
        ops = build_operators(s)
        QL, QR, lg = reconstruct(
            Q, None, meshrt, ops, None,
            force_o1=force_o1 if self.supports_force_o1 else None)
        n = s.n_state
        QL, QR = QL[:n], QR[:n]
        if not with_grad:
            return QL, QR
        if lg is None:
            lg = torch.zeros((4, meshrt.dimension, meshrt.n_cells),
                             dtype=s.dtype, device=s.device)
        return QL, QR, lg[:n]

    def __call__(self, Q, bf_face_values=None, force_o1=None, Qaux=None,
                 parameters=None, bf_aux=None):
        return self._run(Q, force_o1, False)

    def reconstruct_with_grad(self, Q, bf_face_values=None, force_o1=None,
                              Qaux=None, parameters=None, bf_aux=None):
        return self._run(Q, force_o1, True)


class IMEXSourceSolver(HyperbolicSolver):
    """Explicit hyperbolic step, Crank-Nicolson diffusion of the rows the
    model lists in ``diffusion`` and the local implicit source, in the
    reference's order (``zoomy_jax/fvm/solver_imex_jax.py:604-625``;
    ``implicit_maxiter = 6``, :142)."""

    implicit_maxiter = 6
    # the global stage (solver_imex_jax.py:140-146)
    source_mode = "auto"          # "auto" | "local" | "global"
    implicit_tol = 1e-8
    gmres_tol = 1e-7
    gmres_maxiter = 30
    gmres_restart = 20
    _ROWS = {"b": 0, "h": 1, "hu": 2, "q_0": 2, "hv": 3, "q_1": 3}

    def resolved_source_mode(self) -> str:
        """"global" for a source that reads derivative aux rows, else
        "local" (``_resolve_source_mode``, solver_imex_jax.py:182-188)."""
        if self.source_mode in ("local", "global"):
            return self.source_mode
        return "global" if self.model.depth_gradient_source else "local"

    def step(self, dt, time, Q, Qaux=None, parameters=None):
        Q4 = self._full(Q)
        Qe = self._hyperbolic_only(dt, Q4)   # explicit stage, no source
        for row, op in self._diffusion:
            Qe[row] = op.implicit_solve(Qe[row], float(dt))
        out = torch.empty_like(Q4)
        if self._global is not None:
            L.check(L.lib().zj_imex_global_solve(
                self._global, float(dt), _p(Qe), _p(out),
                int(self.implicit_maxiter), float(self.implicit_tol),
                float(self.gmres_tol), int(self.gmres_maxiter), _stream()),
                "zj_imex_global_solve")
            return self._view(out)
        L.check(L.lib().zj_imex_local(self._rt.handle, _p(self._dt_dev(dt)),
                                      int(self.implicit_maxiter), _p(Qe),
                                      _p(out), _stream()), "zj_imex_local")
        return self._view(out)

    def implicit_stats(self):
        """{newton, gmres, fallbacks, rnorm} of the last global stage."""
        if self._global is None:
            return None
        st = np.zeros(3, np.int32)
        rn = ctypes.c_double()
        L.check(L.lib().zj_imex_global_stats(self._global, st.ctypes.data,
                                             ctypes.byref(rn)),
                "zj_imex_global_stats")
        return {"newton": int(st[0]), "gmres": int(st[1]),
                "fallbacks": int(st[2]), "rnorm": float(rn.value)}

    def __del__(self):
        g = getattr(self, "_global", None)
        if g is not None and L is not None and L._lib is not None:
            L._lib.zj_imex_global_destroy(g)
            self._global = None

    def setup_simulation(self, mesh, model):
        out = super().setup_simulation(mesh, model)
        # a second runtime without the explicit source for the hyperbolic
        # stage (the implicit stage owns the friction)
        import copy
        m2 = copy.copy(model)
        m2.manning_n = 0.0
        # the explicit stage keeps MOOD, the positivity guards and the stage
        # aux, as the reference's IMEX explicit stage does
        # (solver_imex_jax.py:604-611)
        self._rt_h = self._make_runtime(m2)
        self._global = None
        if self.source_mode not in ("auto", "local", "global"):
            raise ValueError("source_mode must be auto, local or global")
        if self.resolved_source_mode() == "global":
            if self.dtype != torch.float64:
                raise ValueError("the global implicit source stage is FP64")
            g = ctypes.c_void_p()
            L.check(L.lib().zj_imex_global_create(
                self._rt.handle, float(model.depth_gradient_source),
                int(self.gmres_restart), ctypes.byref(g)),
                "zj_imex_global_create")
            self._global = g
        self._diffusion = []
        if model.diffusion:
            if self.dtype != torch.float64:
                raise ValueError("implicit diffusion runs in FP64")
            from .diffusion import DiffusionOperator
            for key, nu in model.diffusion.items():
                row = self._ROWS[key] if isinstance(key, str) else int(key)
                if not 0 <= row < model.n_state:
                    raise ValueError(f"diffusion row {key!r} out of range")
                self._diffusion.append(
                    (row, DiffusionOperator(mesh, mesh.dim, float(nu))))
        return out

    def _hyperbolic_only(self, dt, Q4):
        out = torch.empty_like(Q4)
        L.check(L.lib().zj_step(self._rt_h.handle, _p(self._dt_dev(dt)),
                                _p(Q4), _p(out), _p(self._troubled),
                                _stream()), "zj_step")
        return out
