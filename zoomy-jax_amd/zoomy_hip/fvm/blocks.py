"""The reference's block contract on the HIP kernels.

Mirrors ``zoomy_jax/fvm/solver2`` -- the same numerics as named blocks over
stored per-face arrays (``blocks.py:36-323``), the runtime mesh view
(``context.py:38-163``) and the march that drives them (``march.py:44-250``):

* ``prepare_mesh(mesh, solver) -> MeshRT`` and ``build_operators(solver) ->
  Ops`` (``context.py:99-163``);
* ``dt_pass(Q, Qaux, p, MeshRT, Ops) -> (lam_lo_f, lam_hi_f)``,
  ``reduce_dt(...)`` and ``assert_dt_admissible(...)`` (the step head);
* ``halo_bc``, ``reconstruct``, ``flux_pass``, ``gather_update``,
  ``mood_resolve`` (the stage blocks) and ``update_aux``;
* ``MarchSolver`` with ``setup_simulation`` / ``march``.

Every block is one kernel launch through the C ABI (``zj_face_speeds``,
``zj_reduce_dt``, ``zj_bc_values``, ``zj_reconstruct``, ``zj_flux_pass``,
``zj_gather_update``); arrays are torch tensors on the GPU in the reference's
layouts: states ``(4, N)``, face arrays ``(4, F)`` in the face order of
``Mesh.faces()``, ``lim_grad (4, dim, N)``.  dt stays a device scalar; the
only host reads are the dt guard and MOOD's troubled test, where the
reference reads too.  The fused stage kernel behind ``HyperbolicSolver`` runs
the same step without storing the face arrays.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from time import perf_counter
from typing import Any, Optional

import numpy as np
import torch

from .. import _lib as L
from .solver import HyperbolicSolver, _p, _stream

# Shu-Osher stage tableaus (alpha, beta) of context.py:34-35
TABLEAU_EULER = ((0.0, 1.0),)
TABLEAU_SSPRK2 = ((0.0, 1.0), (0.5, 1.0))


@dataclass(frozen=True)
class MeshRT:
    """Runtime mesh view (``context.py:38-79``): the reference-layout face
    arrays on the device, bound to the solver context whose kernels use
    them."""
    dimension: int
    n_cells: int
    n_faces: int
    n_boundary_faces: int
    face_owner: Any          # (F,) int32, device
    face_neigh: Any          # (F,) int32, device; == owner on boundary faces
    interior_faces: Any      # (F_int,) int64, device
    boundary_faces: Any      # (F_bnd,) int64, device
    face_normals: Any        # (2, F) float64, device
    face_volumes: Any        # (F,) float64, device
    cell_volumes: Any        # (N,) float64, device
    inradius_f: Any          # (F,) float64, device: min(r_A, r_B)
    solver: Any = None       # the HyperbolicSolver whose context runs the blocks
    mesh: Any = None


@dataclass
class Ops:
    """Operator bundle (``context.py:82-96``): here the solver context and
    the scheme switches it was built with."""
    solver: Any
    order: int
    n_state: int
    h_index: Optional[int]
    tableau: Any
    use_interior_ncp: bool


def _rt(meshrt):
    return meshrt.solver.runtime.handle


def prepare_mesh(mesh, solver: HyperbolicSolver) -> MeshRT:
    """Uploads the face list of ``mesh`` (``Mesh.faces()``) to the solver's
    device mesh (``zj_mesh_set_faces``) and returns the runtime view
    (``context.py:99-163``; a boundary face carries its owner on both
    sides)."""
    fc, fs, bff = mesh.faces()
    fc = np.ascontiguousarray(fc, np.int32)
    fs = np.ascontiguousarray(fs, np.int32)
    F = fc.shape[1]
    L.check(L.lib().zj_mesh_set_faces(solver._dmesh.handle, int(F),
                                      fc.ctypes.data, fs.ctypes.data),
            "zj_mesh_set_faces")
    dev = solver.device
    neigh = np.where(fc[1] >= 0, fc[1], fc[0])
    nrm, length, _ = mesh.slot_geometry()
    r = mesh.cell_inradius()
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a),
                                                    dtype=dt, device=dev)
    return MeshRT(
        dimension=mesh.dim, n_cells=mesh.n_cells, n_faces=F,
        n_boundary_faces=mesh.n_bfaces,
        face_owner=t(fc[0], torch.int32), face_neigh=t(neigh, torch.int32),
        interior_faces=t(np.nonzero(fc[1] >= 0)[0], torch.int64),
        boundary_faces=t(bff, torch.int64),
        face_normals=t(nrm[fc[0], fs[0]].T),
        face_volumes=t(length[fc[0], fs[0]]),
        cell_volumes=t(mesh.cell_volumes()),
        inradius_f=t(np.minimum(r[fc[0]], r[neigh])),
        solver=solver, mesh=mesh)


def build_operators(solver: HyperbolicSolver) -> Ops:
    """``build_operators`` (``context.py``): the numerics of the solver's
    context."""
    m = solver.model
    return Ops(solver=solver, order=int(m.order), n_state=m.n_state,
               h_index=1,
               tableau=TABLEAU_SSPRK2 if m.order >= 2 else TABLEAU_EULER,
               use_interior_ncp=m.order >= 2)


def _hinv_row(Qaux):
    """The aux 1/h row (the last row of ``Qaux``), or None."""
    if Qaux is None:
        return None
    return Qaux[-1].contiguous()


def _dev4(Q, meshrt):
    return meshrt.solver._full(Q)


# ── step head ───────────────────────────────────────────────────────────
def dt_pass(Q, Qaux, p, MeshRT, Ops=None):
    """Per-face bounds ``(lam_lo_f, lam_hi_f)`` of the eigenvalues of both
    cell states (``blocks.py:36-61``)."""
    s = MeshRT.solver
    Q4 = _dev4(Q, MeshRT)
    lo = torch.empty(MeshRT.n_faces, dtype=s.dtype, device=s.device)
    hi = torch.empty_like(lo)
    ha = _hinv_row(Qaux)
    L.check(L.lib().zj_face_speeds(_rt(MeshRT), _p(Q4), _p(ha), _p(lo),
                                   _p(hi), _stream()), "zj_face_speeds")
    return lo, hi


def reduce_dt(lam_lo_f, lam_hi_f, inradius_f, dt_max, *, CFL, dimension,
              degree=0, t_remaining=None, dt_window=None, MeshRT=None):
    """The dt reduction (``blocks.py:64-95``): ``min_f CFL 2 r_f / (d (2k+1)
    |lam|max)``, capped by ``dt_max`` and by EITHER the remaining time OR the
    coupling window.  Returns a device scalar.  ``MeshRT`` names the solver
    context whose kernels run it (the one ``inradius_f`` came from)."""
    if t_remaining is not None and dt_window is not None:
        raise ValueError("reduce_dt takes t_remaining XOR dt_window: "
                         "dt_window replaces the t_end clamp")
    if MeshRT is None:
        raise ValueError("reduce_dt needs the MeshRT of its face arrays")
    s = MeshRT.solver
    clamp = t_remaining if t_remaining is not None else dt_window
    out = torch.empty((), dtype=torch.float64, device=s.device)
    L.check(L.lib().zj_reduce_dt(
        _rt(MeshRT), _p(lam_lo_f.contiguous()), _p(lam_hi_f.contiguous()),
        _p(inradius_f.contiguous()), float(CFL), int(dimension), int(degree),
        float(dt_max) if dt_max is not None else -1.0,
        float(clamp) if clamp is not None else -1.0, _p(out), _stream()),
        "zj_reduce_dt")
    return out


def assert_dt_admissible(dt, lam_lo_f, lam_hi_f, inradius_f, *, time,
                         iteration):
    """The march's dt guard (``blocks.py:98-124``): a non-positive or
    non-finite dt is fatal, with the worst face named."""
    d = float(dt)
    if np.isfinite(d) and d > 0.0:
        return d
    lo = lam_lo_f.double().cpu().numpy()
    hi = lam_hi_f.double().cpu().numpy()
    r = inradius_f.double().cpu().numpy()
    amax = np.maximum(np.abs(lo), np.abs(hi))
    with np.errstate(divide="ignore", invalid="ignore"):
        local = 2.0 * r / amax
    finite = np.isfinite(local)
    bad = (int(np.argmin(np.where(finite, local, np.inf))) if finite.any()
           else int(np.argmax(np.nan_to_num(amax, nan=np.inf))))
    raise FloatingPointError(
        f"reduce_dt returned dt = {d!r} at t = {float(time)} (iteration "
        f"{int(iteration)}) -- FATAL. Worst face f = {bad}: lam_lo = "
        f"{float(lo[bad])!r}, lam_hi = {float(hi[bad])!r}, inradius = "
        f"{float(r[bad])!r}.")


# ── stage blocks ────────────────────────────────────────────────────────
def halo_bc(Q, Qaux, p, MeshRT, t, Ops=None):
    """Ghost values per boundary face (``blocks.py:129-157``): returns
    ``(Q, Qaux, bf_values (4, n_bf))``."""
    s = MeshRT.solver
    Q4 = _dev4(Q, MeshRT)
    bf = torch.zeros((4, max(MeshRT.n_boundary_faces, 1)), dtype=s.dtype,
                     device=s.device)
    if MeshRT.n_boundary_faces:
        L.check(L.lib().zj_bc_values(_rt(MeshRT), _p(Q4), _p(bf), _stream()),
                "zj_bc_values")
    return Q, Qaux, bf


def reconstruct(Q, Qaux, MeshRT, Ops, bf_values, *, o1=False,
                force_o1=None):
    """Face states ``(Q_L, Q_R, lim_grad)`` (``blocks.py:160-182``):
    ``lim_grad`` is the limited cell gradient ``(4, dim, N)`` at order >= 2
    and None at order 1 or with ``o1=True`` (the MOOD redo).  ``bf_values``
    are recomputed on the device from Q (they are halo_bc's values in every
    caller)."""
    s = MeshRT.solver
    Q4 = _dev4(Q, MeshRT)
    F, N = MeshRT.n_faces, MeshRT.n_cells
    QL = torch.empty((4, F), dtype=s.dtype, device=s.device)
    QR = torch.empty_like(QL)
    order1 = bool(o1) or Ops.order < 2
    lg = None if order1 else torch.zeros((4, MeshRT.dimension, N),
                                         dtype=s.dtype, device=s.device)
    mask = _mask(force_o1, N, s.device)
    L.check(L.lib().zj_reconstruct(_rt(MeshRT), _p(Q4), _p(mask),
                                   int(order1), _p(QL), _p(QR), _p(lg),
                                   _stream()), "zj_reconstruct")
    return QL, QR, (lg if Ops.use_interior_ncp else None)


def _mask(force_o1, N, device):
    if force_o1 is None:
        return None
    m = torch.as_tensor(force_o1, device=device)
    if m.numel() != N:
        raise ValueError(f"force_o1 must have {N} entries, got {m.numel()}")
    return m.reshape(N).to(torch.uint8).contiguous()


def flux_pass(Q, Qaux, p, MeshRT, Ops, t_stage, faces):
    """Per-face flux and fluctuations ``(Fface, Dp, Dm)``, each ``(4, F)``
    (``blocks.py:185-238``); a boundary face's Q_R is BC(Q_L)."""
    s = MeshRT.solver
    QL, QR = faces
    Q4 = _dev4(Q, MeshRT)
    Fx = torch.empty_like(QL)
    Dp = torch.empty_like(QL)
    Dm = torch.empty_like(QL)
    L.check(L.lib().zj_flux_pass(_rt(MeshRT), _p(Q4), _p(_hinv_row(Qaux)),
                                 _p(QL.contiguous()), _p(QR.contiguous()),
                                 _p(Fx), _p(Dp), _p(Dm), _stream()),
            "zj_flux_pass")
    return Fx, Dp, Dm


def gather_update(Q0, Fface, Dp, Dm, MeshRT, dt, a_stage, source=None,
                  cell_term=None, h_index=None, lim_grad=None, Qaux=None):
    """``Q_cand = Q0 + a dt (-div(F + D) + cell_term + source)`` gathered
    per cell, and the troubled flags (``blocks.py:241-301``).  ``source``:
    None, ``"manning"`` (the explicit friction of Q0, 1/h from ``Qaux``) or
    a ``(4, N)`` array; ``lim_grad`` instead of ``cell_term`` gives the
    interior nonconservative term ``-B(Q0) . lim_grad``."""
    s = MeshRT.solver
    Q04 = _dev4(Q0, MeshRT)
    N = MeshRT.n_cells
    extra = cell_term
    if source is not None and not isinstance(source, str):
        extra = source if extra is None else extra + source
    Qc = torch.empty_like(Q04)
    tr = torch.empty(N, dtype=torch.uint8, device=s.device)
    dtd = s._dt_dev(dt)
    L.check(L.lib().zj_gather_update(
        _rt(MeshRT), _p(Q04), _p(Fface), _p(Dp), _p(Dm), _p(dtd),
        float(a_stage), _p(None if extra is None else extra.contiguous()),
        _p(None if (extra is not None or lim_grad is None)
           else lim_grad.contiguous()),
        int(source == "manning"), _p(_hinv_row(Qaux)), _p(Qc), _p(tr), None,
        _stream()), "zj_gather_update")
    return Qc, tr.bool()


def mood_resolve(troubled, Q0, Q_cand, Qaux, p, t, dt, MeshRT, Ops,
                 stage_loop):
    """Whole-step order-1 redo when any cell is troubled
    (``blocks.py:304-323``)."""
    if bool(troubled.any()):
        return stage_loop(Q0, Qaux, p, t, dt, True)[0]
    return Q_cand


def update_aux(Q, Qaux, p, t, dt, MeshRT, Ops):
    """Post-step aux refresh (derivative rows and 1/h)."""
    return MeshRT.solver.update_qaux(Q, Qaux)


# ── the march ───────────────────────────────────────────────────────────
class MarchSolver:
    """``zoomy_jax.fvm.solver2.march.MarchSolver``: the step head
    (``dt_pass`` + ``reduce_dt`` + the host dt guard) and the stage loop of
    blocks over the Shu-Osher tableau, with MOOD's whole-step order-1 redo,
    the post-step aux refresh and the max_steps guard."""

    def __init__(self, time_end: float = 1.0, cfl: float = 0.9,
                 dtype: torch.dtype = torch.float64, mood_redo: bool = False,
                 dt_max: Optional[float] = None, source: bool = True,
                 reconstruction_variables: str = "conservative"):
        self.time_end = float(time_end)
        self._cfl = float(cfl)
        self.dtype = dtype
        self.mood_redo = bool(mood_redo)
        self._dt_max = dt_max
        self._source = bool(source)
        self._recon = reconstruction_variables

    def setup_simulation(self, mesh, model, Q=None):
        self.solver = HyperbolicSolver(
            time_end=self.time_end, cfl=self._cfl, dtype=self.dtype,
            reconstruction_variables=self._recon)
        Q0, Qaux = self.solver.setup_simulation(mesh, model)
        if Q is not None:
            Q0 = self.solver._view(self.solver._full(
                torch.as_tensor(Q, dtype=self.dtype,
                                device=self.solver.device)))
            Qaux = self.solver.update_qaux(Q0, None)
        self.MeshRT = prepare_mesh(mesh, self.solver)
        self.Ops = build_operators(self.solver)
        self.time = 0.0
        self.iteration = 0
        return self.solver._full(Q0).clone(), Qaux

    def stage_loop(self, Q0, Qaux, p, t, dt, o1):
        """The explicit stage loop (``march.py:126-157``) with dt frozen at
        the step head."""
        M, O = self.MeshRT, self.Ops
        Qk = Q0
        troubled = torch.zeros(M.n_cells, dtype=torch.bool,
                               device=self.solver.device)
        for k, (alpha, beta) in enumerate(O.tableau):
            t_stage = t + (0.0 if k == 0 else dt)
            Qk, Qaux_s, bf = halo_bc(Qk, Qaux, p, M, t_stage, O)
            QL, QR, lim_grad = reconstruct(Qk, Qaux_s, M, O, bf, o1=o1)
            Fx, Dp, Dm = flux_pass(Qk, Qaux_s, p, M, O, t_stage, (QL, QR))
            Qc, tr = gather_update(
                Qk, Fx, Dp, Dm, M, dt, beta,
                source="manning" if self._source else None,
                lim_grad=lim_grad, Qaux=Qaux_s)
            troubled = troubled | tr
            Qk = Qc if alpha == 0.0 else alpha * Q0 + (1.0 - alpha) * Qc
        return Qk, troubled

    def step_head(self, Q, Qaux, p, t_remaining, dt_window=None):
        lo, hi = dt_pass(Q, Qaux, p, self.MeshRT, self.Ops)
        dt = reduce_dt(lo, hi, self.MeshRT.inradius_f, self._dt_max,
                       CFL=self._cfl, dimension=self.MeshRT.dimension,
                       t_remaining=None if dt_window is not None
                       else t_remaining, dt_window=dt_window,
                       MeshRT=self.MeshRT)
        return lo, hi, dt

    def hyperbolic_step(self, Q, Qaux, dt, p=None):
        Qc, troubled = self.stage_loop(Q, Qaux, p, self.time, dt, False)
        if self.mood_redo:
            Qc = mood_resolve(troubled, Q, Qc, Qaux, p, self.time, dt,
                              self.MeshRT, self.Ops, self.stage_loop)
        Qaux = update_aux(Qc, Qaux, p, self.time, dt, self.MeshRT, self.Ops)
        return Qc, Qaux, troubled

    def march(self, Q, Qaux, *, time_end=None, n_steps=None, dt_window=None,
              parameters=None, max_steps=1_000_000):
        """Runs the march; returns ``(Q, Qaux, stats)``."""
        time_end = self.time_end if time_end is None else time_end
        stats = {"steps": 0, "dt_min": np.inf, "dt_max": 0.0,
                 "troubled_steps": 0, "wall": 0.0}
        t0 = perf_counter()
        while (self.time < time_end if n_steps is None
               else stats["steps"] < n_steps):
            lo, hi, dt = self.step_head(
                Q, Qaux, parameters,
                (time_end - self.time) if np.isfinite(time_end) else None,
                dt_window)
            d = assert_dt_admissible(dt, lo, hi, self.MeshRT.inradius_f,
                                     time=self.time,
                                     iteration=self.iteration)
            Q, Qaux, troubled = self.hyperbolic_step(Q, Qaux, dt,
                                                     parameters)
            self.time += d
            self.iteration += 1
            stats["steps"] += 1
            stats["dt_min"] = min(stats["dt_min"], d)
            stats["dt_max"] = max(stats["dt_max"], d)
            if bool(troubled.any()):
                stats["troubled_steps"] += 1
            if stats["steps"] >= max_steps:
                raise FloatingPointError(
                    f"march exceeded max_steps = {max_steps} at t = "
                    f"{self.time} / {time_end} with dt = {d:.6e}: the "
                    f"timestep collapsed towards zero")
        torch.cuda.synchronize()
        stats["wall"] = perf_counter() - t0
        return Q, Qaux, stats
