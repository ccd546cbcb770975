"""Batched ensemble of independent shallow-water instances on one GPU.

Many dam-break instances of uneven size march together: their meshes are
stored as one disjoint mesh (``Mesh.union``), every SSP-RK2 stage is one
launch of the stage kernel over the whole union's patch list (uneven
instances balance by themselves: a patch is the unit of work), and each
instance keeps its own CFL dt, folded per instance from the stage kernel's
per-wave partials.  ``capture`` records a run of steps as a HIP graph that
``replay`` relaunches.  The numerics per instance are those of
``HyperbolicSolver.march_device`` (LSQ-MUSCL, HLL, SSP-RK2, walls).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional

import numpy as np
import torch

from . import _lib as L
from .fvm.solver import DeviceMesh, ShallowWater, _Runtime, _p, _stream
from .mesh import Mesh, lattice_triangles


@dataclass
class Instance:
    """One member of the ensemble: its seed, target cell count, reservoir
    height and domain side."""
    seed: int
    n_cells: int
    h0: float
    side: float = 1.0


def ensemble_spec(n: int = 256, n_min: int = 16384, n_max: int = 1 << 20,
                  seed: int = 5) -> List[Instance]:
    """``n`` instances with cell counts log-uniform in ``[n_min, n_max]`` and
    reservoir heights ``h0 ~ U[1, 4]`` (seeded)."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(n_min), np.log(n_max), n)))
    h0 = rng.uniform(1.0, 4.0, n)
    return [Instance(seed=int(seed * 1000 + k), n_cells=int(sizes[k]),
                     h0=float(h0[k])) for k in range(n)]


def instance_mesh(inst: Instance) -> Mesh:
    """A seeded triangle mesh of about ``inst.n_cells`` cells on the square
    of side ``inst.side`` (a lattice jittered by 0.15 of its spacing, cut
    into triangles), cells in Morton order, walls."""
    nq = max(inst.n_cells // 2, 4)
    nx = max(int(round(np.sqrt(nq))), 2)
    ny = max(nq // nx, 2)
    L_ = inst.side
    m = lattice_triangles(nx, ny, (0.0, L_, 0.0, L_), 0.15, inst.seed)
    return m.reorder(m.morton_permutation()).set_bc("wall")


def instance_state(mesh: Mesh, inst: Instance) -> np.ndarray:
    """Dam break over a seeded bed: eight Gaussian bumps (amplitude U[0,
    0.3], width U[0.03, 0.1] of the side; where they overlap the bed is
    scaled to stay at or below 0.35), free surface ``h0`` for x < 0.4 side
    and 0.5 beyond -- wet everywhere -- at rest."""
    rng = np.random.default_rng(inst.seed + 1)
    xc = mesh.cell_centers()
    L_ = inst.side
    b = np.zeros(mesh.n_cells)
    for _ in range(8):
        c = rng.uniform(0.0, L_, 2)
        a = rng.uniform(0.0, 0.3)
        w = rng.uniform(0.03, 0.1) * L_
        b += a * np.exp(-((xc - c) ** 2).sum(1) / (2.0 * w * w))
    if b.max() > 0.35:
        b *= 0.35 / b.max()
    eta = np.where(xc[:, 0] < 0.4 * L_, inst.h0, 0.5)
    Q = np.zeros((4, mesh.n_cells))
    Q[0] = b
    Q[1] = np.maximum(eta - b, 0.0)
    return Q


def _build_instance(inst: Instance):
    m = instance_mesh(inst)
    m.extra_stencil()          # computed here, carried to the union
    return m, instance_state(m, inst)


def build_instances(spec: List[Instance], workers: int = 1):
    """Meshes and initial states of the instances, built by ``workers``
    processes (host-side set-up only)."""
    if workers <= 1:
        out = [_build_instance(i) for i in spec]
    else:
        import concurrent.futures as cf
        import multiprocessing as mp
        with cf.ProcessPoolExecutor(max_workers=workers,
                                    mp_context=mp.get_context("fork")) as ex:
            out = list(ex.map(_build_instance, spec, chunksize=4))
    return [m for m, _ in out], [q for _, q in out]


class EnsembleSolver:
    """Device march of an ensemble (``zj_ensemble_*``).

    ``setup(meshes, states)`` uploads the union mesh and the stacked state;
    ``march(n)`` runs n steps eagerly; ``capture(n)`` records n steps as a
    HIP graph and ``replay()`` runs them again; per-instance ``t``, ``dt``
    and ``mass()`` are device arrays; ``split(Q)`` gives each instance's
    columns."""

    def __init__(self, dtype: torch.dtype = torch.float32, cfl: float = 0.9,
                 limiter: str = "minmod", riemann: str = "hll"):
        self.dtype = dtype
        self.cfl = float(cfl)
        self.model = ShallowWater(dimension=2, order=2, limiter=limiter,
                                  riemann=riemann)

    def setup(self, meshes: List[Mesh], states: List[np.ndarray]):
        if len(meshes) != len(states) or not meshes:
            raise ValueError("one state per mesh")
        if not torch.cuda.is_available():
            raise L.ZoomyHipError("the ensemble march needs a GPU")
        self.n = len(meshes)
        self.offsets = np.cumsum([0] + [m.n_cells for m in meshes]).astype(
            np.int32)
        union = Mesh.union(meshes)
        self.mesh = union
        # the per-wave dt partials need patches of consecutive cells
        self._dmesh = DeviceMesh(union, patch_order="consecutive")
        self._rt = _Runtime(self._dmesh, self.model, self.cfl, self.dtype,
                            "conservative")
        h = ctypes.c_void_p()
        L.check(L.lib().zj_ensemble_create(self._rt.handle, int(self.n),
                                           self.offsets.ctypes.data,
                                           ctypes.byref(h)),
                "zj_ensemble_create")
        self.handle = h
        dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.Q = torch.as_tensor(np.concatenate(states, axis=1),
                                 dtype=self.dtype, device=dev).contiguous()
        self.t = torch.zeros(self.n, dtype=torch.float64, device=dev)
        self.dt = torch.zeros(self.n, dtype=torch.float64, device=dev)
        self._mass = torch.zeros(self.n, dtype=torch.float64, device=dev)
        return self.Q

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and L is not None and L._lib is not None:
            L._lib.zj_ensemble_destroy(h)
            self.handle = None

    def compute_dt(self, Q=None):
        out = torch.empty(self.n, dtype=torch.float64, device=self.device)
        L.check(L.lib().zj_ensemble_dt(self.handle, _p(self.Q if Q is None
                                                       else Q), _p(out),
                                       _stream()), "zj_ensemble_dt")
        return out

    def march(self, n_steps: int, t_end: float = -1.0):
        L.check(L.lib().zj_ensemble_march(self.handle, _p(self.Q),
                                          int(n_steps), float(t_end),
                                          _p(self.t), _p(self.dt), _stream()),
                "zj_ensemble_march")
        return self.Q

    def capture(self, n_steps: int, t_end: float = -1.0):
        L.check(L.lib().zj_ensemble_capture(self.handle, _p(self.Q),
                                            int(n_steps), float(t_end),
                                            _p(self.t), _p(self.dt),
                                            _stream()), "zj_ensemble_capture")

    def replay(self):
        L.check(L.lib().zj_ensemble_replay(self.handle, _stream()),
                "zj_ensemble_replay")
        return self.Q

    def mass(self, Q=None):
        L.check(L.lib().zj_ensemble_mass(self.handle, _p(self.Q if Q is None
                                                         else Q),
                                         _p(self._mass), _stream()),
                "zj_ensemble_mass")
        return self._mass.clone()

    def split(self, Q=None) -> List[torch.Tensor]:
        Q = self.Q if Q is None else Q
        o = self.offsets
        return [Q[:, o[k]:o[k + 1]] for k in range(self.n)]
