"""Python side of the FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker.

It builds the reference-layout mesh arrays (what ``convert_mesh_to_jax``
holds, ``zoomy_jax/mesh/mesh.py:186-264``: face owner/neighbour, normals,
face volumes and centres, cell centres/volumes/inradius, LSQ pseudo-inverse
rows) from a topology-only mesh, and calls ``liboracle.so`` (``zj_oracle.c``)
through ctypes.

LSQ rows: the stencil of a cell is its vertex neighbours plus the ghost
image ``c + 2 (x_f - c)`` of each of its boundary faces
(``reconstruction_jax.py:433-522``).  A ghost row enters the fit of gradient
component ``d`` only when its face normal is along axis ``d``; with that rule
the reference's 2-D derivative aux rows (``tests/refs/swe_2d.npz``) are
reproduced in every cell including the corners.
"""
from __future__ import annotations

import ctypes
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

LIM = {"none": 0, "barth_jespersen": 1, "minmod": 2, "venkatakrishnan": 3,
       "vanleer": 4}
RIEMANN = {"rusanov_nc": 0, "rusanov": 1, "hll": 2}
RECON = {"conservative": 0, "eta": 1, "xz": 2, "eta_hu": 3}
POSITIVITY = {"": 0, "zhang_shu": 1}
STAGE_AUX = {"refresh": 0, "frozen": 1}


class _Mesh(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("n_cells", ctypes.c_int),
                ("n_faces", ctypes.c_int), ("n_bf", ctypes.c_int),
                ("max_st", ctypes.c_int)] + [
        (n, ctypes.c_void_p) for n in (
            "face_cells", "face_normals", "face_volumes", "face_centers",
            "cell_centers", "cell_volumes", "cell_inradius", "bface_face",
            "bface_cell", "bface_kind", "bface_partner", "stencil", "lsq_w",
            "gslot", "lsq_wg", "bface_mask", "bface_value")] + [
        ("max_cf", ctypes.c_int), ("cell_faces", ctypes.c_void_p)]


class _Params(ctypes.Structure):
    _fields_ = [("g", ctypes.c_double), ("cfl", ctypes.c_double),
                ("eps_wet", ctypes.c_double), ("dt_max", ctypes.c_double),
                ("manning_n", ctypes.c_double), ("order", ctypes.c_int),
                ("limiter", ctypes.c_int), ("riemann", ctypes.c_int),
                ("recon", ctypes.c_int), ("positivity", ctypes.c_int),
                ("front_tol", ctypes.c_double), ("stage_aux", ctypes.c_int)]


_lib = None


def build(force: bool = False) -> pathlib.Path:
    """Compile ``zj_oracle.c`` into ``liboracle.so`` with gcc."""
    src = HERE / "zj_oracle.c"
    if force or not LIB_PATH.exists() or \
            LIB_PATH.stat().st_mtime < src.stat().st_mtime:
        import subprocess
        subprocess.check_call(["make", "-s", "-C", str(HERE)])
    return LIB_PATH


LIB_OMP_PATH = HERE / "liboracle_omp.so"
_lib_omp = None


def lib(omp: bool = False):
    """The serial checker, or (``omp=True``) the same source built with
    OpenMP -- the multi-threaded CPU baseline, bit-identical results."""
    global _lib, _lib_omp
    if omp:
        if _lib_omp is None:
            if not LIB_OMP_PATH.exists():
                build(force=True)
            _lib_omp = _bind(ctypes.CDLL(str(LIB_OMP_PATH)))
        return _lib_omp
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = _bind(ctypes.CDLL(str(LIB_PATH)))
    return _lib


def set_threads(n: int) -> int:
    """Threads of the OpenMP build (returns the count in effect)."""
    return int(lib(omp=True).zjo_set_threads(int(n)))


def _bind(L):
    P = ctypes.POINTER
    d = ctypes.c_double
    vp = ctypes.c_void_p
    L.zjo_hinv.restype = d
    L.zjo_hinv.argtypes = [d, d]
    L.zjo_compute_dt.restype = d
    L.zjo_compute_dt.argtypes = [P(_Mesh), P(_Params), vp, d]
    L.zjo_step.restype = ctypes.c_int
    L.zjo_step.argtypes = [P(_Mesh), P(_Params), d, vp, vp]
    L.zjo_step_mood.restype = ctypes.c_int
    L.zjo_step_mood.argtypes = [P(_Mesh), P(_Params), d, vp, vp,
                                P(ctypes.c_int)]
    L.zjo_flux_operator.argtypes = [P(_Mesh), P(_Params), vp, vp]
    L.zjo_flux_operator_o1.argtypes = [P(_Mesh), P(_Params), vp, vp, vp]
    L.zjo_reconstruct.argtypes = [P(_Mesh), P(_Params), vp, vp, vp, vp]
    L.zjo_update_aux.argtypes = [P(_Mesh), P(_Params), vp, vp]
    L.zjo_mass.restype = d
    L.zjo_mass.argtypes = [P(_Mesh), vp]
    L.zjo_lsq_grad.argtypes = [P(_Mesh), vp, vp, vp]
    L.zjo_imex_local.argtypes = [P(_Mesh), P(_Params), d, ctypes.c_int,
                                 vp, vp]
    L.zjo_march.restype = ctypes.c_int
    L.zjo_march.argtypes = [P(_Mesh), P(_Params), vp, P(d), d,
                            ctypes.c_int]
    L.zjo_face_flux.argtypes = [P(_Params), vp, vp, d, d, vp, vp, vp, vp]
    L.zjo_set_threads.restype = ctypes.c_int
    L.zjo_set_threads.argtypes = [ctypes.c_int]
    return L


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def lsq_weights(mesh, centers, stencil, fcent_slot, normal_slot):
    """Per-cell LSQ pseudo-inverse rows ``w (N, S, 2)`` and ghost rows
    ``wg (N, 2, 2)`` with ghost face ids ``gslot (N, 2)``."""
    N, S = stencil.shape
    dim = mesh.dim
    valid = stencil >= 0
    sc = np.where(valid, stencil, 0)
    r = centers[sc] - centers[:, None, :]              # (N, S, 2)
    r[~valid] = 0.0
    gslot = np.full((N, 2), -1, np.int64)
    rg = np.zeros((N, 2, 2))
    gn = np.zeros((N, 2, 2))
    cnt = np.zeros(N, np.int64)
    for b in range(mesh.n_bfaces):
        c = mesh.bface_cell[b]
        s = mesh.bface_slot[b]
        k = cnt[c]
        if k >= 2:
            raise ValueError("cell with more than two boundary faces")
        gslot[c, k] = b
        rg[c, k] = 2.0 * (fcent_slot[c, s] - centers[c])
        gn[c, k] = normal_slot[c, s]
        cnt[c] += 1
    w = np.zeros((N, S, 2))
    wg = np.zeros((N, 2, 2))
    for d in range(dim):
        use_g = (gslot >= 0) & (np.abs(gn[..., d]) > 0.5)   # (N, 2)
        rgd = rg * use_g[..., None]
        if dim == 1:
            den = (r[..., 0] ** 2).sum(1) + (rgd[..., 0] ** 2).sum(1)
            w[..., 0] = r[..., 0] / den[:, None]
            wg[..., 0] = rgd[..., 0] / den[:, None]
            continue
        M = np.einsum("nsi,nsj->nij", r, r) + np.einsum("ngi,ngj->nij",
                                                        rgd, rgd)
        Minv = np.linalg.inv(M)
        row = Minv[:, d, :]                                # (N, 2)
        w[..., d] = np.einsum("ni,nsi->ns", row, r)
        wg[..., d] = np.einsum("ni,ngi->ng", row, rgd)
    return w, gslot.astype(np.int32), wg


class RefMesh:
    """Reference-layout arrays for a :class:`zoomy_hip.mesh.Mesh`."""

    def __init__(self, mesh):
        self.mesh = mesh
        N = mesh.n_cells
        fc, fs, bff = mesh.faces()
        F = fc.shape[1]
        nrm, L, fcen = mesh.slot_geometry()
        A = fc[0]
        sA = fs[0]
        self.face_cells = np.ascontiguousarray(fc, np.int32)
        self.face_normals = np.ascontiguousarray(nrm[A, sA].T)
        self.face_volumes = np.ascontiguousarray(L[A, sA])
        self.face_centers = np.ascontiguousarray(fcen[A, sA].T)
        cen = mesh.cell_centers()
        self.cell_centers = np.ascontiguousarray(cen.T)
        self.cell_volumes = np.ascontiguousarray(mesh.cell_volumes())
        self.cell_inradius = np.ascontiguousarray(mesh.cell_inradius())
        self.bface_face = np.ascontiguousarray(bff, np.int32)
        self.bface_cell = np.ascontiguousarray(mesh.bface_cell, np.int32)
        self.bface_kind = np.ascontiguousarray(mesh.bface_kind(), np.int32)
        pp = mesh.periodic_partner
        self.bface_partner = np.ascontiguousarray(
            pp if pp is not None else -np.ones(mesh.n_bfaces), np.int32)
        st = mesh.vertex_stencil()
        self.stencil = np.ascontiguousarray(st, np.int32)
        w, gslot, wg = lsq_weights(mesh, cen, st, fcen, nrm)
        self.lsq_w = np.ascontiguousarray(w)
        self.gslot = np.ascontiguousarray(gslot, np.int32)
        self.lsq_wg = np.ascontiguousarray(wg)
        dm, dv = mesh.bface_dirichlet()
        self.bface_mask = np.ascontiguousarray(dm, np.int32)
        self.bface_value = np.ascontiguousarray(dv)
        # faces of each cell in increasing id, -1 last
        cf = mesh.cell_faces().astype(np.int64)
        cf = np.sort(np.where(cf < 0, np.iinfo(np.int64).max, cf), axis=1)
        self.cell_faces = np.ascontiguousarray(
            np.where(cf == np.iinfo(np.int64).max, -1, cf), np.int32)
        self.n_faces = F
        self.c = _Mesh(mesh.dim, N, F, mesh.n_bfaces, st.shape[1],
                       *[a.ctypes.data for a in (
                           self.face_cells, self.face_normals,
                           self.face_volumes, self.face_centers,
                           self.cell_centers, self.cell_volumes,
                           self.cell_inradius, self.bface_face,
                           self.bface_cell, self.bface_kind,
                           self.bface_partner, self.stencil, self.lsq_w,
                           self.gslot, self.lsq_wg, self.bface_mask,
                           self.bface_value)],
                       self.cell_faces.shape[1], self.cell_faces.ctypes.data)


def params(g=9.81, cfl=0.9, eps_wet=1e-8, dt_max=0.0, manning_n=0.0,
           order=1, limiter="minmod", riemann="rusanov", recon="conservative",
           positivity="", front_tol=None, stage_aux="refresh"):
    return _Params(g, cfl, eps_wet, dt_max, manning_n, int(order),
                   LIM[limiter], RIEMANN[riemann], RECON[recon],
                   POSITIVITY[positivity],
                   -1.0 if front_tol is None else float(front_tol),
                   STAGE_AUX[stage_aux])


class Oracle:
    """Stateless wrapper; ``Q`` is a (4, N) float64 array (b, h, hu, hv)."""

    def __init__(self, mesh, omp: bool = False, **kw):
        self.rm = mesh if isinstance(mesh, RefMesh) else RefMesh(mesh)
        self.p = params(**kw)
        self.L = lib(omp)

    @property
    def n(self):
        return self.rm.mesh.n_cells

    def _q(self, Q):
        Q = np.ascontiguousarray(Q, np.float64)
        assert Q.shape == (4, self.n)
        return Q

    def dt(self, Q, t_remaining=-1.0):
        return self.L.zjo_compute_dt(ctypes.byref(self.rm.c),
                                     ctypes.byref(self.p), _ptr(self._q(Q)),
                                     float(t_remaining))

    def step(self, dt, Q):
        Q = self._q(Q)
        out = np.empty_like(Q)
        bad = self.L.zjo_step(ctypes.byref(self.rm.c), ctypes.byref(self.p),
                              float(dt), _ptr(Q), _ptr(out))
        return out, bad

    def step_mood(self, dt, Q):
        """Order-2 step with the a-posteriori MOOD redo; returns
        (Q, troubled count, number of demoted cells)."""
        Q = self._q(Q)
        out = np.empty_like(Q)
        nd = ctypes.c_int(0)
        bad = self.L.zjo_step_mood(ctypes.byref(self.rm.c),
                                   ctypes.byref(self.p), float(dt), _ptr(Q),
                                   _ptr(out), ctypes.byref(nd))
        return out, bad, nd.value

    def flux_o1(self, Q, force_o1):
        """Flux operator with a per-cell first-order mask."""
        Q = self._q(Q)
        mask = np.ascontiguousarray(force_o1, np.uint8)
        assert mask.shape == (self.n,)
        out = np.empty_like(Q)
        self.L.zjo_flux_operator_o1(ctypes.byref(self.rm.c),
                                    ctypes.byref(self.p), _ptr(Q),
                                    _ptr(mask), _ptr(out))
        return out

    def flux(self, Q):
        Q = self._q(Q)
        out = np.empty_like(Q)
        self.L.zjo_flux_operator(ctypes.byref(self.rm.c), ctypes.byref(self.p),
                                 _ptr(Q), _ptr(out))
        return out

    def reconstruct(self, Q):
        Q = self._q(Q)
        F = self.rm.n_faces
        QL = np.zeros((4, F))
        QR = np.zeros((4, F))
        lg = np.zeros((4, 2, self.n))
        self.L.zjo_reconstruct(ctypes.byref(self.rm.c), ctypes.byref(self.p),
                               _ptr(Q), _ptr(QL), _ptr(QR), _ptr(lg))
        return QL, QR, lg

    def aux(self, Q):
        Q = self._q(Q)
        d = self.rm.mesh.dim
        out = np.zeros((3 * d + 1, self.n))
        self.L.zjo_update_aux(ctypes.byref(self.rm.c), ctypes.byref(self.p),
                              _ptr(Q), _ptr(out))
        return out

    def grad(self, u, ubf=None):
        u = np.ascontiguousarray(u, np.float64)
        out = np.zeros((2, self.n))
        ub = None if ubf is None else np.ascontiguousarray(ubf, np.float64)
        self.L.zjo_lsq_grad(ctypes.byref(self.rm.c), _ptr(u), _ptr(ub),
                            _ptr(out))
        return out

    def mass(self, Q):
        return self.L.zjo_mass(ctypes.byref(self.rm.c), _ptr(self._q(Q)))

    def imex_local(self, dt, Qexp, maxiter=6):
        Qexp = self._q(Qexp)
        out = np.empty_like(Qexp)
        self.L.zjo_imex_local(ctypes.byref(self.rm.c), ctypes.byref(self.p),
                              float(dt), int(maxiter), _ptr(Qexp), _ptr(out))
        return out

    def march(self, Q, t_end=None, n_steps=None, t0=0.0):
        Q = self._q(Q).copy()
        t = ctypes.c_double(t0)
        k = self.L.zjo_march(ctypes.byref(self.rm.c), ctypes.byref(self.p),
                             _ptr(Q), ctypes.byref(t),
                             float(t_end if t_end is not None else 0.0),
                             int(n_steps if n_steps is not None else -1))
        return Q, t.value, k


def face_flux(qL, qR, hinvA, hinvB, n, **kw):
    """One face of the numerics (test helper)."""
    p = params(**kw)
    qL = np.ascontiguousarray(qL, np.float64)
    qR = np.ascontiguousarray(qR, np.float64)
    n = np.ascontiguousarray(n, np.float64)
    F = np.zeros(4)
    Dp = np.zeros(4)
    Dm = np.zeros(4)
    lib().zjo_face_flux(ctypes.byref(p), _ptr(qL), _ptr(qR), float(hinvA),
                        float(hinvB), _ptr(n), _ptr(F), _ptr(Dp), _ptr(Dm))
    return F, Dp, Dm
