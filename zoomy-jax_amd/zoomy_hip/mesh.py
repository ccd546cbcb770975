"""Unstructured finite-volume meshes for the HIP time step.

The reference solver consumes ``zoomy_core``'s ``LSQMesh`` converted to a
``MeshJAX`` pytree (``zoomy_jax/mesh/mesh.py:123-264``) and the solver2
``MeshRT`` contract (``zoomy_jax/fvm/solver2/context.py:99-163``).  This module
builds the same topology natively (no ``zoomy_core``) for the three mesh
families the time step runs on:

* ``Mesh.line(n, x0, x1)``   -- 1-D uniform cells (``LSQMesh.create_1d``),
* ``Mesh.quad(nx, ny, box)`` -- a structured quad grid stored unstructured
  (``LSQMesh.create_2d``); cell ``(i, j)`` (``i`` along x) gets id
  ``i * ny + j``, the order the reference's 2-D golden state is stored in,
* ``Mesh.triangles(points)`` -- a Delaunay triangulation of a point set.

Every mesh is described by per-cell *slot* tables: slot ``k`` of a cell is
the face from its vertex ``k`` to vertex ``k+1`` (counter-clockwise), so the
outward normal, face length and face centre of every cell face can be
recomputed from vertex coordinates.  ``cell_nbr[c, k]`` is the cell across
slot ``k`` or ``-1 - b`` for boundary face ``b``.  The face list (``face_cells``
with the reference's owner/neighbour convention: a boundary face's owner is
its inner cell and its normal points out of the domain) is derived from the
slot tables.

The LSQ stencil of a cell is the set of cells sharing a vertex with it --
the 3x3 block on a quad grid, which is what the reference's derivative aux
rows in ``tests/refs/swe_2d.npz`` are computed over.

Cell ordering: ``Mesh.reorder(perm)`` applies a cell permutation (used for
reverse Cuthill-McKee on triangle meshes so that neighbour gathers coalesce).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

# Boundary-condition kinds understood by the kernels.  The reference's BC
# kernels come from zoomy_core (``zoomy_jax/transformation/jax_runtime.py:
# 837-863``); Wall mirrors the normal momentum and Extrapolation copies the
# inner state (``zoomy_jax/fvm/reconstruction_jax.py:494-510``); Dirichlet
# prescribes the face value of some state rows and extrapolates the others.
BC_EXTRAPOLATION = 0
BC_WALL = 1
BC_PERIODIC = 2
BC_DIRICHLET = 3

_BC_NAMES = {"extrapolation": BC_EXTRAPOLATION, "wall": BC_WALL,
             "periodic": BC_PERIODIC, "dirichlet": BC_DIRICHLET}

# state rows by the reference's variable names (SWE: b, h, q_0, q_1)
_ROWS = {"b": 0, "h": 1, "q_0": 2, "hu": 2, "q_1": 3, "hv": 3}


@dataclass(frozen=True)
class Dirichlet:
    """Prescribed value of one state row on a boundary tag (the reference's
    ``BC.Dirichlet(tag=..., on=..., value=...)``, tests/cases.py:168-183).
    A tag given a list of these prescribes each listed row; the rows left
    out are extrapolated from the inner state."""
    on: str
    value: float

    def row(self) -> int:
        try:
            return _ROWS[self.on]
        except KeyError:
            raise ValueError(f"unknown state row {self.on!r}; "
                             f"known: {sorted(_ROWS)}") from None


def bc_kind(name) -> int:
    if isinstance(name, (int, np.integer)):
        return int(name)
    try:
        return _BC_NAMES[str(name).lower()]
    except KeyError:
        raise ValueError(f"unknown boundary condition {name!r}; "
                         f"known: {sorted(_BC_NAMES)}") from None


@dataclass
class Mesh:
    dim: int
    verts: np.ndarray            # (n_verts, 2) float64
    cell_vertices: np.ndarray    # (N, MF) int32, CCW; 1-D: (left, right)
    cell_nbr: np.ndarray         # (N, MF) int32, cell across slot or -1-b
    bface_cell: np.ndarray       # (n_bf,) int32 inner cell of boundary face
    bface_slot: np.ndarray       # (n_bf,) int32 slot of that face in its cell
    bface_tag: np.ndarray        # (n_bf,) int32 index into ``tags``
    tags: List[str]
    bc: Dict[str, int] = field(default_factory=dict)
    periodic_partner: Optional[np.ndarray] = None   # (n_bf,) partner bface
    bc_values: Dict[str, Dict[int, float]] = field(default_factory=dict)

    # ── construction ──────────────────────────────────────────────────
    @staticmethod
    def line(n: int, x0: float = 0.0, x1: float = 1.0) -> "Mesh":
        """``n`` uniform cells on ``[x0, x1]``; tags ``left``/``right``."""
        xs = np.linspace(x0, x1, n + 1)
        # exact uniform spacing (linspace end points are exact; interior
        # points x0 + i*dx match LSQMesh.create_1d's construction)
        dx = (x1 - x0) / n
        xs = x0 + dx * np.arange(n + 1)
        xs[-1] = x1
        verts = np.zeros((n + 1, 2))
        verts[:, 0] = xs
        cv = np.stack([np.arange(n), np.arange(1, n + 1)], 1).astype(np.int32)
        nbr = np.stack([np.arange(-1, n - 1), np.arange(1, n + 1)], 1)
        nbr = nbr.astype(np.int32)
        nbr[0, 0] = -1 - 0          # boundary face 0 (left)
        nbr[n - 1, 1] = -1 - 1      # boundary face 1 (right)
        return Mesh(dim=1, verts=verts, cell_vertices=cv, cell_nbr=nbr,
                    bface_cell=np.array([0, n - 1], np.int32),
                    bface_slot=np.array([0, 1], np.int32),
                    bface_tag=np.array([0, 1], np.int32),
                    tags=["left", "right"])

    @staticmethod
    def quad(nx: int, ny: int, box=(0.0, 1.0, 0.0, 1.0)) -> "Mesh":
        """``nx x ny`` quads on ``box = (x0, x1, y0, y1)``.

        Cell ``(i, j)`` -> id ``i*ny + j``; vertex ``(i, j)`` -> ``i*(ny+1)+j``.
        Slots: 0 bottom, 1 right, 2 top, 3 left.  Tags: left, right, bottom,
        top (the reference's 2-D tag order, ``tests/cases.py:108-116``).
        """
        x0, x1, y0, y1 = box
        dx = (x1 - x0) / nx
        dy = (y1 - y0) / ny
        xs = x0 + dx * np.arange(nx + 1)
        ys = y0 + dy * np.arange(ny + 1)
        xs[-1] = x1
        ys[-1] = y1
        VX, VY = np.meshgrid(xs, ys, indexing="ij")
        verts = np.stack([VX.ravel(), VY.ravel()], 1)
        I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        I = I.ravel().astype(np.int64)
        J = J.ravel().astype(np.int64)
        vid = lambda a, b: (a * (ny + 1) + b)
        cv = np.stack([vid(I, J), vid(I + 1, J), vid(I + 1, J + 1),
                       vid(I, J + 1)], 1).astype(np.int32)
        cid = lambda a, b: a * ny + b
        N = nx * ny
        nbr = np.empty((N, 4), np.int64)
        nbr[:, 0] = np.where(J > 0, cid(I, J - 1), -1)
        nbr[:, 1] = np.where(I < nx - 1, cid(I + 1, J), -1)
        nbr[:, 2] = np.where(J < ny - 1, cid(I, J + 1), -1)
        nbr[:, 3] = np.where(I > 0, cid(I - 1, J), -1)
        # boundary faces in tag order: left (i=0, slot 3), right (i=nx-1,
        # slot 1), bottom (j=0, slot 0), top (j=ny-1, slot 2)
        b_cells, b_slots, b_tags = [], [], []
        jj = np.arange(ny)
        ii = np.arange(nx)
        for tag, cells, slot in ((0, cid(0, jj), 3), (1, cid(nx - 1, jj), 1),
                                 (2, cid(ii, 0), 0), (3, cid(ii, ny - 1), 2)):
            b_cells.append(cells)
            b_slots.append(np.full(len(cells), slot))
            b_tags.append(np.full(len(cells), tag))
        b_cells = np.concatenate(b_cells)
        b_slots = np.concatenate(b_slots)
        b_tags = np.concatenate(b_tags)
        nbr[b_cells, b_slots] = -1 - np.arange(len(b_cells))
        # diagonal (vertex-only) neighbours of the 3x3 block
        extra = np.full((N, 4), -1, np.int64)
        col = np.zeros(N, np.int64)
        for di, dj in ((-1, -1), (-1, 1), (1, -1), (1, 1)):
            ok = (I + di >= 0) & (I + di < nx) & (J + dj >= 0) & (J + dj < ny)
            cand = np.where(ok, cid(I + di, J + dj), -1)
            rows = np.nonzero(ok)[0]
            extra[rows, col[rows]] = cand[rows]
            col[rows] += 1
        # keep the diagonal neighbours in ascending id order
        ex = np.where(extra < 0, np.iinfo(np.int64).max, extra)
        ex.sort(axis=1)
        ex[ex == np.iinfo(np.int64).max] = -1
        m = Mesh(dim=2, verts=verts, cell_vertices=cv,
                 cell_nbr=nbr.astype(np.int32),
                 bface_cell=b_cells.astype(np.int32),
                 bface_slot=b_slots.astype(np.int32),
                 bface_tag=b_tags.astype(np.int32),
                 tags=["left", "right", "bottom", "top"])
        m._extra = ex.astype(np.int32)
        return m

    @staticmethod
    def triangles(points: np.ndarray, simplices: np.ndarray,
                  box=None) -> "Mesh":
        """Triangle mesh from points ``(n, 2)`` and CCW or CW simplices.

        Boundary faces are tagged by the side of ``box`` they lie on
        (left/right/bottom/top); faces not on an axis-aligned box side get
        tag ``boundary``.
        """
        pts = np.asarray(points, np.float64)
        tri = np.asarray(simplices, np.int64).copy()
        a, b, c = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
        cross = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - \
                (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
        flip = cross < 0
        tri[flip, 1], tri[flip, 2] = tri[flip, 2].copy(), tri[flip, 1].copy()
        N = len(tri)
        # edges: slot k = (v_k, v_{k+1})
        e0 = tri
        e1 = np.roll(tri, -1, axis=1)
        lo = np.minimum(e0, e1).ravel()
        hi = np.maximum(e0, e1).ravel()
        key = lo * (len(pts) + 1) + hi
        order = np.argsort(key, kind="stable")
        ks = key[order]
        owner = order // 3
        same = np.zeros(len(ks), bool)
        same[1:] = ks[1:] == ks[:-1]
        nbr = np.full(3 * N, -1, np.int64)
        # pair consecutive equal keys
        idx = np.nonzero(same)[0]
        i1 = order[idx - 1]
        i2 = order[idx]
        nbr[i1] = i2 // 3
        nbr[i2] = i1 // 3
        nbr = nbr.reshape(N, 3)
        bmask = nbr < 0
        b_cells, b_slots = np.nonzero(bmask)
        # tags
        tags = ["left", "right", "bottom", "top", "boundary"]
        va = pts[tri[b_cells, b_slots]]
        vb = pts[tri[b_cells, (b_slots + 1) % 3]]
        mid = 0.5 * (va + vb)
        if box is None:
            box = (pts[:, 0].min(), pts[:, 0].max(),
                   pts[:, 1].min(), pts[:, 1].max())
        x0, x1, y0, y1 = box
        tol = 1e-9 * max(x1 - x0, y1 - y0)
        t = np.full(len(b_cells), 4)
        t[np.abs(mid[:, 1] - y1) < tol] = 3
        t[np.abs(mid[:, 1] - y0) < tol] = 2
        t[np.abs(mid[:, 0] - x1) < tol] = 1
        t[np.abs(mid[:, 0] - x0) < tol] = 0
        # order boundary faces by tag, then by cell (deterministic)
        o = np.lexsort((b_slots, b_cells, t))
        b_cells, b_slots, t = b_cells[o], b_slots[o], t[o]
        nbr[b_cells, b_slots] = -1 - np.arange(len(b_cells))
        return Mesh(dim=2, verts=pts, cell_vertices=tri.astype(np.int32),
                    cell_nbr=nbr.astype(np.int32),
                    bface_cell=b_cells.astype(np.int32),
                    bface_slot=b_slots.astype(np.int32),
                    bface_tag=t.astype(np.int32), tags=tags)

    # ── sizes ─────────────────────────────────────────────────────────
    @property
    def n_cells(self) -> int:
        return int(self.cell_vertices.shape[0])

    @property
    def max_faces(self) -> int:
        return int(self.cell_vertices.shape[1])

    @property
    def n_bfaces(self) -> int:
        return int(self.bface_cell.shape[0])

    @property
    def n_faces(self) -> int:
        n_int = int(np.count_nonzero(self.cell_nbr >= 0)) // 2
        return n_int + self.n_bfaces

    # ── boundary conditions ───────────────────────────────────────────
    def set_bc(self, spec) -> "Mesh":
        """``spec``: a BC name for every tag, or ``{tag: name}``, where a
        tag may instead take a :class:`Dirichlet` or a list of them."""
        if isinstance(spec, (str, int, Dirichlet)):
            spec = {t: spec for t in self.tags}
        unknown = set(spec) - set(self.tags)
        if unknown:
            raise ValueError(f"unknown boundary tags {sorted(unknown)}; "
                             f"known: {self.tags}")
        self.bc, self.bc_values = {}, {}
        for t in self.tags:
            v = spec.get(t, "extrapolation")
            if isinstance(v, Dirichlet):
                v = [v]
            if isinstance(v, (list, tuple)):
                if not v or not all(isinstance(d, Dirichlet) for d in v):
                    raise ValueError(f"tag {t!r}: expected Dirichlet entries")
                self.bc[t] = BC_DIRICHLET
                self.bc_values[t] = {d.row(): float(d.value) for d in v}
            else:
                self.bc[t] = bc_kind(v)
                if self.bc[t] == BC_DIRICHLET:
                    raise ValueError(f"tag {t!r}: a Dirichlet BC needs its "
                                     f"values (Dirichlet(on, value))")
        if any(k == BC_PERIODIC for k in self.bc.values()):
            self._resolve_periodic()
        return self

    def bface_kind(self) -> np.ndarray:
        if not self.bc:
            self.set_bc("extrapolation")
        table = np.array([self.bc[t] for t in self.tags], np.int32)
        return table[self.bface_tag]

    def bface_dirichlet(self):
        """``(mask (n_bf,) int32, values (4, n_bf) float64)``: bit v of
        mask[b] is set when row v of boundary face b is prescribed."""
        mask = np.zeros(self.n_bfaces, np.int32)
        val = np.zeros((4, self.n_bfaces))
        for ti, t in enumerate(self.tags):
            sel = self.bface_tag == ti
            for row, x in self.bc_values.get(t, {}).items():
                mask[sel] |= 1 << row
                val[row, sel] = x
        return mask, val

    def _resolve_periodic(self):
        """Pair left<->right and bottom<->top boundary faces by position
        (the reference's ``mesh.resolve_periodic_bcs``, REQ-116)."""
        mid = self.bface_midpoints()
        partner = np.full(self.n_bfaces, -1, np.int64)
        for a, b, ax in (("left", "right", 1), ("bottom", "top", 0)):
            if a not in self.tags:
                continue
            ta, tb = self.tags.index(a), self.tags.index(b)
            ia = np.nonzero(self.bface_tag == ta)[0]
            ib = np.nonzero(self.bface_tag == tb)[0]
            if len(ia) == 0:
                continue
            if self.dim == 1:
                partner[ia] = ib
                partner[ib] = ia
                continue
            oa = ia[np.argsort(mid[ia, ax])]
            ob = ib[np.argsort(mid[ib, ax])]
            partner[oa] = ob
            partner[ob] = oa
        self.periodic_partner = partner.astype(np.int32)

    def bface_midpoints(self) -> np.ndarray:
        c, s = self.bface_cell, self.bface_slot
        va = self.verts[self.cell_vertices[c, s]]
        if self.dim == 1:
            return va
        vb = self.verts[self.cell_vertices[c, (s + 1) % self.max_faces]]
        return 0.5 * (va + vb)

    # ── derived geometry (float64, host) ──────────────────────────────
    def cell_centers(self) -> np.ndarray:
        """Vertex average (centroid for triangles and parallelograms)."""
        return self.verts[self.cell_vertices].mean(axis=1)

    def cell_volumes(self) -> np.ndarray:
        v = self.verts[self.cell_vertices]
        if self.dim == 1:
            return v[:, 1, 0] - v[:, 0, 0]
        x, y = v[..., 0], v[..., 1]
        return 0.5 * np.sum(x * np.roll(y, -1, 1) - np.roll(x, -1, 1) * y, 1)

    def cell_inradius(self) -> np.ndarray:
        """2*area/perimeter (exact for triangles and squares); dx/2 in 1-D."""
        if self.dim == 1:
            return 0.5 * self.cell_volumes()
        v = self.verts[self.cell_vertices]
        e = np.roll(v, -1, 1) - v
        per = np.sqrt((e ** 2).sum(-1)).sum(1)
        return 2.0 * self.cell_volumes() / per

    def slot_geometry(self):
        """Per (cell, slot): outward unit normal (N,MF,2), length (N,MF),
        face centre (N,MF,2)."""
        v = self.verts[self.cell_vertices]
        N, MF = self.cell_vertices.shape
        if self.dim == 1:
            n = np.zeros((N, MF, 2))
            n[:, 0, 0] = -1.0
            n[:, 1, 0] = 1.0
            return n, np.ones((N, MF)), v.copy()
        w = np.roll(v, -1, 1)
        e = w - v
        L = np.sqrt((e ** 2).sum(-1))
        n = np.stack([e[..., 1], -e[..., 0]], -1) / L[..., None]
        return n, L, 0.5 * (v + w)

    def faces(self):
        """Reference-layout face list.

        Returns ``face_cells (2, F)`` (neighbour ``-1`` on boundary faces),
        ``face_slot (2, F)`` (slot of the face in A / in B), and
        ``bface_face (n_bf,)`` (face id of each boundary face).  Interior
        faces come first, ordered by owner cell (owner = lower cell id);
        boundary faces follow in boundary-face order.
        """
        N, MF = self.cell_nbr.shape
        c = np.repeat(np.arange(N), MF)
        s = np.tile(np.arange(MF), N)
        nb = self.cell_nbr.ravel().astype(np.int64)
        own = (nb >= 0) & (nb > c)
        A = c[own]
        B = nb[own]
        sA = s[own]
        # slot of the same face in B
        nbrB = self.cell_nbr[B]
        sB = np.argmax(nbrB == A[:, None], axis=1)
        nbf = self.n_bfaces
        fc = np.concatenate([np.stack([A, B]), np.stack(
            [self.bface_cell.astype(np.int64), -np.ones(nbf, np.int64)])], 1)
        fs = np.concatenate([np.stack([sA, sB]), np.stack(
            [self.bface_slot.astype(np.int64), -np.ones(nbf, np.int64)])], 1)
        bface_face = len(A) + np.arange(nbf)
        return fc.astype(np.int32), fs.astype(np.int32), \
            bface_face.astype(np.int32)

    def cell_faces(self) -> np.ndarray:
        """(N, MF) face id of each cell slot (inverse of :meth:`faces`)."""
        fc, fs, bff = self.faces()
        N, MF = self.cell_nbr.shape
        cf = np.full((N, MF), -1, np.int32)
        F = fc.shape[1]
        fid = np.arange(F)
        cf[fc[0], fs[0]] = fid
        inter = fc[1] >= 0
        cf[fc[1, inter], fs[1, inter]] = fid[inter]
        return cf

    def vertex_stencil(self) -> np.ndarray:
        """(N, MS) cells sharing a vertex with each cell, -1 padded.

        Face neighbours come first in slot order, then the remaining vertex
        neighbours by ascending id."""
        N, MF = self.cell_vertices.shape
        ex = getattr(self, "_extra", None)
        if ex is not None and self.dim == 2:
            fn = self.cell_nbr.astype(np.int64)
            full = np.full((N, MF + ex.shape[1]), -1, np.int64)
            col = np.zeros(N, np.int64)
            for k in range(MF):
                m = fn[:, k] >= 0
                full[m, col[m]] = fn[m, k]
                col[m] += 1
            for j in range(ex.shape[1]):
                m = ex[:, j] >= 0
                full[m, col[m]] = ex[m, j]
                col[m] += 1
            return full.astype(np.int32)
        nv = len(self.verts)
        cv = self.cell_vertices.ravel().astype(np.int64)
        cc = np.repeat(np.arange(N), MF)
        o = np.argsort(cv, kind="stable")
        vptr = np.zeros(nv + 1, np.int64)
        np.add.at(vptr, cv + 1, 1)
        vptr = np.cumsum(vptr)
        vcells = cc[o]
        if self.dim == 1:
            # an end cell reaches one cell further: its neighbour, then that
            # neighbour's other neighbour (the reference's bc_wall,
            # bc_periodic and bc_inflow derivative aux rows at the end cells
            # and its boundary_o2_small end state need exactly this)
            nb = self.cell_nbr
            st = np.where(nb >= 0, nb, -1).astype(np.int64)
            for c in np.nonzero((nb < 0).any(1))[0]:
                own = nb[c][nb[c] >= 0]
                if len(own) != 1:
                    continue
                far = [x for x in nb[own[0]] if x >= 0 and x != c]
                st[c] = [own[0], far[0] if far else -1]
            return st.astype(np.int32)
        rows = []
        maxs = 0
        # vectorised: for each cell, gather cells of its vertices
        cnt = vptr[1:] - vptr[:-1]
        per_cell = cnt[self.cell_vertices].sum(1)
        tot = per_cell.sum()
        owner = np.repeat(np.arange(N), per_cell)
        # build list of (owner, candidate)
        starts = vptr[self.cell_vertices]          # (N, MF)
        counts = cnt[self.cell_vertices]
        flat_start = starts.ravel()
        flat_cnt = counts.ravel()
        base = np.repeat(flat_start, flat_cnt)
        off = np.arange(tot) - np.repeat(np.cumsum(flat_cnt) - flat_cnt,
                                         flat_cnt)
        cand = vcells[base + off]
        keep = cand != owner
        owner, cand = owner[keep], cand[keep]
        key = owner * (N + 1) + cand
        key = np.unique(key)
        owner = key // (N + 1)
        cand = key % (N + 1)
        counts = np.bincount(owner, minlength=N)
        MS = int(counts.max())
        out = np.full((N, MS), -1, np.int64)
        # face neighbours first
        fn = self.cell_nbr.astype(np.int64)
        pos = np.cumsum(counts) - counts
        rank = np.arange(len(owner)) - np.repeat(pos, counts)
        out[owner, rank] = cand
        # reorder each row: face neighbours (slot order) first
        res = np.full((N, MS), -1, np.int64)
        isface = np.zeros((N, MS), bool)
        for k in range(MF):
            isface |= (out == fn[:, k:k + 1]) & (fn[:, k:k + 1] >= 0)
        # place face nbrs in slot order
        col = np.zeros(N, np.int64)
        for k in range(MF):
            m = fn[:, k] >= 0
            res[m, col[m]] = fn[m, k]
            col[m] += 1
        # remaining
        rest = np.where(isface, -1, out)
        rest_sorted = np.sort(np.where(rest < 0, np.iinfo(np.int64).max,
                                       rest), axis=1)
        rest_sorted[rest_sorted == np.iinfo(np.int64).max] = -1
        for j in range(MS):
            v = rest_sorted[:, j]
            m = v >= 0
            idx = col[m]
            ok = idx < MS
            rr = np.nonzero(m)[0][ok]
            res[rr, idx[ok]] = v[m][ok]
            col[rr] += 1
        return res.astype(np.int32)

    def extra_stencil(self) -> np.ndarray:
        """(N, MX) vertex neighbours that are not face neighbours, -1 padded
        (the part of the LSQ stencil beyond the face slots)."""
        ex = getattr(self, "_extra", None)
        if ex is not None:
            return ex
        if self.dim == 2:
            self._extra = self._extra_stencil_2d()
            return self._extra
        st = self.vertex_stencil()
        nf = np.count_nonzero(self.cell_nbr >= 0, axis=1)
        N, MS = st.shape
        MF = self.max_faces
        # vertex_stencil puts the face neighbours first
        out = np.full((N, max(MS - 1, 1)), -1, np.int32)
        for c_nf in np.unique(nf):
            rows = np.nonzero(nf == c_nf)[0]
            rest = st[rows, c_nf:]
            out[rows, :rest.shape[1]] = rest[:, :out.shape[1]]
        width = int((out >= 0).sum(1).max()) if out.size else 0
        self._extra = np.ascontiguousarray(out[:, :max(width, 1)])
        return self._extra

    def _extra_stencil_2d(self) -> np.ndarray:
        """The vertex neighbours of each cell that are not its face
        neighbours, ascending, -1 padded -- the same table as the generic
        path through :meth:`vertex_stencil`, built with fixed-width row
        sorts (the cells of each vertex padded to the largest vertex
        degree)."""
        N, MF = self.cell_vertices.shape
        nv = len(self.verts)
        cv = self.cell_vertices.astype(np.int64)
        flat = cv.ravel()
        cells = np.repeat(np.arange(N, dtype=np.int64), MF)
        o = np.argsort(flat, kind="stable")
        deg = np.bincount(flat, minlength=nv)
        W = int(deg.max())
        start = np.zeros(nv + 1, np.int64)
        start[1:] = np.cumsum(deg)
        rank = np.arange(len(flat)) - start[flat[o]]
        vc = np.full((nv, W), -1, np.int64)
        vc[flat[o], rank] = cells[o]
        big = np.iinfo(np.int64).max
        cand = vc[cv].reshape(N, MF * W)
        me = np.arange(N, dtype=np.int64)[:, None]
        drop = (cand < 0) | (cand == me)
        fn = self.cell_nbr.astype(np.int64)
        for k in range(MF):
            drop |= cand == np.where(fn[:, k:k + 1] >= 0, fn[:, k:k + 1], -2)
        cand = np.where(drop, big, cand)
        cand.sort(axis=1)
        dup = np.zeros_like(cand, bool)
        dup[:, 1:] = cand[:, 1:] == cand[:, :-1]
        cand[dup] = big
        cand.sort(axis=1)
        width = max(int((cand != big).sum(1).max()), 1)
        out = cand[:, :width]
        return np.ascontiguousarray(np.where(out == big, -1, out), np.int32)

    # ── reordering ────────────────────────────────────────────────────
    def morton_permutation(self, bits: int = 16) -> np.ndarray:
        """Cell permutation along the Z-order (Morton) curve of the cell
        centres (the reference orders points the same way in
        ``gnn_blueprint/mesh_spectral_geometry.py:34-46``).  On a 2^k x 2^k
        grid every run of 4^m consecutive cells is a 2^m x 2^m tile, so a
        block of 256 cells gathers its neighbours from one 16 x 16 tile and
        its one-cell ring.  Both axes share one scale, so that runs of the
        curve are square in space on any domain."""
        xc = self.cell_centers()
        lo = self.verts.min(0)
        span = max(float((self.verts.max(0) - lo).max()), 1e-300)
        q = np.floor((xc - lo) / span * (1 << bits)).astype(np.int64)
        q = np.clip(q, 0, (1 << bits) - 1)
        key = np.zeros(len(xc), np.int64)
        for b in range(bits):
            key |= ((q[:, 0] >> b) & 1) << (2 * b)
            key |= ((q[:, 1] >> b) & 1) << (2 * b + 1)
        return np.argsort(key, kind="stable")

    def reorder(self, perm: np.ndarray, renumber_vertices: bool = True) -> "Mesh":
        """New mesh with cells in order ``perm`` (new id i = old perm[i]);
        vertices are renumbered in order of first use by the new cell order
        so that vertex gathers follow the cells."""
        perm = np.asarray(perm, np.int64)
        N = self.n_cells
        inv = np.empty(N, np.int64)
        inv[perm] = np.arange(N)
        nbr = self.cell_nbr[perm].astype(np.int64)
        nbr = np.where(nbr >= 0, inv[np.maximum(nbr, 0)], nbr)
        ex = getattr(self, "_extra", None)
        cv = self.cell_vertices[perm]
        verts = self.verts
        if renumber_vertices:
            flat = cv.ravel()
            _, first = np.unique(flat, return_index=True)
            order = np.unique(flat)[np.argsort(first, kind="stable")]
            vnew = np.full(len(self.verts), -1, np.int64)
            vnew[order] = np.arange(len(order))
            cv = vnew[cv]
            verts = self.verts[order]
        m = Mesh(dim=self.dim, verts=np.ascontiguousarray(verts),
                 cell_vertices=np.ascontiguousarray(cv, np.int32),
                 cell_nbr=nbr.astype(np.int32),
                 bface_cell=inv[self.bface_cell].astype(np.int32),
                 bface_slot=self.bface_slot.copy(),
                 bface_tag=self.bface_tag.copy(), tags=list(self.tags),
                 bc=dict(self.bc),
                 bc_values={t: dict(v) for t, v in self.bc_values.items()},
                 periodic_partner=(None if self.periodic_partner is None
                                   else self.periodic_partner.copy()))
        if ex is not None:
            e = ex[perm].astype(np.int64)
            e = np.where(e >= 0, inv[np.maximum(e, 0)], -1)
            e = np.where(e < 0, np.iinfo(np.int64).max, e)
            e.sort(axis=1)
            e[e == np.iinfo(np.int64).max] = -1
            m._extra = e.astype(np.int32)
        return m

    @staticmethod
    def union(meshes) -> "Mesh":
        """Disjoint union of 2-D meshes with the same tags and boundary
        conditions: the cells of mesh k follow those of meshes 0..k-1, no
        face joins two of them (the batched ensemble's layout)."""
        if not meshes:
            raise ValueError("union of no meshes")
        m0 = meshes[0]
        for m in meshes:
            if m.dim != m0.dim or m.tags != m0.tags or m.bc != m0.bc or \
                    m.periodic_partner is not None or m.bc_values:
                raise ValueError("union needs meshes of one dimension, the "
                                 "same tags and non-periodic, non-Dirichlet "
                                 "boundary conditions")
            if m.max_faces != m0.max_faces:
                raise ValueError("union needs one number of faces per cell")
        vo = np.cumsum([0] + [len(m.verts) for m in meshes])
        co = np.cumsum([0] + [m.n_cells for m in meshes])
        bo = np.cumsum([0] + [m.n_bfaces for m in meshes])
        nbr = []
        for k, m in enumerate(meshes):
            nb = m.cell_nbr.astype(np.int64)
            nbr.append(np.where(nb >= 0, nb + co[k], nb - bo[k]))
        ex = [m.extra_stencil().astype(np.int64) for m in meshes]
        w = max(e.shape[1] for e in ex)
        exu = np.full((co[-1], w), -1, np.int64)
        for k, e in enumerate(ex):
            exu[co[k]:co[k + 1], :e.shape[1]] = np.where(e >= 0, e + co[k],
                                                         -1)
        u = Mesh(dim=m0.dim,
                 verts=np.concatenate([m.verts for m in meshes]),
                 cell_vertices=np.concatenate(
                     [m.cell_vertices.astype(np.int64) + vo[k]
                      for k, m in enumerate(meshes)]).astype(np.int32),
                 cell_nbr=np.concatenate(nbr).astype(np.int32),
                 bface_cell=np.concatenate(
                     [m.bface_cell.astype(np.int64) + co[k]
                      for k, m in enumerate(meshes)]).astype(np.int32),
                 bface_slot=np.concatenate([m.bface_slot for m in meshes]),
                 bface_tag=np.concatenate([m.bface_tag for m in meshes]),
                 tags=list(m0.tags), bc=dict(m0.bc))
        u._extra = np.ascontiguousarray(exu, np.int32)
        return u

    def adjacency_csr(self):
        """Face-neighbour graph as CSR (indptr, indices), int32, the column
        indices of each row ascending (the canonical CSR, as scipy's
        sort_indices leaves it)."""
        N, MF = self.cell_nbr.shape
        nb = self.cell_nbr.astype(np.int64)
        m = nb >= 0
        counts = m.sum(1)
        indptr = np.zeros(N + 1, np.int64)
        indptr[1:] = np.cumsum(counts)
        srt = np.sort(np.where(m, nb, np.iinfo(np.int64).max), axis=1)
        indices = srt[srt != np.iinfo(np.int64).max]
        return indptr.astype(np.int32), indices.astype(np.int32)


def jittered_points(nx: int, ny: int, box, jitter: float, seed: int):
    """``nx x ny`` lattice over ``box`` with each interior point moved by
    ``U[-jitter, jitter] * spacing`` (boundary points stay on the box)."""
    x0, x1, y0, y1 = box
    rng = np.random.default_rng(seed)
    xs = np.linspace(x0, x1, nx)
    ys = np.linspace(y0, y1, ny)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    hx, hy = (x1 - x0) / (nx - 1), (y1 - y0) / (ny - 1)
    jx = rng.uniform(-jitter, jitter, X.shape) * hx
    jy = rng.uniform(-jitter, jitter, Y.shape) * hy
    interior_x = np.ones_like(X, bool)
    interior_x[0, :] = interior_x[-1, :] = False
    interior_y = np.ones_like(Y, bool)
    interior_y[:, 0] = interior_y[:, -1] = False
    X = X + np.where(interior_x, jx, 0.0)
    Y = Y + np.where(interior_y, jy, 0.0)
    return np.stack([X.ravel(), Y.ravel()], 1)


def lattice_triangles(nx: int, ny: int, box, jitter: float, seed: int) -> Mesh:
    """Triangle mesh of an ``nx x ny`` jittered lattice of quads over
    ``box`` (the points of :func:`jittered_points`), each quad cut into two
    triangles along its (0,0)-(1,1) diagonal: ``2 nx ny`` triangles without a
    Delaunay pass.  A point can move toward the diagonal opposite it by
    sqrt(2) jitter and so can that diagonal, against a nominal distance of
    spacing / sqrt(2): a jitter of at most 0.15 of the spacing keeps every
    triangle's height above 0.28 of the spacing."""
    if not 0.0 <= jitter <= 0.15:
        raise ValueError("lattice_triangles: jitter must be in [0, 0.15]")
    pts = jittered_points(nx + 1, ny + 1, box, jitter, seed)
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    a = i * (ny + 1) + j
    b = a + (ny + 1)
    c = b + 1
    d = a + 1
    tri = np.stack([np.stack([a, b, c], -1), np.stack([a, c, d], -1)],
                   -2).reshape(-1, 3)
    return Mesh.triangles(pts, tri, box=box)


def delaunay_mesh(points: np.ndarray, box) -> Mesh:
    """Triangle mesh of ``points`` by ``scipy.spatial.Delaunay`` (host-side
    set-up only; the time step never calls it)."""
    from scipy.spatial import Delaunay
    tri = Delaunay(points)
    return Mesh.triangles(points, tri.simplices, box=box)
