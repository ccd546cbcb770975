"""Host-side mesh layer: geometry invariants, RCM against scipy fixtures,
and the LDS patch decomposition (built and validated on the host)."""
import glob
import pathlib

import numpy as np
import pytest

from zoomy_hip import _lib
from zoomy_hip.fvm.solver import patch_plan, patch_stats
from zoomy_hip.mesh import Mesh, delaunay_mesh, jittered_points

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def _meshes():
    q = Mesh.quad(37, 23, (0.0, 2.0, -1.0, 0.5))
    yield "quad", q
    yield "quad_morton", q.reorder(q.morton_permutation())
    t = delaunay_mesh(jittered_points(25, 14, (0.0, 10.0, 0.0, 5.0), 0.3, 3),
                      (0.0, 10.0, 0.0, 5.0))
    yield "tri", t
    ip, ix = t.adjacency_csr()
    perm = np.zeros(t.n_cells, np.int32)
    assert _lib.lib().zj_rcm(t.n_cells, ip.ctypes.data, ix.ctypes.data,
                             perm.ctypes.data) == 0
    yield "tri_rcm", t.reorder(perm)
    yield "line", Mesh.line(300, 0.0, 10.0)


@pytest.mark.parametrize("name,mesh", list(_meshes()))
def test_geometry_invariants(name, mesh):
    mesh.set_bc("wall")
    vol = mesh.cell_volumes()
    assert (vol > 0).all()
    x = mesh.verts
    area = (x[:, 0].max() - x[:, 0].min()) * (
        (x[:, 1].max() - x[:, 1].min()) if mesh.dim == 2 else 1.0)
    assert abs(vol.sum() - area) < 1e-12 * area
    n, L, fc = mesh.slot_geometry()
    assert np.allclose((n ** 2).sum(-1), 1.0)
    # closed cells: sum of n |f| vanishes
    assert np.abs((n * L[..., None]).sum(1)).max() < 1e-12
    # neighbour symmetry and opposite normals across every interior face
    N, MF = mesh.cell_nbr.shape
    for c in range(0, N, max(1, N // 200)):
        for k in range(MF):
            nb = mesh.cell_nbr[c, k]
            if nb < 0:
                continue
            kk = list(mesh.cell_nbr[nb]).index(c)
            assert np.allclose(n[c, k], -n[nb, kk])
            assert np.allclose(fc[c, k], fc[nb, kk])


@pytest.mark.parametrize("name,mesh", list(_meshes()))
@pytest.mark.parametrize("P", [256, 64])
def test_patch_decomposition_valid(name, mesh, P):
    rc, st = patch_stats(mesh, P)
    assert rc == 0, (rc, st)
    assert st["npatch"] == -(-mesh.n_cells // P)
    assert st["max_slope"] >= min(P, mesh.n_cells)


def test_patch_templates_shared_on_regular_mesh():
    m = Mesh.quad(256, 256)
    m = m.reorder(m.morton_permutation())
    rc, st = patch_stats(m, 256)
    assert rc == 0
    # interior 16x16 tiles share one template; only boundary tiles differ
    assert st["ntemplates"] <= 16
    assert st["max_load"] <= 448 and st["max_slope"] <= 320


def _rcm_delaunay(nx, ny):
    t = delaunay_mesh(jittered_points(nx, ny, (0.0, 10.0, 0.0, 5.0), 0.3, 3),
                      (0.0, 10.0, 0.0, 5.0))
    ip, ix = t.adjacency_csr()
    perm = np.zeros(t.n_cells, np.int32)
    assert _lib.lib().zj_rcm(t.n_cells, ip.ctypes.data, ix.ctypes.data,
                             perm.ctypes.data) == 0
    return t.reorder(perm)


def test_patch_plan_clusters_strip_orders():
    """The plan zj_mesh_create follows: a cell order whose id runs are thin
    strips (RCM on a Delaunay mesh, row-major quads) is cut into clusters of
    the Morton order of the centres, at the full patch size, with the ring
    of a Morton-ordered mesh; Morton-ordered meshes keep consecutive ids.
    Every plan validates (each cell owned by exactly one patch)."""
    t = _rcm_delaunay(200, 100)
    rc, cons = patch_plan(t, patch_order="consecutive")
    assert rc == 0 and cons["clustered"] == 0 and cons["P"] < 256
    rc, st = patch_plan(t)
    assert rc == 0 and st["clustered"] == 1 and st["P"] == 256
    assert st["max_load"] <= int(2.6 * 256)
    tm = t.reorder(t.morton_permutation())
    rc, sm = patch_plan(tm)
    assert rc == 0 and sm["clustered"] == 0 and sm["P"] == 256
    q = Mesh.quad(64, 256)        # a patch of consecutive ids is one line
    rc, sq = patch_plan(q)
    assert rc == 0 and sq["clustered"] == 1 and sq["max_load"] <= 448
    qm = q.reorder(q.morton_permutation())
    rc, sqm = patch_plan(qm)
    assert rc == 0 and sqm["clustered"] == 0
    rc, forced = patch_plan(qm, 64, "clusters")
    assert rc == 0 and forced["clustered"] == 1 and forced["P"] == 64
    assert forced["npatch"] == -(-qm.n_cells // 64)


def test_patch_plan_arguments():
    m = Mesh.quad(20, 20)
    assert patch_plan(m, 100)[0] == _lib.ZJ_EINVAL      # not a power of two
    assert patch_plan(m, 512)[0] == _lib.ZJ_EINVAL      # above the block
    assert patch_plan(m, 16)[0] == _lib.ZJ_EINVAL
    assert patch_plan(Mesh.line(100), 0,
                      "clusters")[0] == _lib.ZJ_EUNSUP  # 2-D only


@pytest.mark.parametrize("f", sorted(glob.glob(str(GOLDEN / "rcm_*.npz"))))
def test_rcm_matches_scipy(f):
    d = np.load(f)
    n = len(d["indptr"]) - 1
    out = np.zeros(n, np.int32)
    ip = np.ascontiguousarray(d["indptr"], np.int32)
    ix = np.ascontiguousarray(d["indices"], np.int32)
    assert _lib.lib().zj_rcm(n, ip.ctypes.data, ix.ctypes.data,
                             out.ctypes.data) == 0
    assert np.array_equal(out, d["perm"])


def test_rcm_matches_scipy_with_tied_degrees():
    """scipy's RCM of the adjacency of a reduced wet/dry triangle mesh
    (9,702 Delaunay cells), the BFS levels full of equal degrees and no
    de-tying edges: bit-exact, from the lowest-degree start and, through
    zj_rcm_from, from the start node scipy chose (perm[-1])."""
    d = np.load(GOLDEN / "rcm_tied_wetdry.npz")
    n = len(d["indptr"]) - 1
    ip = np.ascontiguousarray(d["indptr"], np.int32)
    ix = np.ascontiguousarray(d["indices"], np.int32)
    deg = np.diff(ip)
    # many equal degrees inside the levels
    assert np.bincount(deg).max() > n // 2
    out = np.zeros(n, np.int32)
    assert _lib.lib().zj_rcm(n, ip.ctypes.data, ix.ctypes.data,
                             out.ctypes.data) == 0
    assert np.array_equal(out, d["perm"])
    out2 = np.zeros(n, np.int32)
    assert _lib.lib().zj_rcm_from(n, ip.ctypes.data, ix.ctypes.data,
                                  int(d["perm"][-1]), out2.ctypes.data) == 0
    assert np.array_equal(out2, d["perm"])
    # another start gives another valid permutation ending at that start
    other = int(np.argmax(deg))
    assert _lib.lib().zj_rcm_from(n, ip.ctypes.data, ix.ctypes.data, other,
                                  out2.ctypes.data) == 0
    assert out2[-1] == other and sorted(out2) == list(range(n))
    assert _lib.lib().zj_rcm_from(n, ip.ctypes.data, ix.ctypes.data, n,
                                  out2.ctypes.data) == _lib.ZJ_EINVAL


def test_rcm_reduces_bandwidth():
    t = delaunay_mesh(jittered_points(40, 30, (0.0, 10.0, 0.0, 5.0), 0.3, 1),
                      (0.0, 10.0, 0.0, 5.0))
    rng = np.random.default_rng(0)
    t = t.reorder(rng.permutation(t.n_cells))
    ip, ix = t.adjacency_csr()
    perm = np.zeros(t.n_cells, np.int32)
    _lib.lib().zj_rcm(t.n_cells, ip.ctypes.data, ix.ctypes.data,
                      perm.ctypes.data)
    assert sorted(perm) == list(range(t.n_cells))

    def bw(m):
        nb = m.cell_nbr
        c = np.arange(m.n_cells)[:, None]
        return np.abs(np.where(nb >= 0, nb - c, 0)).max()
    assert bw(t.reorder(perm)) < bw(t) / 5


def test_fast_extra_stencil_equals_the_vertex_stencil_path():
    """The fixed-width extra stencil of 2-D meshes is the same table as the
    vertex-stencil path (face neighbours removed, the rest ascending)."""
    import numpy as np
    from zoomy_hip.mesh import (Mesh, delaunay_mesh, jittered_points,
                                lattice_triangles)
    for m in (Mesh.quad(13, 9),
              delaunay_mesh(jittered_points(20, 15, (0, 1, 0, 1), 0.3, 1),
                            (0, 1, 0, 1)),
              lattice_triangles(30, 20, (0, 1, 0, 1), 0.15, 2)):
        st = m.vertex_stencil()
        nf = np.count_nonzero(m.cell_nbr >= 0, axis=1)
        rest = np.full((m.n_cells, st.shape[1]), -1, np.int64)
        for c in range(m.n_cells):
            r = st[c, nf[c]:]
            rest[c, :len(r)] = r
        width = int((rest >= 0).sum(1).max())
        np.testing.assert_array_equal(m._extra_stencil_2d(), rest[:, :width])


def test_lattice_triangles_and_union():
    """The jittered lattice triangulation is a valid closed mesh (positive
    areas summing to the box, every interior slot paired), and the union of
    meshes keeps each member's cells, faces and stencils, shifted."""
    import numpy as np
    import pytest
    from zoomy_hip.mesh import Mesh, lattice_triangles
    m = lattice_triangles(12, 7, (0, 2, 0, 1), 0.15, 5).set_bc("wall")
    assert m.n_cells == 2 * 12 * 7
    assert np.all(m.cell_volumes() > 0)
    # the jitter bound keeps the triangles fat: inradius >= 0.08 spacing
    big = lattice_triangles(200, 200, (0, 1, 0, 1), 0.15, 9)
    assert big.cell_inradius().min() > 0.08 / 200
    with pytest.raises(ValueError):
        lattice_triangles(4, 4, (0, 1, 0, 1), 0.3, 1)
    assert abs(m.cell_volumes().sum() - 2.0) < 1e-12
    nb = m.cell_nbr
    for c in range(m.n_cells):
        for k in range(3):
            if nb[c, k] >= 0:
                assert c in nb[nb[c, k]]
    m2 = Mesh.quad(5, 4).set_bc("wall")
    m3 = lattice_triangles(4, 4, (0, 1, 0, 1), 0.15, 6).set_bc("wall")
    u = Mesh.union([m, m3])
    assert u.n_cells == m.n_cells + m3.n_cells
    assert u.n_bfaces == m.n_bfaces + m3.n_bfaces
    np.testing.assert_allclose(u.cell_volumes()[m.n_cells:], m3.cell_volumes())
    ex = u.extra_stencil()[m.n_cells:]
    e3 = m3.extra_stencil()
    np.testing.assert_array_equal(ex[:, :e3.shape[1]],
                                  np.where(e3 >= 0, e3 + m.n_cells, -1))
    with pytest.raises(ValueError):
        Mesh.union([m, m2])        # quads and triangles do not mix
