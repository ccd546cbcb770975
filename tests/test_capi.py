"""The C-ABI library loads and exports every symbol include/ declares."""
import pathlib
import re

from zoomy_hip import _lib

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(zj_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_declared_symbols():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.exported_symbols())


def test_version_string():
    assert _lib.lib().zj_version().startswith(b"zoomy_hip")


def test_argument_checks_before_any_device_work():
    """Entry points reject bad arguments with ZJ_EINVAL before they touch
    the device (these calls run without a GPU)."""
    import ctypes
    L = _lib.lib()
    null = ctypes.c_void_p(None)
    assert L.zj_diffusion_create(null, null) == _lib.ZJ_EINVAL
    assert L.zj_diffusion_explicit(null, 1.0, null, null, null) == _lib.ZJ_EINVAL
    assert L.zj_diffusion_cn(null, 1.0, 0.1, null, null, 1e-12, 10,
                             null) == _lib.ZJ_EINVAL
    assert L.zj_diffusion_stats(null, null, null) == _lib.ZJ_EINVAL
    assert L.zj_diffusion_destroy(null) == _lib.ZJ_OK
    assert L.zj_solver_set_stage_aux(null, 0) == _lib.ZJ_EINVAL
    assert L.zj_mesh_create(None, null) == _lib.ZJ_EINVAL
    # the block contract and the per-face entry points
    assert L.zj_mesh_set_faces(null, 4, null, null) == _lib.ZJ_EINVAL
    assert L.zj_face_speeds(null, null, null, null, null,
                            null) == _lib.ZJ_EINVAL
    assert L.zj_reduce_dt(null, null, null, null, 0.9, 2, 0, -1.0, -1.0,
                          null, null) == _lib.ZJ_EINVAL
    assert L.zj_bc_values(null, null, null, null) == _lib.ZJ_EINVAL
    assert L.zj_reconstruct(null, null, null, 0, null, null, null,
                            null) == _lib.ZJ_EINVAL
    assert L.zj_flux_pass(null, null, null, null, null, null, null, null,
                          null) == _lib.ZJ_EINVAL
    assert L.zj_gather_update(null, null, null, null, null, null, 1.0, null,
                              null, 0, null, null, null, null,
                              null) == _lib.ZJ_EINVAL
    assert L.zj_flux_operator_o1(null, null, null, null,
                                 null) == _lib.ZJ_EINVAL
    assert L.zj_face_numerics(null, 4, null, null, null, null, null, null,
                              null, null, null) == _lib.ZJ_EINVAL


def test_mesh_size_limit():
    """Meshes beyond 32-bit row offsets (8 rows of N cells) are refused
    before any array is read."""
    import ctypes
    d = _lib.MeshDesc()
    d.dim, d.n_cells, d.n_verts, d.max_faces = 2, 300_000_000, 4, 4
    h = ctypes.c_void_p()
    assert _lib.lib().zj_mesh_create(ctypes.byref(d),
                                     ctypes.byref(h)) == _lib.ZJ_EUNSUP


def test_dirichlet_descriptor_validation():
    """A Dirichlet face needs its tables, and only state rows 0-3 may be
    prescribed (host-side checks of zj_mesh_create / zj_patch_check)."""
    from zoomy_hip.fvm.solver import mesh_desc
    from zoomy_hip.mesh import Dirichlet, Mesh
    m = Mesh.line(8, 0.0, 1.0).set_bc({"left": Dirichlet("h", 1.0)})
    mask, val = m.bface_dirichlet()
    assert mask.tolist() == [2, 0] and val[1, 0] == 1.0
    import pytest
    with pytest.raises(ValueError):
        Mesh.line(8).set_bc({"left": Dirichlet("u", 1.0)})
    with pytest.raises(ValueError):
        Mesh.line(8).set_bc({"left": "dirichlet"})
    with pytest.raises(ValueError):
        Mesh.line(8).set_bc({"middle": "wall"})
