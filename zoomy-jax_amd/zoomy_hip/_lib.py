"""ctypes binding of ``libzoomy_hip.so`` (the C ABI in ``include/zoomy_hip.h``).

The library is built in-tree by ``csrc/build.sh``; there is no fallback: if
it is missing the import fails with an error that says how to build it.
"""
from __future__ import annotations

import ctypes
import pathlib

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libzoomy_hip.so"

ZJ_OK, ZJ_EINVAL, ZJ_EHIP, ZJ_ENOMEM, ZJ_EUNSUP, ZJ_ESOLVE = 0, 1, 2, 3, 4, 5
ZJ_EIO, ZJ_EEXIST = 6, 7
PRECOND = {"mg": 0, "learned": 1, "mixed": 2, "learned_guess": 3}
LIMITERS = {"none": 0, "barth_jespersen": 1, "minmod": 2,
            "venkatakrishnan": 3, "vanleer": 4}
RIEMANN = {"rusanov_nc": 0, "rusanov": 1, "hll": 2}
RECON = {"conservative": 0, "eta": 1, "xz": 2, "eta_hu": 3}
POSITIVITY = {"": 0, "zhang_shu": 1}
STAGE_AUX = {"refresh": 0, "frozen": 1}
DTYPES = {"float64": 0, "float32": 1}
PATCH_ORDER = {"auto": 0, "consecutive": 1, "clusters": 2}


class ZoomyHipError(RuntimeError):
    pass


class MeshDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("n_cells", ctypes.c_int32),
                ("n_verts", ctypes.c_int32), ("max_faces", ctypes.c_int32),
                ("max_extra", ctypes.c_int32), ("n_bf", ctypes.c_int32),
                ("cell_vertices", ctypes.c_void_p),
                ("cell_nbr", ctypes.c_void_p),
                ("cell_extra", ctypes.c_void_p),
                ("verts", ctypes.c_void_p),
                ("cell_center", ctypes.c_void_p),
                ("cell_inradius", ctypes.c_void_p),
                ("bface_cell", ctypes.c_void_p),
                ("bface_slot", ctypes.c_void_p),
                ("bface_kind", ctypes.c_void_p),
                ("bface_partner", ctypes.c_void_p),
                ("bface_mask", ctypes.c_void_p),
                ("bface_value", ctypes.c_void_p),
                ("patch_cells", ctypes.c_int32),
                ("patch_order", ctypes.c_int32)]


class Params(ctypes.Structure):
    _fields_ = [("g", ctypes.c_double), ("cfl", ctypes.c_double),
                ("eps_wet", ctypes.c_double), ("dt_max", ctypes.c_double),
                ("manning_n", ctypes.c_double), ("order", ctypes.c_int32),
                ("limiter", ctypes.c_int32), ("riemann", ctypes.c_int32),
                ("recon", ctypes.c_int32), ("dtype", ctypes.c_int32)]


class ChorinParams(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double),
                ("eps_wet", ctypes.c_double), ("maxiter", ctypes.c_int32),
                ("nu", ctypes.c_int32), ("coarse_sweeps", ctypes.c_int32),
                ("min_coarse", ctypes.c_int32), ("precond", ctypes.c_int32),
                ("time_order", ctypes.c_int32)]


class ChorinStats(ctypes.Structure):
    _fields_ = [("solves", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("fallbacks", ctypes.c_int32),
                ("lp_iterations", ctypes.c_int32),
                ("lp_solves", ctypes.c_int32),
                ("relres", ctypes.c_double),
                ("lp_relres", ctypes.c_double)]


_lib = None

_SIGS = {
    "zj_version": (ctypes.c_char_p, []),
    "zj_mesh_create": (ctypes.c_int, [ctypes.POINTER(MeshDesc),
                                      ctypes.POINTER(ctypes.c_void_p)]),
    "zj_mesh_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "zj_rcm": (ctypes.c_int, [ctypes.c_int32, ctypes.c_void_p,
                              ctypes.c_void_p, ctypes.c_void_p]),
    "zj_rcm_from": (ctypes.c_int, [ctypes.c_int32, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_int32,
                                   ctypes.c_void_p]),
    "zj_face_numerics": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]
                         + [ctypes.c_void_p] * 9),
    "zj_patch_plan": (ctypes.c_int, [ctypes.POINTER(MeshDesc),
                                     ctypes.c_void_p]),
    "zj_patch_check": (ctypes.c_int, [ctypes.POINTER(MeshDesc),
                                      ctypes.c_int32, ctypes.c_void_p]),
    "zj_solver_create": (ctypes.c_int, [ctypes.c_void_p,
                                        ctypes.POINTER(Params),
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "zj_solver_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "zj_flux_operator": (ctypes.c_int, [ctypes.c_void_p] * 4),
    "zj_compute_dt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_double, ctypes.c_void_p,
                                     ctypes.c_void_p]),
    "zj_compute_dt_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_double, ctypes.c_void_p,
                                             ctypes.c_void_p]),
    "zj_step": (ctypes.c_int, [ctypes.c_void_p] * 6),
    "zj_mesh_set_faces": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                         ctypes.c_void_p, ctypes.c_void_p]),
    "zj_face_speeds": (ctypes.c_int, [ctypes.c_void_p] * 6),
    "zj_reduce_dt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_double, ctypes.c_int32,
                                    ctypes.c_int32, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_void_p,
                                    ctypes.c_void_p]),
    "zj_bc_values": (ctypes.c_int, [ctypes.c_void_p] * 4),
    "zj_reconstruct": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_int32,
                                      ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p]),
    "zj_flux_pass": (ctypes.c_int, [ctypes.c_void_p] * 9),
    "zj_gather_update": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_double, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_int32,
                                        ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p]),
    "zj_flux_operator_o1": (ctypes.c_int, [ctypes.c_void_p] * 5),
    "zj_imex_global_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double,
                                             ctypes.c_int32, ctypes.c_void_p]),
    "zj_imex_global_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "zj_imex_global_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double,
                                            ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_int32, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_int32,
                                            ctypes.c_void_p]),
    "zj_imex_global_stats": (ctypes.c_int, [ctypes.c_void_p] * 3),
    "zj_ensemble_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                          ctypes.c_void_p, ctypes.c_void_p]),
    "zj_ensemble_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "zj_ensemble_dt": (ctypes.c_int, [ctypes.c_void_p] * 4),
    "zj_ensemble_march": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_int32, ctypes.c_double,
                                         ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p]),
    "zj_ensemble_capture": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_int32, ctypes.c_double,
                                           ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p]),
    "zj_ensemble_replay": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "zj_ensemble_mass": (ctypes.c_int, [ctypes.c_void_p] * 4),
    "zj_update_aux": (ctypes.c_int, [ctypes.c_void_p] * 4),
    "zj_march": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_void_p,
                                ctypes.c_int32, ctypes.c_double,
                                ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_void_p]),
    "zj_mass": (ctypes.c_int, [ctypes.c_void_p] * 4),
    "zj_time_kernels": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_int32,
                                       ctypes.c_void_p, ctypes.c_void_p]),
    "zj_solver_set_mood": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "zj_solver_set_positivity": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                                ctypes.c_double]),
    "zj_solver_set_stage_aux": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "zj_diffusion_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "zj_diffusion_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "zj_diffusion_explicit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double,
                                             ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p]),
    "zj_diffusion_cn": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_double,
                                       ctypes.c_int32, ctypes.c_void_p]),
    "zj_diffusion_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p]),
    "zj_solver_mood_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "zj_imex_local": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_int32, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p]),
}

_V = ctypes.c_void_p
_D = ctypes.c_double
# the Chorin / Poisson / learned-preconditioner entry points
_SIGS_POISSON = {
    "zj_chorin_create": (ctypes.c_int, [_V, ctypes.c_int32, ctypes.c_int32,
                                        _D, _D, ctypes.POINTER(ChorinParams),
                                        ctypes.POINTER(_V)]),
    "zj_chorin_destroy": (ctypes.c_int, [_V]),
    "zj_chorin_set_precond": (ctypes.c_int, [_V, ctypes.c_int32, _V]),
    "zj_chorin_set_fv_order": (ctypes.c_int, [_V, _V]),
    "zj_chorin_set_mg_fused": (ctypes.c_int, [_V, ctypes.c_int32]),
    "zj_chorin_set_fv32": (ctypes.c_int, [_V, _V]),
    "zj_chorin_step": (ctypes.c_int, [_V, _D, _V, _V, _V]),
    "zj_chorin_stage": (ctypes.c_int, [_V, _D, _V, _V, _V, _V]),
    "zj_chorin_predict": (ctypes.c_int, [_V, _D, _V, _V, _V]),
    "zj_chorin_rhs": (ctypes.c_int, [_V, _D, _V, _V, _V]),
    "zj_chorin_correct": (ctypes.c_int, [_V, _D, _V, _V, _V, _V, _V]),
    "zj_chorin_setup_operator": (ctypes.c_int, [_V, _V, _V, _V]),
    "zj_chorin_apply": (ctypes.c_int, [_V, _V, _V, _V]),
    "zj_chorin_vcycle": (ctypes.c_int, [_V, _V, _V, ctypes.c_int32, _V]),
    "zj_chorin_solve": (ctypes.c_int, [_V, _V, _V, _V]),
    "zj_chorin_levels": (ctypes.c_int, [_V, _V, _V, ctypes.c_int32]),
    "zj_chorin_get_stats": (ctypes.c_int, [_V, ctypes.POINTER(ChorinStats)]),
    "zj_chorin_learned_vcycle": (ctypes.c_int, [_V, _V, _V, ctypes.c_int32,
                                                _V]),
    "zj_chorin_learned_vcycle_backward": (ctypes.c_int, [_V, _V, _V]),
    "zj_mpnn_param_count": (ctypes.c_int64, []),
    "zj_mpnn_create": (ctypes.c_int, [ctypes.POINTER(_V)]),
    "zj_mpnn_destroy": (ctypes.c_int, [_V]),
    "zj_mpnn_set_params": (ctypes.c_int, [_V, _V, ctypes.c_int64]),
    "zj_mpnn_get_params": (ctypes.c_int, [_V, _V, ctypes.c_int64]),
    "zj_mpnn_get_grad": (ctypes.c_int, [_V, _V, ctypes.c_int64]),
    "zj_mpnn_zero_grad": (ctypes.c_int, [_V, _V]),
    "zj_mpnn_set_fused": (ctypes.c_int, [_V, ctypes.c_int32]),
    "zj_mpnn_invalidate_features": (ctypes.c_int, [_V]),
    "zj_mpnn_time_forward": (ctypes.c_int, [_V, ctypes.c_int32, ctypes.c_int32,
                                            _V, _V, _V, _V, _V, ctypes.c_int32,
                                            _V, _V]),
    "zj_mpnn_adam": (ctypes.c_int, [_V, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_float, _V]),
    "zj_mpnn_forward": (ctypes.c_int, [_V, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32, _V, _V, _V, _V, _V,
                                       ctypes.c_int32, ctypes.c_int32, _V]),
    "zj_mpnn_backward": (ctypes.c_int, [_V, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_int32, _V, _V, _V,
                                        ctypes.c_int32, _V]),
}


_I64 = ctypes.c_int64
# the HDF5 snapshot writer (include/zoomy_io.h)
_SIGS_IO = {
    "zj_io_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32,
                                  ctypes.POINTER(_V)]),
    "zj_io_write": (ctypes.c_int, [_V, _I64, _D, _V, ctypes.c_int32, _I64,
                                   _I64, _V, _I64]),
    "zj_io_write_async": (ctypes.c_int, [_V, _I64, _D, _V, ctypes.c_int32,
                                         _I64, _I64, _V, _I64, _V]),
    "zj_io_flush": (ctypes.c_int, [_V]),
    "zj_io_close": (ctypes.c_int, [_V]),
}


def exported_symbols():
    return sorted(list(_SIGS) + list(_SIGS_POISSON) + list(_SIGS_IO))


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ZoomyHipError(
                f"{LIB_PATH} not found: build it with "
                f"`bash zoomy-jax_amd/csrc/build.sh` (or __graft_entry__.build())")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in {**_SIGS, **_SIGS_POISSON,
                                  **_SIGS_IO}.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != ZJ_OK:
        names = {ZJ_EINVAL: "invalid argument", ZJ_EHIP: "HIP runtime error",
                 ZJ_ENOMEM: "out of device memory",
                 ZJ_EUNSUP: "unsupported option",
                 ZJ_ESOLVE: "solve missed its tolerance or accept gate",
                 ZJ_EIO: "HDF5 library error",
                 ZJ_EEXIST: "the snapshot group exists"}
        raise ZoomyHipError(f"{what}: {names.get(rc, rc)} (code {rc})")
