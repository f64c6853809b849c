"""ctypes binding of libfvhip.so (the C ABI declared in include/fvhip.h).

The library is the product: every numerical call goes through it.  If the
shared object is missing this module raises at import time -- there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
# FVHIP_LIB names another build of the library (e.g. one made with a different ARCH)
LIB_PATH = Path(os.environ.get("FVHIP_LIB") or _HERE / "libfvhip.so")

# status codes (include/fvhip.h)
FV_OK = 0
FV_EINVAL = 1
FV_ENOMEM_BUDGET = 2
FV_ENOCONVERGE = 3
FV_EDEVICE = 4
FV_EBUSY = 5
FV_ECANCELLED = 6
FV_ENOTREADY = 7

FV_F64, FV_F32 = 0, 1
FV_LINEAR, FV_UPWIND = 0, 1
FV_BC_FIXED_VALUE, FV_BC_FIXED_GRADIENT, FV_BC_CALCULATED = 0, 1, 2

GEOM_IDS = {
    "V": 0, "C": 1, "owner": 2, "neighbour": 3, "Sf": 4, "Cf": 5, "magSf": 6, "weights": 7,
    "deltaCoeffs": 8, "nonOrthDeltaCoeffs": 9, "faceCells": 10, "Sf_b": 11, "Cf_b": 12,
    "magSf_b": 13, "nf_b": 14, "delta_b": 15, "weights_b": 16, "deltaCoeffs_b": 17,
    "nonOrthDeltaCoeffs_b": 18, "patchOffsets": 19, "cellFaceOffsets": 20, "cellFaces": 21,
}


class FvError(RuntimeError):
    """A libfvhip call returned a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libfvhip error {code}: {msg}")
        self.code = code


class FvPolyMeshDesc(C.Structure):
    _fields_ = [
        ("n_points", C.c_int32), ("n_faces", C.c_int32), ("n_internal_faces", C.c_int32),
        ("n_cells", C.c_int32), ("n_patches", C.c_int32),
        ("h_points", C.c_void_p), ("h_face_offsets", C.c_void_p), ("h_face_points", C.c_void_p),
        ("h_owner", C.c_void_p), ("h_neighbour", C.c_void_p), ("h_patch_start", C.c_void_p),
        ("h_patch_size", C.c_void_p), ("h_patch_empty", C.c_void_p),
    ]


class FvMeshDesc(C.Structure):
    _fields_ = [
        ("n_cells", C.c_int32), ("n_internal_faces", C.c_int32),
        ("n_boundary_faces", C.c_int32), ("n_patches", C.c_int32),
    ] + [(n, C.c_void_p) for n in (
        "h_V", "h_C", "h_owner", "h_neighbour", "h_Sf", "h_Cf", "h_magSf", "h_weights",
        "h_delta_coeffs", "h_nonorth_delta_coeffs", "h_face_cells", "h_Sf_b", "h_Cf_b",
        "h_magSf_b", "h_nf_b", "h_delta_b", "h_weights_b", "h_delta_coeffs_b",
        "h_nonorth_delta_coeffs_b", "h_patch_offsets")]


FV_SOLVE_F64, FV_SOLVE_F32_IR = 0, 1


class FvSolveOpts(C.Structure):
    _fields_ = [("rel_tol", C.c_double), ("max_iter", C.c_int32), ("mode", C.c_int32),
                ("inner_rel_tol", C.c_double), ("max_outer", C.c_int32),
                ("check_every", C.c_int32), ("abs_tol", C.c_double), ("eager", C.c_int32)]


class FvSolverStats(C.Structure):
    _fields_ = [("init_res", C.c_double), ("final_res", C.c_double), ("iters", C.c_int32),
                ("outer_iters", C.c_int32), ("true_res", C.c_double)]


_vp, _i32, _i64, _u64, _d, _sz = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double, C.c_size_t
_pp = C.POINTER(C.c_void_p)

# name -> argtypes (restype int unless listed in _RESTYPES)
_SIGS = {
    "fv_last_error": [],
    "fv_version": [],
    "fv_ctx_create": [C.c_int, _sz, _pp],
    "fv_ctx_destroy": [_vp],
    "fv_sync": [_vp],
    "fv_ctx_stream": [_vp],
    "fv_ctx_bytes_in_use": [_vp],
    "fv_event_record": [_vp, C.c_int],
    "fv_event_elapsed": [_vp, C.c_int, C.c_int, C.POINTER(C.c_float)],
    "fv_device_alloc": [_vp, _sz, _pp],
    "fv_device_free": [_vp, _vp],
    "fv_memcpy_h2d": [_vp, _vp, _vp, _sz],
    "fv_memcpy_d2h": [_vp, _vp, _vp, _sz],
    "fv_memcpy_d2d": [_vp, _vp, _vp, _sz],
    "fv_fill": [_vp, C.c_int, _vp, _i64, _d],
    "fv_convert": [_vp, C.c_int, _vp, C.c_int, _vp, _i64],
    "fv_fill_uniform": [_vp, C.c_int, _vp, _i64, _d, _d, _u64, _u64],
    "fv_stream": [_vp, C.c_int, _i64, _vp, _vp, _vp, _d],
    "fv_component": [_vp, C.c_int, _i64, C.c_int, C.c_int, C.c_int, _vp, _vp],
    "fv_hexmesh_sizes": [_i32, _i32, _i32, _vp],
    "fv_hexmesh_generate": [_i32, _i32, _i32, _d, _d, _d, _d, _u64, _u64,
                            _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "fv_polymesh_renumber": [C.POINTER(FvPolyMeshDesc), C.c_int, _vp, _vp, _vp, _vp, _vp, _vp],
    "fv_permute": [_vp, C.c_int, _i64, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp],
    "fv_mesh_create": [_vp, C.POINTER(FvMeshDesc), _pp],
    "fv_mesh_create_poly": [_vp, C.POINTER(FvPolyMeshDesc), _pp],
    "fv_mesh_destroy": [_vp],
    "fv_mesh_sizes": [_vp, _vp],
    "fv_mesh_get": [_vp, C.c_int, _vp],
    "fv_mesh_set_gather": [_vp, C.c_int],
    "fv_mesh_set_gather_run": [_vp, C.c_int],
    "fv_mesh_get_gather": [_vp, C.POINTER(C.c_int)],
    "fv_mesh_gather_stats": [_vp, C.c_int, _vp],
    "fv_bc_create": [_vp, _vp, _pp],
    "fv_bc_destroy": [_vp],
    "fv_interpolate": [_vp, _vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp],
    "fv_grad_gauss": [_vp, _vp, C.c_int, _d, _vp, _vp, _vp],
    "fv_div_gauss": [_vp, _vp, C.c_int, C.c_int, _d, _vp, _vp, _vp, _vp],
    "fv_laplacian_uncorrected": [_vp, _vp, C.c_int, _d, _vp, _vp, _vp, _vp, _vp, _vp],
    "fv_sngrad_uncorrected": [_vp, _vp, C.c_int, _vp, _vp, _vp, _vp, _vp],
    "fv_flux": [_vp, _vp, C.c_int, _vp, _vp, _vp],
    "fv_correct_boundary": [_vp, _vp, C.c_int, _vp, _vp, _vp, _vp],
    "fv_sparsity_create": [_vp, _pp],
    "fv_sparsity_destroy": [_vp],
    "fv_sparsity_sizes": [_vp, _vp],
    "fv_sparsity_get": [_vp, C.c_int, _vp],
    "fv_linsys_create": [_vp, _pp],
    "fv_linsys_destroy": [_vp],
    "fv_linsys_values": [_vp],
    "fv_linsys_rhs": [_vp],
    "fv_linsys_zero": [_vp],
    "fv_assemble_ddt_euler": [_vp, _vp, _d, _d],
    "fv_assemble_source": [_vp, _vp, _d],
    "fv_assemble_explicit": [_vp, _vp, _d],
    "fv_assemble_laplacian": [_vp, _vp, _vp, _vp, _vp, _d],
    "fv_assemble_div": [_vp, C.c_int, _vp, _vp, _vp, _vp, _d],
    "fv_assemble_ddt_laplacian": [_vp, _vp, _d, _vp, _vp, _vp, _vp, _d],
    "fv_assemble_ddt_laplacian_many": [C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d],
    "fv_set_reference": [_vp, C.c_int, _d],
    "fv_residual": [_vp, _vp, _vp],
    "fv_compute_rau": [_vp, _vp, _vp],
    "fv_compute_hbya": [_vp, _vp, _vp, _vp, _vp, _vp],
    "fv_update_face_velocity": [_vp, _vp, _vp, _vp],
    "fv_update_velocity": [_vp, _vp, _vp, _vp, _vp, _vp],
    "fv_constrain_hbya": [_vp, _vp, _vp],
    "fv_linsys_boundary_coeffs": [_vp, C.POINTER(_vp), C.POINTER(_vp)],
    "fv_courant_number": [_vp, _vp, _vp, _d, C.POINTER(_d)],
    "fv_courant_number_noalloc": [_vp, _vp, _vp, _d, _vp],
    "fv_set_delta_t": [_d, _d, _d, _d, C.POINTER(_d)],
    "fv_solve_cg": [_vp, _vp, C.POINTER(FvSolveOpts), C.POINTER(FvSolverStats)],
    "fv_solve_bicgstab": [_vp, _vp, C.POINTER(FvSolveOpts), C.POINTER(FvSolverStats)],
    "fv_batch_create": [_vp, C.c_int, C.c_int, C.c_int, _pp],
    "fv_batch_destroy": [_vp],
    "fv_batch_add": [_vp, _vp, _vp, _vp, C.POINTER(C.c_int)],
    "fv_batch_clear": [_vp],
    "fv_batch_size": [_vp, C.POINTER(C.c_int)],
    "fv_batch_set_shape": [_vp, C.c_int, C.c_int],
    "fv_batch_timing": [_vp, _vp, _vp],
    "fv_batch_solve": [_vp, C.POINTER(FvSolveOpts), C.c_int, C.POINTER(FvSolverStats)],
    "fv_batch_autotune": [_vp, C.POINTER(FvSolveOpts), C.c_char_p, C.POINTER(C.c_int),
                          C.POINTER(C.c_int)],
}
_RESTYPES = {
    "fv_last_error": C.c_char_p,
    "fv_version": C.c_char_p,
    "fv_ctx_stream": C.c_void_p,
    "fv_ctx_bytes_in_use": C.c_size_t,
    "fv_linsys_values": C.c_void_p,
    "fv_linsys_rhs": C.c_void_p,
}


def _load():
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run `make -C foamadapter_amd/csrc` or "
            "`python -c 'import __graft_entry__ as g; g.build()'` (there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | C.RTLD_GLOBAL)
    for name, args in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
    return lib


lib = _load()


def declared_symbols() -> list[str]:
    return list(_SIGS)


def last_error() -> str:
    return lib.fv_last_error().decode()


def check(rc: int) -> None:
    if rc != FV_OK:
        raise FvError(rc, last_error())


def call(name: str, *args) -> None:
    fn = getattr(lib, name)
    check(fn(*args))
