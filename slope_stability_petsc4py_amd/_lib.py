"""ctypes binding of the ssx HIP library (``libssx.so``, built in-tree for gfx950).

The product path has no CPU fallback: if the library is missing or fails to load, every entry
point raises.  Arguments are device pointers (``tensor.data_ptr()``), int64 sizes and the HIP
stream handle; each call returns 0 or raises ``RuntimeError`` with the library's message.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "libssx.so"

P = C.c_void_p
I64 = C.c_int64
I32 = C.c_int
F64 = C.c_double
F32 = C.c_float

# name -> argument types (all return int status, except where noted)
_SIGS: dict[str, list] = {
    "ssx_reduction": [P, P, P, I64, F64, I32, P, P, P],
    "ssx_constitutive_3d": [P, P, P, P, P, I64, P, P, P, P, P],
    "ssx_constitutive_3d_mat": [P, P, P, I32, I32, I64, P, P, P, P, P],
    "ssx_constitutive_2d": [P, P, P, P, P, P, I64, P, P, P, P, P],
    "ssx_strain": [I32, I32, P, P, P, I64, I32, P, P],
    "ssx_force_elem": [I32, I32, P, P, P, I64, I32, P, P],
    "ssx_force_gather": [P, P, P, P, I64, I32, P, P],
    "ssx_force_gather_nodes": [I32, I32, P, P, P, P, P, I64, P, P],
    "ssx_tangent_rows": [I32, I32, P, P, P, P, P, P, P, P, P, I64, I64, I32, F64, P, P, P, I64, P],
    "ssx_spmv_f64": [I64, P, P, P, P, P, F64, F64, I32, P],
    "ssx_spmv_f32": [I64, P, P, P, P, P, F32, F32, I32, P],
    "ssx_axpby_f64": [I64, F64, P, F64, P, P],
    "ssx_axpby_f32": [I64, F32, P, F32, P, P],
    "ssx_cast_f64_f32": [I64, P, P, P],
    "ssx_cast_f32_f64": [I64, P, P, P],
    "ssx_dot_cols": [I64, I32, P, I64, P, P, P, P],
    "ssx_dot": [I64, P, P, P, P, P],
    "ssx_gemv_cols": [I64, I32, P, I64, P, F64, P, P],
    "ssx_spmm_f64": [I64, P, P, P, I32, P, I64, P, I64, I32, P],
    "ssx_cheb_step_f32": [I64, P, P, P, P, P, P, P, P, F32, F32, I32, I32, P],
    "ssx_cheb_step_f64": [I64, P, P, P, P, P, P, P, P, F64, F64, I32, I32, P],
    "ssx_sym_tiles_sizes": [I64, P, P],
    "ssx_sym_tiles_pack": [I64, P, I32, P, I32, P],
    "ssx_spd_inverse_sizes": [I64, P],
    "ssx_spd_inverse_tiles": [I64, P, I32, P, P, I32, P, P],
    "ssx_sym_tiles_gemv_f32": [I64, P, P, P, P, P],
    "ssx_sym_tiles_gemv_f64": [I64, P, P, P, P, P],
    "ssx_a_orthogonalize": [I64, I32, P, I64, P, P, P, I32, F64, P, P, P, P, P],
    "ssx_a_orthogonalize_noalloc": [I64, I32, P, I64, P, P, P, I32, F64, P, P, P, P],
    "ssx_cheb_update_f32": [I64, P, P, F32, F32, P, P, P],
    "ssx_cheb_update_f64": [I64, P, P, F64, F64, P, P, P],
    "ssx_inv_diag_f32": [I64, I64, P, P, P, P, P],
    "ssx_inv_diag_f64": [I64, I64, P, P, P, P, P],
    "ssx_galerkin_dense": [I64, P, P, P, I64, P, P, P, I64, F64, P, P, P],
    "ssx_galerkin_csr": [I64, P, P, P, I64, P, P, P, P, P, P, I64, P, P, I64, I32, F64, P, P, P],
    "ssx_element_matrices_p4": [P, P, P, I64, F64, P, P, P, P],
    "ssx_element_matrices_3d": [I32, P, P, P, I64, I32, F64, P, P, P, P],
    "ssx_probe_fp64": [I32, I64, I64, P, P, P],
    "ssx_gather_node_rows": [I32, I32, P, P, P, P, P, P, P, P, I64, I32, F64, P, P, P, P],
    "ssx_element_blocks_3d": [I32, P, P, P, I64, I32, F64, P, P, P, P, P],
    "ssx_element_blocks_affine_3d": [I32, P, P, P, I32, I64, I32, F64, P, P, P, P, P],
    "ssx_constitutive_3d_sym": [P, P, P, P, P, I64, P, P, P, P],
    "ssx_ds_unpack_sym": [I64, P, P, P],
    "ssx_strain_affine_3d": [I32, P, P, P, P, I64, I32, P, P],
    "ssx_force_elem_affine_3d": [I32, P, P, P, I64, I32, P, P],
    "ssx_gather_node_blocks": [I32, P, P, P, P, P, P, P, P, I64, I32, F64, P, P, P, P],
    "ssx_gemm_bf16": [I64, I64, I64, P, I64, I64, P, I64, I64, P, I64, P, I32, P, F32, P],
    "ssx_tanh_bwd": [I64, P, P, P, P, P],
    "ssx_colsum_f32": [I64, I64, P, P, P],
    "ssx_adam_f32": [I64, P, P, P, P, F32, F32, F32, F32, I32, P, P],
    "ssx_f32_to_bf16": [I64, P, P, P],
    "ssx_bf16_to_f32": [I64, P, P, P],
    "ssx_mse_f32": [I64, P, P, P, F32, P, P, P],
    "ssx_stream_sync": [P],
}

_lib = None


class SSXError(RuntimeError):
    pass


def lib_path() -> Path:
    return _LIB_PATH


def load() -> C.CDLL:
    """Load libssx.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise SSXError(
            f"{_LIB_PATH} is not built; run `python -m slope_stability_petsc4py_amd.csrc.build` (no CPU fallback)")
    lib = C.CDLL(str(_LIB_PATH))
    lib.ssx_version.restype = C.c_int
    lib.ssx_last_error.restype = C.c_char_p
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return ["ssx_version", "ssx_last_error", *_SIGS.keys()]


def call(name: str, *args) -> None:
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.ssx_last_error().decode(errors="replace")
        if rc == 1:
            raise ValueError(f"{name}: {msg}")
        raise SSXError(f"{name} failed ({rc}): {msg}")
