"""ctypes binding of libjfem_hip.so (the C ABI declared in include/jfem_hip.h).

The library must have been built (``python -m jf_2026_05_25_amd.build``); loading
fails loudly otherwise -- there is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libjfem_hip.so")

JFEM_OK = 0
JFEM_ERR_NOT_SPD = 6
JFEM_ERR_NOT_CONVERGED = 8
STATUS = {1: "HIP error", 2: "degenerate element", 3: "inverted element", 4: "warped element",
          5: "bad argument", 6: "matrix not positive definite", 7: "out of device memory",
          8: "refinement did not converge"}

# every symbol declared in include/jfem_hip.h
EXPORTS = [
    "jfem_ctx_create", "jfem_ctx_destroy", "jfem_ctx_synchronize", "jfem_last_error", "jfem_device_memory",
    "jfem_ctx_kernel_time", "jfem_ctx_reset_kernel_time", "jfem_ctx_enable_timing",
    "jfem_element_matrices", "jfem_q4_local_batch", "jfem_t3_local_batch", "jfem_tet10_batch", "jfem_shell_recover",
    "jfem_mesh_create", "jfem_mesh_update_props", "jfem_mesh_free",
    "jfem_pattern_create", "jfem_pattern_free", "jfem_assemble", "jfem_reassemble", "jfem_matrix_download",
    "jfem_matrix_free", "jfem_bc_partition",
    "jfem_symbolic_create", "jfem_symbolic_free", "jfem_factor_create", "jfem_refactor",
    "jfem_factor_small_pivots", "jfem_factor_pivot_ratios", "jfem_symbolic_order", "jfem_factor_free", "jfem_solve", "jfem_solve_dev",
    "jfem_solve_dev_static",
    "jfem_dev_alloc", "jfem_dev_free", "jfem_memcpy_h2d", "jfem_memcpy_d2h",
    "jfem_factor_device_bytes", "jfem_matrix_device_bytes", "jfem_mesh_device_bytes", "jfem_pattern_device_bytes",
    "jfem_symbolic_device_bytes", "jfem_ctx_device_bytes", "jfem_solve_scratch_bytes", "jfem_selftest_update",
    "jfem_loadset_create", "jfem_loadset_set_gravity", "jfem_load_vector", "jfem_load_vector_dev", "jfem_loadset_free",
    "jfem_batch_create", "jfem_batch_update_props", "jfem_batch_assemble", "jfem_batch_matrix", "jfem_batch_factor",
    "jfem_batch_factor_view", "jfem_batch_free",
    "jfem_shell_kg_batch", "jfem_assemble_kg", "jfem_matrix_shift", "jfem_matrix_spmv", "jfem_block_gram",
    "jfem_block_combine", "jfem_factor_inertia", "jfem_matrix_shift_scaled",
]


class JfemError(RuntimeError):
    def __init__(self, code: int, msg: str = ""):
        super().__init__(f"jfem status {code} ({STATUS.get(code, '?')}): {msg}")
        self.code = code


class SymbolicInfo(C.Structure):
    _fields_ = [("n_free", C.c_int64), ("n_supernodes", C.c_int64), ("n_levels", C.c_int64),
                ("nnz_L", C.c_int64), ("flops", C.c_double), ("max_front", C.c_int64),
                ("update_entries", C.c_int64), ("update_flops", C.c_double), ("update_launches", C.c_int64),
                ("inv_entries", C.c_int64), ("upd_rows", C.c_int64), ("flow_fwd_tasks", C.c_int64),
                ("flow_bwd_tasks", C.c_int64), ("solve_entries", C.c_int64), ("schur_flops", C.c_double),
                ("schur_launches", C.c_int64)]


_lib = None
vp = C.c_void_p
i64 = C.c_int64
i32 = C.c_int32
dbl = C.c_double
P = C.POINTER


def load():
    """Load libjfem_hip.so; raises if it is missing (build it first)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -m jf_2026_05_25_amd.build`")
    L = C.CDLL(LIB_PATH)
    sig = {
        "jfem_ctx_create": ([C.c_int, P(vp)], C.c_int),
        "jfem_ctx_destroy": ([vp], C.c_int),
        "jfem_ctx_synchronize": ([vp], C.c_int),
        "jfem_last_error": ([vp], C.c_char_p),
        "jfem_device_memory": ([vp, P(i64), P(i64)], C.c_int),
        "jfem_ctx_kernel_time": ([vp, C.c_int, P(dbl), P(i64)], C.c_int),
        "jfem_ctx_reset_kernel_time": ([vp], C.c_int),
        "jfem_ctx_enable_timing": ([vp, C.c_int], C.c_int),
        "jfem_element_matrices": ([vp, vp, C.c_int, vp, P(i64)], C.c_int),
        "jfem_q4_local_batch": ([vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, P(i64)], C.c_int),
        "jfem_t3_local_batch": ([vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, P(i64)], C.c_int),
        "jfem_tet10_batch": ([vp, i64, vp, vp, vp, vp, P(i64)], C.c_int),
        "jfem_shell_recover": ([vp, vp, vp, C.c_int, vp, vp, P(i64)], C.c_int),
        "jfem_shell_kg_batch": ([vp, vp, vp, vp, P(i64)], C.c_int),
        "jfem_assemble_kg": ([vp, vp, vp, vp, P(vp), P(i64)], C.c_int),
        "jfem_matrix_shift": ([vp, vp, dbl, vp, P(vp)], C.c_int),
        "jfem_matrix_spmv": ([vp, vp, vp, C.c_int, vp, vp], C.c_int),
        "jfem_block_gram": ([vp, i64, C.c_int, vp, C.c_int, vp, vp], C.c_int),
        "jfem_block_combine": ([vp, i64, C.c_int, vp, C.c_int, vp, vp, C.c_int], C.c_int),
        "jfem_factor_inertia": ([vp, vp, vp, P(i64), P(dbl), P(dbl), P(i64)], C.c_int),
        "jfem_matrix_shift_scaled": ([vp, vp, dbl, vp, P(vp)], C.c_int),
        "jfem_mesh_create": ([vp, i64, vp, vp, vp, i64, vp, vp, i64, vp, vp, i64, vp, vp, dbl, P(vp)], C.c_int),
        "jfem_mesh_update_props": ([vp, vp, vp, vp, vp], C.c_int),
        "jfem_mesh_free": ([vp], C.c_int),
        "jfem_pattern_create": ([vp, vp, P(vp), P(i64)], C.c_int),
        "jfem_pattern_free": ([vp], C.c_int),
        "jfem_assemble": ([vp, vp, vp, P(vp), P(i64)], C.c_int),
        "jfem_reassemble": ([vp, vp, vp, P(i64)], C.c_int),
        "jfem_matrix_download": ([vp, vp, vp, vp, vp], C.c_int),
        "jfem_matrix_free": ([vp], C.c_int),
        "jfem_bc_partition": ([vp, vp, vp, dbl, vp, P(i64)], C.c_int),
        "jfem_symbolic_create": ([vp, vp, vp, vp, P(vp), P(SymbolicInfo)], C.c_int),
        "jfem_symbolic_free": ([vp], C.c_int),
        "jfem_factor_create": ([vp, vp, vp, C.c_int, P(vp), P(i64)], C.c_int),
        "jfem_refactor": ([vp, vp, vp, P(i64)], C.c_int),
        "jfem_factor_small_pivots": ([vp, vp, vp, dbl, vp, i64, P(i64)], C.c_int),
        "jfem_factor_pivot_ratios": ([vp, vp, vp, vp], C.c_int),
        "jfem_symbolic_order": ([vp, vp], C.c_int),
        "jfem_factor_free": ([vp], C.c_int),
        "jfem_solve": ([vp, vp, vp, C.c_int, vp, vp, dbl, C.c_int, vp, vp, vp, vp], C.c_int),
        "jfem_solve_dev": ([vp, vp, vp, C.c_int, vp, vp, dbl, C.c_int, vp, vp, vp, vp], C.c_int),
        "jfem_solve_dev_static": ([vp, vp, vp, C.c_int, vp, vp, dbl, C.c_int, vp, vp, vp, vp], C.c_int),
        "jfem_dev_alloc": ([vp, i64, P(vp)], C.c_int),
        "jfem_dev_free": ([vp, vp], C.c_int),
        "jfem_memcpy_h2d": ([vp, vp, vp, i64], C.c_int),
        "jfem_memcpy_d2h": ([vp, vp, vp, i64], C.c_int),
        "jfem_factor_device_bytes": ([vp], i64),
        "jfem_matrix_device_bytes": ([vp], i64),
        "jfem_mesh_device_bytes": ([vp], i64),
        "jfem_pattern_device_bytes": ([vp], i64),
        "jfem_symbolic_device_bytes": ([vp], i64),
        "jfem_ctx_device_bytes": ([vp], i64),
        "jfem_solve_scratch_bytes": ([vp, i64, C.c_int, C.c_int], i64),
        "jfem_selftest_update": ([vp, C.c_int, P(dbl)], C.c_int),
        "jfem_loadset_create": ([vp, vp, i64, vp, vp, vp, i64, vp, vp, vp, i64, vp, vp, vp, P(vp)], C.c_int),
        "jfem_loadset_set_gravity": ([vp, vp, i64, vp, vp, vp], C.c_int),
        "jfem_load_vector": ([vp, vp, dbl, vp], C.c_int),
        "jfem_load_vector_dev": ([vp, vp, dbl, vp], C.c_int),
        "jfem_loadset_free": ([vp], C.c_int),
        "jfem_batch_create": ([vp, vp, vp, i64, vp, vp, vp, P(vp)], C.c_int),
        "jfem_batch_update_props": ([vp, vp, vp, vp, vp], C.c_int),
        "jfem_batch_assemble": ([vp, vp, P(i64), P(i64)], C.c_int),
        "jfem_batch_matrix": ([vp, i64, P(vp)], C.c_int),
        "jfem_batch_factor": ([vp, vp, vp, C.c_int, P(i64), P(i64)], C.c_int),
        "jfem_batch_factor_view": ([vp, i64, P(vp)], C.c_int),
        "jfem_batch_free": ([vp], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def ptr(a: np.ndarray | None):
    return None if a is None or a.size == 0 else a.ctypes.data


def check(ctx, code: int):
    if code != JFEM_OK:
        msg = load().jfem_last_error(ctx).decode() if ctx else ""
        raise JfemError(code, msg)


def cm(a, dtype):
    """Column-major flattening (Julia layout) of a 2-D array."""
    a = np.asarray(a, dtype=dtype)
    if a.size == 0:
        return np.zeros(0, dtype=dtype)
    return np.ascontiguousarray(a.T).ravel()
