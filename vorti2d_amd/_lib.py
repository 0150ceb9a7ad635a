"""ctypes binding of libvorti2d.so (the C ABI declared in include/vorti2d.h).

The library is built in-tree by ``python vorti2d_amd/_build.py``.  There is no
fallback: if the shared library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvorti2d.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build the HIP library first "
        "(python vorti2d_amd/_build.py, or __graft_entry__.build())")

lib = C.CDLL(LIB_PATH)

i32, i64, f64, vp = C.c_int, C.c_int64, C.c_double, C.c_void_p
P = C.POINTER

_SIGS = {
    "v2d_last_error": ([], C.c_char_p),
    "v2d_version": ([], i32),
    "v2d_node_neighbors": ([i32, i32, i32, vp], i32),
    "v2d_compute_metrics": ([i32, i32, f64, f64, vp, vp, vp], i32),
    "v2d_compute_metrics_dev": ([i32, i32, f64, f64, vp, vp, vp, vp], i32),
    "v2d_assemble_coo": ([i32, i32, f64, f64, f64, f64, f64, f64, vp, vp, vp, vp, vp, i32, i32,
                          i32, i32, f64, f64, vp, vp, vp, P(i32), vp], i32),
    "v2d_compute_metrics_local": ([i32, i32, i32, i32, f64, f64, vp, vp, vp], i32),
    "v2d_assemble_coo_local": ([i32, i32, i32, i32, f64, f64, f64, f64, f64, f64, vp, vp, vp, vp,
                                vp, i32, i32, vp, i32, i32, f64, f64, vp, vp, vp, P(i32), vp], i32),
    "v2d_force_coeffs": ([i32, i32, f64, f64, f64, vp, vp, vp, vp, vp, vp, vp, f64, f64, f64, vp],
                         i32),
    "v2d_force_coeffs_noalloc": ([i32, i32, f64, f64, f64, vp, vp, vp, vp, vp, vp, vp, f64, f64,
                                  f64, vp], i32),
    "v2d_ctx_create": ([P(vp), i32, i32, vp, vp, i32, f64, f64, i32], i32),
    "v2d_ctx_create_like": ([P(vp), vp, i32], i32),
    "v2d_ctx_destroy": ([vp], i32),
    "v2d_ctx_device_bytes": ([vp, P(i64)], i32),
    "v2d_ctx_metrics": ([vp, vp], i32),
    "v2d_ctx_set_state": ([vp, vp, vp, vp, vp], i32),
    "v2d_ctx_get_state": ([vp, vp, vp, vp, vp], i32),
    "v2d_ctx_init_state": ([vp], i32),
    "v2d_ctx_set_solver": ([vp, i32, f64, i32], i32),
    "v2d_ctx_set_forcing": ([vp, f64, f64], i32),
    "v2d_ctx_set_update_kernel": ([vp, i32], i32),
    "v2d_ctx_set_step_pairing": ([vp, i32], i32),
    "v2d_ctx_set_front_fusion": ([vp, i32], i32),
    "v2d_ctx_set_sweep_kernel": ([vp, i32], i32),
    "v2d_ctx_set_sweep_spin_limit": ([vp, i32], i32),
    "v2d_ctx_get_sweep_kernel": ([vp, P(i32), P(i64)], i32),
    "v2d_ctx_set_refactor_policy": ([vp, i32], i32),
    "v2d_ctx_set_graph": ([vp, i32], i32),
    "v2d_ctx_profile_factor": ([vp, vp], i32),
    "v2d_ctx_assemble": ([vp, f64, f64, f64, f64], i32),
    "v2d_ctx_rhs_norm": ([vp, P(f64)], i32),
    "v2d_ctx_nnz": ([vp, P(i64)], i32),
    "v2d_ctx_export_coo": ([vp, vp, vp, vp, vp], i32),
    "v2d_ctx_export_csr": ([vp, vp, vp, vp], i32),
    "v2d_ctx_spmv": ([vp, vp, vp], i32),
    "v2d_ctx_factor": ([vp], i32),
    "v2d_ctx_solve": ([vp, vp, vp], i32),
    "v2d_ctx_newton_step": ([vp, f64, f64, f64, f64, P(f64), P(f64), P(i32)], i32),
    "v2d_ctx_bdf2_shift": ([vp], i32),
    "v2d_ctx_forces": ([vp, f64, f64, f64, f64, vp], i32),
    "v2d_ctx_timers": ([vp, vp, vp, i32], i32),
    "v2d_ctx_time_assembly": ([vp, f64, f64, f64, f64, i32, P(f64)], i32),
    "v2d_ctx_time_spmv": ([vp, i32, P(f64)], i32),
    "v2d_ctx_solve_transpose": ([vp, vp, vp], i32),
    "v2d_ctx_residual_vjp": ([vp, f64, f64, f64, vp, vp, vp], i32),
    "v2d_metrics_vjp": ([i32, i32, f64, f64, vp, vp, vp, vp, vp], i32),
    "v2d_ctx_forces_vjp": ([vp, f64, f64, f64, f64, vp, vp, vp], i32),
    "v2d_ctx_adjoint_forces": ([vp, f64, f64, f64, f64, f64, vp, vp, vp, vp], i32),
    "v2d_batch_create": ([P(vp), i32, i32], i32),
    "v2d_batch_destroy": ([vp], i32),
    "v2d_batch_set_case": ([vp, i32, vp], i32),
    "v2d_batch_newton_step": ([vp, vp, vp, vp, vp, vp, vp, vp, vp], i32),
    "v2d_batch_get_blocks": ([vp, vp], i32),
    "v2d_batch_set_blocks": ([vp, vp], i32),
    "v2d_batch_autotune": ([vp, f64, vp], i32),
    "v2d_ctx_adjoint_bdf2_begin": ([vp], i32),
    "v2d_ctx_adjoint_bdf2_step": ([vp, f64, f64, f64, f64, f64, f64, vp], i32),
    "v2d_ctx_adjoint_bdf2_end": ([vp, vp, vp, vp], i32),
    "v2d_ctx_set_linear_solver": ([vp, i32], i32),
    "v2d_ctx_get_linear_solver": ([vp, P(i32)], i32),
    "v2d_ctx_mf_info": ([vp, vp, P(f64)], i32),
    "v2d_ctx_mf_fronts": ([vp, vp], i32),
    "v2d_ctx_profile_mf": ([vp, vp], i32),
    "v2d_mf_analyse": ([i32, i32, vp, P(vp)], i32),
    "v2d_mf_sizes": ([vp, vp, P(f64)], i32),
    "v2d_mf_export": ([vp, vp, vp, vp, vp, vp, vp, vp], i32),
    "v2d_mf_destroy": ([vp], i32),
}

for _name, (_args, _res) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = _res

EXPORTED = tuple(_SIGS)


class V2DError(RuntimeError):
    """A libvorti2d call failed (message from v2d_last_error)."""


def check(rc: int) -> None:
    if rc != 0:
        msg = lib.v2d_last_error().decode(errors="replace")
        raise V2DError(msg or "libvorti2d call failed")


def ptr(a: np.ndarray | None):
    """Raw data pointer of a contiguous array (None -> NULL)."""
    if a is None:
        return None
    return C.c_void_p(a.ctypes.data)


def f64c(a) -> np.ndarray:
    """C-contiguous float64 view/copy (f2py converts dtype and order the same way)."""
    return np.ascontiguousarray(a, dtype=np.float64)


def ptr_array(arrs) -> C.Array:
    """Array of data pointers (for the `const double *const *` arguments)."""
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
