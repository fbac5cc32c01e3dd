"""ctypes binding of libmpdok_hip.so.

torch is imported first so that its HIP runtime (soname ``libamdhip64.so.7``)
is the one the library binds to: torch tensors, torch streams and our kernels
then share one HIP context. There is no fallback: if the library is missing
or cannot be loaded, every entry point raises.
"""
import ctypes
import os

import torch  # noqa: F401  (must be loaded before the HIP library)

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_LIB = os.path.join(_HERE, "libmpdok_hip.so")

P = ctypes.c_void_p
I = ctypes.c_int
I64 = ctypes.c_int64
SZ = ctypes.c_size_t
D = ctypes.c_double
PI = ctypes.POINTER(ctypes.c_int)
PP = ctypes.POINTER(ctypes.c_void_p)

# name: (restype, [argtypes])
SIGNATURES = {
    # reference-compatible (mpdok_ops.py:102-141, :362-382)
    "py_gmres_ir": (None, [P, P, P, P, I, D, I, I, PI]),
    "py_alloc_managed": (None, [I, PP]),
    "py_free_managed": (None, []),
    "py_lu_ir": (None, [P, P, P, I, D, I, PI]),
    "py_lu_alloc_managed": (None, [I, PP]),
    "py_lu_free_managed": (None, []),
    "py_lu_factor": (None, [P, I, PI]),
    "py_lu_solve_factored": (None, [P, P, P, I, D, I, PI]),
    "py_lu_free_factored": (None, []),
    # library services
    "mpdok_last_error": (ctypes.c_char_p, []),
    "mpdok_set_stream": (I, [P]),
    "mpdok_set_tuning": (I, [I, I]),
    "mpdok_get_tuning": (I, [I]),
    "mpdok_reduce_workspace_bytes": (SZ, [I]),
    "mpdok_gemv_workspace_bytes": (SZ, [I, I, I, I]),
    "mpdok_residual_workspace_bytes": (SZ, [I, I]),
    "mpdok_arnoldi_workspace_bytes": (SZ, [I, I]),
    # BLAS-1 / casts
    "mpdok_dot_f32": (I, [P, P, P, I, P, P]),
    "mpdok_dot_f64": (I, [P, P, P, I, P, P]),
    "mpdok_cast_matrix": (I, [P, P, I, I64, I, P, I, I64, I, I, I]),
    "mpdok_axpby": (I, [P, I, D, P, I, D, P, I, P, I]),
    # GEMV / residual
    "mpdok_gemv_f32": (I, [P, P, I64, I, I, I, P, P, P, SZ, P]),
    "mpdok_gemv_f64": (I, [P, P, I64, I, I, I, P, P, P, SZ, P]),
    "mpdok_residual_f64": (I, [P, P, I64, I, I, P, P, P, I, P, SZ]),
    "mpdok_residual_f64_noalloc": (I, [P, P, I64, I, I, P, P, P, I]),
    "mpdok_read_ceiling_workspace_bytes": (SZ, []),
    "mpdok_read_ceiling": (I, [P, P, SZ, P, P]),
    # RBF assembly
    "mpdok_rbf_assemble": (I, [P, P, I, I, I, D, D, I, I, I, I, I, P, I, I64, I]),
    "mpdok_rbf_mean_d2": (I, [P, P, I, I, P]),
    "mpdok_rbf_diag": (I, [P, P, I, I, I, D, D, I, I, P, I]),
    "mpdok_rbf_matvec": (I, [P, P, I, I, I, D, D, I, P, P, I, I, I]),
    "mpdok_rbf_symv_workspace_bytes": (SZ, [I]),
    "mpdok_rbf_symv": (I, [P, P, I, I, I, D, D, I, P, P, I, P, SZ]),
    "mpdok_rbf_cross_matvec": (I, [P, P, I, P, I, I, I, D, P, P]),
    "mpdok_rbf_dbilinear_workspace_bytes": (SZ, [I]),
    "mpdok_rbf_dbilinear": (I, [P, P, I, P, I, I, I, D, P, P, P, P, SZ]),
    # out-of-core streaming
    "mpdok_stream_gemv_workspace_bytes": (SZ, [I, I]),
    "mpdok_stream_gemv_f32": (I, [P, P, I64, I, I, I, P, P, P, SZ]),
    "mpdok_rbf_build_host": (I, [P, P, I, I, I, D, D, I, I, I, P, I64, P, SZ]),
    # Krylov primitives
    "mpdok_arnoldi_orth_f32": (I, [P, P, I64, I, P, I, P, I, D, I, P, P, SZ]),
    "mpdok_arnoldi_orth_f64": (I, [P, P, I64, I, P, I, P, I, D, I, P, P, SZ]),
    "mpdok_givens_step": (I, [P, P, I, I, P, P, P, P, D, D, P]),
    "mpdok_basis_update_f32": (I, [P, P, I64, I, P, P, I, I]),
    "mpdok_basis_update_f32_into_f64": (I, [P, P, I64, I, P, P, I, I]),
    "mpdok_basis_update_f64": (I, [P, P, I64, I, P, P, I, I]),
    "mpdok_scale_f64_to_f32": (I, [P, P, P, I, P]),
    "mpdok_scale_f32": (I, [P, P, P, I, P]),
    "mpdok_scale_f64": (I, [P, P, P, I, P]),
    "mpdok_cg_step_updates": (I, [P, P, P, P, P, I, P, P, P, P]),
    "mpdok_pcg_workspace_bytes": (SZ, []),
    "mpdok_pcg_step_updates": (I, [P, I, P, P, P, P, P, P, I, P, P, P, P, P]),
    "mpdok_diag_scale": (I, [P, I, P, P, P, I]),
    # LU factorisation / trailing-update kernels
    "mpdok_lu_create": (I, [I, I, I, I, I, PP]),
    "mpdok_lu_factor": (I, [P, P, P, I64, I, PI]),
    "mpdok_lu_solve": (I, [P, P, P, P]),
    "mpdok_lu_destroy": (I, [P]),
    "mpdok_lu_solve_multi": (I, [P, P, P, I64, P, I64, I]),
    "mpdok_residual_multi_workspace_bytes": (SZ, [I]),
    "mpdok_residual_multi_f64": (I, [P, P, I64, I, P, I64, P, I64, P, I64, I, I, P, SZ]),
    "mpdok_lu_ir_multi": (I, [P, P, I64, I, P, I, P, I64, P, I64, I, D, I, I, P, P]),
    "mpdok_lu_resize": (I, [P, I]),
    "mpdok_lu_factors": (I, [P, PP, PP, PI]),
    "mpdok_lu_set_profiling": (I, [P, I]),
    "mpdok_lu_set_lookahead": (I, [P, I]),
    "mpdok_lu_stats": (I, [P, ctypes.POINTER(D), ctypes.POINTER(D), PI, ctypes.POINTER(D)]),
    "mpdok_lu_set_adaptive": (I, [P, I, I, I]),
    "mpdok_lu_slice_counts": (I, [P, PI, I, PI]),
    "mpdok_lu_set_update_panels": (I, [P, I, I, I]),
    "mpdok_lu_step_starts": (I, [P, PI, I, PI]),
    "mpdok_lu_set_row_lookahead": (I, [P, I]),
    "mpdok_gemm": (I, [P, I, I, I, I, D, P, I64, P, I64, D, P, I64, I]),
    "mpdok_ozaki_workspace_bytes": (SZ, [I, I, I, I]),
    "mpdok_ozaki_gemm_sub": (I, [P, I, I, I, I, P, I64, P, I64, P, I64, P]),
    "mpdok_ozaki_apply_sliced": (I, [P, I, I, I, I, P, I64, P, I, P]),
    "mpdok_ozaki_gemm_sub_adaptive": (I, [P, I, I, I, I, P, I64, P, I64, P, I64, P, I, I, I, I, P,
                                          PI, PI]),
    # refinement drivers
    "mpdok_gmres_ir": (I, [P, P, I64, I, P, I64, I, P, I, P, P, D, I, I, D, I, I, I, PI, P, I,
                           PI, PI]),
    "mpdok_lu_ir": (I, [P, P, I64, I, P, I, P, P, D, I, I, PI, P, I, PI]),
    "mpdok_batched_rbf_solve": (I, [I, P, P, I, I, P, D, I, P, P, D, I, I, I, I, I, P, P]),
    # batched plan: grouped launches captured in one hipGraph
    "mpdok_bplan_create": (I, [I, P, P, I, I, P, D, I, P, P, D, I, I, I, PP]),
    "mpdok_bplan_run": (I, [P, P]),
    "mpdok_bplan_results": (I, [P, P, P, P, P, P]),
    "mpdok_bplan_info": (I, [P, P, PI]),
    "mpdok_bplan_destroy": (None, [P]),
}

F64, F32, BF16, F16 = 0, 1, 2, 3
ROW_MAJOR, COL_MAJOR = 0, 1

_lib = None
_lib_path = None


class MPDOKError(RuntimeError):
    """A libmpdok_hip entry point returned a nonzero status."""


class MPDOKMemoryError(MemoryError):
    """Device or host allocation failed inside libmpdok_hip."""


def load(path=None):
    """Load (once) and return the ctypes handle of libmpdok_hip.so."""
    global _lib, _lib_path
    path = os.path.abspath(path or os.environ.get("MPDOK_LIB", DEFAULT_LIB))
    if _lib is not None and _lib_path == path:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"libmpdok_hip.so not found at {path}. Build it with "
            f"'python -m mpdok_amd.build' (hipcc, gfx950).")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    if _lib is None:
        _lib, _lib_path = lib, path
    return lib


def lib():
    return _lib if _lib is not None else load()


def check(status, what="", from_lib=None):
    """Raise for a nonzero status, with the message of the library instance
    that returned it (from_lib; default: the one load() returned first)."""
    if status == 0:
        return
    msg = (from_lib or lib()).mpdok_last_error().decode(errors="replace")
    if status == -2:
        raise MPDOKMemoryError(f"{what}: {msg}")
    raise MPDOKError(f"{what} failed (status {status}): {msg}")
