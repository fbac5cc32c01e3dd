"""Typed wrappers over the libmpdok_hip C-ABI. Every function enqueues HIP
kernels on torch's current stream and returns without synchronising unless it
says otherwise."""
import torch

from . import _dev
from . import _lib
from ._lib import check, lib


def matvec(A, x, out=None, skip=None):
    """y = A @ x for a 2-D FP32/FP64 device matrix (C or F storage). `skip`:
    optional device int32 flag; when nonzero the kernels return at once."""
    order, ld = _dev.layout(A)
    M, N = A.shape
    if x.dtype != A.dtype:
        raise TypeError(f"matvec: x dtype {x.dtype} != A dtype {A.dtype}")
    if x.numel() != N:
        raise ValueError(f"matvec: x has {x.numel()} elements, A has {N} columns")
    y = out if out is not None else torch.empty(M, dtype=A.dtype, device=A.device)
    dt = _dev.dtcode(A)
    wsb = lib().mpdok_gemv_workspace_bytes(M, N, dt, order)
    ws = _dev.workspace(wsb, "gemv") if wsb else None
    fn = lib().mpdok_gemv_f32 if A.dtype == torch.float32 else lib().mpdok_gemv_f64
    check(fn(_dev.stream(), A.data_ptr(), ld, order, M, N, x.data_ptr(), y.data_ptr(),
             _dev.ptr(ws), wsb, _dev.ptr(skip)), "gemv")
    return y


def residual(A64, x64, b64, out=None, mode="fp64"):
    """r = b - A x in FP64 (mode 'fp64') or compensated Dot2 accumulation ('dot2')."""
    order, ld = _dev.layout(A64)
    N = A64.shape[0]
    r = out if out is not None else torch.empty(N, dtype=torch.float64, device=A64.device)
    wsb = lib().mpdok_residual_workspace_bytes(N, order)
    ws = _dev.workspace(wsb, "resid") if wsb else None
    check(lib().mpdok_residual_f64(_dev.stream(), A64.data_ptr(), ld, order, N, x64.data_ptr(),
                                   b64.data_ptr(), r.data_ptr(), 1 if mode == "dot2" else 0,
                                   _dev.ptr(ws), wsb), "residual")
    return r


def dot(x, y, out=None):
    """Device FP64 scalar <x, y> (x, y same dtype, FP32 or FP64)."""
    o = out if out is not None else torch.empty(1, dtype=torch.float64, device=x.device)
    n = x.numel()
    ws = _dev.workspace(lib().mpdok_reduce_workspace_bytes(n), "dot")
    fn = lib().mpdok_dot_f32 if x.dtype == torch.float32 else lib().mpdok_dot_f64
    check(fn(_dev.stream(), x.data_ptr(), y.data_ptr(), n, o.data_ptr(), ws.data_ptr()), "dot")
    return o


def norm(x):
    """||x||_2 as a Python float (synchronises)."""
    return float(dot(x, x).item()) ** 0.5


def axpby(alpha, x, beta=0.0, y=None, out=None, dtype=None):
    """out = alpha x + beta y (FP64 arithmetic, one rounding to out's dtype)."""
    if out is None:
        out = torch.empty(x.numel(), dtype=dtype or x.dtype, device=x.device)
    check(lib().mpdok_axpby(_dev.stream(), x.numel(), float(alpha), x.data_ptr(), _dev.dtcode(x),
                            float(beta), _dev.ptr(y), _dev.dtcode(y) if y is not None else 0,
                            out.data_ptr(), _dev.dtcode(out)), "axpby")
    return out


def cast_matrix(A, dtype, order="same", out=None):
    """Copy a 2-D device matrix into `dtype` with C ('C'), Fortran ('F') or the
    source ('same') storage order."""
    iorder, ild = _dev.layout(A)
    M, N = A.shape
    if order == "same":
        order = "C" if iorder == _lib.ROW_MAJOR else "F"
    if out is None:
        out = _dev.empty((M, N), dtype, order)
    oorder, old = _dev.layout(out)
    check(lib().mpdok_cast_matrix(_dev.stream(), A.data_ptr(), _dev.dtcode(A), ild, iorder,
                                  out.data_ptr(), _dev.dtcode(out), old, oorder, M, N),
          "cast_matrix")
    return out


def rbf_assemble(coords, kind, param, reg, out, row0=0, rows=None, col0=0, cols=None,
                 reg_mode=0):
    """Write kernel values for rows [row0, row0+rows) x cols [col0, col0+cols)
    into `out` (2-D device tensor, C or F storage, FP64/FP32)."""
    N, D = coords.shape
    rows = N - row0 if rows is None else rows
    cols = N - col0 if cols is None else cols
    order, ld = _dev.layout(out)
    if tuple(out.shape) != (rows, cols):
        raise ValueError(f"rbf_assemble: out shape {tuple(out.shape)} != ({rows}, {cols})")
    check(lib().mpdok_rbf_assemble(_dev.stream(), coords.data_ptr(), N, D, int(kind),
                                   float(param), float(reg), int(reg_mode), row0, rows, col0,
                                   cols, out.data_ptr(), _dev.dtcode(out), ld, order),
          "rbf_assemble")
    return out


def rbf_mean_d2(coords, n):
    o = torch.empty(1, dtype=torch.float64, device=coords.device)
    check(lib().mpdok_rbf_mean_d2(_dev.stream(), coords.data_ptr(), int(n), coords.shape[1],
                                  o.data_ptr()), "rbf_mean_d2")
    return o


def pcg_step(x, r, z, p, Ap, dinv, rz_old, pAp, rz_new, rr_new=None):
    """One set of (preconditioned) CG updates on device vectors of one dtype
    (FP32 or FP64) with device FP64 scalars: see mpdok_pcg_step_updates."""
    ws = _dev.workspace(lib().mpdok_pcg_workspace_bytes(), "pcg")
    check(lib().mpdok_pcg_step_updates(_dev.stream(), _dev.dtcode(x), x.data_ptr(), r.data_ptr(),
                                       _dev.ptr(z), p.data_ptr(), Ap.data_ptr(), _dev.ptr(dinv),
                                       x.numel(), rz_old.data_ptr(), pAp.data_ptr(),
                                       rz_new.data_ptr(), _dev.ptr(rr_new), ws.data_ptr()),
          "pcg_step_updates")


def diag_scale(d, x, out=None):
    """out = d .* x (same dtype, FP32 or FP64)."""
    out = out if out is not None else torch.empty_like(x)
    check(lib().mpdok_diag_scale(_dev.stream(), _dev.dtcode(x), d.data_ptr(), x.data_ptr(),
                                 out.data_ptr(), x.numel()), "diag_scale")
    return out


def rbf_diag(coords, kind, param, reg, elem_dtype=torch.float32, out_dtype=torch.float32,
             invert=False):
    """Diagonal of the kernel matrix as the assembly forms it (elem FP32: the
    stored FP32 tile's element), optionally inverted."""
    N, D = coords.shape
    out = torch.empty(N, dtype=out_dtype, device=coords.device)
    check(lib().mpdok_rbf_diag(_dev.stream(), coords.data_ptr(), N, D, int(kind), float(param),
                               float(reg), _DTC[elem_dtype], 1 if invert else 0, out.data_ptr(),
                               _DTC[out_dtype]), "rbf_diag")
    return out


_DTC = {torch.float64: _lib.F64, torch.float32: _lib.F32}
