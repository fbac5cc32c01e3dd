"""
MPDOK on MI355X — dense operator with FP32 device storage and Krylov solvers.

Same entry points, signatures, defaults and return values as the reference's
``dense_krylov.py`` (DenseLinearOperator :18-112, cg :119-205, gmres :212-331,
gmres_ir :338-421). Differences a caller can see:

* arrays come back as torch ROCm tensors (DLPack-exportable) where the
  reference returns CuPy arrays;
* every product, reduction and update runs in this package's HIP kernels
  (libmpdok_hip.so); dot products accumulate in FP64 and stay on the device,
  so an Arnoldi step needs no host synchronisation (the reference performs two);
* ``residual='dot2'`` (keyword, default ``'fp64'`` as in the reference)
  evaluates the FP64 residual with compensated (Dot2) accumulation; plain FP64
  residuals floor near relres 1e-12 on kernel systems with condition ~1e9.

Usage:
    from mpdok_amd.dense_krylov import DenseLinearOperator, cg
    op = DenseLinearOperator(A)          # A: FP64 NumPy or torch (N x N)
    x, hist, ok = cg(op, b)
"""
import numpy as np
import torch

from . import _dev, _ops
from ._lib import check, lib


class DenseLinearOperator:
    """Dense matrix stored on the device in FP32 (Krylov products) and FP64
    (residual checks), row-major as in dense_krylov.py:44-47.

    Call free_fp64() to drop the FP64 copy; residual_norm() raises afterwards.
    """

    def __init__(self, A, residual="fp64"):
        if A.ndim != 2 or A.shape[0] != A.shape[1]:
            raise ValueError(f"A must be square 2-D, got shape {tuple(A.shape)}")
        self.N = int(A.shape[0])
        self.residual_mode = residual
        self.A_fp64 = _dev.as_device(A, torch.float64, "C")
        self.A_fp32 = _ops.cast_matrix(self.A_fp64, torch.float32, "C")
        torch.cuda.current_stream().synchronize()
        mb32 = self.A_fp32.numel() * 4 / 1e6
        mb64 = self.A_fp64.numel() * 8 / 1e6
        print(f"DenseLinearOperator: N={self.N:,}  "
              f"FP32={mb32:.1f} MB  FP64={mb64:.1f} MB  "
              f"total={mb32 + mb64:.1f} MB")

    # -- core operations -------------------------------------------------
    def matvec(self, x):
        """A @ x in FP32. x: FP32 device vector (N,). Returns FP32 (N,)."""
        if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float32):
            x = _dev.as_device(x, torch.float32)
        return _ops.matvec(self.A_fp32, x)

    def residual_norm(self, x, b):
        """||b - A x||_2 in FP64, as a Python float (dense_krylov.py:65-79)."""
        if self.A_fp64 is None:
            raise RuntimeError(
                "FP64 copy was freed; residual_norm() is no longer available. "
                "Reconstruct the operator to recompute residuals.")
        x64 = _as64(x)
        b64 = _as64(b)
        r = _ops.residual(self.A_fp64, x64, b64, mode=self.residual_mode)
        return _ops.norm(r)

    # -- memory management -----------------------------------------------
    def free_fp64(self):
        """Release the FP64 copy; matvec() still works, residual_norm() raises."""
        self.A_fp64 = None

    # -- diagnostics -----------------------------------------------------
    @property
    def shape(self):
        return (self.N, self.N)

    def vram_mb(self):
        total = self.A_fp32.numel() * 4
        if self.A_fp64 is not None:
            total += self.A_fp64.numel() * 8
        return total / 1e6

    def __repr__(self):
        fp64_state = (f"{self.A_fp64.numel() * 8 / 1e6:.1f} MB"
                      if self.A_fp64 is not None else "freed")
        return (f"DenseLinearOperator(N={self.N:,}, "
                f"fp32={self.A_fp32.numel() * 4 / 1e6:.1f} MB, fp64={fp64_state})")


def _as64(v):
    if isinstance(v, torch.Tensor) and v.is_cuda and v.dtype == torch.float64 and v.is_contiguous():
        return v
    if isinstance(v, torch.Tensor) and v.is_cuda and v.dtype == torch.float32:
        return _ops.axpby(1.0, v.contiguous(), dtype=torch.float64)
    return _dev.as_device(v, torch.float64)


def _as32(v):
    if isinstance(v, torch.Tensor) and v.is_cuda and v.dtype == torch.float32 and v.is_contiguous():
        return v.clone()
    if isinstance(v, torch.Tensor) and v.is_cuda and v.dtype == torch.float64:
        return _ops.axpby(1.0, v.contiguous(), dtype=torch.float32)
    return _dev.as_device(v, torch.float32).clone()


# ----------------------------------------------------------------------
# Conjugate gradients (dense_krylov.py:119-205)
# ----------------------------------------------------------------------
def cg(op, b, x0=None, tol=1e-6, maxiter=None, check_every=10, verbose=False, body="fp32",
       precond=None):
    """CG for SPD A: FP32 vectors, FP64 residual check every `check_every`
    iterations before the update, stop on <1% improvement between checks.

    Extensions (defaults reproduce the reference): body='fp64' runs the
    iteration on the FP64 copy with FP64 vectors (the plain FP64 CG);
    precond='jacobi' preconditions with diag(A)^-1.

    Returns (x FP64 (N,), history [(iter, relres)], converged).
    """
    if body not in ("fp32", "fp64"):
        raise ValueError(f"body must be 'fp32' or 'fp64', got {body!r}")
    if precond not in (None, "jacobi"):
        raise ValueError(f"precond must be None or 'jacobi', got {precond!r}")
    N = op.N
    if maxiter is None:
        maxiter = N
    dt = torch.float32 if body == "fp32" else torch.float64
    A = op.A_fp32 if body == "fp32" else op.A_fp64
    if A is None:
        raise RuntimeError("body='fp64' needs the FP64 copy, which was freed")
    b64 = _as64(b)
    bv = _ops.axpby(1.0, b64, dtype=dt)
    b_norm = _ops.norm(b64)
    if b_norm == 0.0:
        return _dev.zeros(N, torch.float64), [(0, 0.0)], True

    if x0 is None:
        x = _dev.zeros(N, dt)
    else:
        x = _as32(x0) if body == "fp32" else _as64(x0).clone()
    r = _ops.axpby(1.0, bv, -1.0, _ops.matvec(A, x))
    dinv = None
    if precond == "jacobi":
        # setup only: the diagonal is read once and inverted on the host
        d = A.diagonal().cpu().numpy().astype(np.float64)
        dinv = _dev.as_device((1.0 / d).astype(np.float32 if body == "fp32" else np.float64), dt)
    z = _ops.diag_scale(dinv, r) if dinv is not None else None
    p = (z if z is not None else r).clone()
    Ap = torch.empty_like(r)
    rs = [_ops.dot(r, z if z is not None else r),
          torch.empty(1, dtype=torch.float64, device=r.device)]
    pAp = torch.empty(1, dtype=torch.float64, device=r.device)
    ws = _dev.workspace(8 * 4096, "cg")
    reference_kernel = body == "fp32" and dinv is None

    history = []
    converged = False
    prev_rel = float("inf")
    stagnation_tol = 1e-2
    k = -1
    for k in range(maxiter):
        if k % check_every == 0:
            rel = op.residual_norm(x, b64) / b_norm
            history.append((k, rel))
            if verbose:
                print(f"  CG iter {k:4d}  rel_res(FP64) = {rel:.3e}")
            if rel < tol:
                converged = True
                break
            if rel > prev_rel * (1.0 - stagnation_tol):
                if verbose:
                    print(f"  CG stagnated at iter {k}: {prev_rel:.3e} → {rel:.3e}")
                break
            prev_rel = rel
        _ops.matvec(A, p, out=Ap)
        _ops.dot(p, Ap, out=pAp)
        cur, nxt = rs[k % 2], rs[(k + 1) % 2]
        if reference_kernel:
            check(lib().mpdok_cg_step_updates(_dev.stream(), x.data_ptr(), r.data_ptr(),
                                              p.data_ptr(), Ap.data_ptr(), N, cur.data_ptr(),
                                              pAp.data_ptr(), nxt.data_ptr(), ws.data_ptr()),
                  "cg_step_updates")
        else:
            _ops.pcg_step(x, r, z, p, Ap, dinv, cur, pAp, nxt)

    if not converged:
        rel = op.residual_norm(x, b64) / b_norm
        history.append((k + 1, rel))
        converged = rel < tol

    return (_ops.axpby(1.0, x, dtype=torch.float64) if body == "fp32" else x), history, converged


def _pcg_inner(matvec, rhs, dtype, dinv=None, tol=1e-6, maxiter=1000, check_every=10):
    """(Jacobi-)PCG on A e = rhs from e = 0 with vectors in `dtype` and device
    scalars; stops when the recursive residual satisfies ||r|| < tol ||rhs||
    (read on the host every `check_every` iterations) or after maxiter
    iterations. matvec(v) returns A v in `dtype`. Returns (e, iterations,
    ||r|| / ||rhs||)."""
    r = _ops.axpby(1.0, rhs, dtype=dtype)
    beta = _ops.norm(r)
    N = r.numel()
    e = _dev.zeros(N, dtype)
    if beta == 0.0:
        return e, 0, 0.0
    z = _ops.diag_scale(dinv, r) if dinv is not None else None
    p = (z if z is not None else r).clone()
    rs = [_ops.dot(r, z if z is not None else r),
          torch.empty(1, dtype=torch.float64, device=r.device)]
    rr = torch.empty(1, dtype=torch.float64, device=r.device)
    pAp = torch.empty(1, dtype=torch.float64, device=r.device)
    rel, it = 1.0, 0
    for it in range(1, maxiter + 1):
        Ap = matvec(p)
        _ops.dot(p, Ap, out=pAp)
        cur, nxt = rs[(it - 1) % 2], rs[it % 2]
        _ops.pcg_step(e, r, z, p, Ap, dinv, cur, pAp, nxt, rr)
        if it % check_every == 0 or it == maxiter:
            rel = float(rr.item()) ** 0.5 / beta
            if rel < tol:
                break
    return e, it, rel


# ----------------------------------------------------------------------
# Restarted GMRES(m) (dense_krylov.py:212-331)
# ----------------------------------------------------------------------
class _Arnoldi:
    """Device workspace for one restart cycle: basis V ((m+1) x N, one vector
    per row), Hessenberg H (stored column by column), control words."""

    def __init__(self, N, restart, dtype):
        dev = _dev.device()
        self.N, self.m, self.dtype = N, restart, dtype
        self.V = torch.empty((restart + 1, N), dtype=dtype, device=dev)
        self.Hc = torch.zeros((restart, restart + 1), dtype=torch.float64, device=dev)
        self.R = torch.zeros_like(self.Hc)
        self.cs = torch.zeros(restart, dtype=torch.float64, device=dev)
        self.sn = torch.zeros(restart, dtype=torch.float64, device=dev)
        self.g = torch.zeros(restart + 1, dtype=torch.float64, device=dev)
        self.ctl = torch.zeros(4, dtype=torch.int32, device=dev)
        self.w = torch.empty(N, dtype=dtype, device=dev)
        self.wsb = lib().mpdok_arnoldi_workspace_bytes(N, restart)
        self.ws = torch.empty(max(self.wsb, 256), dtype=torch.uint8, device=dev)
        self._orth = (lib().mpdok_arnoldi_orth_f32 if dtype == torch.float32
                      else lib().mpdok_arnoldi_orth_f64)

    def reset(self):
        self.Hc.zero_()
        self.ctl.zero_()

    def orth(self, k, passes, stop_now=1, thresh=1e-12):
        check(self._orth(_dev.stream(), self.V.data_ptr(), self.V.stride(0), k, self.w.data_ptr(),
                         self.N, self.Hc[k].data_ptr(), passes, thresh, stop_now,
                         self.ctl.data_ptr(), self.ws.data_ptr(), self.wsb), "arnoldi_orth")

    def givens(self, k, beta, tol):
        check(lib().mpdok_givens_step(_dev.stream(), self.Hc.data_ptr(), self.m + 1, k,
                                      self.R.data_ptr(), self.cs.data_ptr(), self.sn.data_ptr(),
                                      self.g.data_ptr(), float(beta), float(tol),
                                      self.ctl.data_ptr()), "givens_step")

    def steps_done(self):
        return int(self.ctl[1].item())

    def H_host(self, m):
        """(m+1) x m Hessenberg matrix on the host (FP64)."""
        return self.Hc[:m, :m + 1].cpu().numpy().T.copy()


def gmres(op, b, x0=None, tol=1e-6, maxiter=None, restart=50, verbose=False):
    """Restarted GMRES(m): FP32 Arnoldi basis with one-pass batched CGS on the
    device, (m+1) x m least squares on the host, FP64 residual per restart,
    stop on <1% improvement between restarts. `maxiter` counts Arnoldi matvecs.

    Returns (x FP64 (N,), history [(total_iters, relres)], converged).
    """
    N = op.N
    if maxiter is None:
        maxiter = N
    b64 = _as64(b)
    b32 = _ops.axpby(1.0, b64, dtype=torch.float32)
    b_norm = _ops.norm(b64)
    if b_norm == 0.0:
        return _dev.zeros(N, torch.float64), [(0, 0.0)], True

    x = _dev.zeros(N, torch.float32) if x0 is None else _as32(x0)
    history = []
    converged = False
    total_iters = 0
    prev_rel = float("inf")
    stagnation_tol = 1e-2
    max_restarts = (maxiter + restart - 1) // restart
    ar = _Arnoldi(N, restart, torch.float32)

    for outer in range(max_restarts):
        r = _ops.axpby(1.0, b32, -1.0, op.matvec(x))
        beta = _ops.norm(r)
        if beta == 0.0:
            converged = True
            break
        _ops.axpby(1.0 / beta, r, out=ar.V[0])
        ar.reset()
        started = 0
        for k in range(restart):
            if total_iters + started >= maxiter:
                break
            _ops.matvec(op.A_fp32, ar.V[k], out=ar.w, skip=ar.ctl)
            ar.orth(k, passes=1)
            started += 1
        m = ar.steps_done() if started else 0
        total_iters += m

        e1 = np.zeros(m + 1, dtype=np.float64)
        e1[0] = beta
        if m > 0:
            y, _, _, _ = np.linalg.lstsq(ar.H_host(m), e1, rcond=None)
            y_dev = torch.as_tensor(y, dtype=torch.float64).to(ar.V.device)
            check(lib().mpdok_basis_update_f32(_dev.stream(), ar.V.data_ptr(), ar.V.stride(0), m,
                                               y_dev.data_ptr(), x.data_ptr(), N, 0),
                  "basis_update")

        rel = op.residual_norm(x, b64) / b_norm
        history.append((total_iters, rel))
        if verbose:
            print(f"  GMRES restart {outer + 1:3d}  "
                  f"iters={total_iters:4d}  rel_res(FP64)={rel:.3e}")
        if rel < tol:
            converged = True
            break
        if rel > prev_rel * (1.0 - stagnation_tol):
            if verbose:
                print(f"  GMRES stagnated: {prev_rel:.3e} → {rel:.3e}")
            break
        prev_rel = rel

    return _ops.axpby(1.0, x, dtype=torch.float64), history, converged


# ----------------------------------------------------------------------
# GMRES-IR (dense_krylov.py:338-421)
# ----------------------------------------------------------------------
def gmres_ir(op, b, tol=1e-11, maxiter_outer=5, restart=50,
             maxiter_inner=None, verbose=False):
    """FP64 iterative refinement around an FP32 GMRES inner solve (inner
    tolerance 1e-6, no preconditioner), FP64 update x += e.

    The history always ends with an entry (maxiter_outer, relres), also after
    an early exit, as in the reference (dense_krylov.py:414-419).
    Returns (x FP64 (N,), history [(outer, relres)], converged).
    """
    if op.A_fp64 is None:
        raise RuntimeError(
            "GMRES-IR requires op.A_fp64 for FP64 residual computation. "
            "Do not call free_fp64() before gmres_ir().")
    N = op.N
    b64 = _as64(b)
    b_norm = _ops.norm(b64)
    if b_norm == 0.0:
        return _dev.zeros(N, torch.float64), [(0, 0.0)], True

    mode = getattr(op, "residual_mode", "fp64")
    x = _dev.zeros(N, torch.float64)
    r64 = torch.empty_like(x)
    history = []
    converged = False
    for outer in range(maxiter_outer):
        _ops.residual(op.A_fp64, x, b64, out=r64, mode=mode)
        rel = _ops.norm(r64) / b_norm
        history.append((outer, rel))
        if verbose:
            print(f"  GMRES-IR outer {outer:2d}  rel_res(FP64) = {rel:.3e}")
        if rel < tol:
            converged = True
            break
        e, _, _ = gmres(op, r64, tol=1e-6, restart=restart, maxiter=maxiter_inner,
                        verbose=False)
        x = _ops.axpby(1.0, x, 1.0, e)

    _ops.residual(op.A_fp64, x, b64, out=r64, mode=mode)
    rel = _ops.norm(r64) / b_norm
    history.append((maxiter_outer, rel))
    converged = converged or (rel < tol)
    return x, history, converged
