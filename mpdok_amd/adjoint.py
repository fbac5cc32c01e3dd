"""Differentiable RBF interpolation: validation loss and its hand-written
adjoint, behind a torch.autograd.Function.

Model. Training points X (N x D), targets y; validation points Xv (M x D),
targets yv. Gaussian kernel k(r) = exp(-(eps r)^2) (gamma = eps^2), nugget
sigma = exp(log_sigma):

    alpha = (K + sigma I)^-1 y,   f_v = K_v alpha,   L = (1/M) ||f_v - yv||^2

with K = K(X, X), K_v = K(Xv, X). Gradients (K symmetric):

    g       = (2/M) (f_v - yv)
    lambda  = (K + sigma I)^-1 K_v^T g                        (adjoint solve)
    dL/deps = 2 eps [ g^T (dK_v/dgamma) alpha - lambda^T (dK/dgamma) alpha ]
    dL/dlog_sigma = -sigma lambda^T alpha

with dk/dgamma = -r^2 k. The forward solve and the adjoint solve share one
factorisation (int8-sliced LU by default, FP64 storage) as the preconditioner
of GMRES-IR, so the adjoint costs one more refinement run, not a second
factorisation. K_v alpha, K_v^T g and both bilinear forms are matrix-free:
every kernel value and derivative is recomputed in registers
(mpdok_rbf_cross_matvec, mpdok_rbf_dbilinear), so neither K_v nor dK/dtheta is
stored. This path has no counterpart in the reference; it follows the north
star's differentiable-solve configuration (loss = validation MSE).
"""
import math

import torch

from . import _dev, _ops
from ._lib import check, lib

GAUSSIAN = 0


def _cross(R, C, x):
    """y = K(R, C) x (FP64, no nugget)."""
    y = torch.empty(R.shape[0], dtype=torch.float64, device=R.device)
    check(lib().mpdok_rbf_cross_matvec(_dev.stream(), R.data_ptr(), R.shape[0], C.data_ptr(),
                                       C.shape[0], R.shape[1], GAUSSIAN, _cross.gamma,
                                       x.data_ptr(), y.data_ptr()), "rbf_cross_matvec")
    return y


def _dbilinear(R, C, u, w, gamma):
    """u^T (dK(R, C)/dgamma) w."""
    nb = lib().mpdok_rbf_dbilinear_workspace_bytes(R.shape[0])
    ws = _dev.workspace(nb, "dbilinear")
    out = torch.empty(1, dtype=torch.float64, device=R.device)
    check(lib().mpdok_rbf_dbilinear(_dev.stream(), R.data_ptr(), R.shape[0], C.data_ptr(),
                                    C.shape[0], R.shape[1], GAUSSIAN, float(gamma), u.data_ptr(),
                                    w.data_ptr(), out.data_ptr(), ws.data_ptr(), nb),
          "rbf_dbilinear")
    return float(out.item())


class RBFValidationLoss(torch.autograd.Function):
    """L(eps, log_sigma); gradients by the adjoint method (see module doc).

    forward(eps, log_sigma, X, y, Xv, yv, opts): eps and log_sigma are 0-d
    tensors; X, Xv (n, D) and y, yv (n,) FP64 device tensors; opts a dict of
    solver options (tol, slices, precision, maxiter_outer, restart).
    """

    @staticmethod
    def forward(ctx, eps, log_sigma, X, y, Xv, yv, opts):
        from .mpdok_ops import MPDOKSolver
        e = float(eps)
        sigma = math.exp(float(log_sigma))
        gamma = e * e
        N, M = X.shape[0], Xv.shape[0]
        K = _dev.empty((N, N), torch.float64, "F")
        _ops.rbf_assemble(X, GAUSSIAN, gamma, sigma, K)
        solver = MPDOKSolver()
        kw = dict(tol=opts["tol"], maxiter_outer=opts["maxiter_outer"], restart=opts["restart"],
                  precond="lu", factor_precision=opts["precision"], slices=opts["slices"],
                  pivot="none", body="fp64", residual="dot2")
        alpha = solver.solve(K, y, **kw)
        _cross.gamma = gamma
        f = _cross(Xv, X, alpha)
        r = _ops.axpby(1.0, f, -1.0, yv)                        # f_v - yv
        loss = torch.tensor(float(_ops.dot(r, r).item()) / M, dtype=torch.float64,
                            device=alpha.device)
        ctx.state = (solver, K, kw, alpha, r, e, sigma, gamma)
        ctx.save_for_backward(X, Xv)
        ctx.forward_info = solver.last_info
        return loss

    @staticmethod
    def backward(ctx, gL):
        X, Xv = ctx.saved_tensors
        solver, K, kw, alpha, r, e, sigma, gamma = ctx.state
        M = Xv.shape[0]
        g = _ops.axpby(2.0 / M * float(gL), r)
        _cross.gamma = gamma
        q = _cross(X, Xv, g)                                     # K_v^T g
        lam = solver.solve(K, q, reuse_factor=True, **kw)         # adjoint solve, same factor
        t_v = _dbilinear(Xv, X, g, alpha, gamma)                  # g^T dK_v/dgamma alpha
        t_k = _dbilinear(X, X, lam, alpha, gamma)                 # lambda^T dK/dgamma alpha
        d_eps = 2.0 * e * (t_v - t_k)
        d_logsig = -sigma * float(_ops.dot(lam, alpha).item())
        solver.free_fp32_buf()
        ctx.state = None
        dev = alpha.device
        return (torch.tensor(d_eps, dtype=torch.float64, device=dev),
                torch.tensor(d_logsig, dtype=torch.float64, device=dev),
                None, None, None, None, None)


def rbf_validation_loss(eps, log_sigma, X, y, Xv, yv, tol=1e-12, slices=5, precision="ozaki",
                        maxiter_outer=20, restart=50):
    """L(eps, log_sigma) as a differentiable 0-d tensor (see module doc).

    eps, log_sigma: Python floats or 0-d tensors (requires_grad as needed).
    X, Xv: (n, D) point arrays; y, yv: (n,) targets (NumPy or torch).
    """
    dev = _dev.device()

    def scalar(v):
        if isinstance(v, torch.Tensor):
            return v.to(dtype=torch.float64)
        return torch.tensor(float(v), dtype=torch.float64)
    X = _dev.as_device(X, torch.float64).contiguous()
    Xv = _dev.as_device(Xv, torch.float64).contiguous()
    y = _dev.as_device(y, torch.float64)
    yv = _dev.as_device(yv, torch.float64)
    if X.shape[1] != Xv.shape[1] or y.shape[0] != X.shape[0] or yv.shape[0] != Xv.shape[0]:
        raise ValueError("rbf_validation_loss: inconsistent shapes")
    opts = dict(tol=tol, slices=slices, precision=precision, maxiter_outer=maxiter_outer,
                restart=restart)
    return RBFValidationLoss.apply(scalar(eps).to(dev), scalar(log_sigma).to(dev), X, y, Xv, yv,
                                   opts)
