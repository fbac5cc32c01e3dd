"""Adjoint gradients of the RBF validation loss (north star's differentiable
solve) against a dense FP64 closed-form gradient and central differences,
plus the matrix-free cross product and derivative bilinear form kernels."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _problem(N, M, seed, noise=1e-3):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0, 1, (N, 2))
    Xv = rng.uniform(0, 1, (M, 2))
    f = lambda P: np.sin(3 * np.pi * P[:, 0]) * np.exp(-P[:, 1])
    return X, f(X) + noise * rng.standard_normal(N), Xv, f(Xv) + noise * rng.standard_normal(M)


def _d2(A, B):
    return np.maximum((A * A).sum(1)[:, None] + (B * B).sum(1)[None, :] - 2 * A @ B.T, 0.0)


def dense_loss_and_grad(eps, log_sigma, X, y, Xv, yv):
    """FP64 closed form (LAPACK solves) of L, dL/deps, dL/dlog_sigma."""
    gamma, sigma = eps * eps, np.exp(log_sigma)
    D2, D2v = _d2(X, X), _d2(Xv, X)
    E, Ev = np.exp(-gamma * D2), np.exp(-gamma * D2v)
    K = E + sigma * np.eye(len(X))
    alpha = np.linalg.solve(K, y)
    r = Ev @ alpha - yv
    M = len(Xv)
    g = 2.0 * r / M
    lam = np.linalg.solve(K, Ev.T @ g)
    dgamma = g @ ((-D2v * Ev) @ alpha) - lam @ ((-D2 * E) @ alpha)
    return r @ r / M, 2 * eps * dgamma, -sigma * lam @ alpha


def test_cross_matvec_and_dbilinear_match_closed_form():
    from mpdok_amd import adjoint
    rng = np.random.default_rng(3)
    R, C = rng.uniform(0, 1, (300, 3)), rng.uniform(0, 1, (517, 3))
    x, u = rng.standard_normal(517), rng.standard_normal(300)
    gamma = 20.0
    Rt, Ct = torch.as_tensor(R).cuda(), torch.as_tensor(C).cuda()
    adjoint._cross.gamma = gamma
    y = adjoint._cross(Rt, Ct, torch.as_tensor(x).cuda()).cpu().numpy()
    E = np.exp(-gamma * _d2(R, C))
    assert np.allclose(y, E @ x, rtol=1e-12, atol=1e-12 * np.abs(E) @ np.abs(x))
    b = adjoint._dbilinear(Rt, Ct, torch.as_tensor(u).cuda(), torch.as_tensor(x).cuda(), gamma)
    ref = u @ ((-_d2(R, C) * E) @ x)
    assert b == pytest.approx(ref, rel=1e-11, abs=1e-12 * np.abs(u) @ (_d2(R, C) * E) @ np.abs(x))


# (eps, log_sigma, train points, validation points, relative tolerance of the
# gradients). log sigma = -12 is the adjoint workload's own noise level (a
# nugget of e^-12 = 6.1e-6): the system's condition number grows about e^3
# over log sigma = -9, and both the GPU solve and the LAPACK reference lose
# digits with it.
@pytest.mark.parametrize("eps,log_sigma,n,nv,rtol", [(6.0, -6.0, 1024, 256, 1e-7),
                                                     (4.0, -9.0, 1024, 256, 1e-7),
                                                     (6.0, -12.0, 2048, 512, 1e-5)])
def test_adjoint_gradient_matches_dense_fp64(eps, log_sigma, n, nv, rtol):
    from mpdok_amd.adjoint import rbf_validation_loss
    X, y, Xv, yv = _problem(n, nv, 7)
    e = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
    s = torch.tensor(log_sigma, dtype=torch.float64, requires_grad=True)
    L = rbf_validation_loss(e, s, X, y, Xv, yv, tol=1e-13)
    L.backward()
    L_ref, de_ref, ds_ref = dense_loss_and_grad(eps, log_sigma, X, y, Xv, yv)
    assert float(L.detach()) == pytest.approx(L_ref, rel=rtol * 1e-2)
    assert float(e.grad) == pytest.approx(de_ref, rel=rtol)
    assert float(s.grad) == pytest.approx(ds_ref, rel=rtol)


def test_adjoint_gradient_matches_central_differences():
    """Central FD with h = 1e-5 on both parameters, to 1e-6 relative."""
    from mpdok_amd.adjoint import rbf_validation_loss
    X, y, Xv, yv = _problem(768, 192, 11)
    eps, log_sigma, h = 6.0, -6.0, 1e-5
    e = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
    s = torch.tensor(log_sigma, dtype=torch.float64, requires_grad=True)
    rbf_validation_loss(e, s, X, y, Xv, yv, tol=1e-14, maxiter_outer=30).backward()
    Lf = lambda a, b: float(rbf_validation_loss(a, b, X, y, Xv, yv, tol=1e-14, maxiter_outer=30))
    fd_e = (Lf(eps + h, log_sigma) - Lf(eps - h, log_sigma)) / (2 * h)
    fd_s = (Lf(eps, log_sigma + h) - Lf(eps, log_sigma - h)) / (2 * h)
    assert float(e.grad) == pytest.approx(fd_e, rel=1e-6)
    assert float(s.grad) == pytest.approx(fd_s, rel=1e-6)
