"""GPU parity of mpdok_amd.dense_krylov against the oracle restatement of
dense_krylov.py (same seeded inputs, same arguments)."""
import numpy as np
import pytest
import torch

from oracle import krylov as ok
from oracle import rbf as orb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dk():
    from mpdok_amd import dense_krylov
    return dense_krylov


def spd(N, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((N, N))
    return X.T @ X + N * np.eye(N), rng.standard_normal(N)


def nonsym(N, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((N, N)) + N * np.eye(N), rng.standard_normal(N)


def relerr(x, y):
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


def test_operator_matvec_residual_and_free(dk):
    A, b = spd(512, 1)
    op = dk.DenseLinearOperator(A)
    x = np.random.default_rng(2).standard_normal(512)
    y = op.matvec(torch.as_tensor(x, dtype=torch.float32).cuda()).cpu().numpy()
    ref = A.astype(np.float32).astype(np.float64) @ x.astype(np.float32).astype(np.float64)
    assert relerr(y, ref) < 1e-6
    assert op.residual_norm(x, A @ x) < 1e-9 * np.linalg.norm(A @ x)
    assert op.shape == (512, 512) and op.vram_mb() == pytest.approx(512 * 512 * 12 / 1e6)
    op.free_fp64()
    assert op.A_fp64 is None and "freed" in repr(op)
    with pytest.raises(RuntimeError):
        op.residual_norm(x, b)
    with pytest.raises(RuntimeError):
        dk.gmres_ir(op, b)
    with pytest.raises(ValueError):
        dk.DenseLinearOperator(np.zeros((3, 4)))


@pytest.mark.parametrize("N", [256, 2048])
def test_cg_matches_oracle(dk, N):
    A, b = spd(N, 7)
    x, hist, conv = dk.cg(dk.DenseLinearOperator(A), b, tol=1e-6)
    xo, histo, convo = ok.cg(ok.DenseOp(A), b, tol=1e-6)
    assert conv and convo
    assert [h[0] for h in hist] == [h[0] for h in histo]
    assert hist[-1][1] < 1e-6
    xr = np.linalg.solve(A, b)
    assert relerr(x.cpu().numpy(), xr) < 1e-5
    assert x.dtype == torch.float64


def test_cg_zero_rhs_and_x0(dk):
    A, b = spd(128, 3)
    op = dk.DenseLinearOperator(A)
    x, hist, conv = dk.cg(op, np.zeros(128))
    assert conv and hist == [(0, 0.0)] and not x.any()
    xr = np.linalg.solve(A, b)
    x, hist, conv = dk.cg(op, b, x0=xr, tol=1e-6)
    assert conv and hist[0][0] == 0 and len(hist) == 1


@pytest.mark.parametrize("kind,N", [("spd", 512), ("nonsym", 512), ("nonsym", 2048)])
def test_gmres_matches_oracle(dk, kind, N):
    A, b = (spd if kind == "spd" else nonsym)(N, 13)
    x, hist, conv = dk.gmres(dk.DenseLinearOperator(A), b, tol=1e-6, restart=50)
    xo, histo, convo = ok.gmres(ok.DenseOp(A), b, tol=1e-6, restart=50)
    assert conv == convo
    assert [h[0] for h in hist] == [h[0] for h in histo]
    assert relerr(x.cpu().numpy(), xo) < 1e-5


def test_gmres_exact_breakdown(dk):
    N = 40
    A = np.diag(np.arange(1.0, N + 1.0))
    b = np.zeros(N)
    b[3] = 8.0
    x, hist, conv = dk.gmres(dk.DenseLinearOperator(A), b, tol=1e-6, restart=10)
    assert conv and hist[0][0] == 1
    assert float(x[3]) == pytest.approx(2.0, rel=1e-7)


def test_gmres_maxiter_counts_matvecs(dk):
    A, b = nonsym(300, 5)
    x, hist, conv = dk.gmres(dk.DenseLinearOperator(A), b, tol=1e-14, restart=7, maxiter=10)
    assert [h[0] for h in hist][:2] == [7, 10][:len(hist)]


@pytest.mark.parametrize("N", [512, 2048])
def test_gmres_ir_matches_oracle(dk, N):
    A, b = nonsym(N, 42)
    x, hist, conv = dk.gmres_ir(dk.DenseLinearOperator(A), b, tol=1e-11)
    xo, histo, convo = ok.gmres_ir(ok.DenseOp(A), b, tol=1e-11)
    assert conv and convo
    assert hist[-1][0] == 5 and len(hist) == len(histo)
    xr = np.linalg.solve(A, b)
    assert relerr(x.cpu().numpy(), xr) < 1e-12


def test_gmres_ir_dot2_residual_on_rbf(dk):
    X = np.random.default_rng(3).uniform(0, 1, (1024, 2))
    A = orb.build_kernel(X, "gaussian", 4.0, 1e-2)
    b = np.sin(2 * np.pi * X[:, 0])
    op = dk.DenseLinearOperator(A, residual="dot2")
    x, hist, conv = dk.gmres_ir(op, b, tol=1e-12, maxiter_outer=8)
    assert conv
    xr = np.linalg.solve(A, b)
    assert relerr(x.cpu().numpy(), xr) < 1e-9


def test_cg_fp64_body_and_jacobi_match_oracle(dk):
    """The FP64 body and the Jacobi preconditioner (extensions of cg) against
    the oracle's same loop on D (X^T X + N I) D (a spread spectrum, on which
    CG does not amplify rounding differences; D makes the diagonal vary):
    the same check iterations, relres values within 1e-6, LAPACK's solution."""
    N = 512
    rng = np.random.default_rng(7)
    X = rng.standard_normal((N, N))
    b = rng.standard_normal(N)
    d = np.linspace(1.0, 10.0, N)
    A = (X.T @ X + N * np.eye(N)) * d[:, None] * d[None, :]
    op = dk.DenseLinearOperator(A)
    xr = np.linalg.solve(A, b)
    dinv = 1.0 / np.diag(A)
    for pre, dv in ((None, None), ("jacobi", dinv)):
        x, hist, conv = dk.cg(op, b, tol=1e-12, maxiter=4000, check_every=5, body="fp64",
                              precond=pre)
        xo, histo, convo = ok.cg(ok.DenseOp(A), b, tol=1e-12, maxiter=4000, check_every=5,
                                 body="fp64", dinv=dv)
        assert conv and convo
        assert [h[0] for h in hist] == [h[0] for h in histo]
        # rounding differences grow once relres is small (1e-15 perturbations
        # of A move the oracle's own checks by 1e-4 after iteration 65)
        a = np.array([h[1] for h in hist])
        o = np.array([h[1] for h in histo])
        early = (np.array([h[0] for h in hist]) <= 60) & (o > 1e-9)
        assert np.allclose(a[early], o[early], rtol=1e-6, atol=0)
        assert np.all(np.abs(np.log(a[~early] / o[~early])) < np.log(2.0))
        assert relerr(x.cpu().numpy(), xr) < 1e-9
    # FP32 vectors, Jacobi-preconditioned, to 1e-6
    x, hist, conv = dk.cg(op, b, tol=1e-6, maxiter=4000, check_every=5, precond="jacobi")
    xo, histo, convo = ok.cg(ok.DenseOp(A), b, tol=1e-6, maxiter=4000, check_every=5, dinv=dinv)
    assert conv and convo
    assert [h[0] for h in hist] == [h[0] for h in histo]
    a = np.array([h[1] for h in hist])
    o = np.array([h[1] for h in histo])
    assert np.allclose(a[o > 1e-5], o[o > 1e-5], rtol=1e-3, atol=0)   # above the FP32 floor
    assert relerr(x.cpu().numpy(), xr) < 1e-5
    with pytest.raises(ValueError):
        dk.cg(op, b, body="fp16")


def test_cg_fp64_on_the_cpu_sized_2d_case(dk):
    """The CPU-runnable configuration's FP64 CG (N = 4096 in [0,1]^2, eps = 8,
    nugget 1e-8, seeds 42/43): kappa ~ 2e10. The first checks agree with the
    oracle's FP64 CG; then the reference's rule (stop when a check improves by
    less than 1%) ends both runs far from 1e-12 (relres ~ 1e-2). Where it
    stops is chaotic: perturbing A by 2e-16 moves the oracle's own stop from
    iteration 50 to 80."""
    N = 4096
    X = np.random.default_rng(42).uniform(0, 1, (N, 2))
    A = orb.build_kernel(X, "gaussian", 64.0, 1e-8)
    b = (np.sin(2 * np.pi * X[:, 0]) * np.cos(2 * np.pi * X[:, 1])
         + 1e-4 * np.random.default_rng(43).standard_normal(N))
    x, hist, conv = dk.cg(dk.DenseLinearOperator(A), b, tol=1e-12, body="fp64")
    xo, histo, convo = ok.cg(ok.DenseOp(A), b, tol=1e-12, body="fp64")
    assert not conv and not convo
    assert [h[0] for h in hist[:3]] == [0, 10, 20] == [h[0] for h in histo[:3]]
    assert np.allclose([h[1] for h in hist[:3]], [h[1] for h in histo[:3]], rtol=1e-6)
    assert 1e-3 < hist[-1][1] < 0.1 and 1e-3 < histo[-1][1] < 0.1
    r = b - A @ x.cpu().numpy()
    assert np.linalg.norm(r) / np.linalg.norm(b) == pytest.approx(hist[-1][1], rel=1e-6)


def _agree(hist, histo):
    """CG histories that start alike: the checks at iterations 0-20 within
    1e-5, the later ones within a factor of 4."""
    a = np.array([h[1] for h in hist])
    o = np.array([h[1] for h in histo])
    assert np.allclose(a[:3], o[:3], rtol=1e-5)
    assert np.all(np.abs(np.log(a[3:] / o[3:])) < np.log(4.0))
