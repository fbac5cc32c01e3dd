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
