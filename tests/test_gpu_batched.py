"""Batched uneven systems (mpdok_batched_rbf_solve) against the CPU oracle:
every right-hand side of every system to relres <= 1e-12, results independent
of how systems are spread over streams, and the tuner's choice."""
import numpy as np
import pytest
import torch

from oracle import lu as olu
from oracle import rbf as orb

pytestmark = pytest.mark.gpu

SIZES = [64, 1100, 100, 257, 513, 700, 1024, 300, 80, 512]


def _systems(seed=5, nrhs=3):
    """Smooth right-hand sides (random combinations of 8 Fourier modes), as in
    the batched configuration: with rough ones (white noise) the solutions of
    these kappa ~ 1e8 systems are so large that rounding x itself to FP64
    keeps relres above 1e-12 for any solver."""
    rng = np.random.default_rng(seed)
    a, c = rng.uniform(-3, 3, (8, 2)), rng.uniform(0, 2 * np.pi, 8)
    out = []
    for N in SIZES:
        X = rng.uniform(0, 1, (N, 2))
        eps = float(rng.uniform(2, 12))
        out.append((X, eps, np.cos(np.pi * X @ a.T + c) @ rng.standard_normal((8, nrhs))))
    return out


@pytest.mark.parametrize("precision", ["fp64", "ozaki"])
def test_every_system_reaches_1e12(precision):
    from mpdok_amd.batched import solve_batched
    systems = _systems()
    xs, info = solve_batched(systems, tol=1e-12, reg=1e-6, precision=precision, nstreams=4)
    assert info["converged"].shape == (len(SIZES), 3) and info["converged"].all()
    assert (info["relres"] <= 1e-12).all()
    from mpdok_amd import _ops
    for (X, eps, B), x in zip(systems, xs):
        # the matrix the solver was given (GPU assembly; its parity with the
        # oracle's assembly is tested on its own): a few ulps of assembly
        # rounding times ||x|| / ||b|| ~ 1e4 would otherwise dominate relres
        N = X.shape[0]
        Ad = _ops.rbf_assemble(torch.as_tensor(X).cuda(), 0, eps * eps, 1e-6,
                               torch.empty((N, N), dtype=torch.float64, device="cuda"))
        A = Ad.cpu().numpy()
        assert np.allclose(A, orb.build_kernel(X, "gaussian", eps * eps, 1e-6), rtol=1e-13)
        xh = x.cpu().numpy()
        assert xh.shape == B.shape
        for r in range(B.shape[1]):
            rel = np.linalg.norm(olu.residual_longdouble(A, xh[:, r], B[:, r])) / \
                np.linalg.norm(B[:, r])
            assert rel <= 1.5e-12


def test_results_do_not_depend_on_stream_count():
    from mpdok_amd.batched import solve_batched
    systems = _systems(seed=9, nrhs=2)
    x1, _ = solve_batched(systems, nstreams=1)
    x4, _ = solve_batched(systems, nstreams=4)
    x8, _ = solve_batched(systems, nstreams=8)
    for a, b, c in zip(x1, x4, x8):
        assert torch.equal(a, b) and torch.equal(a, c)


def test_c4_workload_shapes_and_tuner_choice():
    from mpdok_amd.batched import BatchTuner, c4_workload
    sysw = c4_workload(seed=1, nsys=6, nrhs=4, nmin=64, nmax=600)
    assert len(sysw) == 6
    for X, eps, B in sysw:
        assert 64 <= X.shape[0] <= 600 and X.shape[1] == 2 and B.shape == (X.shape[0], 4)
        assert 2.0 <= eps <= 12.0
    tuner = BatchTuner()
    choice = tuner.choose(sysw)
    assert choice["nstreams"] in (4, 8, 16) and choice["nb"] in (256, 512)
    assert tuner.choose(sysw) is choice            # cached per size profile
    xs, info = tuner.solve(sysw)
    assert info["converged"].all()
