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


# ---------------------------------------------------------------------------
# Grouped launches captured in one hipGraph (BatchedPlan)
# ---------------------------------------------------------------------------
def _assembled(X, eps):
    from mpdok_amd import _ops
    N = X.shape[0]
    Ad = _ops.rbf_assemble(torch.as_tensor(X).cuda(), 0, eps * eps, 1e-6,
                           torch.empty((N, N), dtype=torch.float64, device="cuda"))
    return Ad.cpu().numpy()


def test_graph_plan_every_system_reaches_1e12_with_lapack_forward_error():
    """Each right-hand side of each system (N not a multiple of 64 included:
    64, 80, 100, 257, ...) to long-double relres <= 1.5e-12, with a forward
    error within 10x of LAPACK's FP64 solve (both against LAPACK plus one
    long-double refinement step)."""
    import scipy.linalg as sla
    from mpdok_amd.batched import BatchedPlan
    systems = _systems(seed=5, nrhs=3)
    plan = BatchedPlan(systems, tol=1e-12, reg=1e-6, maxiter_outer=4)
    xs = plan.run()
    res = plan.results()
    assert res["converged"].all() and (res["info"] == 0).all()
    assert (res["relres"] < 1e-12).all()
    # history: step 0 is ||b|| / ||b|| = 1; a column stops once below tol
    assert np.all(res["hist"][0] == 1.0)
    assert (res["steps"] >= 1).all() and (res["steps"] <= 4).all()
    for (X, eps, B), x in zip(systems, xs):
        A = _assembled(X, eps)
        xh = x.cpu().numpy()
        for r in range(B.shape[1]):
            b = B[:, r]
            rel = np.linalg.norm(olu.residual_longdouble(A, xh[:, r], b)) / np.linalg.norm(b)
            assert rel <= 1.5e-12
            x64 = sla.solve(A, b)
            xl = x64 + sla.solve(A, olu.residual_longdouble(A, x64, b))
            e64 = np.linalg.norm(x64 - xl) / np.linalg.norm(xl)
            assert np.linalg.norm(xh[:, r] - xl) / np.linalg.norm(xl) <= 10 * max(e64, 1e-15)
    plan.close()


def test_graph_replays_equal_the_eager_launches_bit_for_bit():
    """The captured graph replayed three times, with other GPU work and a
    host synchronisation between replays, gives the eager sequence's
    solution bit for bit every time; a new right-hand side written into the
    plan's buffer is picked up by the next replay."""
    from mpdok_amd.batched import BatchedPlan
    systems = _systems(seed=9, nrhs=2)
    eager = BatchedPlan(systems, use_graph=False)
    ref = [x.clone() for x in eager.run()]
    graph = BatchedPlan(systems, use_graph=True)
    for rep in range(3):
        xs = graph.run()
        for a, b in zip(ref, xs):
            assert torch.equal(a, b), rep
        assert graph.results()["converged"].all()
        junk = torch.randn(1 << 20, device="cuda")
        (junk * 2).sum().item()
    # a changed right-hand side for system 3, in both plans
    B3 = systems[3][2][:, ::-1].copy()
    eager.set_rhs(3, B3)
    graph.set_rhs(3, B3)
    e3 = eager.run()[3].clone()
    g3 = graph.run()[3]
    assert torch.equal(e3, g3)
    assert not torch.equal(e3, ref[3])
    assert graph.results()["converged"].all()
    eager.close()
    graph.close()


@pytest.mark.parametrize("update_rows", [0, 1 << 30])
def test_graph_plan_tile_choice_does_not_change_convergence(update_rows):
    """All trailing updates on 128 x 128 tiles (0) or all on 64 x 64 tiles:
    the same tile body, so the same factors and solutions bit for bit."""
    from mpdok_amd.batched import BatchedPlan
    systems = _systems(seed=11, nrhs=1)
    a = BatchedPlan(systems, update_rows=update_rows)
    b = BatchedPlan(systems, update_rows=512)
    xa, xb = a.run(), b.run()
    assert a.results()["converged"].all()
    for u, v in zip(xa, xb):
        assert torch.equal(u, v)
    a.close()
    b.close()


def test_tuner_picks_a_plan_threshold_and_solves_with_it():
    """The graph path's tuner times each candidate threshold on the sample,
    caches the choice per size profile, and the tuned plan reaches 1e-12 with
    the same solutions as a plan built with that threshold by hand."""
    from mpdok_amd.batched import BatchedPlan, BatchTuner
    systems = _systems(seed=17, nrhs=2)
    tuner = BatchTuner()
    choice = tuner.choose_plan(systems)
    assert choice["update_rows"] in BatchTuner.PLAN_ROWS
    assert tuner.choose_plan(systems) is choice
    assert sum(t.get("path") == "graph" for t in tuner.trials) == len(BatchTuner.PLAN_ROWS)
    xs, info = tuner.solve_plan(systems)
    assert info["converged"].all() and (info["relres"] < 1e-12).all()
    assert info["config"] == choice and info["plan"]["launches"] > 0
    ref = BatchedPlan(systems, update_rows=choice["update_rows"])
    for a, b in zip(xs, ref.run()):
        assert torch.equal(a, b)
    ref.close()


def test_graph_plan_matches_the_stream_path_and_reports_zero_pivots():
    """Against the host-threaded path (FP64 LU, 512-wide panels): both reach
    1e-12 and agree to the accuracy the conditioning allows; a matrix with a
    zero leading entry stops with info = 1 and converged = 0 without touching
    the other systems."""
    from mpdok_amd.batched import BatchedPlan, solve_batched
    systems = _systems(seed=13, nrhs=2)
    xs_s, _ = solve_batched(systems, precision="fp64", nstreams=4)
    plan = BatchedPlan(systems)
    xs_g = plan.run()
    for u, v in zip(xs_s, xs_g):
        # ||x1 - x2|| <= ||inv(K)|| (||r1|| + ||r2||) <= 1e6 * 2e-12 ||b||
        assert float(torch.linalg.vector_norm(u - v) / torch.linalg.vector_norm(u)) < 1e-5
    plan.close()
    # a zero pivot: all points of system 0 at one location and no nugget, so
    # K = ones (rank one) and the second pivot is exactly 1 - 1 = 0
    X0 = np.zeros((70, 2))
    bad = [(X0, 3.0, np.ones((70, 2)))] + systems[1:4]
    plan = BatchedPlan(bad, reg=0.0)
    plan.run()
    res = plan.results()
    assert res["info"][0] == 2 and not res["converged"][0].any()
    plan.close()
