"""Pins the CPU oracle: each check is chosen so that a plausible mistake (a
dropped term, a wrong sign or index, a transposed operand) fails it, and none
of them retypes the oracle's own formula."""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse.linalg as spla

from oracle import krylov as ok
from oracle import lu as olu
from oracle import rbf as orb


# ----------------------------------------------------------------- assembly
def test_gram_d2_matches_brute_force_distances():
    X = np.random.default_rng(0).uniform(-3, 5, (37, 3))
    brute = ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1)
    assert np.allclose(orb.gram_d2(X), brute, rtol=1e-12, atol=1e-12)
    # row/column sub-blocks index the right points
    D2 = orb.gram_d2(X, rows=np.arange(5, 9), cols=np.arange(30, 37))
    assert np.allclose(D2, brute[5:9, 30:37], rtol=1e-12, atol=1e-12)


def test_gaussian_closed_form_values_and_structure():
    X = np.array([[0.0, 0.0], [3.0, 4.0], [1.0, 0.0]])
    A = orb.build_kernel(X, "gaussian", 0.1, reg=1e-6)
    assert A[0, 1] == pytest.approx(np.exp(-0.1 * 25.0), rel=1e-15)
    assert A[0, 2] == pytest.approx(np.exp(-0.1), rel=1e-15)
    assert np.all(np.diag(A) == 1.0 + 1e-6)
    assert np.array_equal(A, A.T)


def test_matern32_closed_form():
    X = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.0, 0.3, 0.4]])
    ell = 0.1
    A = orb.build_kernel(X, "matern32", ell, reg=1e-4)
    for (i, j, r) in [(0, 1, 0.1), (0, 2, 0.5)]:
        a = np.sqrt(3.0) * r / ell
        assert A[i, j] == pytest.approx((1 + a) * np.exp(-a), rel=1e-13)
    assert A[0, 0] == pytest.approx(1.0 + 1e-4, rel=1e-15)
    # Matern-3/2 is decreasing in r and bounded by 1
    assert A[0, 1] > A[0, 2] > 0


def test_fp32_tile_adds_nugget_after_cast():
    X = np.random.default_rng(1).uniform(0, 1, (64, 2))
    reg = 1e-9  # below FP32 resolution of 1.0: vanishes when added after the cast
    t = orb.build_fp32_tile(X, 16, 8, "gaussian", 25.0, reg)
    assert t.dtype == np.float32 and t.shape == (8, 64)
    assert np.all(t[np.arange(8), 16 + np.arange(8)] == np.float32(1.0))
    full = orb.build_kernel(X, "gaussian", 25.0, 0.0)
    off = np.ones((8, 64), bool)
    off[np.arange(8), 16 + np.arange(8)] = False
    assert np.array_equal(t[off], full[16:24][off].astype(np.float32))


def test_synthetic_coords_bounds_moments_and_seeds():
    X = orb.synthetic_coords(200000, D=3, seed=42, scale=100.0)
    assert X.shape == (200000, 3) and X.dtype == np.float64
    assert X.min() >= 0.0 and X.max() < 100.0
    # uniform on [0, L]: mean L/2, variance L^2/12 (5 sigma of the sample error)
    assert np.allclose(X.mean(axis=0), 50.0, atol=5 * 100 / np.sqrt(12 * 200000))
    assert np.allclose(X.var(axis=0), 100.0 ** 2 / 12, rtol=0.01)
    assert np.array_equal(orb.synthetic_coords(50, seed=7), orb.synthetic_coords(50, seed=7))
    assert not np.array_equal(orb.synthetic_coords(50, seed=7), orb.synthetic_coords(50, seed=8))
    assert orb.synthetic_coords(10, D=2, scale=1.0).max() < 1.0


def test_weather_front_closed_form_points_bounds_and_noise():
    """rbf_kernel.py:230-252: 15 + 5 sin((x + y)/20 - t/2) + noise."""
    pi = np.pi
    # (x + y)/20 - t/2 = pi/2 -> 20, = 3 pi/2 -> 10, = 0 -> 15
    P = np.array([[5 * pi, 5 * pi], [15 * pi, 15 * pi], [0.0, 0.0], [7.0, -7.0],
                  [10.0 + 5 * pi, 10.0 + 5 * pi]])
    f0 = orb.weather_front(P, 0, noise_std=0.0)
    assert np.allclose(f0[:4], [20.0, 10.0, 15.0, 15.0], rtol=0, atol=1e-13)
    # t = 2 shifts the phase by -1: x + y = 20 + 10 pi peaks at t = 2
    assert orb.weather_front(P, 2, noise_std=0.0)[4] == pytest.approx(20.0, abs=1e-13)
    # symmetric in x and y; columns after the second are ignored
    X = orb.synthetic_coords(1000, D=3, seed=3)
    f = orb.weather_front(X, 4, noise_std=0.0)
    assert np.array_equal(f, orb.weather_front(X[:, [1, 0, 2]], 4, noise_std=0.0))
    assert np.array_equal(f, orb.weather_front(X[:, :2], 4, noise_std=0.0))
    assert f.min() >= 10.0 and f.max() <= 20.0
    # period 4 pi in t
    assert np.allclose(orb.weather_front(X, 4 + 4 * pi, noise_std=0.0), f, atol=1e-12)
    # noise: seed defaults to t, std noise_std, zero mean
    n = orb.weather_front(X, 4) - f
    assert np.array_equal(orb.weather_front(X, 4), orb.weather_front(X, 4, seed=4))
    assert not np.array_equal(orb.weather_front(X, 4), orb.weather_front(X, 5) - (
        orb.weather_front(X, 5, noise_std=0.0) - f))
    big = orb.synthetic_coords(100000, seed=1)
    nb = orb.weather_front(big, 1, noise_std=0.3) - orb.weather_front(big, 1, noise_std=0.0)
    assert np.std(nb) == pytest.approx(0.3, rel=0.02)
    assert abs(np.mean(nb)) < 5 * 0.3 / np.sqrt(100000)
    assert np.std(n) == pytest.approx(0.05, rel=0.1)


def test_estimate_gamma_known_answer_uniform_square():
    # E[mean d2] for uniform [0, L]^D with the diagonal zeros is D L^2 / 6 (1 - 1/n)
    vals = []
    for seed in range(8):
        X = np.random.default_rng(seed).uniform(0, 100.0, (2000, 2))
        vals.append(orb.estimate_gamma(X))
    g = orb.uniform_cube_gamma(2, 100.0, 512)
    assert np.mean(vals) == pytest.approx(g, rel=0.02)
    # the reference printed gamma = 1.505e-4 for this setting (rbf_demo.ipynb:103)
    assert g == pytest.approx(1.505e-4, rel=0.01)
    # degenerate input -> 1.0
    assert orb.estimate_gamma(np.zeros((10, 2))) == 1.0


def test_matvec_fp64_tiled_equals_dense():
    X = np.random.default_rng(2).uniform(0, 1, (300, 3))
    v = np.random.default_rng(3).standard_normal(300)
    A = orb.build_kernel(X, "gaussian", 25.0, 1e-6)
    assert np.allclose(orb.matvec_fp64(X, v, "gaussian", 25.0, 1e-6, tile_rows=64), A @ v,
                       rtol=1e-13, atol=1e-13)


# ------------------------------------------------------------------- Krylov
def _spd(N, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((N, N))
    return X.T @ X + N * np.eye(N), rng.standard_normal(N)


def test_cg_matches_direct_solve_and_history_rules():
    A, b = _spd(200, 7)
    x, hist, conv = ok.cg(ok.DenseOp(A), b, tol=1e-6)
    assert conv
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-5
    # checks fire at multiples of check_every, before the update
    assert [h[0] % 10 for h in hist] == [0] * len(hist)
    assert hist[0] == (0, pytest.approx(1.0))


def test_cg_identity_one_step_and_zero_rhs():
    N = 50
    b = np.random.default_rng(0).standard_normal(N)
    x, hist, conv = ok.cg(ok.DenseOp(np.eye(N)), b, tol=1e-6, check_every=1)
    assert conv and hist[-1][0] == 1
    assert np.allclose(x, b, rtol=1e-6)
    x, hist, conv = ok.cg(ok.DenseOp(np.eye(N)), np.zeros(N))
    assert conv and hist == [(0, 0.0)] and not x.any()


def test_pcg_jacobi_solves_a_diagonal_system_in_one_iteration():
    d = np.linspace(1.0, 50.0, 40)
    b = np.random.default_rng(1).standard_normal(40)
    e, it = ok.pcg(lambda v: d * v, b, dinv=1.0 / d, tol=1e-14, maxiter=10, check_every=1,
                   dtype=np.float64)
    assert it == 1
    assert np.allclose(e, b / d, rtol=1e-14, atol=0)
    # without the preconditioner the 40 distinct eigenvalues need many steps
    _, it2 = ok.pcg(lambda v: d * v, b, tol=1e-14, maxiter=100, check_every=1, dtype=np.float64)
    assert it2 > 10


def test_pcg_three_distinct_eigenvalues_and_scipy_agreement():
    rng = np.random.default_rng(2)
    Q, _ = np.linalg.qr(rng.standard_normal((30, 30)))
    A = Q @ np.diag(np.repeat([1.0, 4.0, 9.0], 10)) @ Q.T
    b = rng.standard_normal(30)
    e, it = ok.pcg(lambda v: A @ v, b, tol=1e-12, maxiter=30, check_every=1, dtype=np.float64)
    assert it == 3
    assert np.allclose(A @ e, b, atol=1e-10)
    # Jacobi-preconditioned, against SciPy's cg with M = diag(A)^-1
    X = rng.uniform(0, 1, (300, 3))
    K = orb.build_kernel(X, "matern32", 0.1, 1e-2) * np.linspace(1.0, 3.0, 300)[:, None]
    K = (K + K.T) / 2
    dinv = 1.0 / np.diag(K)
    bb = rng.standard_normal(300)
    e, _ = ok.pcg(lambda v: K @ v, bb, dinv=dinv, tol=1e-12, maxiter=2000, check_every=1,
                  dtype=np.float64)
    M = spla.LinearOperator((300, 300), matvec=lambda v: dinv * v)
    es, info = spla.cg(K, bb, M=M, rtol=1e-12, maxiter=2000)
    assert info == 0
    ref = np.linalg.solve(K, bb)
    assert np.linalg.norm(e - ref) / np.linalg.norm(ref) < 1e-8
    assert np.linalg.norm(es - ref) / np.linalg.norm(ref) < 1e-8


def test_cg_fp64_body_and_jacobi_match_scipy():
    rng = np.random.default_rng(3)
    X = rng.uniform(0, 1, (400, 2))
    A = orb.build_kernel(X, "gaussian", 4.0, 1e-1)
    rng.standard_normal(400)
    b = rng.standard_normal(400)
    op = ok.DenseOp(A)
    ref = np.linalg.solve(A, b)
    x, hist, conv = ok.cg(op, b, tol=1e-12, maxiter=4000, body="fp64")
    assert conv and hist[-1][1] < 1e-12
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-9
    # the FP32 body cannot reach 1e-12: the FP64 checks see it stagnate
    _, hist32, conv32 = ok.cg(op, b, tol=1e-12, maxiter=4000)
    assert not conv32 and hist32[-1][1] > 1e-9
    # a diagonal system is solved by one Jacobi-preconditioned step
    d = np.linspace(1.0, 9.0, 50)
    opd = ok.DenseOp(np.diag(d))
    bd = rng.standard_normal(50)
    xd, hd, cd = ok.cg(opd, bd, tol=1e-13, body="fp64", dinv=1.0 / d, check_every=1)
    assert cd and hd[-1][0] == 1
    assert np.allclose(xd, bd / d, rtol=1e-14, atol=0)


def test_cg_ir_reaches_1e10_with_fp32_inner_cg():
    X = np.random.default_rng(4).uniform(0, 1, (600, 3))
    A = orb.build_kernel(X, "matern32", 0.1, 1e-4)
    A32 = orb.build_fp32_tile(X, 0, 600, "matern32", 0.1, 1e-4)
    b = np.sin(2 * np.pi * X[:, 0])
    dinv = (np.float32(1.0) / np.diag(A32)).astype(np.float32)
    x, conv, hist, inner = ok.cg_ir(lambda v: A32 @ v.astype(np.float32), lambda v: A @ v, b,
                                    dinv=dinv, tol=1e-10)
    assert conv and hist[-1] < 1e-10 and len(hist) >= 2
    assert all(h1 < h0 for h0, h1 in zip(hist, hist[1:]))
    ref = np.linalg.solve(A, b)
    kappa = np.linalg.cond(A)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < kappa * 1e-10


def test_gmres_three_distinct_eigenvalues_solve_in_three_steps():
    # The Krylov space of A with 3 distinct eigenvalues has dimension 3: one
    # restart cycle of 3 steps reaches the FP32 floor.
    N = 60
    Q, _ = np.linalg.qr(np.random.default_rng(4).standard_normal((N, N)))
    lam = np.repeat([1.0, 2.0, 5.0], N // 3)
    A = (Q * lam) @ Q.T
    b = np.random.default_rng(5).standard_normal(N)
    x, hist, conv = ok.gmres(ok.DenseOp(A), b, tol=1e-5, restart=3)
    assert conv and hist[0][0] == 3
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-5
    # with only 2 steps it cannot
    x, hist, conv = ok.gmres(ok.DenseOp(A), b, tol=1e-5, restart=2, maxiter=2)
    assert not conv and hist[0][1] > 1e-3


def test_gmres_exact_breakdown_on_eigenvector():
    # b along an eigenvector: A v0 = lambda v0 exactly in FP32, w - h v0 = 0,
    # lucky breakdown after one matvec and the exact solution b / lambda.
    N = 40
    A = np.diag(np.arange(1.0, N + 1.0))
    b = np.zeros(N)
    b[3] = 8.0
    x, hist, conv = ok.gmres(ok.DenseOp(A), b, tol=1e-6, restart=10)
    assert conv and hist[0][0] == 1
    assert x[3] == pytest.approx(2.0, rel=1e-7) and np.count_nonzero(x) == 1


def test_gmres_nonsymmetric_matches_scipy():
    N = 300
    rng = np.random.default_rng(13)
    A = rng.standard_normal((N, N)) + N * np.eye(N)
    b = rng.standard_normal(N)
    x, hist, conv = ok.gmres(ok.DenseOp(A), b, tol=1e-6, restart=50)
    xs, info = spla.gmres(A, b, rtol=1e-6, restart=50, maxiter=N)
    assert conv and info == 0
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-5
    assert np.linalg.norm(xs - xr) / np.linalg.norm(xr) < 1e-5


def test_gmres_ir_reaches_tol_and_keeps_trailing_entry():
    N = 300
    rng = np.random.default_rng(42)
    A = rng.standard_normal((N, N)) + N * np.eye(N)
    b = rng.standard_normal(N)
    x, hist, conv = ok.gmres_ir(ok.DenseOp(A), b, tol=1e-11, maxiter_outer=5)
    assert conv
    assert hist[-1][0] == 5          # trailing (maxiter_outer, rel) entry
    assert hist[-2][1] < 1e-11       # the converged check before it
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-12


def test_inner_gmres_and_ooc_solve():
    X = np.random.default_rng(9).uniform(0, 1, (400, 2))
    A = orb.build_kernel(X, "gaussian", 4.0, 1e-2)
    A32 = A.astype(np.float32)
    b = np.random.default_rng(10).standard_normal(400)
    e = ok.inner_gmres(lambda v: A32 @ v, b, restart=50)
    assert np.linalg.norm(b - A @ e) / np.linalg.norm(b) < 1e-2  # one cycle, kappa ~ 3e4
    x, conv = ok.ooc_solve(lambda v: A32 @ v, lambda v: A @ v, b, tol=1e-10)
    assert conv
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-10


def test_inner_gmres_breakdown_is_exact():
    # identity: breakdown at k = 0, e equals the right-hand side
    b = np.random.default_rng(0).standard_normal(30)
    e = ok.inner_gmres(lambda v: v.copy(), b, restart=10)
    assert np.allclose(e, b, rtol=1e-6)
    # 2 distinct eigenvalues: breakdown at k = 1 must still return the solution
    Q, _ = np.linalg.qr(np.random.default_rng(1).standard_normal((30, 30)))
    A = (Q * np.repeat([1.0, 3.0], 15)) @ Q.T
    e = ok.inner_gmres(lambda v: (A @ v).astype(np.float32), b, restart=10)
    assert np.linalg.norm(b - A @ e) / np.linalg.norm(b) < 1e-5


# ----------------------------------------------------------------------- LU
def test_lu_ir_fp32_factor_reaches_fp64_accuracy():
    N = 256
    rng = np.random.default_rng(42)
    A = rng.standard_normal((N, N)) + N * np.eye(N)
    b = rng.standard_normal(N)
    x, conv, hist = olu.lu_ir(A, b, tol=1e-11, maxiter_outer=3)
    assert conv and hist[-1] < 1e-11
    assert olu.backward_error(A, x, b) < 2e-14  # a few units of FP64 roundoff


def test_lu_nopiv_reconstructs_and_matches_lapack_on_spd():
    A, b = _spd(150, 3)
    LU = olu.lu_nopiv(A, nb=32)
    L, U = olu.unpack(LU)
    assert np.allclose(L @ U, A, rtol=1e-12, atol=1e-10)
    lu, piv = sla.lu_factor(A)
    assert np.all(piv == np.arange(150))  # no pivoting happens on this matrix
    assert np.allclose(LU, lu, rtol=1e-10, atol=1e-10)
    x = olu.lu_solve_packed(LU, b)
    assert np.allclose(x, sla.solve(A, b), rtol=1e-12)


def test_lu_nopiv_block_starts_match_lapack_and_uniform_blocking():
    """Blocks of several widths (starts) give LAPACK's unpivoted factors like
    uniform blocks do; starts equal to the uniform ones give the same bits;
    and the trailing-update callable sees each block's K = its width."""
    A, _ = _spd(150, 3)
    lu, piv = sla.lu_factor(A)
    assert np.all(piv == np.arange(150))
    starts = [0, 64, 128, 136]
    LU = olu.lu_nopiv(A, starts=starts)
    assert np.allclose(LU, lu, rtol=1e-10, atol=1e-10)
    assert np.array_equal(olu.lu_nopiv(A, nb=32, starts=range(0, 150, 32)), olu.lu_nopiv(A, nb=32))
    widths = []
    olu.lu_nopiv(A, starts=starts, trailing=lambda L, U: widths.append(L.shape[1]) or L @ U)
    assert widths == [64, 64, 8]


def test_lu_solve_packed_with_pivots_matches_lapack():
    rng = np.random.default_rng(11)
    A = rng.standard_normal((80, 80))
    b = rng.standard_normal(80)
    lu, piv = sla.lu_factor(A)
    assert not np.all(piv == np.arange(80))
    assert np.allclose(olu.lu_solve_packed(lu, b, piv), np.linalg.solve(A, b), rtol=1e-10)


# -------------------------------------------------------------------- Ozaki
def test_ozaki_slices_represent_input_to_truncation_bound():
    M = np.random.default_rng(0).standard_normal((20, 30)) * np.logspace(-3, 3, 30)
    for s in (2, 4, 7):
        d, e = olu.ozaki_slices(M, s, axis=1)
        P = 8 * s - 1
        # integer reconstruction: V = sum_q d_q 256^(s - q) exactly (Python ints)
        V = np.zeros(M.shape, dtype=object)
        for di in d:
            V = V * 256 + di.astype(np.int64).astype(object)
        rec = np.array([[float(np.ldexp(float(v), int(ee) - P)) for v in row]
                        for row, ee in zip(V, e)])
        # round-to-nearest of the scaled element: half a unit of the last digit
        bound = np.ldexp(1.0, e[:, None] - P - 1)
        assert np.all(np.abs(rec - M) <= bound)
        assert all(di.dtype == np.int8 for di in d)
        assert np.abs(d[0].astype(int)).max() <= 127
        # the scale is the smallest power of two that keeps the leading digit in range
        assert np.all(np.abs(d[0].astype(int)).max(axis=1) >= 63)


def test_ozaki_slices_digit_ranges_at_the_int8_limits():
    # a row maximum whose leading 8 bits round to 128/128 bumps the exponent
    # (no digit may reach +128); low digits use the full [-128, 127] range
    M = np.array([[1.0 - 2.0 ** -9, 0.25, -0.5],
                  [1.0 - 2.0 ** -12, -(1.0 - 2.0 ** -12), 0.0],
                  [0.5 + 2.0 ** -9, -2.0 ** -9, 2.0 ** -10]])
    d, e = olu.ozaki_slices(M, 3, axis=1)
    assert e[0] == 1 and e[1] == 1  # 127.75 and 127.97 exceed 127
    assert e[2] == 0
    allv = np.concatenate([di.ravel().astype(int) for di in d[1:]])
    M2 = np.random.default_rng(5).uniform(-1, 1, (200, 64))
    d2, _ = olu.ozaki_slices(M2, 4, axis=1)
    low = np.concatenate([di.ravel().astype(int) for di in d2[1:]])
    assert low.min() == -128 and low.max() == 127
    assert np.abs(d2[0].astype(int)).max() <= 127
    # a matrix of zeros slices to zeros with exponent 0
    dz, ez = olu.ozaki_slices(np.zeros((3, 5)), 3, axis=0)
    assert np.all(ez == olu.ZERO_EXP) and all(np.all(di == 0) for di in dz)
    assert allv.size > 0


def test_ozaki_matmul_exact_on_small_integers_and_error_scaling():
    rng = np.random.default_rng(1)
    A = rng.integers(-100, 100, (17, 40)).astype(np.float64)
    B = rng.integers(-100, 100, (40, 23)).astype(np.float64)
    # integers below 2^7 in magnitude fit the 7-bit leading digit after the
    # power-of-two scaling, so one slice (s = 1) already gives the exact product;
    # s = 3 keeps it (the extra digits are zero)
    assert np.array_equal(olu.ozaki_matmul(A, B, 1), A @ B)
    assert np.array_equal(olu.ozaki_matmul(A, B, 3), A @ B)
    A = rng.standard_normal((64, 128))
    B = rng.standard_normal((128, 48))
    ref = A @ B
    errs = [np.abs(olu.ozaki_matmul(A, B, s) - ref).max() for s in (2, 3, 4, 5)]
    for e1, e2 in zip(errs, errs[1:]):
        assert e2 < e1 * 2.0 ** -6  # each slice buys ~8 bits
    # against the exact rational product: error <= (few) * 2^(-(8 s - 1)) * K * max|a| max|b|
    scale = np.abs(A).max() * np.abs(B).max() * 128
    for s, err in zip((2, 3, 4, 5), errs):
        assert err < 4.0 * 2.0 ** -(8 * s - 1) * scale
    assert errs[-1] < 1e-9


def test_ozaki_matmul_drops_exactly_the_low_pairs():
    # one row / one column with known digits: a = 2^e (d1 2^-7 + d2 2^-15),
    # b likewise; s = 2 keeps d1*e1 (level 2) and d1*e2 + d2*e1 (level 3),
    # drops d2*e2 (level 4)
    a = np.array([[ (100 * 256 + 7) * 2.0 ** -15 ]])
    b = np.array([[ (-90 * 256 - 100) * 2.0 ** -15 ]])
    got = olu.ozaki_matmul(a, b, 2)[0, 0]
    want = (100 * -90 * 2.0 ** -14 + (100 * -100 + 7 * -90) * 2.0 ** -22)
    assert got == want
    assert olu.ozaki_matmul(a, b, 3)[0, 0] == a[0, 0] * b[0, 0]


def test_ozaki_decide_rule_on_known_exponents():
    # A's rows peak at 1.0 (exponent 1), B's columns at 0.75 (exponent 0):
    # E = 1. Scaling B by 2^-k lowers E by k, and the rule drops one slice per
    # whole 8 bits below the first update's E (ceil), within [smin, smax].
    rng = np.random.default_rng(3)
    A = rng.uniform(-1, 1, (40, 16))
    A[:, 0] = 1.0
    B = rng.uniform(-0.5, 0.5, (16, 30))
    B[0, :] = 0.75
    assert olu.ozaki_scales(A, 1).max() == 1 and olu.ozaki_scales(B, 0).max() == 0
    S0, e0 = olu.ozaki_decide(A, B, 5, 2, 0, None)
    assert (S0, e0) == (5, 1)
    for k, want in [(0, 5), (1, 5), (7, 5), (8, 4), (9, 4), (16, 3), (17, 3), (24, 2), (40, 2)]:
        S, _ = olu.ozaki_decide(A, np.ldexp(B, -k), 5, 2, 0, e0)
        assert S == want, (k, S, want)
    # growth above the first update never exceeds smax; delta tightens by bits
    assert olu.ozaki_decide(A, B * 64, 5, 2, 0, e0)[0] == 5
    assert olu.ozaki_decide(A, np.ldexp(B, -8), 5, 2, 4, e0)[0] == 5   # -8 + 4 -> ceil(-0.5) = 0
    assert olu.ozaki_decide(A, np.ldexp(B, -16), 5, 3, 0, e0)[0] == 3
    assert olu.ozaki_decide(A, np.ldexp(B, -40), 5, 3, 0, e0)[0] == 3  # clamped at smin
    # all-zero operands carry ZERO_EXP: the smallest count
    assert olu.ozaki_decide(np.zeros((4, 16)), B, 5, 2, 0, e0)[0] == 2


def test_ozaki_leading_digits_truncate_like_rounding():
    # the leading s of 5 digits err by at most about half a unit of digit s,
    # like an s-digit split; the product error scales the same way
    rng = np.random.default_rng(4)
    A = rng.standard_normal((48, 96))
    B = rng.standard_normal((96, 40))
    ref = A @ B
    for s in (2, 3, 4):
        e_lead = np.abs(olu.ozaki_matmul_lead(A, B, 5, s) - ref).max()
        e_own = np.abs(olu.ozaki_matmul(A, B, s) - ref).max()
        assert e_lead < 4 * e_own and e_own < 4 * e_lead
        scale = np.abs(A).max() * np.abs(B).max() * 128
        assert e_lead < 8.0 * 2.0 ** -(8 * s - 1) * scale
    assert np.array_equal(olu.ozaki_matmul_lead(A, B, 5, 5), olu.ozaki_matmul(A, B, 5))


def test_ozaki_adaptive_lu_keeps_the_first_updates_error():
    # decaying Schur complements of an RBF matrix: later updates take fewer
    # slices, and the factors still precondition LU-IR to the FP64 floor
    X = np.random.default_rng(42).uniform(0, 1, (512, 2))
    A = orb.build_kernel(X, "gaussian", 25.0, 1e-6)
    tr = olu.OzakiAdaptiveTrailing(5)
    LU = olu.lu_nopiv(A, 64, trailing=tr)
    assert tr.counts == [5, 5, 5, 4, 4, 3, 3]
    Lm, Um = olu.unpack(LU)
    LU5 = olu.lu_nopiv(A, 64, trailing=lambda L, U: olu.ozaki_matmul(L, U, 5))
    L5, U5 = olu.unpack(LU5)
    be = np.linalg.norm(Lm @ Um - A) / np.linalg.norm(A)
    be5 = np.linalg.norm(L5 @ U5 - A) / np.linalg.norm(A)
    assert be < 4 * be5 + 1e-15
    b = np.sin(2 * np.pi * X[:, 0])
    x = np.zeros(512)
    for _ in range(4):
        x += olu.lu_solve_packed(LU, b - A @ x)
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-12


def test_ozaki_lu_converges_in_lu_ir():
    X = np.random.default_rng(42).uniform(0, 1, (512, 3))
    A = orb.build_kernel(X, "gaussian", 25.0, 1e-6)
    b = np.sin(2 * np.pi * X[:, 0])
    LU = olu.lu_nopiv(A, 64, trailing=lambda L, U: olu.ozaki_matmul(L, U, 6))
    LUf = olu.lu_nopiv(A, 64)
    assert np.abs(LU - LUf).max() < 1e-6 * np.abs(LUf).max()
    x = np.zeros(512)
    for _ in range(4):
        x += olu.lu_solve_packed(LU, b - A @ x)
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-12


# ------------------------------------------------------------ low precision
def test_bf16_and_e4m3_rounding_match_torch_and_fix_every_code():
    import torch
    rng = np.random.default_rng(9)
    x = (rng.standard_normal(20000) * np.exp(rng.uniform(-20, 20, 20000))).astype(np.float32)
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(olu.round_bf16(x), ref)
    # E4M3 within range (torch's float8_e4m3fn rounds to nearest even and has
    # no saturation, so compare inside +-448)
    y = rng.uniform(-448, 448, 20000) * np.exp(rng.uniform(-14, 0, 20000))
    y = y.astype(np.float32).astype(np.float64)
    ref8 = torch.tensor(y, dtype=torch.float32).to(torch.float8_e4m3fn).to(torch.float64).numpy()
    assert np.array_equal(olu.round_e4m3(y), ref8)
    # every finite E4M3 code is a fixed point; 448 is the largest, larger
    # values saturate
    codes = torch.arange(256, dtype=torch.uint8).view(torch.float8_e4m3fn).to(torch.float64).numpy()
    codes = codes[np.isfinite(codes)]
    assert np.array_equal(olu.round_e4m3(codes), codes)
    assert np.max(codes) == 448.0 and olu.round_e4m3(np.array([1e6]))[0] == 448.0
    assert olu.round_e4m3(np.array([2.0 ** -10]))[0] == 0.0          # a tie: to even, 0
    assert olu.round_e4m3(np.array([3 * 2.0 ** -10]))[0] == 2.0 ** -8
    # scale: max |M| times it lands in (224, 448]
    M = rng.standard_normal((30, 20)) * 1e-5
    sc = olu.e4m3_scale(M)
    assert 224 < np.max(np.abs(M)) * sc <= 448 and np.log2(sc) == int(np.log2(sc))


def test_lowp_matmul_error_shrinks_with_the_format():
    rng = np.random.default_rng(10)
    A = rng.standard_normal((64, 200))
    B = rng.standard_normal((200, 48))
    exact = A.astype(np.float32).astype(np.float64) @ B.astype(np.float32).astype(np.float64)
    errs = {k: np.abs(olu.lowp_matmul(A, B, k) - exact).max() / np.abs(exact).max()
            for k in ("fp8", "bf16", "fp16")}
    assert errs["fp16"] < errs["bf16"] < errs["fp8"] < 0.2
    assert 2.0 ** -14 < errs["fp16"] < 2.0 ** -7 and 2.0 ** -11 < errs["bf16"] < 2.0 ** -4


# ---------------------------------------------------------------- residuals
def test_residual_exact_catches_cancellation_and_longdouble_agrees():
    # b - A x with massive cancellation: exact answer known by construction
    A = np.array([[1e16, 1.0, -1e16], [3.0, 1e-8, 2.0]])
    x = np.array([1.0, 1.0, 1.0])
    b = np.array([1.0, 5.0])
    r = olu.residual_exact(A, x, b)
    assert r[0] == 0.0  # 1 - (1e16 + 1 - 1e16) = 0 exactly
    assert r[1] == pytest.approx(-1e-8, rel=1e-15)
    # FP64 summed left to right loses it: (1e16 + 1) rounds to 1e16
    naive0 = b[0] - ((A[0, 0] * x[0] + A[0, 1] * x[1]) + A[0, 2] * x[2])
    assert naive0 == 1.0
    rl = olu.residual_longdouble(A, x, b)
    assert rl[0] == 0.0  # a 64-bit significand holds 1e16 + 1 exactly
    assert abs(rl[1] - r[1]) < 1e-18


def test_forward_and_backward_error_closed_forms():
    x_ref = np.array([3.0, -4.0, 12.0])            # ||x_ref|| = 13
    assert olu.forward_error(x_ref, x_ref) == 0.0
    assert olu.forward_error(2.0 * x_ref, x_ref) == pytest.approx(1.0, rel=1e-15)
    assert olu.forward_error(x_ref + np.array([0.0, 0.0, 1.3]), x_ref) == pytest.approx(
        0.1, rel=1e-14)
    # backward error: exact solution -> 0; x = 0 -> ||b|| / ||b|| = 1
    A = np.array([[2.0, 0.0], [0.0, 3.0]])
    b = np.array([2.0, 3.0])
    assert olu.backward_error(A, np.array([1.0, 1.0]), b) == 0.0
    assert olu.backward_error(A, np.zeros(2), b) == pytest.approx(1.0, rel=1e-15)
    # a perturbation r = b - A x of norm 1 with ||A||_F = sqrt(13), ||x|| = sqrt(2)
    x = np.array([1.0, 1.0 + 1.0 / 3.0])
    assert olu.backward_error(A, x, b) == pytest.approx(
        1.0 / (np.sqrt(13.0) * np.linalg.norm(x) + np.sqrt(13.0)), rel=1e-14)
