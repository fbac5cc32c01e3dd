"""GPU parity of the LU factorisations, the int8-sliced trailing update and
the refinement drivers against the oracle and LAPACK."""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sla
import torch

from oracle import lu as olu
from oracle import rbf as orb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from mpdok_amd import _lib
    return _lib.load()


def dev(a, dtype=torch.float64, order="F"):
    from mpdok_amd import _dev
    return _dev.as_device(a, dtype, order)


def stream():
    from mpdok_amd import _dev
    return _dev.stream()


def gemm(L, A, B, C, alpha, beta, round_=0):
    m, k = A.shape
    n = B.shape[1]
    dt = 0 if A.dtype == torch.float64 else 1
    st = L.mpdok_gemm(stream(), dt, m, n, k, alpha, A.data_ptr(), A.stride(1), B.data_ptr(),
                      B.stride(1), beta, C.data_ptr(), C.stride(1), round_)
    assert st == 0, L.mpdok_last_error()
    return C


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 33, 5), (128, 128, 16), (300, 257, 129),
                                   (1000, 64, 64), (64, 1000, 512)])
@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_gemm_bit_exact_integer_inputs(L, m, n, k, dt):
    rng = np.random.default_rng(m + 7 * n + 13 * k)
    A = rng.integers(-8, 9, (m, k)).astype(np.float64)
    B = rng.integers(-8, 9, (k, n)).astype(np.float64)
    C = rng.integers(-100, 100, (m, n)).astype(np.float64)
    out = gemm(L, dev(A, dt), dev(B, dt), dev(C, dt), -1.0, 1.0).cpu().numpy()
    assert np.array_equal(out, C - A @ B)
    out = gemm(L, dev(A, dt), dev(B, dt), dev(np.full((m, n), np.nan), dt), 2.0, 0.0).cpu().numpy()
    assert np.array_equal(out, 2 * (A @ B))   # beta = 0 never reads C


def test_gemm_fp64_random_close(L):
    rng = np.random.default_rng(3)
    A, B = rng.standard_normal((513, 200)), rng.standard_normal((200, 301))
    out = gemm(L, dev(A), dev(B), dev(np.zeros((513, 301))), 1.0, 0.0).cpu().numpy()
    assert np.allclose(out, A @ B, rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("round_,kind", [(1, "bf16"), (2, "fp16"), (3, "fp8")])
@pytest.mark.parametrize("m,n,k", [(257, 300, 129), (256, 384, 512), (64, 64, 16)])
def test_gemm_low_precision_matrix_cores_match_oracle(L, round_, kind, m, n, k):
    """BF16 / FP16 operands on their matrix cores, and scaled E4M3 values on
    the BF16 ones: the products of the rounded operands (oracle lowp_matmul,
    exact) to within FP32 accumulation, C = -A B + C as in the LU."""
    rng = np.random.default_rng(m + n + round_)
    A = (rng.standard_normal((m, k)) * 3e-3).astype(np.float32)
    B = (rng.standard_normal((k, n)) * 0.7).astype(np.float32)
    C = rng.standard_normal((m, n)).astype(np.float32)
    out = gemm(L, dev(A, torch.float32), dev(B, torch.float32), dev(C, torch.float32), -1.0, 1.0,
               round_).cpu().numpy()
    ref = C.astype(np.float64) - olu.lowp_matmul(A, B, kind)
    bound = 2.0 ** -23 * (k + 2) * (np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
                                     + np.abs(C))
    assert np.all(np.abs(out - ref) <= 4 * bound + 1e-30)
    # and it is the low precision, not FP32 products
    assert not np.allclose(out, C - A.astype(np.float64) @ B, rtol=1e-6, atol=1e-7)


def test_gemm_bf16_rounding_matches_emulation(L):
    rng = np.random.default_rng(4)
    A = rng.standard_normal((64, 128)).astype(np.float32)
    B = rng.standard_normal((128, 64)).astype(np.float32)
    out = gemm(L, dev(A, torch.float32), dev(B, torch.float32),
               dev(np.zeros((64, 64)), torch.float32), 1.0, 0.0, 1).cpu().numpy()
    ab = torch.tensor(A).to(torch.bfloat16).to(torch.float64).numpy()
    bb = torch.tensor(B).to(torch.bfloat16).to(torch.float64).numpy()
    ref = ab @ bb
    assert np.allclose(out, ref, rtol=1e-5, atol=1e-4)
    assert not np.allclose(out, A.astype(np.float64) @ B, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("S", [3, 4, 5, 6, 7])
@pytest.mark.parametrize("m,n,K", [(1, 1, 1), (100, 70, 64), (257, 129, 300), (640, 512, 512),
                                   (64, 130, 2048), (96, 80, 2100)])
def test_ozaki_gemm_bit_exact_vs_oracle(L, S, m, n, K):
    """K = 2048 is the longest K whose column slicing stages a column in LDS;
    K = 2100 takes the slicing kernel that reads the columns directly."""
    rng = np.random.default_rng(S * 1000 + m)
    A = rng.standard_normal((m, K)) * np.exp(rng.uniform(-5, 5, (m, 1)))
    B = rng.standard_normal((K, n)) * np.exp(rng.uniform(-5, 5, (1, n)))
    C = rng.standard_normal((m, n))
    ws = torch.empty(L.mpdok_ozaki_workspace_bytes(m, n, K, S), dtype=torch.uint8, device="cuda")
    # keep every operand referenced until the kernels have run: a temporary
    # freed early is recycled by torch's caching allocator for the next one
    Ad, Bd, Cd = dev(A), dev(B), dev(C)
    st = L.mpdok_ozaki_gemm_sub(stream(), S, m, n, K, Ad.data_ptr(), m, Bd.data_ptr(), K,
                                Cd.data_ptr(), m, ws.data_ptr())
    assert st == 0, L.mpdok_last_error()
    ref = C - olu.ozaki_matmul(A, B, S)
    assert np.array_equal(Cd.cpu().numpy(), ref)


@pytest.mark.parametrize("shape", [0, 1, 2, 3])
@pytest.mark.parametrize("epi", [0, 1])
@pytest.mark.parametrize("smax,m,n,K", [(5, 257, 129, 300), (3, 130, 260, 96), (4, 300, 520, 640),
                                        (5, 640, 384, 1024)])
def test_trailing_update_shapes_and_c_handling_bit_exact(L, shape, epi, smax, m, n, K):
    """Every block shape (tuning key 0) and both ways of handling C (key 5:
    read into registers at the start of a tile, or after its K loop) give the
    oracle's bits for every slice count: they change tiling, pipelining,
    operand order and when C is read, never the sums."""
    rng = np.random.default_rng(shape * 7 + smax + 31 * epi)
    A = rng.standard_normal((m, K)) * np.exp(rng.uniform(-3, 3, (m, 1)))
    B = rng.standard_normal((K, n)) * np.exp(rng.uniform(-3, 3, (1, n)))
    C = rng.standard_normal((m, n))
    assert L.mpdok_set_tuning(0, shape) == 0 and L.mpdok_set_tuning(5, epi) == 0
    try:
        for S in range(2, smax + 1):
            out, used, _ = _ozaki_adaptive(L, A, B, C, smax, s_fixed=S)
            assert used == S
            assert np.array_equal(out, C - olu.ozaki_matmul_lead(A, B, smax, S)), (S, shape, epi)
    finally:
        L.mpdok_set_tuning(0, 0)
        L.mpdok_set_tuning(5, 0)


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("S,m,n,K", [(5, 700, 90, 512), (3, 130, 64, 200), (6, 64, 64, 1)])
def test_ozaki_row_slicing_variants_bit_exact(L, variant, S, m, n, K):
    """Both row-slicing kernels (tuning key 4: two passes over a 2-D grid, or
    one workgroup per 64 rows) give the oracle's bits, including all-zero
    rows and rows whose maximum sits in the last 64-column block."""
    rng = np.random.default_rng(11 * S + m + variant)
    A = rng.standard_normal((m, K)) * np.exp(rng.uniform(-6, 6, (m, 1)))
    A[::7] = 0.0
    A[3::11, -1] = 1e6
    B = rng.standard_normal((K, n))
    C = rng.standard_normal((m, n))
    ws = torch.empty(L.mpdok_ozaki_workspace_bytes(m, n, K, S), dtype=torch.uint8, device="cuda")
    Ad, Bd, Cd = dev(A), dev(B), dev(C)
    assert L.mpdok_set_tuning(4, variant) == 0
    try:
        st = L.mpdok_ozaki_gemm_sub(stream(), S, m, n, K, Ad.data_ptr(), m, Bd.data_ptr(), K,
                                    Cd.data_ptr(), m, ws.data_ptr())
        torch.cuda.synchronize()
    finally:
        L.mpdok_set_tuning(4, 0)
    assert st == 0, L.mpdok_last_error()
    assert np.array_equal(Cd.cpu().numpy(), C - olu.ozaki_matmul(A, B, S))


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("N", [17, 64, 130, 512, 2048])
def test_diag_block_kernels_factor_and_inverses(L, variant, N):
    """Both 64 x 64 diagonal-block kernels (tuning key 3: blocked with matrix-core
    doubling, or unblocked): with nb = 64 every panel step is one diagonal
    block, so the packed factors must match the oracle's unpivoted LU, and the
    solves, which apply the stored 64 x 64 inverses (at N = 512 the 512-wide
    ones doubled from them, at N = 2048 the 1024-wide ones doubled once more),
    must reproduce numpy's solution."""
    rng = np.random.default_rng(N + 10 * variant)
    A = rng.standard_normal((N, N)) + N * np.eye(N)
    assert L.mpdok_set_tuning(3, variant) == 0
    try:
        h, info = _factor(L, A, "fp64", "none", 64)
        assert info == 0
        LU, _ = _factors(L, h, N)
        b = rng.standard_normal(N)
        x = h.solve(dev(b)).cpu().numpy()
    finally:
        L.mpdok_set_tuning(3, 0)
    ref = olu.lu_nopiv(A, nb=64)
    assert np.abs(LU - ref).max() < 1e-13 * np.abs(ref).max()
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("precision", ["ozaki", "fp64"])
@pytest.mark.parametrize("N,nb", [(2048, 128), (5000, 512), (4096, 512)])
def test_row_lookahead_gives_the_same_factors(L, precision, N, nb):
    """Moving block-row solves, U12 slicing and slice-count decisions to the
    lookahead stream (and updating the next step's rows first) changes where
    and in which launch each product is computed, not its operands or
    summation order: the factors, slice counts and solutions match the plain
    schedule bit for bit (also with a ragged last panel group, and at N = 4096
    with 512-column panel pairs, where the main stream starts a pair's first
    block row after its first panel and the solves use 1024-wide blocks)."""
    X = np.random.default_rng(5).uniform(0, 1, (N, 2))
    A = orb.build_kernel(X, "gaussian", 25.0, 1e-6)
    from mpdok_amd.mpdok_ops import _LUHandle
    out = []
    for rows in (0, 1024):
        h = _LUHandle(L, N, precision, "none", nb, 5, True)
        h.set_row_lookahead(rows)
        assert h.factor(dev(A)) == 0
        LU, _ = _factors(L, h, N)
        b = np.sin(3 * X[:, 0])
        out.append((LU.copy(), h.slice_counts() if precision == "ozaki" else None,
                    h.solve(dev(b)).cpu().numpy()))
        h.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("N,nb,rows", [(4096, 512, 1024), (3000, 512, 1024), (2048, 128, 0)])
def test_first_update_reading_the_input_gives_the_same_factors(L, N, nb, rows):
    """The int8-sliced variant's first trailing update reads a column-major
    FP64 input where it lies (its leading dimension padded or not) instead of
    from a copy in the handle; a row-major input is copied first. Both give
    the same factors and slice counts bit for bit, with panel pairs, row
    lookahead, and a ragged N whose edge tiles are partly outside the matrix;
    the input is left unchanged."""
    rng = np.random.default_rng(8)
    A = rng.standard_normal((N, N)) + N * np.eye(N)   # nonsymmetric, no pivoting needed
    from mpdok_amd.mpdok_ops import _LUHandle
    inputs = {"F": dev(A), "C": dev(A, order="C")}
    # column-major view with leading dimension N + 37
    inputs["F, ld = N + 37"] = torch.empty((N, N + 37), dtype=torch.float64,
                                           device="cuda").t()[:N]
    inputs["F, ld = N + 37"].copy_(torch.as_tensor(A))
    out = {}
    for name, Ad in inputs.items():
        before = Ad.clone()
        h = _LUHandle(L, N, "ozaki", "none", nb, 5, True)
        h.set_update_panels(2)
        h.set_row_lookahead(rows)
        assert h.factor(Ad) == 0
        LU, _ = _factors(L, h, N)
        out[name] = (LU.copy(), h.slice_counts())
        h.close()
        assert torch.equal(Ad, before), name
    assert len(out["F"][1]) > 1
    for name in ("C", "F, ld = N + 37"):
        assert np.array_equal(out["F"][0], out[name][0]), name
        assert out["F"][1] == out[name][1], name


@pytest.mark.parametrize("rows", [0, 1024])
def test_wide_block_solve_matches_own_factors(L, rows):
    """Panels of 512 columns in pairs, N a multiple of 1024: the 512-wide U
    inverses are built on the main stream (rows = 0) or on the lookahead
    stream (rows = 1024, row lookahead throughout), the 1024-wide ones after
    the factorisation, and the solve that applies them must match triangular
    solves with the handle's own factors."""
    N = 4096
    rng = np.random.default_rng(11)
    X = rng.standard_normal((N, N))
    A = X.T @ X / N + np.eye(N)
    from mpdok_amd.mpdok_ops import _LUHandle
    h = _LUHandle(L, N, "ozaki", "none", 512, 5, True)
    h.set_row_lookahead(rows)
    assert h.factor(dev(A)) == 0
    assert h.step_starts() == list(range(0, N, 1024))
    LU, _ = _factors(L, h, N)
    b = rng.standard_normal(N)
    x = h.solve(dev(b)).cpu().numpy()
    h.close()
    Lm, Um = olu.unpack(LU)
    xr = sla.solve_triangular(Um, sla.solve_triangular(Lm, b, lower=True, unit_diagonal=True))
    assert np.linalg.norm(x - xr) <= 1e-12 * np.linalg.norm(xr)


def _factor(L, A, precision, pivot, nb=128, slices=6, adaptive=False):
    from mpdok_amd.mpdok_ops import _LUHandle
    h = _LUHandle(L, A.shape[0], precision, pivot, nb, slices, adaptive)
    info = h.factor(dev(A))
    return h, info


def _ozaki_adaptive(L, A, B, C, smax, s_fixed=0, smin=2, delta=0, e0=None):
    m, K = A.shape
    n = B.shape[1]
    ws = torch.empty(L.mpdok_ozaki_workspace_bytes(m, n, K, smax), dtype=torch.uint8,
                     device="cuda")
    scratch = torch.zeros(2, dtype=torch.int32, device="cuda")
    Ad, Bd, Cd = dev(A), dev(B), dev(C)
    s_used, e_ref = ctypes.c_int(0), ctypes.c_int(0)
    st = L.mpdok_ozaki_gemm_sub_adaptive(
        stream(), smax, m, n, K, Ad.data_ptr(), m, Bd.data_ptr(), K, Cd.data_ptr(), m,
        ws.data_ptr(), s_fixed, smin, delta, -2 ** 31 if e0 is None else e0, scratch.data_ptr(),
        ctypes.byref(s_used), ctypes.byref(e_ref))
    assert st == 0, L.mpdok_last_error()
    return Cd.cpu().numpy(), s_used.value, e_ref.value


@pytest.mark.parametrize("smax", [3, 4, 5])
@pytest.mark.parametrize("m,n,K", [(100, 70, 64), (257, 129, 300), (640, 512, 512),
                                   (2047, 1100, 64), (300, 260, 1000), (50, 70, 1024),
                                   (1100, 700, 1024)])
def test_ozaki_leading_slices_bit_exact_vs_oracle(L, smax, m, n, K):
    """The runtime-count kernel multiplies the leading S of smax digits: bit for
    bit the oracle's ozaki_matmul_lead for every S in [2, smax]."""
    rng = np.random.default_rng(smax * 100 + m)
    A = rng.standard_normal((m, K)) * np.exp(rng.uniform(-4, 4, (m, 1)))
    B = rng.standard_normal((K, n)) * np.exp(rng.uniform(-4, 4, (1, n)))
    C = rng.standard_normal((m, n))
    for S in range(2, smax + 1):
        out, used, _ = _ozaki_adaptive(L, A, B, C, smax, s_fixed=S)
        assert used == S
        assert np.array_equal(out, C - olu.ozaki_matmul_lead(A, B, smax, S)), S


def test_ozaki_decision_matches_oracle(L):
    """The device decision (first update records E0; later ones scale with the
    exponent drop) gives the oracle's counts, and the product that follows is
    bit-exact."""
    rng = np.random.default_rng(9)
    m, n, K = 300, 260, 256
    A = rng.uniform(-1, 1, (m, K))
    B = rng.uniform(-1, 1, (K, n))
    C = rng.standard_normal((m, n))
    out, used, e0 = _ozaki_adaptive(L, A, B, C, 5)
    S0, e0_ref = olu.ozaki_decide(A, B, 5, 2, 0, None)
    assert (used, e0) == (S0, e0_ref) == (5, e0_ref)
    for k, delta, smin in [(3, 0, 2), (8, 0, 2), (13, 0, 2), (20, 0, 3), (30, 0, 2), (12, 6, 2)]:
        Bk = np.ldexp(B, -k)
        out, used, _ = _ozaki_adaptive(L, A, Bk, C, 5, smin=smin, delta=delta, e0=e0)
        want, _ = olu.ozaki_decide(A, Bk, 5, smin, delta, e0_ref)
        assert used == want, (k, used, want)
        assert np.array_equal(out, C - olu.ozaki_matmul_lead(A, Bk, 5, want))


def _groups(N, nb, panels, early, early_rows):
    """First columns of the library's panel groups (one trailing update each)."""
    out, k = [], 0
    while k < N:
        out.append(k)
        k += (early if early > panels and N - k >= early_rows else panels) * nb
    return out


@pytest.mark.parametrize("panels,early,early_rows,nb",
                         [(1, 1, 0, 256), (2, 2, 0, 128), (2, 4, 1024, 128)])
def test_ozaki_adaptive_lu_counts_follow_its_own_panels(L, panels, early, early_rows, nb):
    """Adaptive factorisation of a decaying RBF matrix with one, two, or four
    then two panels per trailing update: each update's count is the oracle's
    decision on the very L21 and U12 the library sliced (read back from its
    factors; K = the group's width), the counts drop below 5, and the
    backward error stays within a small factor of the fixed 5-slice
    factorisation's."""
    N = 2048
    X = np.random.default_rng(42).uniform(0, 1, (N, 2))
    A = orb.build_kernel(X, "gaussian", 25.0, 1e-6)
    from mpdok_amd.mpdok_ops import _LUHandle
    h = _LUHandle(L, N, "ozaki", "none", nb, 5, True)
    h.set_update_panels(panels, early, early_rows)
    info = h.factor(dev(A))
    assert info == 0
    counts = h.slice_counts()
    LU, _ = _factors(L, h, N)
    e0 = None
    want = []
    gs = _groups(N, nb, panels, early, early_rows) + [N]
    for k, e in zip(gs[:-2], gs[1:-1]):
        S, e0 = olu.ozaki_decide(LU[e:, k:e], LU[k:e, e:], 5, 2, 0, e0)
        want.append(S)
    assert len(counts) == len(gs) - 1
    assert counts[:len(want)] == want
    assert want[0] == 5 and min(want) < 5
    Lm, Um = olu.unpack(LU)
    be = np.linalg.norm(Lm @ Um - A) / np.linalg.norm(A)
    h5, _ = _factor(L, A, "ozaki", "none", nb, 5, adaptive=False)
    LU5, _ = _factors(L, h5, N)
    L5, U5 = olu.unpack(LU5)
    be5 = np.linalg.norm(L5 @ U5 - A) / np.linalg.norm(A)
    assert be < 4 * be5 + 1e-15, (be, be5)


def _factors(L, h, N):
    LUp, permp, es = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_int()
    assert L.mpdok_lu_factors(h.ptr, ctypes.byref(LUp), ctypes.cast(ctypes.byref(permp),
                                                                    ctypes.POINTER(ctypes.c_void_p)),
                              ctypes.byref(es)) == 0
    torch.cuda.synchronize()
    dt = torch.float64 if es.value == 8 else torch.float32
    buf = torch.empty((N, N), dtype=dt, device="cuda")
    from mpdok_amd.mpdok_ops import _CAI
    t = torch.as_tensor(_CAI(LUp.value, (N, N), (es.value, es.value * N),
                             "<f8" if es.value == 8 else "<f4"), device="cuda")
    buf.copy_(t)
    perm = None
    if permp.value:
        perm = torch.as_tensor(_CAI(permp.value, (N,), (4,), "<i4"), device="cuda").cpu().numpy()
    return buf.cpu().numpy().astype(np.float64), perm


@pytest.mark.parametrize("N,nb", [(64, 64), (200, 64), (777, 128), (1500, 512)])
def test_lu_nopivot_fp64_matches_oracle(L, N, nb):
    rng = np.random.default_rng(N)
    X = rng.standard_normal((N, N))
    A = X.T @ X + N * np.eye(N)
    h, info = _factor(L, A, "fp64", "none", nb)
    assert info == 0
    LU, perm = _factors(L, h, N)
    ref = olu.lu_nopiv(A, nb=nb)
    assert perm is None
    assert np.allclose(LU, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())
    b = rng.standard_normal(N)
    x = h.solve(dev(b)).cpu().numpy()
    xr = sla.solve(A, b)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-12


@pytest.mark.parametrize("N", [130, 700, 2500])
def test_lu_partial_pivot_matches_lapack(L, N):
    rng = np.random.default_rng(N + 1)
    A = rng.standard_normal((N, N))
    h, info = _factor(L, A, "fp64", "partial", 128)
    assert info == 0
    LU, perm = _factors(L, h, N)
    lu, piv = sla.lu_factor(A)
    # LAPACK's ipiv -> permutation
    p = np.arange(N)
    for i, q in enumerate(piv):
        p[[i, q]] = p[[q, i]]
    assert np.array_equal(perm, p)
    assert np.allclose(LU, lu, rtol=1e-9, atol=1e-9)
    b = rng.standard_normal(N)
    x = h.solve(dev(b)).cpu().numpy()
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-10)


def test_lu_singular_reports_info(L):
    A = np.eye(100)
    A[37, 37] = 0.0
    h, info = _factor(L, A, "fp64", "none", 64)
    assert info == 38
    A2 = np.random.default_rng(0).standard_normal((100, 100))
    A2[:, 50] = 0.0
    h, info = _factor(L, A2, "fp64", "partial", 64)
    assert info == 51


@pytest.mark.parametrize("S", [4, 5, 6])
@pytest.mark.parametrize("panels,early", [(1, 1), (2, 2), (2, 4)])
def test_ozaki_lu_matches_oracle_on_well_conditioned_spd(L, S, panels, early):
    # Factors of a well-conditioned matrix are well-conditioned functions of
    # the rounding, so the GPU factors must match the oracle's restatement.
    # A group of panels is factored in FP64 and its update takes K = the
    # group's width: the oracle's blocked LU with the same block starts.
    N, nb = 1024, 128
    rng = np.random.default_rng(S)
    X = rng.standard_normal((N, N))
    A = X.T @ X / N + np.eye(N)
    from mpdok_amd.mpdok_ops import _LUHandle
    h = _LUHandle(L, N, "ozaki", "none", nb, S, False)
    h.set_update_panels(panels, early, 768)
    assert h.factor(dev(A)) == 0
    LU, _ = _factors(L, h, N)
    ref = olu.lu_nopiv(A, trailing=lambda a, b: olu.ozaki_matmul(a, b, S),
                       starts=_groups(N, nb, panels, early, 768))
    # panel roundings differ (FP64 order), and a last digit may flip where an
    # input sits on a rounding boundary: allow a few units of 2^(-(8 S - 1))
    assert np.abs(LU - ref).max() < 2.0 ** (-(8 * S - 1) + 6) * np.abs(ref).max()
    exact = olu.lu_nopiv(A, nb=256)
    # and S slices bound the distance to the FP64 factors: ~2^(-(8 S - 1)) relative
    assert np.abs(LU - exact).max() < 2.0 ** (-(8 * S - 1) + 12) * np.abs(exact).max()


@pytest.mark.parametrize("S", [4, 5, 6])
def test_ozaki_lu_backward_error_on_rbf(L, S):
    # kappa ~ 1e8: compare the backward error ||L U - A|| / ||A||, not the factors
    N = 1024
    X = np.random.default_rng(42).uniform(0, 1, (N, 3))
    A = orb.build_kernel(X, "gaussian", 25.0, 1e-6)
    h, info = _factor(L, A, "ozaki", "none", 256, S)
    assert info == 0
    LU, _ = _factors(L, h, N)
    Lm, Um = olu.unpack(LU)
    be = np.linalg.norm(Lm @ Um - A) / np.linalg.norm(A)
    # the oracle's factors of this matrix: 1.6e-8 (S = 4), 7.5e-11 (5), 3.4e-13 (6)
    assert be < {4: 6e-8, 5: 3e-10, 6: 1.5e-12}[S]
    # and it preconditions LU-IR to the FP64 floor within a few steps
    b = np.sin(2 * np.pi * X[:, 0])
    x = np.zeros(N)
    for _ in range(6):
        x = x + h.solve(dev(b - A @ x)).cpu().numpy()
    assert np.linalg.norm(olu.residual_longdouble(A, x, b)) / np.linalg.norm(b) < 1e-11


def test_luir_solver_default_fp32_partial(L):
    from mpdok_amd.mpdok_ops import LUIRSolver
    N = 1024
    rng = np.random.default_rng(42)
    A = rng.standard_normal((N, N)) + N * np.eye(N)
    b = rng.standard_normal(N)
    s = LUIRSolver()
    x = s.solve(A, b, tol=1e-11).cpu().numpy()
    xo, conv, hist = olu.lu_ir(A, b, tol=1e-11)
    assert conv and s.last_info["converged"]
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-11
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-12
    # factor once, solve several right-hand sides
    s.factor(A)
    for seed in range(3):
        bb = np.random.default_rng(seed).standard_normal(N)
        xx = s.solve_factored(bb).cpu().numpy()
        assert np.linalg.norm(bb - A @ xx) / np.linalg.norm(bb) < 1e-11
    s.free_factored()
    with pytest.raises(RuntimeError):
        s.solve_factored(b)


def test_mpdok_solver_default_path_nonsym(L):
    from mpdok_amd.mpdok_ops import MPDOKSolver
    N = 2048
    rng = np.random.default_rng(42)
    A = rng.standard_normal((N, N)) + N * np.eye(N)
    b = rng.standard_normal(N)
    x = MPDOKSolver().solve(A, b, tol=1e-11).cpu().numpy()
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-12
    with pytest.raises(ValueError):
        MPDOKSolver().solve(np.zeros((4, 5)), np.zeros(4))


@pytest.mark.parametrize("fp,S", [("ozaki", 6), ("fp64", 0)])
def test_lu_preconditioned_gmres_ir_rbf_to_1e12(L, fp, S):
    from mpdok_amd.mpdok_ops import MPDOKSolver
    N = 2048
    X = np.random.default_rng(42).uniform(0, 1, (N, 3))
    A = orb.build_kernel(X, "gaussian", 25.0, 1e-6)
    b = (np.sin(2 * np.pi * X[:, 0]) * np.cos(2 * np.pi * X[:, 1]) * np.cos(2 * np.pi * X[:, 2])
         + 1e-4 * np.random.default_rng(43).standard_normal(N))
    solver = MPDOKSolver()
    x = solver.solve(A, b, tol=1e-12, maxiter_outer=10, precond="lu", factor_precision=fp,
                     slices=S or 6).cpu().numpy()
    info = solver.last_info
    assert info["converged"], info
    r = olu.residual_longdouble(A, x, b)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-12
    x64 = sla.solve(A, b)
    xl = x64 + sla.solve(A, olu.residual_longdouble(A, x64, b))   # refined reference
    err = np.linalg.norm(x - xl) / np.linalg.norm(xl)
    err64 = np.linalg.norm(x64 - xl) / np.linalg.norm(xl)
    assert err <= 10 * max(err64, 1e-15)


def test_reference_symbols_through_raw_ctypes(L):
    """Bind py_lu_* / py_gmres_ir the way mpdok_ops.py:101-141,361-382 does."""
    from mpdok_amd import _lib
    lib = ctypes.CDLL(_lib.DEFAULT_LIB)
    lib.py_lu_ir.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                      ctypes.POINTER(ctypes.c_int)]
    lib.py_lu_factor.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.py_lu_solve_factored.argtypes = lib.py_lu_ir.argtypes[:3] + lib.py_lu_ir.argtypes[3:]
    lib.py_gmres_ir.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_double,
                                                        ctypes.c_int, ctypes.c_int,
                                                        ctypes.POINTER(ctypes.c_int)]
    N = 512
    rng = np.random.default_rng(1)
    A = rng.standard_normal((N, N)) + N * np.eye(N)
    b = rng.standard_normal(N)
    Ad, bd = dev(A), dev(b)
    x = torch.zeros(N, dtype=torch.float64, device="cuda")
    conv = ctypes.c_int(0)
    lib.py_lu_free_factored()
    lib.py_lu_solve_factored(Ad.data_ptr(), bd.data_ptr(), x.data_ptr(), N, 1e-11, 3,
                             ctypes.byref(conv))
    assert conv.value == -1
    lib.py_lu_ir(Ad.data_ptr(), bd.data_ptr(), x.data_ptr(), N, 1e-11, 3, ctypes.byref(conv))
    assert conv.value == 1
    assert np.linalg.norm(b - A @ x.cpu().numpy()) / np.linalg.norm(b) < 1e-11
    A32 = torch.empty((N, N), dtype=torch.float32, device="cuda")
    x.zero_()
    lib.py_gmres_ir(Ad.data_ptr(), bd.data_ptr(), x.data_ptr(), A32.data_ptr(), N, 1e-11, 5, 50,
                    ctypes.byref(conv))
    assert conv.value == 1
    xg = x.cpu().numpy()
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(b - A @ xg) / np.linalg.norm(b) < 1e-11
    assert np.linalg.norm(xg - xr) / np.linalg.norm(xr) < 1e-12
    # the FP32 scratch holds A rounded to FP32, column-major like A
    assert np.array_equal(A32.cpu().numpy().reshape(N, N).T, A.astype(np.float32))
    lib.py_lu_factor(Ad.data_ptr(), N, ctypes.byref(conv))
    assert conv.value == 0
    x.zero_()
    lib.py_lu_solve_factored(Ad.data_ptr(), bd.data_ptr(), x.data_ptr(), N, 1e-11, 3,
                             ctypes.byref(conv))
    assert conv.value == 1
    assert np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr) < 1e-12
    lib.py_lu_free_factored()


def _refined_reference(A, b):
    """LAPACK FP64 solve plus one step of long-double residual refinement,
    and the forward error of the unrefined solve against it."""
    x64 = sla.solve(A, b)
    xl = x64 + sla.solve(A, olu.residual_longdouble(A, x64, b))
    return xl, np.linalg.norm(x64 - xl) / np.linalg.norm(xl)


def test_headline_solver_settings_on_the_3d_generator(L):
    """MPDOKSolver.solve with the exact keywords of the N = 32768 benchmark
    (at most 5 slices, adaptive counts with margin -4, nb = 512 panel pairs,
    the default row lookahead) on the same generator at N = 8192: long-double
    relres <= 1e-12 and forward error within 10x of LAPACK's."""
    from mpdok_amd.mpdok_ops import MPDOKSolver
    N = 8192
    X = np.random.default_rng(42).uniform(0, 1, (N, 3))
    A = orb.build_kernel(X, "gaussian", 25.0, 1e-6)
    b = (np.sin(2 * np.pi * X[:, 0]) * np.cos(2 * np.pi * X[:, 1]) * np.cos(2 * np.pi * X[:, 2])
         + 1e-4 * np.random.default_rng(43).standard_normal(N))
    solver = MPDOKSolver()
    x = solver.solve(A, b, tol=1e-12, maxiter_outer=20, restart=50, precond="lu",
                     factor_precision="ozaki", slices=5, pivot="none", body="fp64",
                     residual="dot2", nb=512).cpu().numpy()
    info = solver.last_info
    assert info["converged"], info
    counts = info["slice_counts"]
    assert counts[0] == 5 and min(counts[:-1]) >= 2 and max(counts) <= 5
    starts = info["step_starts"]
    assert all(b_ - a == 1024 for a, b_ in zip(starts[:-1], starts[1:]))   # panel pairs
    r = olu.residual_longdouble(A, x, b)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-12
    xl, err64 = _refined_reference(A, b)
    assert np.linalg.norm(x - xl) / np.linalg.norm(xl) <= 10 * max(err64, 1e-15)


def test_cpu_sized_2d_case_int8_lu_gmres_ir_and_reference_lu_ir(L):
    """The CPU-runnable configuration: N = 4096 points in [0,1]^2 (seed 42),
    eps = 8 (gamma = 64), nugget 1e-8, b = sin(2 pi x) cos(2 pi y) + 1e-4
    noise (seed 43). kappa ~ 2e10: 6 fixed slices precondition GMRES-IR,
    which reaches the FP64 floor of this system (||x||/||b|| ~ 2e4: rounding
    the exact solution alone leaves relres ~ 2e-11), with a forward error
    within 10x of LAPACK's. The reference's FP32 LU-IR (py_lu_ir) cannot
    precondition kappa >> 1/u32 and reports non-convergence."""
    from mpdok_amd.mpdok_ops import LUIRSolver, MPDOKSolver
    N = 4096
    X = np.random.default_rng(42).uniform(0, 1, (N, 2))
    A = orb.build_kernel(X, "gaussian", 64.0, 1e-8)
    b = (np.sin(2 * np.pi * X[:, 0]) * np.cos(2 * np.pi * X[:, 1])
         + 1e-4 * np.random.default_rng(43).standard_normal(N))
    solver = MPDOKSolver()
    x = solver.solve(A, b, tol=1e-11, maxiter_outer=20, precond="lu", factor_precision="ozaki",
                     slices=6).cpu().numpy()
    assert solver.last_info["converged"], solver.last_info
    rel = np.linalg.norm(olu.residual_longdouble(A, x, b)) / np.linalg.norm(b)
    assert rel <= 2e-11
    xl, err64 = _refined_reference(A, b)
    assert np.linalg.norm(x - xl) / np.linalg.norm(xl) <= 10 * max(err64, 1e-15)
    with pytest.warns(RuntimeWarning, match="did not converge"):
        LUIRSolver().solve(A, b, tol=1e-12, maxiter_outer=10)


def _handle(L, N, precision, pivot, nb=512, slices=5):
    h = ctypes.c_void_p()
    assert L.mpdok_lu_create(N, precision, pivot, nb, slices, ctypes.byref(h)) == 0
    return h


@pytest.mark.parametrize("N,precision,pivot", [(2048, 4, 0), (1536, 1, 1), (1000, 0, 0),
                                               (3072, 0, 1)])
def test_lu_solve_multi_equals_single_solves(L, N, precision, pivot):
    """mpdok_lu_solve_multi: every column bit for bit the single solve's, for
    the 1024-wide (N = 2048, 3072) and 512-wide (1536) inverse-based solves,
    pivoted or not, and for the per-column fallback (N = 1000); 11 columns
    exercise the blocks of 8, 2 and 1."""
    rng = np.random.default_rng(N + precision)
    X = rng.uniform(0, 1, (N, 2))
    A = orb.build_kernel(X, "gaussian", 10.0, 1e-3) + (0 if not pivot else
                                                      0.01 * rng.standard_normal((N, N)))
    h = _handle(L, N, precision, pivot)
    try:
        info = ctypes.c_int(-1)
        Ad = dev(A)
        assert L.mpdok_lu_factor(h, stream(), Ad.data_ptr(), N, 1, ctypes.byref(info)) == 0
        assert info.value == 0
        k = 11
        B = dev(rng.standard_normal((N, k)))
        Xm = torch.full((N, k), float("nan"), dtype=torch.float64, device="cuda").t().contiguous().t()
        st = L.mpdok_lu_solve_multi(h, stream(), B.data_ptr(), N, Xm.data_ptr(), N, k)
        assert st == 0, L.mpdok_last_error()
        for q in range(k):
            xs = torch.empty(N, dtype=torch.float64, device="cuda")
            assert L.mpdok_lu_solve(h, stream(), B[:, q].contiguous().data_ptr(), xs.data_ptr()) == 0
            torch.cuda.synchronize()
            assert torch.equal(Xm[:, q], xs), q
    finally:
        L.mpdok_lu_destroy(h)


@pytest.mark.parametrize("mode", [0, 1])
def test_residual_multi_equals_single_residuals(L, mode):
    N, k = 1500, 13
    rng = np.random.default_rng(mode)
    A = dev(rng.standard_normal((N, N)))
    Xv = dev(rng.standard_normal((N, k)))
    Bv = dev(rng.standard_normal((N, k)))
    R = torch.empty((k, N), dtype=torch.float64, device="cuda").t()
    nb = L.mpdok_residual_multi_workspace_bytes(N)
    ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
    st = L.mpdok_residual_multi_f64(stream(), A.data_ptr(), N, N, Xv.data_ptr(), N, Bv.data_ptr(), N,
                                    R.data_ptr(), N, k, mode, ws.data_ptr(), nb)
    assert st == 0, L.mpdok_last_error()
    nb1 = L.mpdok_residual_workspace_bytes(N, 1)
    ws1 = torch.empty(nb1, dtype=torch.uint8, device="cuda")
    for q in range(k):
        r = torch.empty(N, dtype=torch.float64, device="cuda")
        x = Xv[:, q].contiguous()
        b = Bv[:, q].contiguous()
        assert L.mpdok_residual_f64(stream(), A.data_ptr(), N, 1, N, x.data_ptr(), b.data_ptr(),
                                    r.data_ptr(), mode, ws1.data_ptr(), nb1) == 0
        torch.cuda.synchronize()
        assert torch.equal(R[:, q], r), q


@pytest.mark.parametrize("precision,pivot,residual", [("fp32", "partial", "fp64"),
                                                      ("ozaki", "none", "dot2")])
def test_solve_factored_block_equals_column_by_column(L, precision, pivot, residual):
    """Factor once, then a block of 10 right-hand sides (the weather-front time
    series of rbf_kernel.py:230-252): the block equals ten single
    solve_factored calls bit for bit and every column reaches the tolerance."""
    from mpdok_amd.mpdok_ops import LUIRSolver
    N = 2048
    X = orb.synthetic_coords(N, seed=3)
    A = orb.build_kernel(X, "gaussian", 1.5e-4, 1e-2)
    B = np.stack([orb.weather_front(X, t) for t in range(10)], axis=1)
    s = LUIRSolver(precision=precision, pivot=pivot, residual=residual)
    s.factor(A)
    Xb = s.solve_factored(B, tol=1e-11, maxiter_outer=5)
    assert tuple(Xb.shape) == (N, 10) and all(s.last_info["converged"])
    for t in range(10):
        xt = s.solve_factored(B[:, t], tol=1e-11, maxiter_outer=5)
        assert torch.equal(Xb[:, t], xt), t
        assert np.linalg.norm(B[:, t] - A @ xt.cpu().numpy()) / np.linalg.norm(B[:, t]) < 1e-11
    s.free_factored()


def test_precision_ladder_on_the_3d_generator(L):
    """Which LU preconditioners carry GMRES-IR (FP64 body, Dot2 residual) to
    1e-12 on the headline generator at N = 8192 (3-D points, gamma 25,
    nugget 1e-6, kappa ~ 1e9): the int8-sliced LU does; FP32, BF16, FP16 and
    scaled-FP8 trailing updates do not (u kappa >> 1: the refinement stalls or
    the factors break down), and neither does an FP32 Krylov body with the
    int8-sliced LU (kappa u32 ~ 60). The FP64 LU is the comparator."""
    import warnings
    from mpdok_amd.mpdok_ops import MPDOKSolver
    N = 8192
    X = np.random.default_rng(42).uniform(0, 1, (N, 3))
    A = orb.build_kernel(X, "gaussian", 25.0, 1e-6)
    b = (np.sin(2 * np.pi * X[:, 0]) * np.cos(2 * np.pi * X[:, 1]) * np.cos(2 * np.pi * X[:, 2])
         + 1e-4 * np.random.default_rng(43).standard_normal(N))
    Ad = dev(A)
    out = {}
    for prec, body in [("ozaki", "fp64"), ("fp64", "fp64"), ("ozaki", "fp32"), ("fp32", "fp64"),
                       ("bf16", "fp64"), ("fp16", "fp64"), ("fp8", "fp64")]:
        s = MPDOKSolver()
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            x = s.solve(Ad, b, tol=1e-12, maxiter_outer=4, restart=50, precond="lu",
                        factor_precision=prec, slices=5, pivot="none", body=body,
                        residual="dot2").cpu().numpy()
        out[(prec, body)] = (s.last_info["converged"], s.last_info["history"][-1], x)
        s.free_fp32_buf()
    for key in [("ozaki", "fp64"), ("fp64", "fp64")]:
        conv, rel, x = out[key]
        assert conv and rel <= 1e-12, key
        assert np.linalg.norm(olu.residual_longdouble(A, x, b)) / np.linalg.norm(b) <= 1e-12
    for key in [("ozaki", "fp32"), ("fp32", "fp64"), ("bf16", "fp64"), ("fp16", "fp64"),
                ("fp8", "fp64")]:
        conv, rel, _ = out[key]
        assert not conv and not (rel <= 1e-8), (key, rel)
