"""GPU parity of the HIP kernels against the CPU oracle, called through the
C-ABI (via the package's ctypes wrappers)."""
import numpy as np
import pytest
import torch

from oracle import lu as olu
from oracle import rbf as orb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ops():
    from mpdok_amd import _ops
    return _ops


def dev(a, dtype=torch.float64, order="C"):
    from mpdok_amd import _dev
    return _dev.as_device(a, dtype, order)


SHAPES = [(1, 1), (7, 13), (64, 64), (129, 1000), (1000, 129), (2048, 2051)]


@pytest.mark.parametrize("M,N", SHAPES)
@pytest.mark.parametrize("order", ["C", "F"])
@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_gemv_bit_exact_on_integer_inputs(ops, M, N, order, dt):
    # |A|, |x| <= 8 and N <= 2051: every partial sum is an integer below 2^24,
    # exact in FP32 and FP64 whatever the summation order.
    rng = np.random.default_rng(M * 7 + N)
    A = rng.integers(-8, 9, (M, N)).astype(np.float64)
    x = rng.integers(-8, 9, N).astype(np.float64)
    y = ops.matvec(dev(A, dt, order), dev(x, dt)).cpu().numpy()
    assert np.array_equal(y, A @ x)


@pytest.mark.parametrize("order", ["C", "F"])
def test_gemv_f32_random_within_error_bound(ops, order):
    rng = np.random.default_rng(5)
    M, N = 3000, 4097
    A = rng.standard_normal((M, N)).astype(np.float32)
    x = rng.standard_normal(N).astype(np.float32)
    y = ops.matvec(dev(A, torch.float32, order), dev(x, torch.float32)).cpu().numpy()
    ref = A.astype(np.float64) @ x.astype(np.float64)
    bound = 2 * N * 2.0 ** -24 * (np.abs(A) @ np.abs(x))
    assert np.all(np.abs(y - ref) <= bound)


def test_gemv_unaligned_views_and_padded_lda(ops):
    rng = np.random.default_rng(6)
    big = rng.integers(-4, 5, (300, 517)).astype(np.float64)
    A = dev(big, torch.float32)[1:200, 3:503]   # odd offset: scalar path
    x = dev(rng.integers(-4, 5, 500).astype(np.float64), torch.float32)
    y = ops.matvec(A, x).cpu().numpy()
    assert np.array_equal(y, big[1:200, 3:503] @ x.cpu().numpy().astype(np.float64))
    AF = dev(big, torch.float64, "F")[5:290, 7:400]  # column-major view, lda = 300
    x2 = dev(rng.integers(-4, 5, 393).astype(np.float64))
    y2 = ops.matvec(AF, x2).cpu().numpy()
    assert np.array_equal(y2, big[5:290, 7:400] @ x2.cpu().numpy())


@pytest.mark.parametrize("order", ["C", "F"])
def test_residual_fp64_and_dot2_against_exact(ops, order):
    rng = np.random.default_rng(7)
    N = 48
    A = rng.standard_normal((N, N)) * 10.0 ** rng.integers(-6, 7, (N, N))
    x = rng.standard_normal(N)
    b = A @ x + rng.standard_normal(N) * 1e-9   # heavy cancellation in b - A x
    exact = olu.residual_exact(A, x, b)
    r0 = ops.residual(dev(A, order=order), dev(x), dev(b), mode="fp64").cpu().numpy()
    r1 = ops.residual(dev(A, order=order), dev(x), dev(b), mode="dot2").cpu().numpy()
    scale = np.abs(A) @ np.abs(x) + np.abs(b)
    u = 2.0 ** -53
    # Dot2: error <= u |r| + 2 (N u)^2 sum|a x| (Ogita, Rump, Oishi 2005, Prop. 4.5 form)
    assert np.all(np.abs(r1 - exact) <= u * np.abs(exact) + 2 * (N * u) ** 2 * scale)
    # plain FP64: the classic gamma_N bound
    assert np.all(np.abs(r0 - exact) <= (N + 1) * u * scale)
    # and Dot2 is much closer than plain FP64 on this input
    assert np.abs(r1 - exact).max() < 1e-3 * max(np.abs(r0 - exact).max(), 1e-300)


@pytest.mark.parametrize("order", ["C", "F"])
def test_residual_dot2_large_matches_longdouble(ops, order):
    rng = np.random.default_rng(8)
    N = 1500
    A = rng.standard_normal((N, N))
    x = rng.standard_normal(N)
    b = A @ x
    ref = olu.residual_longdouble(A, x, b)
    r1 = ops.residual(dev(A, order=order), dev(x), dev(b), mode="dot2").cpu().numpy()
    scale = np.abs(A) @ np.abs(x)
    assert np.all(np.abs(r1 - ref) <= 1e-16 * np.abs(ref) + 1e-18 * scale)


def test_dot_and_norm(ops):
    rng = np.random.default_rng(9)
    for n in (1, 255, 256, 257, 100000):
        x = rng.integers(-100, 100, n).astype(np.float64)
        y = rng.integers(-100, 100, n).astype(np.float64)
        assert float(ops.dot(dev(x, torch.float32), dev(y, torch.float32)).item()) == float(x @ y)
        assert float(ops.dot(dev(x), dev(y)).item()) == float(x @ y)
    assert float(ops.dot(dev(np.zeros(0)), dev(np.zeros(0))).item()) == 0.0


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("order", ["C", "F"])
def test_rbf_assembly_gaussian_fp64_and_fp32(ops, D, order):
    N = 777
    X = np.random.default_rng(D).uniform(0, 1, (N, D))
    gamma, reg = 25.0, 1e-6
    ref = orb.build_kernel(X, "gaussian", gamma, reg)
    out = dev(np.zeros((N, N)), torch.float64, order)
    A = ops.rbf_assemble(dev(X), 0, gamma, reg, out).cpu().numpy()
    # d2 carries a few ulps of (sq_i + sq_j) <= 2 D of absolute error from the
    # Gram identity (summation order differs from BLAS); exp turns it into a
    # relative error of gamma * 4u * 2D on the kernel value, plus 1 ulp of exp.
    rtol = gamma * 4 * 2.0 ** -53 * 2 * D + 2.0 ** -52
    assert np.all(np.abs(A - ref) <= rtol * np.abs(ref))
    assert np.array_equal(A, A.T)                       # symmetric bit for bit
    assert np.all(np.abs(np.diag(A) - (1 + reg)) <= 2.0 ** -52)
    out32 = dev(np.zeros((N, N)), torch.float32, order)
    A32 = ops.rbf_assemble(dev(X), 0, gamma, reg, out32).cpu().numpy()
    ulp = np.spacing(np.abs(ref).astype(np.float32))
    assert np.all(np.abs(A32.astype(np.float64) - ref.astype(np.float32)) <= ulp)


@pytest.mark.parametrize("D", [1, 2, 3, 5])
@pytest.mark.parametrize("order", ["C", "F"])
@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_rbf_symmetric_block_equals_general_tiles(ops, D, order, dt):
    """A principal block (row0 == col0, rows == cols) is evaluated one tile
    pair at a time and written twice; any other tile is evaluated element by
    element. Both give the same values bit for bit, with a partial last tile,
    with an offset principal block and for both nugget modes."""
    N = 300
    X = np.random.default_rng(10 + D).uniform(0, 1, (N, D))
    Xd = dev(X)

    def asm(kind, p, rm, rows, cols, row0=0, col0=0):
        out = dev(np.zeros((rows, cols)), dt, order)
        return ops.rbf_assemble(Xd, kind, p, 1e-3, out, row0=row0, rows=rows, col0=col0,
                                cols=cols, reg_mode=rm).cpu().numpy()

    for kind, p in ((0, 25.0), (1, 0.1)):
        for rm in (0, 1):
            full = asm(kind, p, rm, N, N)
            assert np.array_equal(full, full.T)
            assert np.array_equal(full[:, :N - 1], asm(kind, p, rm, N, N - 1))
            assert np.array_equal(full[:, N - 1:], asm(kind, p, rm, N, 1, col0=N - 1))
            assert np.array_equal(full[37:237, 37:237], asm(kind, p, rm, 200, 200, 37, 37))
            assert np.array_equal(full[37:237, 40:240], asm(kind, p, rm, 200, 200, 37, 40))


def test_rbf_tile_fp32_nugget_after_cast_and_offsets(ops):
    N = 500
    X = np.random.default_rng(3).uniform(0, 1, (N, 3))
    i, rows = 128, 100
    ref = orb.build_fp32_tile(X, i, rows, "gaussian", 25.0, 1e-4)
    out = dev(np.zeros((rows, N)), torch.float32, "C")
    t = ops.rbf_assemble(dev(X), 0, 25.0, 1e-4, out, row0=i, rows=rows, reg_mode=1).cpu().numpy()
    ulp = np.spacing(np.abs(ref))
    assert np.all(np.abs(t - ref) <= ulp)
    k = np.arange(rows)
    assert np.all(np.abs(t[k, i + k] - (np.float32(1.0) + np.float32(1e-4))) <= 2 ** -23)


def test_rbf_assembly_matern32(ops):
    N = 600
    X = np.random.default_rng(4).uniform(0, 1, (N, 3))
    ref = orb.build_kernel(X, "matern32", 0.1, 1e-4)
    out = dev(np.zeros((N, N)), torch.float64, "F")
    A = ops.rbf_assemble(dev(X), 1, 0.1, 1e-4, out).cpu().numpy()
    # Each side's d2 carries up to ~4u (sq_i + sq_j) <= 8 D u of Gram-identity
    # rounding. With a = sqrt(3 d2) / l, the kernel (1 + a) e^-a moves by
    # |dk/dd2| = 1.5 e^-a / l^2 <= 1.5 / l^2 per unit of d2 (finite at
    # d2 = 0, where sqrt alone would not be), so the values differ by at most
    # 1.5 / l^2 * 2 * 8 D u in absolute terms, plus a few ulps relative.
    ell, u = 0.1, 2.0 ** -53
    atol = 1.5 / ell ** 2 * 2 * 8 * 3 * u
    assert np.allclose(A, ref, rtol=1e-14, atol=atol)
    assert np.array_equal(A, A.T)


def test_host_array_build_matches_device_build(ops):
    """_build_rbf_cpu_into (rbf_kernel.py:126-155): the host F-order array
    gets, chunk by chunk, exactly the values of the device build."""
    from mpdok_amd.rbf_kernel import _build_rbf_cpu_into
    N, gamma, reg = 301, 25.0, 1e-6
    X = np.random.default_rng(5).uniform(0, 1, (N, 3))
    A_np = np.zeros((N, N), order="F")
    _build_rbf_cpu_into(X, gamma, reg, A_np, chunk=64)
    full = ops.rbf_assemble(dev(X), 0, gamma, reg, dev(np.zeros((N, N)), torch.float64, "F"))
    assert np.array_equal(A_np, full.cpu().numpy())
    ref = orb.build_kernel(X, "gaussian", gamma, reg)
    assert np.allclose(A_np, ref, rtol=gamma * 4 * 2.0 ** -53 * 6 + 2.0 ** -52, atol=0)


def test_estimate_gamma_matches_oracle():
    from mpdok_amd.rbf_kernel import _estimate_gamma
    X = np.random.default_rng(11).uniform(0, 100, (3000, 2))
    assert _estimate_gamma(dev(X)) == pytest.approx(orb.estimate_gamma(X), rel=1e-12)


def test_cast_matrix_orders(ops):
    X = np.random.default_rng(12).standard_normal((130, 77))
    for oi in "CF":
        for oo in "CF":
            y = ops.cast_matrix(dev(X, order=oi), torch.float32, oo).cpu().numpy()
            assert np.array_equal(y, X.astype(np.float32))


def test_synthetic_coords_matches_oracle():
    """rbf_kernel.py:41-54: (N, D) FP64 points uniform in [0, scale]^D on the
    device, drawn from NumPy PCG64 default_rng(seed) (the reference's CuPy
    stream cannot be reproduced; the oracle states the same generator)."""
    from mpdok_amd.rbf_kernel import synthetic_coords
    for N, D, seed, scale in [(1, 2, 42, 100.0), (1000, 2, 0, 100.0), (777, 3, 9, 1.0)]:
        c = synthetic_coords(N, D=D, seed=seed, scale=scale)
        assert c.is_cuda and c.dtype == torch.float64 and tuple(c.shape) == (N, D)
        assert np.array_equal(c.cpu().numpy(), orb.synthetic_coords(N, D, seed, scale))
    assert tuple(synthetic_coords(5).shape) == (5, 2)


def test_weather_front_matches_oracle():
    """rbf_kernel.py:230-252 on device coordinates, with and without noise,
    the noise seed defaulting to t."""
    from mpdok_amd.rbf_kernel import synthetic_coords, weather_front
    c = synthetic_coords(2048, seed=0)
    X = c.cpu().numpy()
    for t, noise, seed in [(0, 0.05, None), (3, 0.05, None), (3, 0.2, 11), (7, 0.0, None)]:
        b = weather_front(c, t, noise_std=noise, seed=seed)
        assert b.is_cuda and b.dtype == torch.float64 and tuple(b.shape) == (2048,)
        ref = orb.weather_front(X, t, noise_std=noise, seed=seed)
        assert np.allclose(b.cpu().numpy(), ref, rtol=0, atol=1e-13)
    # host coordinates are accepted too
    assert np.allclose(weather_front(X, 1).cpu().numpy(), orb.weather_front(X, 1), rtol=0,
                       atol=1e-13)
