"""Oracle: kernel-matrix assembly (test infrastructure only).

Restates rbf_kernel.py:41-54 (synthetic_coords, NumPy PCG64 in place of
CuPy's generator), rbf_kernel.py:57-68 (_estimate_gamma), rbf_kernel.py:158-227
(build_rbf_kernel, GPU path), rbf_kernel.py:230-252 (weather_front), mpdok_ooc.py:39-46,118-123 (_build_fp32_tile and
the FP32-after-cast nugget) and mpdok_ooc.py:205-227 (FP64 tile recompute).
"""
import numpy as np


def gram_d2(X, rows=None, cols=None):
    """D2[i, j] = (sq_i + sq_j) - 2 c_i.c_j clamped at 0 (rbf_kernel.py:219-220)."""
    X = np.asarray(X, np.float64)
    sq = np.sum(X ** 2, axis=1)
    R = X if rows is None else X[rows]
    C = X if cols is None else X[cols]
    sr = sq if rows is None else sq[rows]
    sc = sq if cols is None else sq[cols]
    D2 = sr[:, None] + sc[None, :] - 2.0 * (R @ C.T)
    np.maximum(D2, 0.0, out=D2)
    return D2


def kernel_from_d2(D2, kind="gaussian", param=1.0):
    """Gaussian exp(-gamma d2) (rbf_kernel.py:221) or Matern-3/2
    (1 + a) exp(-a), a = sqrt(3 d2)/l (north-star kernel, no reference code)."""
    if kind == "gaussian":
        return np.exp(-param * D2)
    a = np.sqrt(3.0 * D2) / param
    return (1.0 + a) * np.exp(-a)


def build_kernel(X, kind="gaussian", param=1.0, reg=0.0):
    """Full FP64 matrix with the nugget added in FP64 (rbf_kernel.py:219-224)."""
    A = kernel_from_d2(gram_d2(X), kind, param)
    A[np.diag_indices_from(A)] += reg
    return A


def build_fp32_tile(X, i, rows, kind="gaussian", param=1.0, reg=0.0):
    """Rows [i, i+rows) in FP64, cast to FP32, nugget added in FP32 after the
    cast (mpdok_ooc.py:39-46 then :121-122)."""
    N = X.shape[0]
    D2 = gram_d2(X, rows=np.arange(i, i + rows))
    tile = kernel_from_d2(D2, kind, param).astype(np.float32)
    k = np.arange(rows)
    tile[k, i + k] += np.float32(reg)
    assert tile.shape == (rows, N)
    return tile


def estimate_gamma(X, max_sample=512):
    """1 / (2 mean d2) over the first min(512, N) points including the zero
    diagonal; 1.0 when the mean is <= 1e-30 (rbf_kernel.py:57-68)."""
    n = min(max_sample, X.shape[0])
    D2 = gram_d2(np.asarray(X)[:n])
    m = float(np.mean(D2))
    return 1.0 / (2.0 * m) if m > 1e-30 else 1.0


def uniform_cube_gamma(D, L, n):
    """Known answer: for n points uniform in [0, L]^D, E[d^2] over distinct
    pairs is D L^2 / 6; with the n zero diagonal terms in the mean,
    E[mean d2] = D L^2 / 6 * (1 - 1/n), so gamma ~ 3 / (D L^2 (1 - 1/n))."""
    return 3.0 / (D * L * L * (1.0 - 1.0 / n))


def matvec_fp64(X, v, kind="gaussian", param=1.0, reg=0.0, tile_rows=4096):
    """y = A v with FP64 tiles recomputed from the coordinates, nugget in FP64
    (mpdok_ooc.py:205-227)."""
    N = X.shape[0]
    y = np.zeros(N)
    for i in range(0, N, tile_rows):
        rows = min(tile_rows, N - i)
        tile = kernel_from_d2(gram_d2(X, rows=np.arange(i, i + rows)), kind, param)
        k = np.arange(rows)
        tile[k, i + k] += reg
        y[i:i + rows] = tile @ v
    return y


def synthetic_coords(N, D=2, seed=42, scale=100.0):
    """N points uniform in [0, scale]^D (rbf_kernel.py:41-54). The reference
    draws them from CuPy's default_rng, whose stream NumPy cannot reproduce;
    this restatement (and the product) use NumPy's PCG64 default_rng(seed)."""
    return np.random.default_rng(seed).uniform(0.0, scale, size=(N, D)).astype(np.float64)


def weather_front(X, t, noise_std=0.05, seed=None):
    """15 + 5 sin((x+y)/20 - 0.5 t) + noise (rbf_kernel.py:246-252), noise
    from NumPy default_rng(t or seed)."""
    X = np.asarray(X)
    f = 15.0 + 5.0 * np.sin((X[:, 0] + X[:, 1]) / 20.0 - 0.5 * t)
    if noise_std > 0:
        f = f + np.random.default_rng(t if seed is None else seed).standard_normal(f.shape) * noise_std
    return f
