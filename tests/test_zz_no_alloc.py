"""The allocation-free residual (mpdok_residual_f64_noalloc) against the
workspace entry point (mpdok_residual_f64), through the C-ABI: same bits on
two calls in a row (the second smaller, so stale scratch contents from the
first would show), both storage orders and both accumulation modes; and the
host-side rejection of a size beyond the scratch array."""
import ctypes
import os

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "mpdok_amd", "libmpdok_hip.so")
SCRATCH = 16 << 20


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_rejects_sizes_beyond_the_scratch_before_any_gpu_work():
    from mpdok_amd import _lib
    L = _lib.load()
    # column-major N = 2^21 needs 2^21 * 16 B = 32 MiB of partial sums
    N = 1 << 21
    assert L.mpdok_residual_workspace_bytes(N, 1) > SCRATCH
    fake = 0x1000          # never dereferenced: the size check comes first
    st = L.mpdok_residual_f64_noalloc(None, fake, N, 1, N, fake, fake, fake, 1)
    assert st != 0
    assert b"needs" in L.mpdok_last_error()
    assert L.mpdok_residual_workspace_bytes(1 << 20, 1) <= SCRATCH


def _noalloc(L, A, order, x, b, mode):
    from mpdok_amd import _dev
    N = b.numel()
    r = torch.full((N,), float("nan"), dtype=torch.float64, device=b.device)
    ld = A.stride(1) if order == 1 else A.stride(0)
    st = L.mpdok_residual_f64_noalloc(_dev.stream(), A.data_ptr(), ld, order, N, x.data_ptr(),
                                      b.data_ptr(), r.data_ptr(), mode)
    assert st == 0, L.mpdok_last_error()
    return r


def _with_workspace(L, A, order, x, b, mode):
    from mpdok_amd import _dev
    N = b.numel()
    r = torch.full((N,), float("nan"), dtype=torch.float64, device=b.device)
    ld = A.stride(1) if order == 1 else A.stride(0)
    nb = L.mpdok_residual_workspace_bytes(N, order)
    ws = torch.empty(max(nb, 8), dtype=torch.uint8, device=b.device)
    st = L.mpdok_residual_f64(_dev.stream(), A.data_ptr(), ld, order, N, x.data_ptr(),
                              b.data_ptr(), r.data_ptr(), mode, ws.data_ptr(), nb)
    assert st == 0, L.mpdok_last_error()
    return r


@pytest.mark.gpu
@pytest.mark.parametrize("order", [1, 0])
@pytest.mark.parametrize("mode", [0, 1])
def test_noalloc_residual_matches_workspace_entry_point(order, mode):
    from mpdok_amd import _dev, _lib
    L = _lib.load()
    for N, seed in [(3000, 1), (1000, 2), (513, 3)]:
        rng = np.random.default_rng(seed)
        A_h = rng.standard_normal((N, N)) + np.eye(N) * 4
        x_h, b_h = rng.standard_normal(N), rng.standard_normal(N)
        A = _dev.as_device(A_h, torch.float64, "F" if order == 1 else "C")
        x, b = _dev.as_device(x_h), _dev.as_device(b_h)
        r1 = _noalloc(L, A, order, x, b, mode)
        r2 = _with_workspace(L, A, order, x, b, mode)
        torch.cuda.synchronize()
        assert torch.equal(r1, r2), (N, order, mode)
        ref = b_h - A_h @ x_h
        assert np.allclose(r1.cpu().numpy(), ref, rtol=0, atol=1e-11)


@pytest.mark.gpu
def test_noalloc_residual_exact_on_integer_data_twice():
    """Integer data whose products and sums are exact in FP64: the
    allocation-free residual equals b - A x exactly, at N = 8192 (64 column
    chunks, 8 MiB of partial sums) and then N = 4096 on the same scratch."""
    from mpdok_amd import _dev, _lib
    L = _lib.load()
    for n in (8192, 4096):
        rng = np.random.default_rng(n)
        A_h = rng.integers(-3, 4, size=(n, n)).astype(np.float64)
        x_h = rng.integers(-3, 4, size=n).astype(np.float64)
        b_h = rng.integers(-100, 100, size=n).astype(np.float64)
        A = _dev.as_device(A_h, torch.float64, "F")
        x, b = _dev.as_device(x_h), _dev.as_device(b_h)
        r = _noalloc(L, A, 1, x, b, 0)
        assert np.array_equal(r.cpu().numpy(), b_h - A_h @ x_h)
        assert torch.equal(r, _with_workspace(L, A, 1, x, b, 0))
