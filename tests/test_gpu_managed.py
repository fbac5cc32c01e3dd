"""Managed-memory entry points (mpdok_ops.py:34-66,147-174,388-418 and the
managed branch of build_rbf_kernel, rbf_kernel.py:187-205): allocate, build
into the managed matrix, solve, free; the one-slot-per-library rule; and the
host build into the managed pages. Sizes stay far below device memory (no
oversubscription: XNACK is off on these GPUs)."""
import gc

import numpy as np
import pytest
import scipy.linalg as sla
import torch

from oracle import rbf as orb

pytestmark = pytest.mark.gpu


def _system(N, seed=3):
    X = np.random.default_rng(seed).uniform(0, 1, (N, 2))
    b = np.sin(2 * np.pi * X[:, 0]) * np.cos(2 * np.pi * X[:, 1])
    return X, b


@pytest.mark.parametrize("which", ["gmres", "lu"])
def test_alloc_build_solve_free(which):
    from mpdok_amd.mpdok_ops import LUIRSolver, MPDOKSolver, _ManagedArray
    from mpdok_amd.rbf_kernel import build_rbf_kernel
    N, gamma, reg = 700, 10.0, 1e-2
    X, b = _system(N)
    s = MPDOKSolver() if which == "gmres" else LUIRSolver()
    A = s.alloc_managed(N)
    assert isinstance(A, _ManagedArray)
    assert tuple(A.shape) == (N, N) and A.dtype == torch.float64
    assert A.array.stride() == (1, N)                          # Fortran order
    A_out, g = build_rbf_kernel(X, gamma=gamma, reg=reg, out=A)
    assert g == gamma and A_out.data_ptr() == A.array.data_ptr()
    ref = orb.build_kernel(X, "gaussian", gamma, reg)
    assert np.allclose(A.array.cpu().numpy(), ref, rtol=1e-13, atol=1e-300)
    x = s.solve(A, b, tol=1e-11).cpu().numpy()
    xr = sla.solve(ref, b)
    assert np.linalg.norm(b - ref @ x) / np.linalg.norm(b) < 1e-11
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-9
    s.free_managed()
    assert A._freed
    A.free()                                                   # second free: no-op
    if which == "gmres":
        s.free_fp32_buf()


def test_second_allocation_disarms_the_first_wrapper():
    """The library frees the previous buffer when it allocates the next: the
    old wrapper is marked freed, so dropping it never frees the live one."""
    from mpdok_amd.mpdok_ops import MPDOKSolver
    from mpdok_amd.rbf_kernel import build_rbf_kernel
    s1, s2 = MPDOKSolver(), MPDOKSolver()
    A1 = s1.alloc_managed(64)
    A2 = s1.alloc_managed(96)
    assert A1._freed and not A2._freed
    del A1
    gc.collect()
    X, _ = _system(96, seed=5)
    build_rbf_kernel(X, gamma=4.0, reg=1e-3, out=A2)
    assert np.allclose(A2.array.cpu().numpy(), orb.build_kernel(X, "gaussian", 4.0, 1e-3),
                       rtol=1e-13, atol=1e-300)
    # another solver on the same library shares the slot
    A3 = s2.alloc_managed(80)
    assert A2._freed and not A3._freed
    s1.free_managed()                  # its wrapper was disarmed: frees nothing
    X3, _ = _system(80, seed=6)
    build_rbf_kernel(X3, gamma=4.0, reg=1e-3, out=A3)
    assert np.allclose(A3.array.cpu().numpy(), orb.build_kernel(X3, "gaussian", 4.0, 1e-3),
                       rtol=1e-13, atol=1e-300)
    s2.free_managed()
    assert A3._freed
    # the LU slot is separate from the GMRES slot
    from mpdok_amd.mpdok_ops import LUIRSolver
    lu = LUIRSolver()
    B = lu.alloc_managed(32)
    C = s2.alloc_managed(32)
    assert not B._freed and not C._freed
    lu.free_managed()
    s2.free_managed()


def test_host_build_branch_into_managed_pages(monkeypatch):
    """Forced through the threshold (the reference switches when the FP64
    matrix exceeds 90% of device memory): the host view of the managed
    buffer receives exactly the device build's values."""
    from mpdok_amd import rbf_kernel
    from mpdok_amd.mpdok_ops import LUIRSolver
    N, gamma, reg = 600, 25.0, 1e-6
    X = np.random.default_rng(8).uniform(0, 1, (N, 3))
    s = LUIRSolver()
    A = s.alloc_managed(N)
    monkeypatch.setattr(rbf_kernel, "_HOST_BUILD_FRACTION", 0.0)
    A_out, g = rbf_kernel.build_rbf_kernel(X, gamma=gamma, reg=reg, out=A)
    torch.cuda.synchronize()
    dev_build, _ = rbf_kernel.build_rbf_kernel(X, gamma=gamma, reg=reg)
    assert np.array_equal(A.array.cpu().numpy(), dev_build.cpu().numpy())
    s.free_managed()
