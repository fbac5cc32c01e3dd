"""Out-of-core solver (mpdok_ooc.py) on the GPU against the CPU oracle:
tile build (RAM and on-disk stores), streamed and matrix-free products, the
DGKS inner GMRES and the outer refinement."""
import os

import numpy as np
import pytest
import torch

from oracle import krylov as okr
from oracle import rbf as orb

pytestmark = pytest.mark.gpu


def _solver(tile_rows):
    from mpdok_amd.mpdok_ops import _dev  # noqa: F401  (loads the library first)
    from mpdok_amd.mpdok_ooc import MPDOKOOCSolver
    return MPDOKOOCSolver(tile_rows=tile_rows)


def _pts(N, D, seed):
    return np.random.default_rng(seed).uniform(0.0, 1.0, (N, D))


@pytest.mark.parametrize("N,tile_rows", [(1000, 256), (777, 1024), (513, 100)])
def test_ram_build_matches_oracle_tiles(N, tile_rows):
    """FP32 tiles (nugget after the cast, mpdok_ooc.py:118-123) within one
    FP32 ulp of the oracle's; ragged last tile included."""
    X = _pts(N, 2, N)
    s = _solver(tile_rows)
    g = s.build(X, gamma=30.0, reg=1e-3, store="ram", verbose=False)
    assert g == 30.0 and s.N == N and s._ram_buf.shape == (N, N)
    assert s._ram_buf.dtype == np.float32 and s._ram_buf.flags["C_CONTIGUOUS"]
    ref = orb.build_fp32_tile(X, 0, N, "gaussian", 30.0, 1e-3)
    ulp = np.spacing(np.abs(ref))
    assert np.all(np.abs(s._ram_buf - ref) <= ulp)
    s.free()


def test_ssd_store_is_raw_row_major_fp32(tmp_path):
    """The on-disk store is raw FP32, row-major, no header (mpdok_ooc.py:142),
    byte-identical to the RAM store."""
    N = 600
    X = _pts(N, 3, 7)
    ram = _solver(128)
    ram.build(X, gamma=20.0, reg=1e-4, store="ram", verbose=False)
    path = str(tmp_path / "A_fp32.bin")
    ssd = _solver(128)
    ssd.build(X, gamma=20.0, reg=1e-4, store="ssd", path=path, verbose=False)
    assert os.path.getsize(path) == N * N * 4
    disk = np.fromfile(path, dtype=np.float32).reshape(N, N)
    assert np.array_equal(disk, ram._ram_buf)
    v = np.random.default_rng(1).standard_normal(N).astype(np.float32)
    y_ram = ram._tiled_sgemv(torch.as_tensor(v).cuda()).cpu().numpy()
    y_ssd = ssd._tiled_sgemv(torch.as_tensor(v).cuda()).cpu().numpy()
    assert np.array_equal(y_ram, y_ssd)       # same tiles, same kernel, same order
    ref = ram._ram_buf.astype(np.float64) @ v.astype(np.float64)
    assert np.allclose(y_ram, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
    ram.free()
    ssd.free()


def test_bring_your_own_matrix_recipe():
    """The README recipe (RBF_README.md:229-248): attributes set by hand, a
    plain (pageable) NumPy matrix, FP64 residual from the coordinates."""
    N = 700
    X = _pts(N, 2, 3)
    A32 = orb.build_fp32_tile(X, 0, N, "gaussian", 15.0, 1e-2)
    s = _solver(200)
    s.N, s.gamma, s.reg = N, 15.0, 1e-2
    s._store = "ram"
    s._ram_buf = np.ascontiguousarray(A32)
    s._coords = torch.as_tensor(X).cuda()
    s._sq = (s._coords ** 2).sum(dim=1)
    v = np.random.default_rng(2).standard_normal(N).astype(np.float32)
    y = s._tiled_sgemv(torch.as_tensor(v).cuda()).cpu().numpy()
    ref = A32.astype(np.float64) @ v.astype(np.float64)
    assert np.allclose(y, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
    b = np.sin(3 * X[:, 0]) + X[:, 1]
    x = s.solve(b, tol=1e-9, maxiter_outer=20, restart=50, verbose=False).cpu().numpy()
    A64 = orb.build_kernel(X, "gaussian", 15.0, 1e-2)
    assert np.linalg.norm(b - A64 @ x) / np.linalg.norm(b) < 1e-9


@pytest.mark.parametrize("N,D", [(1, 3), (255, 2), (256, 3), (513, 1), (700, 3), (300, 5)])
@pytest.mark.parametrize("kind,param", [(0, 25.0), (1, 0.1)])
def test_symmetric_free_product_matches_row_product(N, D, kind, param):
    """mpdok_rbf_symv (each kernel value evaluated once, used for E_ij and
    E_ji) against mpdok_rbf_matvec (each row on its own): the same elements
    summed in another order, within the summation bound; the same result
    bit for bit on every call; a short workspace refused."""
    _solver(256)
    from mpdok_amd import _dev
    from mpdok_amd._lib import F32, F64, check, lib
    X = _pts(N, D, 5)
    c = torch.as_tensor(X).cuda().contiguous()
    E = orb.build_kernel(X, "gaussian" if kind == 0 else "matern32", param, 1e-4)
    rng = np.random.default_rng(N)
    nb = lib().mpdok_rbf_symv_workspace_bytes(N)
    assert nb == ((N + 255) // 256) ** 2 * 256 * 8
    ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
    for elem in (F64, F32):
        for dt in (torch.float64, torch.float32):
            xy = F64 if dt == torch.float64 else F32
            v = torch.as_tensor(rng.standard_normal(N)).to(dt).cuda()
            ys = [torch.empty(N, dtype=dt, device="cuda") for _ in range(3)]
            for y in ys[:2]:
                check(lib().mpdok_rbf_symv(_dev.stream(), c.data_ptr(), N, D, kind, param, 1e-4,
                                           elem, v.data_ptr(), y.data_ptr(), xy, ws.data_ptr(),
                                           nb), "rbf_symv")
            check(lib().mpdok_rbf_matvec(_dev.stream(), c.data_ptr(), N, D, kind, param, 1e-4,
                                         elem, v.data_ptr(), ys[2].data_ptr(), xy, 0, N),
                  "rbf_matvec")
            a, b, r = (y.double().cpu().numpy() for y in ys)
            assert np.array_equal(a, b)
            # two FP64 sums of the same N products: (N + 1) u each; then the
            # rounding of each result to the output type
            uo = 2.0 ** -53 if dt == torch.float64 else 2.0 ** -24
            bound = np.abs(E) @ np.abs(v.double().cpu().numpy())
            assert np.all(np.abs(a - r) <= (2 * (N + 1) * 2.0 ** -53 + 2 * uo) * bound * 1.01)
    assert lib().mpdok_rbf_symv(_dev.stream(), c.data_ptr(), N, D, kind, param, 1e-4, F64,
                                v.data_ptr(), ys[0].data_ptr(), F32, ws.data_ptr(), nb - 8) != 0


@pytest.mark.parametrize("kind,param", [("gaussian", 25.0), ("matern32", 0.1)])
def test_matrix_free_products_match_oracle(kind, param):
    """FP64 product (mpdok_ooc.py:205-227) and the FP32-element product agree
    with the oracle within summation-order bounds."""
    N = 900
    X = _pts(N, 3, 11)
    s = _solver(256)
    if kind == "gaussian":
        s.build(X, gamma=param, reg=1e-4, store="free", verbose=False)
    else:
        s.build(X, reg=1e-4, store="free", verbose=False, kernel="matern32", length_scale=param)
    v = np.random.default_rng(4).standard_normal(N)
    y64 = s._tiled_dgemv_fp64(torch.as_tensor(v).cuda()).cpu().numpy()
    ref64 = orb.matvec_fp64(X, v, kind, param, 1e-4)
    absA = np.abs(orb.build_kernel(X, kind, param, 1e-4)) @ np.abs(v)
    # element error <= 4 u gamma 2 D (Gram identity) + 1 ulp (exp); sum error N u
    atol = (4 * 2.0 ** -53 * (param if kind == "gaussian" else 1e3) * 6 + (N + 2) * 2.0 ** -53)
    assert np.all(np.abs(y64 - ref64) <= atol * absA + 1e-300)
    y32 = s._tiled_sgemv(torch.as_tensor(v.astype(np.float32)).cuda()).cpu().numpy()
    T = orb.build_fp32_tile(X, 0, N, kind, param, 1e-4).astype(np.float64)
    ref32 = T @ v.astype(np.float32).astype(np.float64)
    assert np.all(np.abs(y32 - ref32) <= 2.0 ** -23 * np.abs(T) @ np.abs(v) * 4 + 1e-30)
    s.free()


def test_inner_gmres_matches_oracle_and_solve_converges():
    # unpreconditioned FP32 inner GMRES(50), as the reference's: a system it
    # can refine to 1e-10 (kappa ~ 1e3; the reference's default tol is 1e-5)
    N = 1500
    X = _pts(N, 2, 21)
    gamma, reg = 40.0, 5e-2
    s = _solver(512)
    s.build(X, gamma=gamma, reg=reg, store="ram", verbose=False)
    A32 = s._ram_buf.copy()
    A64 = orb.build_kernel(X, "gaussian", gamma, reg)
    b = np.cos(4 * X[:, 0]) * X[:, 1] + 0.1
    # one inner solve from the exact FP64 residual of x = 0
    e = s._inner_gmres(torch.as_tensor(b).cuda(), 50).cpu().numpy()
    e_ref = okr.inner_gmres(lambda v: A32 @ v, b, 50)
    assert np.linalg.norm(e - e_ref) <= 1e-3 * np.linalg.norm(e_ref)
    calls = []
    orig = s._tiled_dgemv_fp64

    def counted(v):
        calls.append(1)
        return orig(v)
    s._tiled_dgemv_fp64 = counted          # instance patch, called once per outer step
    x = s.solve(b, tol=1e-10, maxiter_outer=20, restart=50, verbose=False).cpu().numpy()
    rel = np.linalg.norm(b - A64 @ x) / np.linalg.norm(b)
    assert rel < 1e-10
    x_ref = np.linalg.solve(A64, b)
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-6
    assert 2 <= len(calls) <= 20
    s.free()


def test_zero_rhs_and_not_converged_warning():
    N = 300
    X = _pts(N, 2, 5)
    s = _solver(128)
    s.build(X, gamma=10.0, reg=1e-2, store="ram", verbose=False)
    x = s.solve(np.zeros(N), verbose=False)
    assert torch.count_nonzero(x).item() == 0 and x.dtype == torch.float64
    with pytest.warns(RuntimeWarning, match="did not converge"):
        s.solve(np.ones(N), tol=1e-30, maxiter_outer=2, restart=5, verbose=False)
    s.free()


@pytest.mark.parametrize("kind,param,reg", [("matern32", 0.1, 1e-4), ("gaussian", 25.0, 1e-6)])
def test_kernel_diagonal_matches_the_stored_tiles(kind, param, reg):
    """mpdok_rbf_diag: the FP32 tile's diagonal (nugget after the cast) bit
    for bit as the assembly writes it, and its reciprocal."""
    from mpdok_amd import _dev, _ops
    N = 1000
    X = _pts(N, 3, 21)
    c = _dev.as_device(X, torch.float64)
    k = 0 if kind == "gaussian" else 1
    tile = _ops.rbf_assemble(c, k, param, reg, _dev.empty((N, N), torch.float32, "C"),
                             reg_mode=1)
    d = _ops.rbf_diag(c, k, param, reg, torch.float32, torch.float32).cpu().numpy()
    assert np.array_equal(d, np.diag(tile.cpu().numpy()))
    ref = np.diag(orb.build_fp32_tile(X, 0, N, kind, param, reg))
    assert np.array_equal(d, ref)
    dinv = _ops.rbf_diag(c, k, param, reg, torch.float32, torch.float32, invert=True)
    assert np.array_equal(dinv.cpu().numpy(), np.float32(1.0) / ref)
    d64 = _ops.rbf_diag(c, k, param, reg, torch.float64, torch.float64).cpu().numpy()
    assert np.allclose(d64, 1.0 + reg, rtol=1e-15, atol=0)     # k(0) = 1 for both kernels


@pytest.mark.parametrize("store", ["free", "ram"])
def test_jacobi_cg_refinement_matches_oracle(store):
    """Jacobi-preconditioned FP32 CG inside the FP64 refinement loop
    (mpdok_ooc.py:315-373), Matern-3/2 with length scale 0.1 and nugget 1e-4
    in [0,1]^3 (the out-of-core configuration's kernel) at N = 4096, to relres
    1e-10: the same number of outer steps as the oracle's cg_ir, inner
    iteration counts within 10%, and the two solutions within the bound that
    relres 1e-10 leaves (kappa * 2e-10)."""
    N, ell, reg = 4096, 0.1, 1e-4
    X = _pts(N, 3, 42)
    b = np.sin(2 * np.pi * X[:, 0]) * np.cos(2 * np.pi * X[:, 1]) * np.cos(2 * np.pi * X[:, 2])
    s = _solver(1024)
    s.build(X, reg=reg, store=store, verbose=False, kernel="matern32", length_scale=ell)
    x = s.solve(b, tol=1e-10, maxiter_outer=10, inner="cg", precond="jacobi",
                verbose=False).cpu().numpy()
    info = s.last_info
    assert info["converged"] and info["history"][-1] < 1e-10
    A = orb.build_kernel(X, "matern32", ell, reg)
    A32 = orb.build_fp32_tile(X, 0, N, "matern32", ell, reg)
    dinv = np.float32(1.0) / np.diag(A32)
    xo, convo, histo, innero = okr.cg_ir(lambda v: A32 @ v.astype(np.float32), lambda v: A @ v,
                                          b, dinv=dinv, tol=1e-10)
    assert convo
    assert len(info["history"]) == len(histo)
    assert all(abs(a - o) <= 0.1 * o for a, o in zip(info["inner_iterations"], innero))
    r = b - A @ x
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-10
    # kappa <= ||A||_inf / nugget (the smallest eigenvalue is at least the nugget)
    kappa = np.abs(A).sum(axis=1).max() / reg
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < kappa * 2e-10
    s.free()


def test_jacobi_cg_with_a_caller_supplied_matrix():
    """The "bring your own matrix" recipe: _ram_buf set by the caller, so the
    Jacobi diagonal comes from that host matrix."""
    N = 700
    X = _pts(N, 2, 5)
    sc = np.sqrt(np.linspace(1.0, 2.0, N))              # S K S stays SPD, diag not constant
    A = orb.build_kernel(X, "gaussian", 30.0, 1e-3) * sc[:, None] * sc[None, :]
    b = np.cos(3 * X[:, 0])
    s = _solver(256)
    s.build(X, gamma=30.0, reg=1e-3, store="ram", verbose=False)
    s._ram_buf = A.astype(np.float32)
    s._tiled_dgemv_fp64 = lambda v: torch.as_tensor(A @ v.cpu().numpy(), device=v.device)
    x = s.solve(b, tol=1e-10, inner="cg", precond="jacobi", verbose=False).cpu().numpy()
    assert s.last_info["converged"]
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-10
    s.free()
