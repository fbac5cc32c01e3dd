"""Out-of-core GMRES-IR for kernel systems larger than the GPU (mpdok_ooc.py).

Same class, methods, attributes and defaults as the reference's
``MPDOKOOCSolver`` (mpdok_ooc.py:49-383). The FP32 working matrix lives in
host memory (``store='ram'``: a page-locked N x N array; ``store='ssd'``: a raw
row-major FP32 file with no header, byte-compatible with the reference's) and
crosses PCIe tile by tile. The FP64 operator of the outer residual is never
stored: its tiles are recomputed from the coordinates by a fused kernel.

What runs where:
  build             GPU assembles each FP32 tile (nugget added after the cast,
                    mpdok_ooc.py:118-123) and copies it to host memory, the copy
                    of one tile overlapping the assembly of the next
                    (mpdok_rbf_build_host).
  _tiled_sgemv_*    double-buffered host-to-device streaming GEMV: the copy of
                    tile t+1 on a library copy stream overlaps the product of
                    tile t (mpdok_stream_gemv_f32). The on-disk store is read by
                    a background thread into two page-locked buffers.
  _tiled_dgemv_fp64 matrix-free FP64 product, kernel values recomputed in
                    registers, each once for both symmetric entries
                    (mpdok_rbf_symv, FP64 elements).
  _inner_gmres      FP32 GMRES with two classical Gram-Schmidt passes (DGKS) and
                    Givens rotations on the device; stop at |g_k+1|/beta < 1e-6.

Extension (not in the reference): ``store='free'`` keeps nothing but the
coordinates and recomputes the FP32 tile values inside the inner product as
well (mpdok_rbf_symv, FP32 elements: bit for bit the stored tile's values),
``kernel='matern32'`` with ``length_scale`` selects the Matern-3/2 kernel, and
``solve(..., inner='cg', precond='jacobi')`` refines with a Jacobi-preconditioned
FP32 CG as the inner solver (SPD kernels; the FP64 residual, update and stopping
test stay as in the reference's outer loop).

Deviation kept from the other solvers here: on an Arnoldi breakdown the
rotation of the last Hessenberg column is applied before stopping (the
reference exits first, mpdok_ooc.py:278-282, and back-substitutes with an
unrotated diagonal entry).
"""
import os
import threading
import time
import warnings

import numpy as np
import torch

from . import _dev, _ops
from ._lib import check, lib, F32, F64

GAUSSIAN, MATERN32 = 0, 1


class MPDOKOOCSolver:
    """Out-of-core GMRES-IR for N beyond what fits in device memory.

    VRAM during solve: two FP32 tiles (2 x tile_rows x N x 4 B) and the Krylov
    basis ((restart + 1) x N x 4 B).
    """

    def __init__(self, tile_rows=4096):
        self.tile_rows = tile_rows
        self._store = None       # 'ram' | 'ssd' | 'free'
        self._ram_buf = None     # numpy (N, N) C-order FP32 (page-locked when built here)
        self._ssd_path = None    # str: raw FP32 file, row-major, no header
        self.N = None
        self.gamma = None
        self.reg = None
        self._coords = None      # device (N, D) FP64
        # The reference keeps the squared norms of the coordinates here (its
        # "bring your own matrix" recipe sets them). The fused kernels here
        # recompute them per tile, so the attribute is accepted and unused.
        self._symv_ws = None
        self._kind = GAUSSIAN
        self._pinned = None      # torch tensor owning the page-locked _ram_buf
        self._dev_tiles = None   # device workspace: two tiles
        self._stage = None       # two page-locked tile buffers for the on-disk store
        self.last_info = None    # outer history and inner iterations of the last solve

    # ---- build (mpdok_ooc.py:75-154) ----------------------------------------
    def build(self, coords, gamma=None, reg=1e-6, store='ram', path=None, verbose=True,
              kernel='gaussian', length_scale=None):
        """Compute and cache the FP32 kernel matrix; returns gamma (for
        ``kernel='matern32'`` the length scale)."""
        from .rbf_kernel import _estimate_gamma
        if store not in ('ram', 'ssd', 'free'):
            raise ValueError(f"store must be 'ram' or 'ssd', got {store!r}")
        if store == 'ssd' and path is None:
            raise ValueError("store='ssd' requires a path argument")
        coords = _dev.as_device(coords, torch.float64)
        N = coords.shape[0]
        if kernel == 'gaussian':
            if gamma is None:
                gamma = _estimate_gamma(coords)
            self._kind, param = GAUSSIAN, float(gamma)
        elif kernel == 'matern32':
            if length_scale is None:
                raise ValueError("kernel='matern32' needs length_scale")
            self._kind, param = MATERN32, float(length_scale)
            gamma = param
        else:
            raise ValueError(f"unknown kernel {kernel!r}")

        self.free()
        self.N, self.gamma, self.reg = N, gamma, reg
        self._store = store
        self._coords = coords
        tr = self.tile_rows
        n_tiles = (N + tr - 1) // tr
        if verbose:
            print(f"  OOC build: N={N:,}  gamma={gamma:.3e}  "
                  f"FP32={N * N * 4 / 1024**3:.2f} GB  tiles={n_tiles}  store={store}")
        t0 = time.perf_counter()
        if store == 'ram':
            self._pinned = torch.empty((N, N), dtype=torch.float32, pin_memory=True)
            self._ram_buf = self._pinned.numpy()
            self._build_rows(0, N, self._pinned.data_ptr(), N)
            torch.cuda.current_stream().synchronize()
            self._ssd_path = None
        elif store == 'ssd':
            self._ssd_path = path
            self._ram_buf = None
            stage = self._staging()
            with open(path, 'wb') as fh:
                for ci, i in enumerate(range(0, N, tr)):
                    rows = min(tr, N - i)
                    buf = stage[ci % 2]
                    self._build_rows(i, rows, buf.data_ptr(), N)
                    torch.cuda.current_stream().synchronize()
                    fh.write(memoryview(buf.numpy()[:rows]).cast('B'))
                    if verbose:
                        print(f"    tile {ci + 1}/{n_tiles}  ({(ci + 1) * 100 // n_tiles}%)",
                              end='\r', flush=True)
        else:
            self._ssd_path = None
            self._ram_buf = None
        if verbose:
            print(f"\n  build done in {time.perf_counter() - t0:.1f} s")
        return gamma

    def _param(self):
        if self.gamma is None:
            raise TypeError("gamma is None: set it, or override _tiled_dgemv_fp64")
        return float(self.gamma)

    def _coords_dev(self):
        c = _dev.as_device(self._coords, torch.float64)
        if c.dim() != 2 or c.shape[0] != self.N:
            raise ValueError(f"_coords must be ({self.N}, D), got {tuple(c.shape)}")
        return c.contiguous()

    def _workspace(self):
        N, tr = self.N, min(self.tile_rows, self.N)
        nbytes = lib().mpdok_stream_gemv_workspace_bytes(tr, N)
        if self._dev_tiles is None or self._dev_tiles.numel() < nbytes:
            self._dev_tiles = torch.empty(nbytes, dtype=torch.uint8, device=_dev.device())
        return self._dev_tiles, nbytes, tr

    def _staging(self):
        N, tr = self.N, min(self.tile_rows, self.N)
        if self._stage is None or self._stage[0].shape != (tr, N):
            self._stage = [torch.empty((tr, N), dtype=torch.float32, pin_memory=True)
                           for _ in range(2)]
        return self._stage

    def _build_rows(self, row0, rows, host_ptr, ldh):
        ws, nbytes, tr = self._workspace()
        c = self._coords_dev()
        check(lib().mpdok_rbf_build_host(_dev.stream(), c.data_ptr(), self.N, c.shape[1],
                                         self._kind, self._param(), float(self.reg), row0, rows,
                                         tr, host_ptr, ldh, ws.data_ptr(), nbytes),
              "rbf_build_host")

    # ---- FP32 products (mpdok_ooc.py:158-201) ------------------------------
    def _stream_rows(self, host_ptr, ld, rows, v32, y32):
        ws, nbytes, tr = self._workspace()
        check(lib().mpdok_stream_gemv_f32(_dev.stream(), host_ptr, ld, rows, self.N, tr,
                                          v32.data_ptr(), y32.data_ptr(), ws.data_ptr(), nbytes),
              "stream_gemv")

    def _tiled_sgemv_ram(self, v_fp32):
        """y_fp32 = A_fp32 v, streaming the RAM copy through two device tiles."""
        N = self.N
        v = _dev.as_device(v_fp32, torch.float32).contiguous()
        y = torch.empty(N, dtype=torch.float32, device=v.device)
        buf = self._ram_buf
        if (self._pinned is not None and isinstance(buf, np.ndarray)
                and buf.ctypes.data == self._pinned.data_ptr()):
            self._stream_rows(self._pinned.data_ptr(), N, N, v, y)
            return y
        # a caller-supplied array (the reference's "bring your own matrix"
        # recipe): staged through page-locked buffers, chunk by chunk
        A = np.ascontiguousarray(buf, dtype=np.float32)
        if A.shape != (N, N):
            raise ValueError(f"_ram_buf must be ({N}, {N}), got {A.shape}")
        return self._stream_from(lambda i, rows, out: np.copyto(out[:rows], A[i:i + rows]), v, y)

    def _tiled_sgemv_ssd(self, v_fp32):
        """y_fp32 = A_fp32 v, reading the on-disk copy (raw FP32, row-major)
        with a background reader thread into two page-locked buffers."""
        N = self.N
        v = _dev.as_device(v_fp32, torch.float32).contiguous()
        y = torch.empty(N, dtype=torch.float32, device=v.device)
        fd = os.open(self._ssd_path, os.O_RDONLY)
        try:
            size = os.fstat(fd).st_size
            if size != N * N * 4:
                raise ValueError(f"{self._ssd_path}: {size} bytes, expected {N * N * 4}")

            def read(i, rows, out):
                view = memoryview(out[:rows]).cast('B')
                off, done = i * N * 4, 0
                while done < len(view):
                    n = os.preadv(fd, [view[done:]], off + done)
                    if n <= 0:
                        raise IOError(f"short read from {self._ssd_path}")
                    done += n
            return self._stream_from(read, v, y)
        finally:
            os.close(fd)

    def _stream_from(self, fill, v, y):
        """Host-side double buffering: a reader thread fills page-locked
        buffer (t + 1) % 2 while tile t is copied and multiplied; a buffer is
        refilled only after the GPU has finished reading it."""
        N = self.N
        tr = min(self.tile_rows, N)
        stage = self._staging()
        host = [s.numpy() for s in stage]
        done_evt = [None, None]
        starts = list(range(0, N, tr))
        err = []

        def reader(t):
            try:
                i = starts[t]
                fill(i, min(tr, N - i), host[t % 2])
            except BaseException as exc:  # surfaced in the caller
                err.append(exc)

        th = threading.Thread(target=reader, args=(0,))
        th.start()
        for t, i in enumerate(starts):
            rows = min(tr, N - i)
            th.join()
            if err:
                raise err[0]
            b = t % 2
            # next tile into the other buffer, once the GPU is done with it
            if t + 1 < len(starts):
                if done_evt[1 - b] is not None:
                    done_evt[1 - b].synchronize()
                th = threading.Thread(target=reader, args=(t + 1,))
                th.start()
            self._stream_rows(stage[b].data_ptr(), N, rows, v, y[i:i + rows])
            ev = torch.cuda.Event()
            ev.record(torch.cuda.current_stream())
            done_evt[b] = ev
        torch.cuda.current_stream().synchronize()
        return y

    def _free_product(self, v, elem):
        """y = E v (v's dtype) with E's values recomputed from the coordinates
        (elem F64: FP64 values, nugget in FP64; F32: the stored FP32 tile's
        values). Each value is evaluated once for E_ij and E_ji, with a device
        workspace of N^2 / 32 bytes kept between calls; beyond an eighth of
        the device's memory, every value is evaluated per row instead."""
        N = self.N
        y = torch.empty(N, dtype=v.dtype, device=v.device)
        c = self._coords_dev()
        xy = F64 if v.dtype == torch.float64 else F32
        nbytes = lib().mpdok_rbf_symv_workspace_bytes(N)
        cap = torch.cuda.get_device_properties(v.device).total_memory // 8
        if nbytes <= cap:
            if self._symv_ws is None or self._symv_ws.numel() < nbytes:
                self._symv_ws = None
                self._symv_ws = torch.empty(nbytes, dtype=torch.uint8, device=v.device)
            check(lib().mpdok_rbf_symv(_dev.stream(), c.data_ptr(), N, c.shape[1], self._kind,
                                       self._param(), float(self.reg), elem, v.data_ptr(),
                                       y.data_ptr(), xy, self._symv_ws.data_ptr(), nbytes),
                  "rbf_symv")
        else:
            check(lib().mpdok_rbf_matvec(_dev.stream(), c.data_ptr(), N, c.shape[1], self._kind,
                                         self._param(), float(self.reg), elem, v.data_ptr(),
                                         y.data_ptr(), xy, 0, N), "rbf_matvec")
        return y

    def _tiled_sgemv_free(self, v_fp32):
        """y_fp32 = A_fp32 v with the FP32 tile values recomputed on the fly."""
        return self._free_product(_dev.as_device(v_fp32, torch.float32).contiguous(), F32)

    def _tiled_sgemv(self, v_fp32):
        """Dispatch on the store."""
        if self._store == 'ram':
            return self._tiled_sgemv_ram(v_fp32)
        if self._store == 'free':
            return self._tiled_sgemv_free(v_fp32)
        return self._tiled_sgemv_ssd(v_fp32)

    # ---- FP64 product, matrix-free (mpdok_ooc.py:205-227) -------------------
    def _tiled_dgemv_fp64(self, v_fp64):
        """y_fp64 = A_fp64 v with every FP64 tile value recomputed from the
        coordinates (nugget added in FP64)."""
        return self._free_product(_dev.as_device(v_fp64, torch.float64).contiguous(), F64)

    # ---- inner FP32 GMRES (mpdok_ooc.py:231-311) ---------------------------
    def _inner_gmres(self, rhs_fp64, restart):
        """FP32 GMRES on A_fp32 e = rhs: DGKS orthogonalisation (two CGS
        passes) and Givens rotations on the device; returns e (FP32)."""
        from .dense_krylov import _Arnoldi
        N = self.N
        rhs = _ops.axpby(1.0, _dev.as_device(rhs_fp64, torch.float64), dtype=torch.float32)
        beta = _ops.norm(rhs)
        if beta < 1e-30:
            return _dev.zeros(N, torch.float32)
        ar = _Arnoldi(N, restart, torch.float32)
        ar.reset()
        _ops.axpby(1.0 / beta, rhs, out=ar.V[0])
        ar.g.zero_()
        ar.g[0] = beta
        for k in range(restart):
            w = self._tiled_sgemv(ar.V[k])
            ar.w.copy_(w)
            ar.orth(k, passes=2, stop_now=0)
            ar.givens(k, beta, 1e-6)
            if int(ar.ctl[0].item()):
                break
        m = ar.steps_done()
        if m == 0:
            return _dev.zeros(N, torch.float32)
        R = ar.R[:m, :m].cpu().numpy().T          # row i, column k: rotated H
        g = ar.g[:m].cpu().numpy()
        y = np.zeros(m)
        for j in range(m - 1, -1, -1):
            y[j] = (g[j] - R[j, j + 1:m] @ y[j + 1:m]) / R[j, j]
        y_dev = torch.as_tensor(y, dtype=torch.float64, device=ar.V.device)
        e = torch.empty(N, dtype=torch.float32, device=ar.V.device)
        check(lib().mpdok_basis_update_f32(_dev.stream(), ar.V.data_ptr(), ar.V.stride(0), m,
                                           y_dev.data_ptr(), e.data_ptr(), N, 1), "basis_update")
        return e

    # ---- inner FP32 Jacobi-CG (extension) -----------------------------------
    def _jacobi_dinv(self):
        """diag(A_fp32)^-1 in FP32: from the caller's host matrix for a
        user-supplied _ram_buf, otherwise recomputed from the coordinates
        exactly as the stored FP32 tiles hold it (nugget after the cast)."""
        buf = self._ram_buf
        own = (self._pinned is not None and isinstance(buf, np.ndarray)
               and buf.ctypes.data == self._pinned.data_ptr())
        if self._store == 'ram' and buf is not None and not own:
            d = np.asarray(np.diagonal(buf), dtype=np.float32)
            return _dev.as_device(np.float32(1.0) / d, torch.float32)
        c = self._coords_dev()
        return _ops.rbf_diag(c, self._kind, self._param(), self.reg, torch.float32,
                             torch.float32, invert=True)

    def _inner_cg(self, rhs_fp64, tol=1e-6, maxiter=None, precond='jacobi', check_every=10):
        """FP32 conjugate gradients on A_fp32 e = rhs from e = 0 (CG updates
        of dense_krylov.py:191-198 with z = M^-1 r, M = diag(A_fp32) for
        precond='jacobi'), products through _tiled_sgemv. Stops when the
        recursive residual ||r|| < tol ||rhs||. Returns (e FP32, iterations)."""
        from .dense_krylov import _pcg_inner
        dinv = self._jacobi_dinv() if precond == 'jacobi' else None
        rhs = _dev.as_device(rhs_fp64, torch.float64)
        e, it, _ = _pcg_inner(self._tiled_sgemv, rhs, torch.float32, dinv, tol,
                              self.N if maxiter is None else int(maxiter), check_every)
        return e, it

    # ---- outer FP64 refinement (mpdok_ooc.py:315-373) ----------------------
    def solve(self, b, tol=1e-5, maxiter_outer=20, restart=50, verbose=True, inner='gmres',
              precond=None, inner_tol=1e-6, inner_maxiter=None):
        """Iterative refinement: FP64 residual (matrix-free) + inner FP32 solve.
        Returns x (N,) FP64 on the device.

        inner='gmres' (default, the reference's): GMRES(restart) with DGKS.
        inner='cg' (extension, SPD kernels): CG with vectors in FP32, Jacobi
        preconditioned for precond='jacobi', to ||r|| < inner_tol ||rhs|| or
        inner_maxiter iterations (default N). last_info records the outer
        relres history and the inner iteration counts."""
        if self.N is None:
            raise RuntimeError("Call build() before solve()")
        if inner not in ('gmres', 'cg'):
            raise ValueError(f"inner must be 'gmres' or 'cg', got {inner!r}")
        if precond not in (None, 'jacobi'):
            raise ValueError(f"precond must be None or 'jacobi', got {precond!r}")
        b_gpu = _dev.as_device(b, torch.float64)
        b_norm = _ops.norm(b_gpu)
        self.last_info = {"history": [], "inner_iterations": [], "converged": False}
        if b_norm == 0.0:
            self.last_info["converged"] = True
            return _dev.zeros(self.N, torch.float64)
        x = _dev.zeros(self.N, torch.float64)
        converged = False
        rel = float('nan')
        for outer in range(maxiter_outer):
            t0 = time.perf_counter()
            r = _ops.axpby(1.0, b_gpu, -1.0, self._tiled_dgemv_fp64(x))
            rel = _ops.norm(r) / b_norm
            self.last_info["history"].append(rel)
            t_dgemv = time.perf_counter() - t0
            if verbose:
                print(f"  outer {outer}: rel_res={rel:.2e}  (FP64 DGEMV {t_dgemv:.1f}s)",
                      flush=True)
            if rel < tol:
                converged = True
                break
            t0 = time.perf_counter()
            if inner == 'cg':
                e, its = self._inner_cg(r, inner_tol, inner_maxiter, precond)
                self.last_info["inner_iterations"].append(its)
                if verbose:
                    print(f"           inner CG {its} iterations "
                          f"{time.perf_counter() - t0:.1f}s", flush=True)
            else:
                e = self._inner_gmres(r, restart)
                if verbose:
                    print(f"           inner GMRES {time.perf_counter() - t0:.1f}s", flush=True)
            x = _ops.axpby(1.0, x, 1.0, e, dtype=torch.float64)
        self.last_info["converged"] = converged
        if not converged:
            warnings.warn(
                f"MPDOKOOCSolver did not converge within {maxiter_outer} outer "
                f"iterations (final rel_res={rel:.2e})", RuntimeWarning, stacklevel=2)
        return x

    # ---- cleanup (mpdok_ooc.py:377-383) ------------------------------------
    def free(self):
        """Release the host copy, the device tiles and the coordinates."""
        self._ram_buf = None
        self._pinned = None
        self._stage = None
        self._dev_tiles = None
        self._coords = None
        self._symv_ws = None
        if torch.cuda.is_available():
            torch.cuda.empty_cache()
