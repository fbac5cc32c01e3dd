"""Many independent kernel systems of uneven size in one call.

Each system i: points X_i (N_i x D), Gaussian kernel exp(-eps_i^2 r^2) plus a
nugget, nrhs right-hand sides; solved to relres <= tol per right-hand side
(kernel assembly, LU without pivoting, LU-IR with the compensated residual).
The whole batch runs inside the library (mpdok_batched_rbf_solve): systems
are spread over several streams, each driven by its own host thread, largest
first by the longest-processing-time rule, so small systems fill the gaps left
by large ones.

`BatchTuner` picks the number of streams, the LU block size and the factor
precision by timing candidates on a sample of the batch (every k-th system by
size, about 24), and caches the choice per size profile. This path has no
counterpart in the reference; it follows the north star's batched
configuration (256 systems, N log-uniform in [256, 8192], 4 right-hand sides).
"""
import ctypes
import itertools
import math
import time

import numpy as np
import torch

from . import _dev
from ._lib import check, lib

PRECISIONS = {"fp64": 0, "ozaki": 4}


def c4_workload(seed=0, nsys=256, nrhs=4, nmin=256, nmax=8192, eps_range=(2.0, 12.0),
                D=2):
    """The batched configuration's synthetic systems: N log-uniform in
    [nmin, nmax], points uniform in [0,1]^D, eps uniform in eps_range, and
    nrhs right-hand sides sum_k w_k phi_k(x) with w ~ N(0,1) over a fixed
    Fourier basis phi_k(x) = cos(pi (a_k . x) + c_k) (8 terms)."""
    rng = np.random.default_rng(seed)
    basis_a = rng.uniform(-3.0, 3.0, (8, D))
    basis_c = rng.uniform(0.0, 2 * np.pi, 8)
    out = []
    for _ in range(nsys):
        N = int(math.floor(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        X = rng.uniform(0.0, 1.0, (N, D))
        eps = float(rng.uniform(*eps_range))
        phi = np.cos(np.pi * X @ basis_a.T + basis_c)           # N x 8
        B = phi @ rng.standard_normal((8, nrhs))                 # N x nrhs
        out.append((X, eps, B))
    return out


def solve_batched(systems, tol=1e-12, reg=1e-6, precision="fp64", slices=6, nb=256,
                  nstreams=4, maxiter_outer=10):
    """Solve every system; returns (list of X_i (N_i, nrhs) FP64 device
    tensors, info) with info['converged'] and info['relres'] (nsys x nrhs
    NumPy arrays) and info['seconds']."""
    nsys = len(systems)
    if nsys == 0:
        return [], {"converged": np.zeros((0, 0), bool), "relres": np.zeros((0, 0)),
                    "seconds": 0.0}
    nrhs = None
    keep = []
    Ns, cps, bps, xps, params = [], [], [], [], []
    D = None
    for X, eps, B in systems:
        Xd = _dev.as_device(X, torch.float64).contiguous()
        Bd = _dev.as_device(np.asarray(B, dtype=np.float64).reshape(Xd.shape[0], -1), torch.float64,
                            "F") if not isinstance(B, torch.Tensor) else \
            _dev.as_device(B.reshape(Xd.shape[0], -1), torch.float64, "F")
        if nrhs is None:
            nrhs, D = Bd.shape[1], Xd.shape[1]
        if Bd.shape[1] != nrhs or Xd.shape[1] != D:
            raise ValueError("solve_batched: every system needs the same D and nrhs")
        Xo = _dev.empty(tuple(Bd.shape), torch.float64, "F")
        keep += [Xd, Bd, Xo]
        Ns.append(Xd.shape[0])
        cps.append(Xd.data_ptr())
        bps.append(Bd.data_ptr())
        xps.append(Xo.data_ptr())
        params.append(float(eps) ** 2)
    conv = np.zeros(nsys * nrhs, dtype=np.int32)
    rel = np.zeros(nsys * nrhs, dtype=np.float64)
    P = ctypes.c_void_p
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    check(lib().mpdok_batched_rbf_solve(
        nsys, (ctypes.c_int * nsys)(*Ns), (P * nsys)(*cps), D, 0, (ctypes.c_double * nsys)(*params),
        float(reg), nrhs, (P * nsys)(*bps), (P * nsys)(*xps), float(tol), int(maxiter_outer),
        PRECISIONS[precision], int(slices), int(nb), int(nstreams),
        conv.ctypes.data_as(P), rel.ctypes.data_as(P)), "batched_rbf_solve")
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    xs = [keep[3 * i + 2] for i in range(nsys)]
    return xs, {"converged": conv.reshape(nsys, nrhs).astype(bool),
                "relres": rel.reshape(nsys, nrhs), "seconds": secs}


class BatchTuner:
    """Picks (nstreams, nb, precision) for a batch by timing candidates on a
    sample of it; remembers the choice per size profile (sizes rounded to
    powers of two)."""

    CANDIDATES = {"nstreams": (4, 8, 16), "nb": (256, 512), "precision": ("fp64", "ozaki")}

    def __init__(self):
        self.cache = {}
        self.trials = []

    @staticmethod
    def profile(systems):
        return tuple(sorted({int(round(math.log2(len(s[0])))) for s in systems}))

    def choose(self, systems, tol=1e-12, reg=1e-6):
        key = self.profile(systems)
        if key in self.cache:
            return self.cache[key]
        # every k-th system by size: the sample keeps the batch's size mix
        by_n = sorted(systems, key=lambda s: len(s[0]))
        step = max(1, len(by_n) // 24)
        sample = by_n[::step]
        best = None
        for ns, nb, prec in itertools.product(*self.CANDIDATES.values()):
            _, info = solve_batched(sample, tol=tol, reg=reg, precision=prec, nb=nb,
                                    nstreams=ns)
            ok = bool(info["converged"].all())
            self.trials.append({"nstreams": ns, "nb": nb, "precision": prec,
                                "seconds": info["seconds"], "converged": ok})
            if ok and (best is None or info["seconds"] < best[0]):
                best = (info["seconds"], {"nstreams": ns, "nb": nb, "precision": prec})
        choice = best[1] if best else {"nstreams": 4, "nb": 256, "precision": "fp64"}
        self.cache[key] = choice
        return choice

    def solve(self, systems, tol=1e-12, reg=1e-6, **kw):
        cfg = dict(self.choose(systems, tol, reg))
        cfg.update(kw)
        return solve_batched(systems, tol=tol, reg=reg, **cfg)
