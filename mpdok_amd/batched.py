"""Many independent kernel systems of uneven size in one call.

Each system i: points X_i (N_i x D), Gaussian kernel exp(-eps_i^2 r^2) plus a
nugget, nrhs right-hand sides; solved to relres <= tol per right-hand side
(kernel assembly, LU without pivoting, LU-IR with the compensated residual).
Two paths run the whole batch inside the library:

* `BatchedPlan` (mpdok_bplan_*): grouped launches, one per kind of work and
  step across every system that still has work, captured once into a
  hipGraph; a solve is one graph launch, with convergence tracked on the
  device.
* `solve_batched` (mpdok_batched_rbf_solve): systems spread over several
  streams, each driven by its own host thread, largest first by the
  longest-processing-time rule, so small systems fill the gaps left by large
  ones.

`BatchTuner` times candidates on a sample of the batch (every k-th system by
size, about 24) and caches the choice per size profile: the tile-shape
threshold of the graph path, and the number of streams, LU block size and
factor precision of the stream path. This path has no
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


class BatchedPlan:
    """All systems of a batch solved by grouped launches captured in one
    hipGraph (mpdok_bplan_create): a run is one graph launch, with the
    assembly, the FP64 no-pivot LU (64-wide steps), LU-IR with the Dot2
    residual and the copy-out inside it; convergence is tracked on the device.

    systems: list of (X_i (N_i, D) points, eps_i, B_i (N_i, nrhs)) as from
    c4_workload. The plan keeps device copies of X_i and B_i and one output
    block per system (self.xs); set_rhs() writes new right-hand sides into
    those buffers, so the next run() solves them without re-capturing.
    update_rows: trailing updates with at least this many rows use 128 x 128
    tiles, smaller ones 64 x 64. use_graph=False launches the same kernels
    eagerly (for comparison)."""

    def __init__(self, systems, tol=1e-12, reg=1e-6, maxiter_outer=4, update_rows=512,
                 use_graph=True, kernel="gaussian"):
        if len(systems) == 0:
            raise ValueError("BatchedPlan: no systems")
        kind = {"gaussian": 0, "matern32": 1}[kernel]
        self._keep, self.xs, self.bs = [], [], []
        Ns, cps, bps, xps, params = [], [], [], [], []
        nrhs = D = None
        for X, eps, B in systems:
            Xd = _dev.as_device(X, torch.float64).contiguous()
            Bd = _dev.as_device(np.asarray(B, dtype=np.float64).reshape(Xd.shape[0], -1),
                                torch.float64, "F")
            if nrhs is None:
                nrhs, D = Bd.shape[1], Xd.shape[1]
            if Bd.shape[1] != nrhs or Xd.shape[1] != D:
                raise ValueError("BatchedPlan: every system needs the same D and nrhs")
            Xo = _dev.empty(tuple(Bd.shape), torch.float64, "F")
            self._keep += [Xd]
            self.bs.append(Bd)
            self.xs.append(Xo)
            Ns.append(Xd.shape[0])
            cps.append(Xd.data_ptr())
            bps.append(Bd.data_ptr())
            xps.append(Xo.data_ptr())
            params.append(float(eps) ** 2 if kind == 0 else float(eps))
        self.nsys, self.nrhs, self.maxiter_outer = len(systems), nrhs, maxiter_outer
        self.N = np.array(Ns)
        P = ctypes.c_void_p
        torch.cuda.synchronize()
        h = ctypes.c_void_p()
        check(lib().mpdok_bplan_create(
            self.nsys, (ctypes.c_int * self.nsys)(*Ns), (P * self.nsys)(*cps), D, kind,
            (ctypes.c_double * self.nsys)(*params), float(reg), nrhs, (P * self.nsys)(*bps),
            (P * self.nsys)(*xps), float(tol), int(maxiter_outer), int(update_rows),
            1 if use_graph else 0, ctypes.byref(h)), "bplan_create")
        self._h = h

    def set_rhs(self, i, B):
        """Overwrite system i's right-hand sides (same shape) for the next run."""
        Bd = _dev.as_device(np.asarray(B, dtype=np.float64).reshape(self.bs[i].shape),
                            torch.float64, "F") if not isinstance(B, torch.Tensor) else B
        self.bs[i].copy_(Bd)

    def run(self, sync=True):
        """One solve of every system (one graph launch on torch's stream)."""
        if self._h is None:
            raise RuntimeError("BatchedPlan: closed")
        stream = torch.cuda.current_stream().cuda_stream
        check(lib().mpdok_bplan_run(self._h, ctypes.c_void_p(stream)), "bplan_run")
        if sync:
            torch.cuda.synchronize()
        return self.xs

    def results(self):
        """Outcome of the last run: converged, relres, steps (nsys x nrhs),
        info (nsys), hist ((maxiter_outer + 1) x nsys x nrhs)."""
        n, r = self.nsys, self.nrhs
        conv = np.zeros(n * r, np.int32)
        rel = np.zeros(n * r)
        steps = np.zeros(n * r, np.int32)
        info = np.zeros(n, np.int32)
        hist = np.zeros((self.maxiter_outer + 1) * n * r)
        P = ctypes.c_void_p
        check(lib().mpdok_bplan_results(self._h, conv.ctypes.data_as(P), rel.ctypes.data_as(P),
                                        steps.ctypes.data_as(P), info.ctypes.data_as(P),
                                        hist.ctypes.data_as(P)), "bplan_results")
        return {"converged": conv.reshape(n, r).astype(bool), "relres": rel.reshape(n, r),
                "steps": steps.reshape(n, r), "info": info,
                "hist": hist.reshape(self.maxiter_outer + 1, n, r)}

    def info(self):
        b = ctypes.c_size_t(0)
        k = ctypes.c_int(0)
        check(lib().mpdok_bplan_info(self._h, ctypes.byref(b), ctypes.byref(k)), "bplan_info")
        return {"device_bytes": int(b.value), "launches": int(k.value)}

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().mpdok_bplan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_batched_graph(systems, tol=1e-12, reg=1e-6, maxiter_outer=4, update_rows=512,
                        use_graph=True):
    """Build a BatchedPlan, run it once; returns (xs, info) like solve_batched
    (info also has 'steps', 'hist' and 'seconds' of the run)."""
    plan = BatchedPlan(systems, tol=tol, reg=reg, maxiter_outer=maxiter_outer,
                       update_rows=update_rows, use_graph=use_graph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xs = plan.run()
    secs = time.perf_counter() - t0
    out = plan.results()
    out["seconds"] = secs
    plan.close()
    return xs, out


class BatchTuner:
    """Picks launch parameters for a batch by timing candidates on a sample of
    it (every k-th system by size, about 24) and remembers the choice per size
    profile (sizes rounded to powers of two).

    choose(): (nstreams, nb, precision) of the stream path (solve_batched).
    choose_plan(): the tile-shape threshold of the graph path (BatchedPlan's
    update_rows: the trailing updates of systems with at least that many rows
    left run 128 x 128 tiles of 4 wavefronts, smaller ones 64 x 64 tiles)."""

    CANDIDATES = {"nstreams": (4, 8, 16), "nb": (256, 512), "precision": ("fp64", "ozaki")}
    PLAN_ROWS = (256, 512, 1024, 2048)

    def __init__(self):
        self.cache = {}
        self.plan_cache = {}
        self.trials = []

    @staticmethod
    def profile(systems):
        return tuple(sorted({int(round(math.log2(len(s[0])))) for s in systems}))

    @staticmethod
    def _sample(systems):
        # every k-th system by size: the sample keeps the batch's size mix
        by_n = sorted(systems, key=lambda s: len(s[0]))
        return by_n[::max(1, len(by_n) // 24)]

    def choose_plan(self, systems, tol=1e-12, reg=1e-6, maxiter_outer=4):
        """update_rows for a BatchedPlan of this batch: a plan per candidate
        on the sample, one warm-up replay, then the time of one replay; the
        fastest candidate whose systems all converged wins."""
        key = self.profile(systems)
        if key in self.plan_cache:
            return self.plan_cache[key]
        sample = self._sample(systems)
        best = None
        for rows in self.PLAN_ROWS:
            plan = BatchedPlan(sample, tol=tol, reg=reg, maxiter_outer=maxiter_outer,
                               update_rows=rows)
            plan.run()
            t0 = time.perf_counter()
            plan.run()
            secs = time.perf_counter() - t0
            ok = bool(plan.results()["converged"].all())
            plan.close()
            self.trials.append({"update_rows": rows, "seconds": secs, "converged": ok,
                                "path": "graph"})
            if ok and (best is None or secs < best[0]):
                best = (secs, {"update_rows": rows})
        choice = best[1] if best else {"update_rows": 512}
        self.plan_cache[key] = choice
        return choice

    def solve_plan(self, systems, tol=1e-12, reg=1e-6, maxiter_outer=4, **kw):
        """Build the tuned BatchedPlan for the whole batch and run it once:
        returns (xs, info) as solve_batched, with info['seconds'] the run (one
        graph launch), info['build_seconds'] the plan's set-up and capture,
        info['config'] the choice and info['plan'] its device bytes and
        captured launches."""
        cfg = dict(self.choose_plan(systems, tol, reg, maxiter_outer))
        cfg.update(kw)
        t0 = time.perf_counter()
        plan = BatchedPlan(systems, tol=tol, reg=reg, maxiter_outer=maxiter_outer, **cfg)
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        t0 = time.perf_counter()
        xs = plan.run()
        secs = time.perf_counter() - t0
        out = plan.results()
        out.update(seconds=secs, build_seconds=tb, config=cfg, plan=plan.info())
        plan.close()
        return xs, out

    def choose(self, systems, tol=1e-12, reg=1e-6):
        key = self.profile(systems)
        if key in self.cache:
            return self.cache[key]
        sample = self._sample(systems)
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
