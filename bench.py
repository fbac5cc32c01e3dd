#!/usr/bin/env python3
"""Headline benchmark: time to relres <= 1e-12 for mixed-precision GMRES-IR vs
the plain FP64 path on one MI355X, N = 32768 Gaussian RBF system.

Workload (synthetic, seeded): N points uniform in [0,1]^3 (NumPy PCG64 seed 42),
K_ij = exp(-25 ||x_i - x_j||^2) (eps = 5) + 1e-6 delta_ij, FP64 matrix resident
in HBM (column-major); b = sin(2 pi x) cos(2 pi y) cos(2 pi z) + 1e-4 N(0,1)
(seed 43). relres = ||b - A x|| / ||b|| evaluated on the GPU with compensated
(Dot2) FP64 accumulation.

One step = one complete solve from the resident matrix: LU factorisation
(int8-sliced trailing updates, FP64 storage, no pivoting: the matrix is SPD)
plus right-preconditioned GMRES-IR (FP64 body, Dot2 residual) until relres <=
tol. The plain FP64 path (FP64 matrix-core LU + FP64 refinement with the same
residual; FP64 residuals cannot reach 1e-12 here), torch's FP64 solver and an
unpreconditioned FP64 GMRES are timed in the same run.

Prints one JSON line (rank 0, N = 1 GPU).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = ("Time to relres≤1e-12: mixed-precision GMRES-IR vs plain FP64 on GPU, N=32768")
PEAK_INT8_TOPS = 5000.0     # MI355X dense int8, tera-operations/s
TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "ozaki_gemmd_traffic_n32768.md")
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rhs_of(X, seed=43):
    return (np.sin(2 * np.pi * X[:, 0]) * np.cos(2 * np.pi * X[:, 1]) * np.cos(2 * np.pi * X[:, 2])
            + 1e-4 * np.random.default_rng(seed).standard_normal(X.shape[0]))


def measured_traffic(N, slices, nb):
    """HBM bytes per launch of the int8-sliced trailing update, from the
    committed FETCH_SIZE / WRITE_SIZE counter summary (its own run; the
    counters cannot be read inside a timed run). Only for the workload the
    summary was measured on."""
    if N != 32768 or slices != 5 or nb != 512 or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(N_full, n_sample):
    """SciPy FP64 LAPACK solve (getrf/getrs, the reference's CPU comparator) on
    the first n_sample points of the same workload, extrapolated by N^3."""
    import scipy.linalg as sla
    from oracle import rbf as orb
    X = np.random.default_rng(42).uniform(0.0, 1.0, (N_full, 3))[:n_sample]
    A = orb.build_kernel(X, "gaussian", 25.0, 1e-6)
    b = rhs_of(np.random.default_rng(42).uniform(0.0, 1.0, (N_full, 3)))[:n_sample]
    t0 = time.perf_counter()
    x = sla.solve(A, b, check_finite=False)
    t = time.perf_counter() - t0
    rel = float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
    cores = len(os.sched_getaffinity(0))
    return {"value": t * (N_full / n_sample) ** 3, "unit": "s", "cores": cores, "kind": "port",
            "sample": (f"scipy.linalg.solve FP64 (LAPACK getrf/getrs) on the first {n_sample} "
                       f"points: {t:.2f} s measured, relres {rel:.1e} (FP64 residual), value "
                       f"extrapolated x(N/{n_sample})^3 to N={N_full}")}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--slices", type=int, default=5)
    ap.add_argument("--tol", type=float, default=1e-12)
    ap.add_argument("--nb", type=int, default=512)
    ap.add_argument("--cpu-sample", type=int, default=12288)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip torch solve and plain GMRES")
    ap.add_argument("--configs", action="store_true",
                    help="also measure the other configurations (CPU-runnable N = 4096 case, "
                         "out-of-core stream, batched systems, adjoint) into 'configs'")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): this build runs the "
              "solver on a single GPU and launches no multi-GPU layout.", file=sys.stderr)
        return 2

    import torch
    from mpdok_amd import _dev, _ops
    from mpdok_amd.mpdok_ops import LUIRSolver, MPDOKSolver

    N, tol = args.n, args.tol
    dev = _dev.device()
    log(f"[bench] device {torch.cuda.get_device_name(0)}  N={N}  slices={args.slices}")
    X = np.random.default_rng(42).uniform(0.0, 1.0, (N, 3))
    b_h = rhs_of(X)
    Xd = _dev.as_device(X, torch.float64)
    b = _dev.as_device(b_h, torch.float64)
    A = _dev.empty((N, N), torch.float64, "F")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _ops.rbf_assemble(Xd, 0, 25.0, 1e-6, A)
    torch.cuda.synchronize()
    t_assemble = time.perf_counter() - t0
    log(f"[bench] assembled K in {t_assemble * 1e3:.1f} ms")

    def dot2_relres(x):
        r = _ops.residual(A, x, b, mode="dot2")
        return _ops.norm(r) / _ops.norm(b)

    # ---- mixed: int8-sliced LU + GMRES-IR ---------------------------------
    mixed = MPDOKSolver()
    kw = dict(tol=tol, maxiter_outer=20, restart=50, precond="lu", factor_precision="ozaki",
              slices=args.slices, pivot="none", body="fp64", residual="dot2", nb=args.nb)
    for w in range(args.warmup):
        mixed.solve(A, b, **kw)
        log(f"[bench] warmup {w}: {mixed.last_info['history']}")
    times = []
    for k in range(args.steps):
        mixed.profile = (k == args.steps - 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x_mix = mixed.solve(A, b, **kw)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        log(f"[bench] step {k}: {times[-1] * 1e3:.1f} ms  outer relres {mixed.last_info['history']}"
            f"  inner {mixed.last_info['inner_iterations']}")
    info_mix = mixed.last_info
    t_mix = float(np.mean(times))
    rel_mix = dot2_relres(x_mix)
    mixed.free_fp32_buf()

    # ---- plain FP64: FP64 LU + FP64 refinement (Dot2 residual) -------------
    plain = LUIRSolver(precision="fp64", pivot="none", nb=args.nb, residual="dot2")
    plain.solve(A, b, tol=tol, maxiter_outer=6)
    ptimes = []
    for k in range(max(1, min(args.steps, 2))):
        plain.profile = True
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x_plain = plain.solve(A, b, tol=tol, maxiter_outer=6)
        torch.cuda.synchronize()
        ptimes.append(time.perf_counter() - t0)
    t_plain = float(np.mean(ptimes))
    plain_hist = plain.last_info["history"]
    plain_stats = plain.factor_stats
    log(f"[bench] plain FP64 LU-IR: {t_plain * 1e3:.1f} ms  history {plain_hist}")
    # the same factors with FP64 residuals: the floor of the plain path
    plain.residual = "fp64"
    x_p64 = plain.solve_factored(b, tol=tol, maxiter_outer=6)
    floor_fp64 = dot2_relres(x_p64)
    plain.residual = "dot2"
    # direct FP64 solve without refinement (error reference for the forward error)
    x_direct = plain._handle.solve(b)
    plain.free_factored()
    del plain
    torch.cuda.empty_cache()

    fwd_mix = float(torch.linalg.vector_norm(x_mix - x_plain) / torch.linalg.vector_norm(x_plain))
    fwd_direct = float(torch.linalg.vector_norm(x_direct - x_plain)
                       / torch.linalg.vector_norm(x_plain))
    a_fro = float(_ops.dot(A.t().reshape(-1), A.t().reshape(-1)).item()) ** 0.5
    rnorm = _ops.norm(_ops.residual(A, x_mix, b, mode="dot2"))
    bwd_mix = rnorm / (a_fro * float(torch.linalg.vector_norm(x_mix)) + _ops.norm(b))

    extra = {}
    if not args.no_extra:
        # torch's FP64 dense solver (rocSOLVER getrf/getrs, partial pivoting)
        try:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x_t = torch.linalg.solve(A, b)
            torch.cuda.synchronize()
            t_t = time.perf_counter() - t0
            extra["torch_linalg_solve_fp64"] = {"s": t_t, "relres": dot2_relres(x_t)}
            del x_t
        except Exception as exc:  # report, do not fail the benchmark
            extra["torch_linalg_solve_fp64"] = {"error": str(exc)[:200]}
        torch.cuda.empty_cache()
        # unpreconditioned FP64 GMRES(50), 4 restart cycles
        g = MPDOKSolver()
        import warnings
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x_g = _plain_gmres(g, A, b, tol)
            torch.cuda.synchronize()
            t_g = time.perf_counter() - t0
        extra["fp64_gmres_unpreconditioned"] = {
            "s": t_g, "iterations": g.last_info["inner_iterations"],
            "relres": dot2_relres(x_g), "converged": g.last_info["converged"]}
        log(f"[bench] extra {extra}")

    st = info_mix["factor_stats"] or {}
    counts = info_mix.get("slice_counts") or []
    roofline = None
    if st.get("launches"):
        avg_ms = st["trailing_ms"] / st["launches"]
        achieved = st["trailing_ops"] / (st["trailing_ms"] * 1e-3) / 1e12
        # per launch: the larger of its int8-MFMA time and its HBM time (FP64 C
        # read + write) at peak, over the measured time: the fraction of the
        # mixed roofline (the low slice counts are HBM-bound)
        # steps: groups of panels, one trailing update (two launches: the next
        # group's columns, then the rest) with K = the group's width each
        t_roof = 0.0
        starts = list(info_mix.get("step_starts") or range(0, N, args.nb)) + [N]
        for step in range(len(starts) - 2):
            k, kb, nxt = starts[step], starts[step + 1] - starts[step], starts[step + 2] - starts[step + 1]
            n3 = N - k - kb
            S = counts[step] if step < len(counts) else args.slices
            for nc in ([nxt] + ([n3 - nxt] if n3 > nxt else [])):
                ops = 2.0 * n3 * nc * kb * S * (S + 1) / 2
                byts = 16.0 * n3 * nc
                t_roof += max(ops / (PEAK_INT8_TOPS * 1e12), byts / (PEAK_HBM_GBS * 1e9))
        roofline = {"bound": "mfma", "kernel": f"ozaki_gemmd_kernel<{args.slices}, ...>",
                    "achieved": achieved, "peak": PEAK_INT8_TOPS, "unit": "TFLOP/s",
                    "frac": achieved / PEAK_INT8_TOPS,
                    "traffic": measured_traffic(N, args.slices, args.nb),
                    "mixed_roofline_frac": t_roof / (st["trailing_ms"] * 1e-3),
                    "note": ("int8 tera-operations/s of the trailing updates: 2 m n k x "
                             "S_k (S_k + 1) / 2 slice pairs per launch, S_k the adaptive slice "
                             f"count of step k (slice_counts), {st['launches']} launches, avg "
                             f"{avg_ms:.3f} ms (HIP events on the factorisation's stream); "
                             "mixed_roofline_frac: sum over launches of max(int8 ops / 5.0 POP/s, "
                             "16 B per C element / 8 TB/s) over the measured time; traffic: HBM "
                             "bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE "
                             "(profiles/ozaki_gemmd_traffic_n32768.md)")}
    plain_roof = None
    if plain_stats and plain_stats.get("launches"):
        a = plain_stats["trailing_ops"] / (plain_stats["trailing_ms"] * 1e-3) / 1e12
        plain_roof = {"kernel": "gemm_kernel<double>", "achieved": a, "peak": PEAK_FP64_TFLOPS,
                      "unit": "TFLOP/s", "frac": a / PEAK_FP64_TFLOPS}

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline ...")
        cpu = cpu_baseline(N, min(args.cpu_sample, N))
        log(f"[bench] CPU baseline {cpu}")

    out = {
        "metric": METRIC, "value": t_mix, "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_mix * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 (int8-sliced trailing updates, adaptive slice counts)",
        "data": "synthetic",
        "config": {"workload": "gaussian_rbf_3d_gmres_ir", "N": N, "D": 3, "eps": 5.0,
                   "nugget": 1e-6, "tol": tol, "slices": args.slices, "nb": args.nb,
                   "seeds": {"points": 42, "noise": 43}, "parallelism": "single_gpu"},
        "relres": rel_mix, "converged": info_mix["converged"],
        "outer_history": info_mix["history"], "inner_iterations": info_mix["inner_iterations"],
        "slice_counts": counts[:-1] if counts else None,
        "forward_error_vs_fp64_refined": fwd_mix, "forward_error_fp64_direct": fwd_direct,
        "backward_error": bwd_mix,
        "plain_fp64": {"s": t_plain, "history": plain_hist, "method": "FP64 LU + FP64 IR (Dot2)",
                       "fp64_residual_floor": floor_fp64, "trailing": plain_roof},
        "speedup_vs_plain_fp64": t_plain / t_mix,
        "assembly_ms": t_assemble * 1e3,
        "roofline": roofline, "cpu_baseline": cpu, **extra,
    }
    if args.configs:
        torch.cuda.empty_cache()
        out["configs"] = other_configs(log)
    print(json.dumps(out), flush=True)
    return 0


def other_configs(log):
    """Secondary configurations, one measurement each (seconds, relres, rates)."""
    import warnings
    import torch
    from mpdok_amd import _dev, _ops
    from mpdok_amd.mpdok_ops import LUIRSolver, MPDOKSolver
    res = {}
    warnings.simplefilter("ignore")

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        v = fn()
        torch.cuda.synchronize()
        return v, time.perf_counter() - t0

    # CPU-runnable case: N = 4096 in [0,1]^2, eps = 8, nugget 1e-8
    N = 4096
    X = np.random.default_rng(42).uniform(0.0, 1.0, (N, 2))
    b = (np.sin(2 * np.pi * X[:, 0]) * np.cos(2 * np.pi * X[:, 1])
         + 1e-4 * np.random.default_rng(43).standard_normal(N))
    A = _dev.empty((N, N), torch.float64, "F")
    _ops.rbf_assemble(_dev.as_device(X, torch.float64), 0, 64.0, 1e-8, A)
    bd = _dev.as_device(b, torch.float64)
    rel = lambda x: _ops.norm(_ops.residual(A, x, bd, mode="dot2")) / _ops.norm(bd)
    s = MPDOKSolver()
    # kappa ~ 2e10 (nugget 1e-8): 47 bits (6 slices, fixed) to precondition
    kw = dict(tol=1e-12, maxiter_outer=20, precond="lu", factor_precision="ozaki", slices=6)
    s.solve(A, bd, **kw)
    x, t = timed(lambda: s.solve(A, bd, **kw))
    lu = LUIRSolver()
    x2, t2 = timed(lambda: lu.solve(A, bd, tol=1e-12, maxiter_outer=10))
    res["n4096_2d"] = {"gmres_ir_int8_lu": {"s": t, "relres": rel(x)},
                       "lu_ir_fp32_lu_reference_symbol": {"s": t2, "relres": rel(x2)}}
    log(f"[configs] n4096 {res['n4096_2d']}")
    del A, s, lu
    torch.cuda.empty_cache()

    # out-of-core: Matern-3/2, length scale 0.1, nugget 1e-4, N = 65536 (the
    # 262144-point FP32 operator, 275 GB, exceeds this process's host memory)
    from mpdok_amd.mpdok_ooc import MPDOKOOCSolver
    N = 65536
    X = np.random.default_rng(42).uniform(0.0, 1.0, (N, 3))
    o = MPDOKOOCSolver(tile_rows=4096)
    _, tb = timed(lambda: o.build(X, reg=1e-4, store="ram", verbose=False, kernel="matern32",
                                  length_scale=0.1))
    v = torch.rand(N, dtype=torch.float32, device=_dev.device())
    o._tiled_sgemv(v)
    _, ts = timed(lambda: o._tiled_sgemv(v))
    v64 = v.double()
    _, tf = timed(lambda: o._tiled_dgemv_fp64(v64))
    res["ooc_matern_n65536"] = {"build_s": tb, "stream_product_s": ts,
                                "stream_GBps": N * N * 4 / ts / 1e9,
                                "matrix_free_fp64_product_s": tf}
    log(f"[configs] ooc {res['ooc_matern_n65536']}")
    o.free()
    del o
    torch.cuda.empty_cache()

    # batched uneven systems: 256 systems, N log-uniform in [256, 8192], 4 RHS
    from mpdok_amd.batched import BatchTuner, c4_workload
    sysw = c4_workload(seed=0)
    tuner = BatchTuner()
    choice = tuner.choose(sysw)
    _, info = tuner.solve(sysw)
    Ns = np.array([len(q[0]) for q in sysw], dtype=float)
    res["batched_256"] = {"s": info["seconds"], "config": choice,
                          "all_converged": bool(info["converged"].all()),
                          "max_relres": float(np.nanmax(info["relres"])),
                          "lu_flop": float((2.0 / 3.0 * Ns ** 3).sum())}
    log(f"[configs] batched {res['batched_256']}")
    torch.cuda.empty_cache()

    # adjoint: N = 16384 training, 4096 validation, eps 6, log nugget -12
    from mpdok_amd.adjoint import rbf_validation_loss
    rng = np.random.default_rng(7)
    Xt, Xv = rng.uniform(0, 1, (16384, 2)), rng.uniform(0, 1, (4096, 2))
    f = lambda P: np.sin(3 * np.pi * P[:, 0]) * np.exp(-P[:, 1])
    yt = f(Xt) + 1e-3 * rng.standard_normal(16384)
    yv = f(Xv) + 1e-3 * rng.standard_normal(4096)
    e = torch.tensor(6.0, dtype=torch.float64, requires_grad=True)
    ls = torch.tensor(-12.0, dtype=torch.float64, requires_grad=True)

    def fwd_bwd():
        L = rbf_validation_loss(e, ls, Xt, yt, Xv, yv, tol=1e-12, maxiter_outer=30)
        L.backward()
        return L
    L, tadj = timed(fwd_bwd)
    h = 1e-5
    Lf = lambda a, c: float(rbf_validation_loss(a, c, Xt, yt, Xv, yv, tol=1e-12,
                                                maxiter_outer=30))
    fd_e = (Lf(6.0 + h, -12.0) - Lf(6.0 - h, -12.0)) / (2 * h)
    fd_s = (Lf(6.0, -12.0 + h) - Lf(6.0, -12.0 - h)) / (2 * h)
    # dense FP64 reference gradient (closed form, rocSOLVER FP64 solves through torch)
    dev = _dev.device()
    Xt_d, Xv_d = (torch.as_tensor(Xt, device=dev), torch.as_tensor(Xv, device=dev))
    K = torch.empty((16384, 16384), dtype=torch.float64, device=dev)
    _ops.rbf_assemble(Xt_d, 0, 36.0, float(np.exp(-12.0)), K)
    D2v = torch.cdist(Xv_d, Xt_d) ** 2
    Ev = torch.exp(-36.0 * D2v)
    yt_d, yv_d = torch.as_tensor(yt, device=dev), torch.as_tensor(yv, device=dev)
    alpha = torch.linalg.solve(K, yt_d)
    r = Ev @ alpha - yv_d
    g = 2.0 * r / 4096
    lam = torch.linalg.solve(K, Ev.T @ g)
    K.fill_diagonal_(0.0)                       # K - sigma I = exp(-gamma D2)
    D2 = torch.cdist(Xt_d, Xt_d) ** 2
    dgamma = g @ ((-D2v * Ev) @ alpha) - lam @ ((-D2 * K) @ alpha)
    ref_e, ref_s = float(2 * 6.0 * dgamma), float(-np.exp(-12.0) * (lam @ alpha))
    del K, D2, D2v, Ev
    torch.cuda.empty_cache()
    res["adjoint_n16384"] = {"forward_backward_s": tadj, "loss": float(L),
                             "dL_deps": float(e.grad), "dense_fp64_deps": ref_e,
                             "rel_diff_dense_eps": abs(float(e.grad) - ref_e) / abs(ref_e),
                             "fd_deps": fd_e,
                             "rel_diff_fd_eps": abs(float(e.grad) - fd_e) / abs(fd_e),
                             "dL_dlogsigma": float(ls.grad), "dense_fp64_dlogsigma": ref_s,
                             "rel_diff_dense_logsigma": abs(float(ls.grad) - ref_s) / abs(ref_s),
                             "fd_dlogsigma": fd_s,
                             "rel_diff_fd_logsigma": abs(float(ls.grad) - fd_s) / abs(fd_s)}
    log(f"[configs] adjoint {res['adjoint_n16384']}")
    return res


def _plain_gmres(solver, A, b, tol):
    """Unpreconditioned FP64 restarted GMRES(50), 4 cycles (200 matvecs)."""
    import ctypes
    import torch
    from mpdok_amd import _dev
    from mpdok_amd._lib import check
    N = A.shape[0]
    order, ld = _dev.layout(A)
    x = _dev.zeros(N, torch.float64)
    hist = (ctypes.c_double * 8)()
    n_hist, conv, inner = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    check(solver._lib.mpdok_gmres_ir(_dev.stream(), A.data_ptr(), ld, order, None, 0, order, None,
                                     N, b.data_ptr(), x.data_ptr(), float(tol), 4, 50, 1e-14, 1,
                                     1, 1, ctypes.byref(conv), hist, 8, ctypes.byref(n_hist),
                                     ctypes.byref(inner)), "gmres")
    torch.cuda.synchronize()
    solver.last_info = {"history": list(hist[:n_hist.value]), "converged": bool(conv.value),
                        "inner_iterations": inner.value}
    return x


if __name__ == "__main__":
    sys.exit(main())
