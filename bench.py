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


def _cores():
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    return min(n, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else n


def cpu_baseline(N_full, n_sample, A_dev=None, b_full=None, gmres_iters=20):
    """The reference's CPU path (SciPy FP64) on a bounded sample of the same
    workload: LAPACK getrf/getrs (scipy.linalg.solve, rbf_demo.ipynb:159) on
    the first n_sample points, extrapolated by (N / n_sample)^3 to the full
    system; and, on the full N x N matrix (copied from the GPU, where it was
    assembled), SciPy GMRES(restart) for a fixed iteration budget
    (dense_krylov.py:617,689 call it with restart=50, maxiter=N), reporting the
    time per iteration and the relres reached."""
    import scipy.linalg as sla
    import scipy.sparse.linalg as spla
    from oracle import rbf as orb
    X = np.random.default_rng(42).uniform(0.0, 1.0, (N_full, 3))
    b_all = rhs_of(X)
    Xs = X[:n_sample]
    A = orb.build_kernel(Xs, "gaussian", 25.0, 1e-6)
    b = b_all[:n_sample]
    t0 = time.perf_counter()
    x = sla.solve(A, b, check_finite=False)
    t = time.perf_counter() - t0
    rel = float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
    del A
    out = {"value": t * (N_full / n_sample) ** 3, "unit": "s", "cores": _cores(), "kind": "port",
           "sample": (f"scipy.linalg.solve FP64 (LAPACK getrf/getrs) on the first {n_sample} of the "
                      f"{N_full} points: {t:.2f} s measured, relres {rel:.1e} (FP64 residual); value "
                      f"= that time x (N/{n_sample})^3")}
    if A_dev is not None and gmres_iters > 0:
        A_h = A_dev.cpu().numpy()                      # column-major N x N, FP64
        it = [0]

        def cb(_):
            it[0] += 1
        t0 = time.perf_counter()
        xg, _ = spla.gmres(A_h, b_all, rtol=1e-12, restart=gmres_iters, maxiter=1,
                           callback=cb, callback_type="pr_norm")
        tg = time.perf_counter() - t0
        relg = float(np.linalg.norm(b_all - A_h @ xg) / np.linalg.norm(b_all))
        out["gmres_full_n"] = {"iterations": it[0], "s": tg, "s_per_iteration": tg / max(it[0], 1),
                               "relres": relg, "restart": gmres_iters,
                               "note": "scipy.sparse.linalg.gmres FP64 on the full matrix, one "
                                       "restart cycle (capped)"}
        del A_h
    return out


def cpu_full(N):
    """scipy.linalg.solve on the full N-point system (minutes of CPU)."""
    import scipy.linalg as sla
    from oracle import rbf as orb
    X = np.random.default_rng(42).uniform(0.0, 1.0, (N, 3))
    A = orb.build_kernel(X, "gaussian", 25.0, 1e-6)
    b = rhs_of(X)
    log(f"[bench] CPU full solve N={N} on {_cores()} cores ...")
    t0 = time.perf_counter()
    x = sla.solve(A, b, check_finite=False, overwrite_a=False)
    t = time.perf_counter() - t0
    rel = float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
    return {"s": t, "relres_fp64": rel, "cores": _cores(), "N": N}


def gemv_profile(N, reps=10):
    """HIP-event times of the bandwidth-bound kernels on the N-point headline
    matrix: FP32 GEMV (row- and column-major A, 4 N^2 B), FP64 GEMV and the
    FP64 / Dot2 residual (column-major, 8 N^2 B), and the read-only ceiling
    (mpdok_read_ceiling) over the same buffers."""
    import torch
    from mpdok_amd import _dev, _ops, _lib
    L = _lib.load()
    X = np.random.default_rng(42).uniform(0.0, 1.0, (N, 3))
    A64 = _dev.empty((N, N), torch.float64, "F")
    _ops.rbf_assemble(_dev.as_device(X, torch.float64), 0, 25.0, 1e-6, A64)
    A32c = _ops.cast_matrix(A64, torch.float32, "F")
    A32r = _ops.cast_matrix(A64, torch.float32, "C")
    x64 = torch.rand(N, dtype=torch.float64, device=A64.device)
    x32 = x64.float()
    b64 = torch.rand(N, dtype=torch.float64, device=A64.device)
    ws = torch.empty(L.mpdok_read_ceiling_workspace_bytes(), dtype=torch.uint8, device=A64.device)
    out = torch.empty(1, dtype=torch.float64, device=A64.device)

    def ceiling(t):
        st = L.mpdok_read_ceiling(_dev.stream(), t.data_ptr(), t.numel() * t.element_size(),
                                  out.data_ptr(), ws.data_ptr())
        assert st == 0

    cases = [("read_ceiling_f64", lambda: ceiling(A64), 8.0 * N * N),
             ("read_ceiling_f32", lambda: ceiling(A32r), 4.0 * N * N),
             ("gemv_f32_row_major", lambda: _ops.matvec(A32r, x32), 4.0 * N * N),
             ("gemv_f32_col_major", lambda: _ops.matvec(A32c, x32), 4.0 * N * N),
             ("gemv_f64_col_major", lambda: _ops.matvec(A64, x64), 8.0 * N * N),
             ("residual_f64_col_major", lambda: _ops.residual(A64, x64, b64, mode="fp64"),
              8.0 * N * N),
             ("residual_dot2_col_major", lambda: _ops.residual(A64, x64, b64, mode="dot2"),
              8.0 * N * N)]
    res = {}
    for name, fn, byts in cases:
        fn()
        fn()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
        torch.cuda.synchronize()
        for r in range(reps):
            ev[2 * r].record()
            fn()
            ev[2 * r + 1].record()
        torch.cuda.synchronize()
        ms = min(ev[2 * r].elapsed_time(ev[2 * r + 1]) for r in range(reps))
        gbs = byts / (ms * 1e-3) / 1e9
        res[name] = {"ms": ms, "GBps": gbs, "frac_of_peak": gbs / PEAK_HBM_GBS, "bytes": byts}
    for name in res:
        if not name.startswith("read_ceiling"):
            c = res["read_ceiling_f32" if "f32" in name else "read_ceiling_f64"]["GBps"]
            res[name]["frac_of_read_ceiling"] = res[name]["GBps"] / c
    return {"metric": "bandwidth-bound kernels at N=%d" % N, "N": N, "unit": "GB/s",
            "peak_GBps": PEAK_HBM_GBS, "kernels": res}


def rbf_profile(N=32768, NF=65536, reps=5):
    """HIP-event times of kernel-matrix assembly at N points (FP64 Gaussian and
    Matern-3/2 principal blocks, FP32, and a block that is not principal, so
    every entry is evaluated) against a plain fill of the same FP64 matrix,
    and of the matrix-free FP64 products at NF points (symmetric: each value
    once for both entries; by rows: every entry)."""
    import torch
    from mpdok_amd import _dev, _ops, _lib
    L = _lib.load()
    X = _dev.as_device(np.random.default_rng(42).uniform(0.0, 1.0, (N, 3)), torch.float64)
    A64 = _dev.empty((N, N), torch.float64, "F")
    A32 = _dev.empty((N, N), torch.float32, "C")
    G = A64[:, :N - 1]
    XF = _dev.as_device(np.random.default_rng(1).uniform(0.0, 1.0, (NF, 3)), torch.float64)
    x = _dev.as_device(np.random.default_rng(2).standard_normal(NF), torch.float64)
    y = torch.empty_like(x)
    nb = L.mpdok_rbf_symv_workspace_bytes(NF)
    ws = torch.empty(nb, dtype=torch.uint8, device=x.device)

    def symv(kind, p):
        assert L.mpdok_rbf_symv(_dev.stream(), XF.data_ptr(), NF, 3, kind, p, 1e-4, 0,
                                x.data_ptr(), y.data_ptr(), 0, ws.data_ptr(), nb) == 0

    def rows(kind, p):
        assert L.mpdok_rbf_matvec(_dev.stream(), XF.data_ptr(), NF, 3, kind, p, 1e-4, 0,
                                  x.data_ptr(), y.data_ptr(), 0, 0, NF) == 0

    w64, w32, e = 8.0 * N * N, 4.0 * N * N, float(NF) * NF
    cases = [("fill_f64", lambda: A64.fill_(1.0), w64, "B"),
             ("assemble_gaussian_f64", lambda: _ops.rbf_assemble(X, 0, 25.0, 1e-6, A64), w64, "B"),
             ("assemble_matern_f64", lambda: _ops.rbf_assemble(X, 1, 0.1, 1e-4, A64), w64, "B"),
             ("assemble_gaussian_f32", lambda: _ops.rbf_assemble(X, 0, 25.0, 1e-6, A32), w32, "B"),
             ("assemble_gaussian_f64_not_principal",
              lambda: _ops.rbf_assemble(X, 0, 25.0, 1e-6, G, cols=N - 1), 8.0 * N * (N - 1), "B"),
             ("matrix_free_matern_f64_symmetric", lambda: symv(1, 0.1), e, "entries"),
             ("matrix_free_matern_f64_rows", lambda: rows(1, 0.1), e, "entries"),
             ("matrix_free_gaussian_f64_symmetric", lambda: symv(0, 25.0), e, "entries"),
             ("matrix_free_gaussian_f64_rows", lambda: rows(0, 25.0), e, "entries")]
    res = {}
    for name, fn, amount, unit in cases:
        fn()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
        torch.cuda.synchronize()
        for r in range(reps):
            ev[2 * r].record()
            fn()
            ev[2 * r + 1].record()
        torch.cuda.synchronize()
        ms = min(ev[2 * r].elapsed_time(ev[2 * r + 1]) for r in range(reps))
        rate = amount / (ms * 1e-3)
        res[name] = ({"ms": ms, "TBps_written": rate / 1e12, "frac_of_peak": rate / 1e9 / PEAK_HBM_GBS}
                     if unit == "B" else {"ms": ms, "matrix_entries_per_s": rate})
    return {"metric": "RBF assembly at N=%d and matrix-free products at N=%d" % (N, NF),
            "kernels": res}


def update_profile(shapes=((30720, 1024), (15360, 1024)), reps=5):
    """HIP-event times of the int8-sliced trailing-update kernel alone
    (mpdok_ozaki_apply_sliced after one slicing call) per slice count S = 2 .. 5,
    at the LU's panel-pair shape (m = n, K = 1024) on random FP64 operands."""
    import ctypes
    import torch
    from mpdok_amd import _dev, _lib
    L = _lib.load()
    out = {"metric": "int8-sliced trailing update alone", "unit": "ms", "peak_int8_TOPS":
           PEAK_INT8_TOPS, "peak_hbm_GBps": PEAK_HBM_GBS, "cases": []}
    dev = _dev.device()
    for m, K in shapes:
        g = torch.Generator(device=dev).manual_seed(0)
        A = torch.randn(K, m, dtype=torch.float64, device=dev, generator=g).t()   # m x K, col-major
        B = torch.randn(m, K, dtype=torch.float64, device=dev, generator=g).t()   # K x m, col-major
        C = torch.randn(m, m, dtype=torch.float64, device=dev, generator=g).t()
        ws = torch.empty(L.mpdok_ozaki_workspace_bytes(m, m, K, 5), dtype=torch.uint8, device=dev)
        sc = torch.zeros(2, dtype=torch.int32, device=dev)
        su, er = ctypes.c_int(0), ctypes.c_int(0)
        st = L.mpdok_ozaki_gemm_sub_adaptive(_dev.stream(), 5, m, m, K, A.data_ptr(), m,
                                             B.data_ptr(), K, C.data_ptr(), m, ws.data_ptr(), 2, 2,
                                             0, -2 ** 31, sc.data_ptr(), ctypes.byref(su),
                                             ctypes.byref(er))
        assert st == 0, L.mpdok_last_error()
        for S in (2, 3, 4, 5):
            def run():
                r = L.mpdok_ozaki_apply_sliced(_dev.stream(), 5, m, m, K, C.data_ptr(), m,
                                               ws.data_ptr(), S, sc.data_ptr())
                assert r == 0, L.mpdok_last_error()
            run()
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
            torch.cuda.synchronize()
            for r in range(reps):
                ev[2 * r].record()
                run()
                ev[2 * r + 1].record()
            torch.cuda.synchronize()
            ms = min(ev[2 * r].elapsed_time(ev[2 * r + 1]) for r in range(reps))
            ops = 2.0 * m * m * K * S * (S + 1) / 2
            roof = max(ops / (PEAK_INT8_TOPS * 1e12), 16.0 * m * m / (PEAK_HBM_GBS * 1e9))
            out["cases"].append({"m": m, "n": m, "K": K, "S": S, "ms": ms,
                                 "ps_per_element": ms * 1e9 / (m * m),
                                 "int8_TOPS": ops / (ms * 1e-3) / 1e12,
                                 "algorithmic_TFLOPs": 2.0 * m * m * K / (ms * 1e-3) / 1e12,
                                 "frac_of_mixed_roofline": roof / (ms * 1e-3)})
        del A, B, C, ws
        torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--slices", type=int, default=5)
    ap.add_argument("--tol", type=float, default=1e-12)
    ap.add_argument("--nb", type=int, default=512)
    ap.add_argument("--cpu-sample", type=int, default=24576,
                    help="points of the bounded CPU sample (about 30 s of LAPACK on 16 cores)")
    ap.add_argument("--cpu-gmres-iters", type=int, default=20,
                    help="SciPy GMRES iterations on the full matrix in the CPU baseline (0: skip)")
    ap.add_argument("--cpu-full", action="store_true",
                    help="also time scipy.linalg.solve on the full system (minutes of CPU)")
    ap.add_argument("--gemv-profile", action="store_true",
                    help="only measure the GEMV / residual kernels and the read-only ceiling "
                         "at N (HIP events) and print them as one JSON line")
    ap.add_argument("--update-profile", action="store_true",
                    help="only time the int8-sliced trailing-update kernel per slice count at "
                         "the LU's panel-pair shape and print one JSON line")
    ap.add_argument("--rbf-profile", action="store_true",
                    help="time RBF assembly and the matrix-free products, then exit")
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
    if args.gemv_profile:
        print(json.dumps(gemv_profile(N)), flush=True)
        return 0
    if args.update_profile:
        print(json.dumps(update_profile()), flush=True)
        return 0
    if args.rbf_profile:
        print(json.dumps(rbf_profile()), flush=True)
        return 0
    log(f"[bench] device {torch.cuda.get_device_name(0)}  N={N}  slices={args.slices}")
    X = np.random.default_rng(42).uniform(0.0, 1.0, (N, 3))
    b_h = rhs_of(X)
    Xd = _dev.as_device(X, torch.float64)
    b = _dev.as_device(b_h, torch.float64)
    A = _dev.empty((N, N), torch.float64, "F")
    _ops.rbf_assemble(Xd, 0, 25.0, 1e-6, A)  # the first call also loads the kernel code
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _ops.rbf_assemble(Xd, 0, 25.0, 1e-6, A)
    torch.cuda.synchronize()
    t_assemble = time.perf_counter() - t0
    log(f"[bench] assembled K in {t_assemble * 1e3:.2f} ms "
        f"({8.0 * N * N / t_assemble / 1e12:.2f} TB/s of FP64 written)")

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
        t_tr = st["trailing_ms"] * 1e-3
        # algorithmic work: 2 m n k flops per update (the FP64 GEMM the
        # sliced product stands for), against the dense int8 peak
        achieved = st["trailing_flops"] / t_tr / 1e12
        int8_rate = st["trailing_ops"] / t_tr / 1e12
        # per launch: the larger of its int8-MFMA time and its HBM time (FP64
        # C read + write) at peak, over the measured time
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
        roofline = {"bound": "mfma", "kernel": "ozaki_gemmd_kernel<5, 2, 4, 2, 1, 4, 1, 1>",
                    "achieved": achieved, "peak": PEAK_INT8_TOPS, "unit": "TFLOP/s",
                    "frac": achieved / PEAK_INT8_TOPS,
                    "traffic": measured_traffic(N, args.slices, args.nb),
                    "algorithmic_bytes_per_launch": 16.0 * st["trailing_flops"] / (2 * 1024) / st["launches"],
                    "launches": st["launches"], "avg_launch_ms": avg_ms,
                    "int8_ops_rate_TOPS": int8_rate,
                    "int8_ops_frac": int8_rate / PEAK_INT8_TOPS,
                    "mixed_roofline_frac": t_roof / t_tr,
                    "note": ("achieved: algorithmic flops of the trailing updates, 2 m n k per "
                             "launch (k = 1024, a panel pair; the FP64 product the int8-sliced "
                             "one replaces), over their summed HIP-event time on the "
                             "factorisation's stream, against the 5.0 POP/s dense int8 peak. "
                             "int8_ops_rate: the int8 operations actually done, 2 m n k "
                             "S_k (S_k + 1) / 2 per launch (slice_counts). mixed_roofline_frac: "
                             "sum over launches of max(int8 ops / 5.0 POP/s, 16 B per C element "
                             "/ 8 TB/s) over the measured time. traffic: HBM bytes per launch "
                             "(rocprofv3 FETCH_SIZE + WRITE_SIZE, profiles/"
                             "ozaki_gemmd_traffic_n32768.md)")}
    plain_roof = None
    if plain_stats and plain_stats.get("launches"):
        a = plain_stats["trailing_ops"] / (plain_stats["trailing_ms"] * 1e-3) / 1e12
        plain_roof = {"kernel": "gemm_kernel<double>", "achieved": a, "peak": PEAK_FP64_TFLOPS,
                      "unit": "TFLOP/s", "frac": a / PEAK_FP64_TFLOPS}

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline ...")
        cpu = cpu_baseline(N, min(args.cpu_sample, N), A_dev=A, b_full=b_h,
                           gmres_iters=args.cpu_gmres_iters)
        if args.cpu_full:
            cpu["full_n_solve"] = cpu_full(N)
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
        "assembly_ms": t_assemble * 1e3, "assembly_write_TBps": 8.0 * N * N / t_assemble / 1e12,
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
    from mpdok_amd.dense_krylov import DenseLinearOperator, cg
    import contextlib
    with contextlib.redirect_stdout(sys.stderr):     # the operator prints its sizes
        op = DenseLinearOperator(A)
    (x3, h3, c3), t3 = timed(lambda: cg(op, bd, tol=1e-12, body="fp64"))
    res["n4096_2d"] = {"gmres_ir_int8_lu": {"s": t, "relres": rel(x)},
                       "lu_ir_fp32_lu_reference_symbol": {"s": t2, "relres": rel(x2)},
                       "cg_fp64": {"s": t3, "relres": rel(x3), "iterations": h3[-1][0],
                                   "converged": bool(c3),
                                   "note": "stops by the reference's rule (<1% improvement "
                                           "between checks every 10 iterations)"}}
    del op
    log(f"[configs] n4096 {res['n4096_2d']}")
    del A, s, lu
    torch.cuda.empty_cache()

    # the precision ladder on the headline system (N = 32768, 3-D, gamma 25,
    # nugget 1e-6): LU preconditioners of each precision, GMRES-IR with the
    # FP64 body and Dot2 residual (and the FP32 body with the int8-sliced LU)
    N = 32768
    X = np.random.default_rng(42).uniform(0.0, 1.0, (N, 3))
    b = _dev.as_device(rhs_of(X), torch.float64)
    A = _dev.empty((N, N), torch.float64, "F")
    _ops.rbf_assemble(_dev.as_device(X, torch.float64), 0, 25.0, 1e-6, A)
    ladder = {}
    for prec, body in [("ozaki", "fp64"), ("ozaki", "fp32"), ("fp64", "fp64"), ("fp32", "fp64"),
                       ("bf16", "fp64"), ("fp16", "fp64"), ("fp8", "fp64")]:
        s = MPDOKSolver()
        s.profile = True
        kwl = dict(tol=1e-12, maxiter_outer=4, restart=50, precond="lu", factor_precision=prec,
                   slices=5, pivot="none", body=body, residual="dot2", nb=512)
        _, tl = timed(lambda: s.solve(A, b, **kwl))
        st = s.last_info["factor_stats"] or {}
        ladder[f"{prec}_lu_{body}_body"] = {
            "s": tl, "converged": s.last_info["converged"],
            "inner_iterations": s.last_info["inner_iterations"],
            "history": [float(h) for h in s.last_info["history"]],
            "trailing_update_ms": st.get("trailing_ms"),
            "trailing_update_TFLOPs": (st["trailing_flops"] / st["trailing_ms"] / 1e9
                                       if st.get("trailing_ms") else None)}
        s.free_fp32_buf()
        del s
        torch.cuda.empty_cache()
    res["precision_ladder_n32768"] = ladder
    log(f"[configs] ladder {ladder}")
    del A
    torch.cuda.empty_cache()

    # factor once, then 72 right-hand sides (the reference's weather-front time
    # series, run_demo.py:75-109 with its TOL = 1e-5 and maxiter_outer = 3;
    # rbf_demo.ipynb:247): the reference-default LUIRSolver (FP32 LU with
    # partial pivoting, FP64 refinement) at N = 8192 and 32768, 72
    # solve_factored calls one at a time and as one block
    from mpdok_amd import rbf_kernel
    for N in (8192, 32768):
        coords = rbf_kernel.synthetic_coords(N, seed=0)
        A, gam = rbf_kernel.build_rbf_kernel(coords, reg=1e-2)
        Bs = torch.stack([rbf_kernel.weather_front(coords, t) for t in range(72)], dim=1)
        lu = LUIRSolver()
        lu.factor(A)
        _, tfac = timed(lambda: lu.factor(A))
        kw72 = dict(tol=1e-5, maxiter_outer=3)
        xs = lu.solve_factored(Bs[:, 0].contiguous(), **kw72)
        _, tone = timed(lambda: [lu.solve_factored(Bs[:, t].contiguous(), **kw72)
                                 for t in range(72)])
        Xb = lu.solve_factored(Bs, **kw72)
        _, tblk = timed(lambda: lu.solve_factored(Bs, **kw72))
        rels = [float(_ops.norm(_ops.residual(A, Xb[:, t].contiguous(), Bs[:, t].contiguous(),
                                              mode="fp64")) / _ops.norm(Bs[:, t].contiguous()))
                for t in range(72)]
        res[f"factor_once_72_rhs_n{N}"] = {
            "factor_s": tfac, "solve_one_at_a_time_s_per_rhs": tone / 72,
            "solve_block_s_per_rhs": tblk / 72, "max_relres": max(rels),
            "all_converged": bool(all(lu.last_info["converged"])), "gamma": gam,
            "method": "LUIRSolver() defaults: FP32 getrf-style LU with partial pivoting, FP64 "
                      "residual and update; solve_factored(tol=1e-5, maxiter_outer=3) as "
                      "run_demo.py:35,90"}
        log(f"[configs] factor once N={N} {res[f'factor_once_72_rhs_n{N}']}")
        lu.free_factored()
        del A, Bs, Xb, lu
        torch.cuda.empty_cache()

    # out-of-core: Matern-3/2, length scale 0.1, nugget 1e-4. N = 262144
    # (275 GB in FP32, more host memory than this process may use) runs
    # matrix-free: Jacobi-preconditioned FP32 CG inside the FP64 refinement,
    # capped at one outer step of 50 CG iterations (the full solve needs tens
    # of thousands of iterations at this condition number)
    from mpdok_amd.mpdok_ooc import MPDOKOOCSolver
    N = 262144
    X = np.random.default_rng(42).uniform(0.0, 1.0, (N, 3))
    bo = np.sin(2 * np.pi * X[:, 0]) * np.cos(2 * np.pi * X[:, 1]) * np.cos(2 * np.pi * X[:, 2])
    o = MPDOKOOCSolver(tile_rows=4096)
    o.build(X, reg=1e-4, store="free", verbose=False, kernel="matern32", length_scale=0.1)
    v64 = torch.rand(N, dtype=torch.float64, device=_dev.device())
    o._tiled_dgemv_fp64(v64)
    _, tf64 = timed(lambda: o._tiled_dgemv_fp64(v64))
    v32 = v64.float()
    o._tiled_sgemv(v32)
    _, tf32 = timed(lambda: o._tiled_sgemv(v32))
    its = 50
    _, tcg = timed(lambda: o.solve(bo, tol=1e-10, maxiter_outer=1, inner="cg", precond="jacobi",
                                   inner_maxiter=its, verbose=False))
    res["ooc_matern_n262144_matrix_free_cg"] = {
        "matrix_free_fp64_product_s": tf64, "matrix_free_fp32_product_s": tf32,
        "kernel_evals_per_s_fp32": N * N / tf32, "kernel_evals_per_s_fp64": N * N / tf64,
        "capped_solve_s": tcg, "inner_cg_iterations": o.last_info["inner_iterations"],
        "s_per_cg_iteration": tcg / max(1, sum(o.last_info["inner_iterations"])),
        "relres_history": o.last_info["history"],
        "note": "one outer FP64 step, 50 Jacobi-CG iterations in FP32; capped, not converged"}
    log(f"[configs] ooc c3 {res['ooc_matern_n262144_matrix_free_cg']}")
    o.free()
    del o
    torch.cuda.empty_cache()

    # streamed FP32 operator from page-locked host memory (17 GB at N = 65536)
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
    res[f"ooc_matern_stream_n{N}"] = {"build_s": tb, "stream_product_s": ts,
                                      "stream_GBps": N * N * 4 / ts / 1e9,
                                      "matrix_free_fp64_product_s": tf}
    log(f"[configs] ooc stream {res[f'ooc_matern_stream_n{N}']}")
    o.free()
    del o
    torch.cuda.empty_cache()

    # batched uneven systems: 256 systems, N log-uniform in [256, 8192], 4 RHS
    from mpdok_amd.batched import BatchTuner, c4_workload
    sysw = c4_workload(seed=0)
    tuner = BatchTuner()
    Ns = np.array([len(q[0]) for q in sysw], dtype=float)
    # grouped launches captured in one hipGraph, tile-shape threshold tuned
    _, ginfo = tuner.solve_plan(sysw)
    res["batched_256_graph"] = {
        "s": ginfo["seconds"], "build_and_capture_s": ginfo["build_seconds"],
        "config": ginfo["config"], "captured_launches": ginfo["plan"]["launches"],
        "device_bytes": ginfo["plan"]["device_bytes"],
        "all_converged": bool(ginfo["converged"].all()),
        "max_relres": float(np.nanmax(ginfo["relres"])),
        "max_refinement_steps": int(ginfo["steps"].max()),
        "lu_flop": float((2.0 / 3.0 * Ns ** 3).sum()),
        "method": "one graph launch: assembly, FP64 no-pivot LU in 64-wide steps, LU-IR with "
                  "the Dot2 residual, device-side convergence flags"}
    log(f"[configs] batched graph {res['batched_256_graph']}")
    # the stream path (host threads, one stream each), tuned the same way
    choice = tuner.choose(sysw)
    _, info = tuner.solve(sysw)
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
