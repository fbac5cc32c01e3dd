"""Oracle: Krylov solvers (test infrastructure only).

CPU restatements of the reference algorithms with the reference's precisions:
FP32 Krylov vectors, FP64 residual checks and updates.
  cg        dense_krylov.py:119-205 (plus the FP64 body and the diagonal
            preconditioner the product offers as extensions)
  pcg / cg_ir  FP32 Jacobi-PCG and its FP64 refinement loop (the outer loop
            of mpdok_ooc.py:315-373 around a CG inner solver)
  gmres     dense_krylov.py:212-331 (one-pass batched CGS, lstsq on (m+1) x m H)
  gmres_ir  dense_krylov.py:338-421 (trailing (maxiter_outer, rel) history entry)
  inner_gmres / ooc_solve   mpdok_ooc.py:231-373 (two-pass DGKS, Givens,
            inner stop |g[k+1]|/beta < 1e-6, outer FP64 IR returning x only)
One intentional difference is shared with the product: on an Arnoldi
breakdown inner_gmres still rotates column k before stopping (the reference,
mpdok_ooc.py:278-282, exits before that rotation).
"""
import numpy as np


class DenseOp:
    """Row-major FP64 matrix with an FP32 copy (dense_krylov.py:44-47)."""

    def __init__(self, A):
        self.A64 = np.ascontiguousarray(A, dtype=np.float64)
        self.A32 = self.A64.astype(np.float32)
        self.N = self.A64.shape[0]

    def matvec(self, x32):
        return self.A32 @ x32.astype(np.float32)

    def residual_norm(self, x, b64):
        return float(np.linalg.norm(b64 - self.A64 @ x.astype(np.float64)))


def cg(op, b, x0=None, tol=1e-6, maxiter=None, check_every=10, body="fp32", dinv=None):
    """dense_krylov.py:119-205. body='fp64' runs the same loop on the FP64
    matrix with FP64 vectors; dinv (a vector) preconditions with z = dinv r
    (p = z + beta p, scalars r.z): the product's extensions."""
    N = op.N
    maxiter = N if maxiter is None else maxiter
    dt = np.float32 if body == "fp32" else np.float64
    mv = op.matvec if body == "fp32" else (lambda v: op.A64 @ v)
    b64 = np.asarray(b, np.float64)
    bv = b64.astype(dt)
    b_norm = float(np.linalg.norm(b64))
    if b_norm == 0.0:
        return np.zeros(N), [(0, 0.0)], True
    x = np.zeros(N, dt) if x0 is None else np.asarray(x0, dt).copy()
    dv = None if dinv is None else np.asarray(dinv, dt)
    r = bv - mv(x)
    z = r if dv is None else r * dv
    p = z.copy()
    rs = float(np.dot(r.astype(np.float64), z))
    history, converged, prev_rel = [], False, float("inf")
    k = -1
    for k in range(maxiter):
        if k % check_every == 0:
            rel = op.residual_norm(x, b64) / b_norm
            history.append((k, rel))
            if rel < tol:
                converged = True
                break
            if rel > prev_rel * (1.0 - 1e-2):
                break
            prev_rel = rel
        Ap = mv(p)
        alpha = dt(rs / float(np.dot(p.astype(np.float64), Ap)))
        x = x + alpha * p
        r = r - alpha * Ap
        z = r if dv is None else r * dv
        rs_new = float(np.dot(r.astype(np.float64), z))
        beta = dt(rs_new / rs)
        p = z + beta * p
        rs = rs_new
    if not converged:
        rel = op.residual_norm(x, b64) / b_norm
        history.append((k + 1, rel))
        converged = rel < tol
    return x.astype(np.float64), history, converged


def pcg(matvec, b, dinv=None, tol=1e-6, maxiter=1000, check_every=10, dtype=np.float32):
    """(Jacobi-)PCG on A e = b from e = 0 with vectors in `dtype` and alpha,
    beta rounded to it (the CG updates of dense_krylov.py:191-198 with z =
    dinv r). Stops when the recursive residual ||r|| < tol ||b||, tested
    every `check_every` iterations and at maxiter. Returns (e, iterations)."""
    r = np.asarray(b, np.float64).astype(dtype)
    beta0 = float(np.linalg.norm(r.astype(np.float64)))
    e = np.zeros(r.shape[0], dtype)
    if beta0 == 0.0:
        return e, 0
    dv = None if dinv is None else np.asarray(dinv, dtype)
    z = r if dv is None else r * dv
    p = z.copy()
    rz = float(np.dot(r.astype(np.float64), z))
    it = 0
    for it in range(1, maxiter + 1):
        Ap = np.asarray(matvec(p), dtype)
        alpha = dtype(rz / float(np.dot(p.astype(np.float64), Ap)))
        e = e + alpha * p
        r = r - alpha * Ap
        z = r if dv is None else r * dv
        rz_new = float(np.dot(r.astype(np.float64), z))
        p = z + dtype(rz_new / rz) * p
        rz = rz_new
        if it % check_every == 0 or it == maxiter:
            if float(np.linalg.norm(r.astype(np.float64))) < tol * beta0:
                break
    return e, it


def cg_ir(matvec32, matvec64, b, dinv=None, tol=1e-10, maxiter_outer=20, inner_tol=1e-6,
          inner_maxiter=None, check_every=10):
    """The outer FP64 refinement of mpdok_ooc.py:315-373 (r = b - A x in FP64,
    stop when ||r||/||b|| < tol, x += e in FP64) with pcg in FP32 as the
    inner solver. Returns (x, converged, outer relres history, inner
    iteration counts)."""
    b = np.asarray(b, np.float64)
    b_norm = float(np.linalg.norm(b))
    x = np.zeros_like(b)
    hist, inner = [], []
    if b_norm == 0.0:
        return x, True, hist, inner
    mi = b.shape[0] if inner_maxiter is None else inner_maxiter
    for _ in range(maxiter_outer):
        r = b - matvec64(x)
        rel = float(np.linalg.norm(r)) / b_norm
        hist.append(rel)
        if rel < tol:
            return x, True, hist, inner
        e, it = pcg(matvec32, r, dinv, inner_tol, mi, check_every, np.float32)
        inner.append(it)
        x = x + e.astype(np.float64)
    return x, False, hist, inner


def gmres(op, b, x0=None, tol=1e-6, maxiter=None, restart=50):
    N = op.N
    maxiter = N if maxiter is None else maxiter
    b64 = np.asarray(b, np.float64)
    b32 = b64.astype(np.float32)
    b_norm = float(np.linalg.norm(b64))
    if b_norm == 0.0:
        return np.zeros(N), [(0, 0.0)], True
    x = np.zeros(N, np.float32) if x0 is None else np.asarray(x0, np.float32).copy()
    history, converged, total, prev_rel = [], False, 0, float("inf")
    V = np.empty((restart + 1, N), np.float32)
    H = np.zeros((restart + 1, restart))
    for outer in range((maxiter + restart - 1) // restart):
        r = b32 - op.matvec(x)
        beta = float(np.linalg.norm(r.astype(np.float64)))
        if beta == 0.0:
            converged = True
            break
        V[0] = r / np.float32(beta)
        H[:] = 0.0
        m = 0
        for k in range(restart):
            if total >= maxiter:
                break
            w = op.matvec(V[k]).astype(np.float64)
            total += 1
            h = V[:k + 1].astype(np.float64) @ w
            H[:k + 1, k] = h
            w = (w - V[:k + 1].astype(np.float64).T @ h).astype(np.float32)
            nw = float(np.linalg.norm(w.astype(np.float64)))
            H[k + 1, k] = nw
            m = k + 1
            if nw < 1e-12:
                break
            V[k + 1] = w / np.float32(nw)
        e1 = np.zeros(m + 1)
        e1[0] = beta
        y = np.linalg.lstsq(H[:m + 1, :m], e1, rcond=None)[0]
        x = (x.astype(np.float64) + V[:m].astype(np.float64).T @ y).astype(np.float32)
        rel = op.residual_norm(x, b64) / b_norm
        history.append((total, rel))
        if rel < tol:
            converged = True
            break
        if rel > prev_rel * (1.0 - 1e-2):
            break
        prev_rel = rel
    return x.astype(np.float64), history, converged


def gmres_ir(op, b, tol=1e-11, maxiter_outer=5, restart=50, maxiter_inner=None):
    N = op.N
    b64 = np.asarray(b, np.float64)
    b_norm = float(np.linalg.norm(b64))
    if b_norm == 0.0:
        return np.zeros(N), [(0, 0.0)], True
    x = np.zeros(N)
    history, converged = [], False
    for outer in range(maxiter_outer):
        r64 = b64 - op.A64 @ x
        rel = float(np.linalg.norm(r64)) / b_norm
        history.append((outer, rel))
        if rel < tol:
            converged = True
            break
        e, _, _ = gmres(op, r64, tol=1e-6, restart=restart, maxiter=maxiter_inner)
        x = x + e
    r64 = b64 - op.A64 @ x
    rel = float(np.linalg.norm(r64)) / b_norm
    history.append((maxiter_outer, rel))
    return x, history, converged or rel < tol


def inner_gmres(matvec32, rhs64, restart, inner_tol=1e-6):
    """FP32 GMRES with DGKS and Givens (mpdok_ooc.py:231-311); returns e (FP32)."""
    N = rhs64.shape[0]
    rhs = np.asarray(rhs64).astype(np.float32)
    beta = float(np.linalg.norm(rhs.astype(np.float64)))
    if beta < 1e-30:
        return np.zeros(N, np.float32)
    V = np.empty((N, restart + 1), np.float32, order="F")
    V[:, 0] = rhs * np.float32(1.0 / beta)
    H = np.zeros((restart + 1, restart))
    cs = np.zeros(restart)
    sn = np.zeros(restart)
    g = np.zeros(restart + 1)
    g[0] = beta
    m = restart
    for k in range(restart):
        w = matvec32(V[:, k]).astype(np.float32)
        Vk = V[:, :k + 1].astype(np.float64)
        h = Vk.T @ w.astype(np.float64)
        H[:k + 1, k] = h
        w = (w.astype(np.float64) - Vk @ h).astype(np.float32)
        h2 = Vk.T @ w.astype(np.float64)
        H[:k + 1, k] += h2
        w = (w.astype(np.float64) - Vk @ h2).astype(np.float32)
        nrm = float(np.linalg.norm(w.astype(np.float64)))
        H[k + 1, k] = nrm
        brk = not nrm > 1e-12
        if not brk:
            V[:, k + 1] = w * np.float32(1.0 / nrm)
        for j in range(k):
            t = cs[j] * H[j, k] + sn[j] * H[j + 1, k]
            H[j + 1, k] = -sn[j] * H[j, k] + cs[j] * H[j + 1, k]
            H[j, k] = t
        rho = np.hypot(H[k, k], H[k + 1, k])
        c, s = (H[k, k] / rho, H[k + 1, k] / rho) if rho > 0 else (1.0, 0.0)
        cs[k], sn[k] = c, s
        H[k, k], H[k + 1, k] = rho, 0.0
        g[k + 1] = -s * g[k]
        g[k] = c * g[k]
        if brk or abs(g[k + 1]) / beta < inner_tol:
            m = k + 1
            break
    y = np.zeros(m)
    y[m - 1] = g[m - 1] / H[m - 1, m - 1]
    for j in range(m - 2, -1, -1):
        y[j] = (g[j] - np.dot(H[j, j + 1:m], y[j + 1:m])) / H[j, j]
    return (V[:, :m].astype(np.float64) @ y).astype(np.float32)


def ooc_solve(matvec32, matvec64, b, tol=1e-5, maxiter_outer=20, restart=50):
    """Outer FP64 IR of mpdok_ooc.py:315-373; returns x only."""
    b = np.asarray(b, np.float64)
    b_norm = float(np.linalg.norm(b))
    if b_norm == 0.0:
        return np.zeros_like(b), True
    x = np.zeros_like(b)
    converged = False
    for outer in range(maxiter_outer):
        r = b - matvec64(x)
        if float(np.linalg.norm(r)) / b_norm < tol:
            converged = True
            break
        x = x + inner_gmres(matvec32, r, restart).astype(np.float64)
    return x, converged
