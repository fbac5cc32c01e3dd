"""Oracle: LU factorisations, sliced-integer products and LU-IR (test
infrastructure only).

  lu_ir          LUIRSolver algorithm as documented in benchmark_mpdok.ipynb:983-996
                 (FP32 getrf, FP32 getrs, FP64 residual, FP64 update), with the
                 outer-loop contract of mpdok_ops.py:204-243 / :256-320.
  lu_nopiv       blocked right-looking LU without pivoting (FP64 panel), the
                 trailing update supplied as a callable.
  ozaki_*        Ozaki-scheme splitting of FP64 matrices into int8 digit slices
                 (a 7-bit signed leading digit, then balanced 8-bit digits) with
                 power-of-two row/column scales (Ozaki, Ogita, Oishi, Rump,
                 Numer. Algorithms 59 (2012)); products of slice pairs are exact
                 integers; pairs with i + j <= s + 1 are kept.
  lowp_matmul    products of BF16 / FP16 / scaled E4M3 operands (OCP FP8
                 specification's E4M3FN format) for the precision ladder.
  residual_*     exact and Dot2 residuals (Ogita, Rump, Oishi, SIAM J. Sci.
                 Comput. 26 (2005)).
"""
from fractions import Fraction

import numpy as np
import scipy.linalg as sla


# ---------------------------------------------------------------------------
# LU-IR (reference: LUIRSolver / lu_ir.so)
# ---------------------------------------------------------------------------
def lu_ir(A, b, tol=1e-11, maxiter_outer=3, factor_dtype=np.float32):
    """Returns (x FP64, converged, history [relres per outer step])."""
    A = np.asarray(A, np.float64)
    b = np.asarray(b, np.float64)
    lu = sla.lu_factor(A.astype(factor_dtype), check_finite=False)
    b_norm = float(np.linalg.norm(b))
    x = np.zeros_like(b)
    hist = []
    if b_norm == 0.0:
        return x, True, [0.0]
    for _ in range(maxiter_outer):
        r = b - A @ x
        rel = float(np.linalg.norm(r)) / b_norm
        hist.append(rel)
        if rel < tol:
            return x, True, hist
        x = x + sla.lu_solve(lu, r.astype(factor_dtype)).astype(np.float64)
    rel = float(np.linalg.norm(b - A @ x)) / b_norm
    hist.append(rel)
    return x, rel < tol, hist


# ---------------------------------------------------------------------------
# Sliced-integer (Ozaki) products
# ---------------------------------------------------------------------------
def ozaki_bits(s):
    """Bits kept per element by s slices: a signed first digit of 7 bits, then
    s - 1 balanced 8-bit digits."""
    return 8 * s - 1


ZERO_EXP = -1000  # exponent of an all-zero row / column (its digits are all zero)


def ozaki_scales(M, axis):
    """Power-of-two scale per row (axis=1) or per column (axis=0): 2^e with
    max|.| * 2^(7 - e) <= 127, i.e. e = floor(log2 max|.|) + 1, plus one when
    the maximum's leading 8 bits round above 127/128; zero rows get
    ZERO_EXP (they scale nothing and stay below every real exponent)."""
    mx = np.max(np.abs(M), axis=axis)
    # frexp: mx = f 2^e with f in [0.5, 1), i.e. e = floor(log2 mx) + 1 exactly
    # (log2 itself can round up to an integer just below a power of two)
    _, e = np.frexp(mx)
    e = np.where(np.ldexp(mx, 7 - e) > 127.0, e + 1, e)
    return np.where(mx > 0, e, ZERO_EXP).astype(np.int64)


def ozaki_slices(M, s, axis):
    """Integer digits d_1..d_s with M ~= 2^(e - P) * sum_q d_q 256^(s - q),
    P = 8 s - 1: the element scaled by 2^(P - e) is rounded to the nearest
    integer V (ties to even), and V is written in balanced base 256 from the
    least significant digit (d_q in [-128, 127] for q >= 2; |d_1| <= 127
    because |V| <= 127 * 256^(s - 1)). Error <= 2^(e - P - 1) per element.
    Returns (list of int8 arrays, exponent array)."""
    M = np.asarray(M, np.float64)
    e = ozaki_scales(M, axis)
    P = ozaki_bits(s)
    sh = (P - e)[:, None] if axis == 1 else (P - e)[None, :]
    V = np.rint(np.ldexp(M, np.minimum(sh, 1100))).astype(np.int64)
    out = [None] * s
    for q in range(s - 1, 0, -1):
        d = ((V + 128) & 0xFF) - 128
        V = (V - d) >> 8
        out[q] = d.astype(np.int8)
    assert np.abs(V).max(initial=0) <= 127
    out[0] = V.astype(np.int8)
    return out, e


def _ozaki_fold(da, ea, db, eb, s):
    """C = sum over kept levels of the digit products: for l = 2..s+1,
    acc_l = sum_{i+j=l} A_i B_j (exact integers), c = sum_l acc_l 2^(2 - 8 l)
    in FP64 from l = 2 upwards, result c * 2^(eA_row + eB_col)."""
    m, n = da[0].shape[0], db[0].shape[1]
    c = np.zeros((m, n))
    for l in range(2, s + 2):
        acc = np.zeros((m, n), np.int64)
        for i in range(1, l):
            j = l - i
            if j <= s:
                acc += da[i - 1].astype(np.int64) @ db[j - 1].astype(np.int64)
        c = c + np.ldexp(acc.astype(np.float64), 2 - 8 * l)
    return np.ldexp(np.ldexp(c, ea[:, None]), eb[None, :])


def ozaki_matmul(A, B, s):
    """Sliced product of A (m x k) and B (k x n) with s digits per element:
    pairs of digits with i + j <= s + 1 are kept (see _ozaki_fold)."""
    return ozaki_matmul_lead(A, B, s, s)


def ozaki_matmul_lead(A, B, smax, s):
    """The product of the leading s of smax digits (s <= smax): the digits
    come from the smax-digit split, so dropping the trailing ones truncates
    each element to within about half a unit of digit s."""
    da, ea = ozaki_slices(A, smax, 1)
    db, eb = ozaki_slices(B, smax, 0)
    return _ozaki_fold(da[:s], ea, db[:s], eb, s)


def ozaki_decide(A, B, smax, smin, delta, e0):
    """Slice count of one trailing update (the library's adaptive rule):
    E = max row exponent of A + max column exponent of B; the first update
    (e0 None) takes smax and records E; a later one takes
    S = smax + ceil((E - e0 + delta) / 8) clamped to [smin, smax], the fewest
    slices whose error bound 2^(E - (8 S - 1)) stays within the first
    update's minus delta bits. Returns (S, e0)."""
    E = int(ozaki_scales(A, 1).max(initial=ZERO_EXP)) + int(ozaki_scales(B, 0).max(initial=ZERO_EXP))
    if e0 is None:
        return smax, E
    d = E - e0 + delta
    S = smax + (-((-d) // 8))  # ceil(d / 8) for either sign
    return min(max(S, smin), smax), e0


class OzakiAdaptiveTrailing:
    """Trailing-update callable for lu_nopiv with adaptive slice counts;
    .counts lists the count of every update in order."""

    def __init__(self, smax, smin=2, delta=0):
        self.smax, self.smin, self.delta = smax, smin, delta
        self.e0 = None
        self.counts = []

    def __call__(self, L21, U12):
        S, self.e0 = ozaki_decide(L21, U12, self.smax, self.smin, self.delta, self.e0)
        self.counts.append(S)
        return ozaki_matmul_lead(L21, U12, self.smax, S)


# ---------------------------------------------------------------------------
# Low-precision operands (the precision ladder)
# ---------------------------------------------------------------------------
def round_bf16(x):
    """Nearest BF16 value of FP32 x (8 significant bits), ties to even: the
    FP32 bit pattern rounded at bit 16."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def round_e4m3(x):
    """Nearest OCP FP8 E4M3 value (E4M3FN: 3 fraction bits, exponent bias 7,
    smallest normal 2^-6, subnormal step 2^-9, largest 448; larger magnitudes
    saturate to 448), ties to even."""
    x = np.asarray(x, np.float64)
    a = np.abs(x)
    _, e = np.frexp(a)                       # a = f 2^e, f in [0.5, 1)
    q = np.maximum(e - 4, -9)
    r = np.ldexp(np.rint(np.ldexp(a, -q)), q)
    r = np.minimum(r, 448.0)
    return np.where(a > 0, np.copysign(r, x), x)


def e4m3_scale(M):
    """Power-of-two scale 2^(e-1) with 448 / max|M| = f 2^e (f in [0.5, 1)),
    so that max|M| times the scale stays within 448; 1 for an all-zero M."""
    mx = float(np.max(np.abs(M), initial=0.0))
    if mx == 0.0:
        return 1.0
    _, e = np.frexp(448.0 / mx)
    return float(np.ldexp(1.0, int(e) - 1))


def lowp_matmul(A, B, kind):
    """A B with the operands rounded as the matrix cores see them ('bf16',
    'fp16', or 'fp8': E4M3 after one power-of-two scale per operand, the
    result scaled back), products and sums exact (FP64) for comparison with an
    FP32-accumulated device result."""
    A32 = np.asarray(A, np.float32)
    B32 = np.asarray(B, np.float32)
    if kind == "bf16":
        return round_bf16(A32).astype(np.float64) @ round_bf16(B32).astype(np.float64)
    if kind == "fp16":
        return A32.astype(np.float16).astype(np.float64) @ B32.astype(np.float16).astype(np.float64)
    sa, sb = e4m3_scale(A32), e4m3_scale(B32)
    Aq = round_e4m3(A32.astype(np.float64) * sa)
    Bq = round_e4m3(B32.astype(np.float64) * sb)
    return (Aq @ Bq) / (sa * sb)


# ---------------------------------------------------------------------------
# Blocked LU without pivoting
# ---------------------------------------------------------------------------
def lu_nopiv(A, nb=256, trailing=None, starts=None):
    """Right-looking blocked LU (no pivoting) in FP64; L unit-lower and U packed
    in one array. trailing(L21, U12) returns the product used for
    A22 -= L21 @ U12 (default: the FP64 product). starts: first columns of
    the blocks when they are not all nb wide (ascending, starting at 0)."""
    A = np.array(A, np.float64, order="F", copy=True)
    n = A.shape[0]
    mm = trailing if trailing is not None else (lambda L, U: L @ U)
    ks = list(starts) if starts is not None else list(range(0, n, nb))
    for b, k in enumerate(ks):
        e = min(ks[b + 1], n) if b + 1 < len(ks) else n
        for j in range(k, e):
            if A[j, j] == 0.0:
                raise ZeroDivisionError(f"zero pivot at {j}")
            A[j + 1:, j] /= A[j, j]
            A[j + 1:, j + 1:e] -= np.outer(A[j + 1:, j], A[j, j + 1:e])
        if e < n:
            L11 = np.tril(A[k:e, k:e], -1) + np.eye(e - k)
            A[k:e, e:] = sla.solve_triangular(L11, A[k:e, e:], lower=True, unit_diagonal=True)
            A[e:, e:] -= mm(A[e:, k:e], A[k:e, e:])
    return A


def lu_solve_packed(LU, b, piv=None):
    """Solve with packed factors (unit-lower L, upper U); piv in LAPACK
    0-based ipiv form (row i swapped with piv[i], in order) or None."""
    x = np.array(b, np.float64, copy=True)
    if piv is not None:
        for i, p in enumerate(piv):
            if p != i:
                x[[i, p]] = x[[p, i]]
    x = sla.solve_triangular(LU, x, lower=True, unit_diagonal=True)
    return sla.solve_triangular(LU, x, lower=False)


def unpack(LU):
    n = LU.shape[0]
    return np.tril(LU, -1) + np.eye(n), np.triu(LU)


# ---------------------------------------------------------------------------
# Residuals
# ---------------------------------------------------------------------------
def residual_exact(A, x, b):
    """r = b - A x in exact rational arithmetic, rounded once to FP64
    (small inputs only)."""
    A = np.asarray(A, np.float64)
    n = A.shape[0]
    out = np.empty(n)
    xf = [Fraction(float(v)) for v in x]
    for i in range(n):
        s = Fraction(float(b[i]))
        for j in range(A.shape[1]):
            s -= Fraction(float(A[i, j])) * xf[j]
        out[i] = float(s)
    return out


def residual_longdouble(A, x, b):
    """r = b - A x in x87 extended precision (64-bit significand)."""
    AL = np.asarray(A, np.longdouble)
    return (np.asarray(b, np.longdouble) - AL @ np.asarray(x, np.longdouble)).astype(np.float64)


def forward_error(x, x_ref):
    return float(np.linalg.norm(np.asarray(x) - x_ref) / np.linalg.norm(x_ref))


def backward_error(A, x, b, r=None):
    """Normwise backward error ||b - A x|| / (||A|| ||x|| + ||b||) (2-norms,
    ||A|| estimated by the Frobenius norm, an upper bound)."""
    A = np.asarray(A)
    r = b - A @ x if r is None else r
    return float(np.linalg.norm(r) / (np.linalg.norm(A) * np.linalg.norm(x) + np.linalg.norm(b)))
