// Iterative-refinement drivers and the reference-compatible plugin symbols.
//
// gmres_ir: outer refinement in FP64 (residual r = b - A x with FP64 or Dot2
// accumulation, update x += e in FP64, stopping test ||r||/||b|| < tol) around
// one restart cycle of GMRES per outer step:
//   * body FP32 or FP64 (Arnoldi vectors and the matvec with A);
//   * optional right preconditioner M = P L U (an LU handle): the cycle solves
//     A inv(M) u = r and returns e = inv(M) u;
//   * CGS with one or two passes, device Givens rotations, inner stop at
//     |g[k+1]| / beta < inner_tol (mpdok_ooc.py:231-311, the port of the
//     native inner_gmres named at mpdok_ooc.py:229-236).
// lu_ir: x += inv(PLU) r with r in FP64 (benchmark_mpdok.ipynb:983-996).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <string>
#include <thread>
#include <vector>
#include "common.h"
#include "gemv.h"
#include "krylov.h"
#include "lu.h"
#include "rbf.h"

namespace mpdok {

static hipStream_t g_stream = nullptr;  // stream of the py_* symbols

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  int get(size_t bytes) {
    if (bytes <= n) return OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      set_error("workspace: hipMalloc of %zu bytes failed", bytes);
      return ENOMEM_;
    }
    n = bytes;
    return OK;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

struct IRWork {
  DevBuf V, Z, w, z, r, e, H, R, cs, sn, g, ctl, y, dots, orth, gemv, resid;
  int* ctl_host = nullptr;
  ~IRWork() {
    if (ctl_host) (void)hipHostFree(ctl_host);
  }
};

struct IROpts {
  double tol = 1e-11;
  int maxiter_outer = 5;
  int restart = 50;
  double inner_tol = 1e-6;
  int passes = 2;
  int resid_mode = 0;  // 0 FP64, 1 Dot2
  int body = 0;        // 0 FP32, 1 FP64
};

struct IRResult {
  int converged = 0;
  int outer = 0;
  int inner_total = 0;
  std::vector<double> hist;
};

static int host_norm(hipStream_t s, const double* v, int n, DevBuf& dots, double* out) {
  MPDOK_TRY(dots.get(reduce_workspace_bytes(n) + 64));
  double* res = reinterpret_cast<double*>(static_cast<char*>(dots.p) + reduce_workspace_bytes(n));
  MPDOK_TRY(dot_f64(s, v, v, n, res, dots.p));
  double h = 0.0;
  MPDOK_HIP(hipMemcpyAsync(&h, res, sizeof(double), hipMemcpyDeviceToHost, s));
  MPDOK_HIP(hipStreamSynchronize(s));
  *out = std::sqrt(h);
  return OK;
}

// One cycle of (right-preconditioned) GMRES for A e = r. Body type TB.
// With a preconditioner the cycle keeps z_k = inv(M) v_k (FP64) beside the
// basis, so the correction is e = Z y (flexible GMRES's update form): the
// same vector as inv(M) (V y) in exact arithmetic, with one triangular-solve
// pass less per cycle.
template <typename TB>
static int gmres_cycle(hipStream_t s, const void* Abody, int64_t ldab, int orderb,
                       LUHandle* M, int N, const double* r, double rnorm, double* e,
                       const IROpts& o, IRWork& W, int* steps) {
  const int m = o.restart;
  const int ldh = m + 1;
  TB* V = static_cast<TB*>(W.V.p);
  TB* w = static_cast<TB*>(W.w.p);
  double* H = static_cast<double*>(W.H.p);
  double* R = static_cast<double*>(W.R.p);
  double* cs = static_cast<double*>(W.cs.p);
  double* sn = static_cast<double*>(W.sn.p);
  double* g = static_cast<double*>(W.g.p);
  int* ctl = static_cast<int*>(W.ctl.p);
  MPDOK_HIP(hipMemsetAsync(H, 0, sizeof(double) * ldh * m, s));
  MPDOK_HIP(hipMemsetAsync(g, 0, sizeof(double) * ldh, s));
  MPDOK_HIP(hipMemsetAsync(ctl, 0, 4 * sizeof(int), s));
  MPDOK_HIP(hipMemcpyAsync(g, &rnorm, sizeof(double), hipMemcpyHostToDevice, s));
  // V_0 = r / ||r||
  MPDOK_TRY(axpby(s, N, 1.0 / rnorm, r, F64, 0.0, nullptr, F64, V, sizeof(TB) == 8 ? F64 : F32));
  const int dt = sizeof(TB) == 8 ? F64 : F32;
  int done = 0;
  for (int k = 0; k < m; ++k) {
    TB* vk = V + (size_t)k * N;
    const TB* xin = vk;
    if (M != nullptr) {
      // z_k = inv(M) v_k in FP64, kept for the update; the matvec takes it in
      // the body type
      double* zk = static_cast<double*>(W.Z.p) + (size_t)k * N;
      MPDOK_TRY(axpby(s, N, 1.0, vk, dt, 0.0, nullptr, F64, zk, F64));
      MPDOK_TRY(lu_solve(M, s, zk, zk));
      if (sizeof(TB) == 8) {
        xin = reinterpret_cast<const TB*>(zk);
      } else {
        MPDOK_TRY(axpby(s, N, 1.0, zk, F64, 0.0, nullptr, F64, w, dt));
        xin = w;
      }
    }
    TB* y = (M != nullptr) ? reinterpret_cast<TB*>(W.e.p) : w;  // matvec output
    if (sizeof(TB) == 4)
      MPDOK_TRY(gemv_f32(s, (const float*)Abody, ldab, orderb, N, N, (const float*)xin, (float*)y,
                         1.0f, W.gemv.p, W.gemv.n, ctl));
    else
      MPDOK_TRY(gemv_f64(s, (const double*)Abody, ldab, orderb, N, N, (const double*)xin,
                         (double*)y, 1.0, W.gemv.p, W.gemv.n, ctl));
    MPDOK_TRY(arnoldi_orth<TB>(s, V, N, k, y, N, H + (size_t)k * ldh, o.passes, 1e-12, 0, ctl,
                               W.orth.p, W.orth.n));
    MPDOK_TRY(givens_step(s, H, ldh, k, R, cs, sn, g, rnorm, o.inner_tol, ctl));
    done = k + 1;
    // The host learns whether the cycle stopped before issuing the next (more
    // expensive, when preconditioned) step: one 4-byte read per step.
    MPDOK_HIP(hipMemcpyAsync(W.ctl_host, ctl, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
    MPDOK_HIP(hipStreamSynchronize(s));
    if (W.ctl_host[0]) break;
  }
  // back substitution on the rotated R (ldh x m, column-major) and g
  std::vector<double> Rh((size_t)ldh * m), gh(ldh), yh(done);
  MPDOK_HIP(hipMemcpyAsync(Rh.data(), R, sizeof(double) * ldh * done, hipMemcpyDeviceToHost, s));
  MPDOK_HIP(hipMemcpyAsync(gh.data(), g, sizeof(double) * ldh, hipMemcpyDeviceToHost, s));
  MPDOK_HIP(hipStreamSynchronize(s));
  for (int j = done - 1; j >= 0; --j) {
    double acc = gh[j];
    for (int i = j + 1; i < done; ++i) acc -= Rh[(size_t)i * ldh + j] * yh[i];
    yh[j] = acc / Rh[(size_t)j * ldh + j];
  }
  double* ydev = static_cast<double*>(W.y.p);
  MPDOK_HIP(hipMemcpyAsync(ydev, yh.data(), sizeof(double) * done, hipMemcpyHostToDevice, s));
  // e = V y, or e = Z y with a preconditioner (FP64)
  if (M != nullptr)
    MPDOK_TRY((basis_update<double, double>(s, static_cast<const double*>(W.Z.p), N, done, ydev,
                                            e, N, 1)));
  else
    MPDOK_TRY((basis_update<TB, double>(s, V, N, done, ydev, e, N, 1)));
  *steps = done;
  return OK;
}

// A64: FP64 matrix for the residual (and the FP64 body); A32: FP32 copy for
// the FP32 body (may be null when body = FP64).
static int gmres_ir_impl(hipStream_t s, const double* A64, int64_t lda64, int order64,
                         const float* A32, int64_t lda32, int order32, LUHandle* M, int N,
                         const double* b, double* x, const IROpts& o, IRResult& res) {
  MPDOK_REQUIRE(N >= 1 && b && x && A64, "gmres_ir: bad arguments");
  MPDOK_REQUIRE(o.restart >= 1 && o.restart <= MAX_RESTART, "gmres_ir: restart must be in [1, %d]",
                MAX_RESTART);
  MPDOK_REQUIRE(o.body == 1 || A32 != nullptr, "gmres_ir: FP32 body needs the FP32 matrix");
  MPDOK_REQUIRE(M == nullptr || lu_n(M) == N, "gmres_ir: preconditioner size mismatch");
  static thread_local IRWork W;
  const size_t tb = o.body == 1 ? 8 : 4;
  const int m = o.restart;
  MPDOK_TRY(W.V.get(tb * (size_t)(m + 1) * N));
  if (M != nullptr) MPDOK_TRY(W.Z.get(8 * (size_t)m * N));
  MPDOK_TRY(W.w.get(8 * (size_t)N));
  MPDOK_TRY(W.z.get(8 * (size_t)N));
  MPDOK_TRY(W.r.get(8 * (size_t)N));
  MPDOK_TRY(W.e.get(8 * (size_t)N));
  MPDOK_TRY(W.H.get(8 * (size_t)(m + 1) * m));
  MPDOK_TRY(W.R.get(8 * (size_t)(m + 1) * m));
  MPDOK_TRY(W.cs.get(8 * (size_t)m));
  MPDOK_TRY(W.sn.get(8 * (size_t)m));
  MPDOK_TRY(W.g.get(8 * (size_t)(m + 1)));
  MPDOK_TRY(W.ctl.get(64));
  MPDOK_TRY(W.y.get(8 * (size_t)(m + 1)));
  MPDOK_TRY(W.orth.get(arnoldi_workspace_bytes(N, m)));
  const int ob = o.body == 1 ? order64 : order32;
  MPDOK_TRY(W.gemv.get(gemv_workspace_bytes(N, N, o.body == 1 ? F64 : F32, ob) + 64));
  MPDOK_TRY(W.resid.get(resid_workspace_bytes(N, order64) + 64));
  if (!W.ctl_host) MPDOK_HIP(hipHostMalloc((void**)&W.ctl_host, 64, 0));

  double* r = static_cast<double*>(W.r.p);
  double* e = static_cast<double*>(W.z.p);  // reuse: z is free between cycles
  double bnorm = 0.0;
  MPDOK_TRY(host_norm(s, b, N, W.dots, &bnorm));
  MPDOK_HIP(hipMemsetAsync(x, 0, sizeof(double) * N, s));
  res = IRResult();
  if (bnorm == 0.0) {
    res.converged = 1;
    res.hist.push_back(0.0);
    return OK;
  }
  double rel = 0.0;
  for (int outer = 0; outer < o.maxiter_outer; ++outer) {
    double rn = 0.0;
    if (outer == 0) {
      // x = 0: the residual is b exactly, without a pass over A
      MPDOK_HIP(hipMemcpyAsync(r, b, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
      rn = bnorm;
    } else {
      MPDOK_TRY(residual_f64(s, A64, lda64, order64, N, x, b, r, o.resid_mode, W.resid.p,
                             W.resid.n));
      MPDOK_TRY(host_norm(s, r, N, W.dots, &rn));
    }
    rel = rn / bnorm;
    res.hist.push_back(rel);
    res.outer = outer;
    if (rel < o.tol) {
      res.converged = 1;
      return OK;
    }
    if (!(rn > 0.0) || !std::isfinite(rn)) break;
    int steps = 0;
    double* ecyc = static_cast<double*>(W.e.p);
    if (o.body == 1)
      MPDOK_TRY(gmres_cycle<double>(s, A64, lda64, order64, M, N, r, rn, e, o, W, &steps));
    else
      MPDOK_TRY(gmres_cycle<float>(s, A32, lda32, order32, M, N, r, rn, e, o, W, &steps));
    (void)ecyc;
    res.inner_total += steps;
    MPDOK_TRY(axpby(s, N, 1.0, x, F64, 1.0, e, F64, x, F64));
  }
  MPDOK_TRY(residual_f64(s, A64, lda64, order64, N, x, b, r, o.resid_mode, W.resid.p, W.resid.n));
  double rn = 0.0;
  MPDOK_TRY(host_norm(s, r, N, W.dots, &rn));
  rel = rn / bnorm;
  res.hist.push_back(rel);
  res.outer = o.maxiter_outer;
  res.converged = rel < o.tol;
  return OK;
}

// LU-IR: x = 0; repeat { r = b - A x; stop if ||r||/||b|| < tol; x += inv(PLU) r }.
static int lu_ir_impl(hipStream_t s, const double* A64, int64_t lda, int order, LUHandle* M, int N,
                      const double* b, double* x, double tol, int maxiter_outer, int resid_mode,
                      IRResult& res) {
  static thread_local IRWork W;
  MPDOK_TRY(W.r.get(8 * (size_t)N));
  MPDOK_TRY(W.e.get(8 * (size_t)N));
  MPDOK_TRY(W.resid.get(resid_workspace_bytes(N, order) + 64));
  double* r = static_cast<double*>(W.r.p);
  double* d = static_cast<double*>(W.e.p);
  double bnorm = 0.0;
  MPDOK_TRY(host_norm(s, b, N, W.dots, &bnorm));
  MPDOK_HIP(hipMemsetAsync(x, 0, sizeof(double) * N, s));
  res = IRResult();
  if (bnorm == 0.0) {
    res.converged = 1;
    res.hist.push_back(0.0);
    return OK;
  }
  for (int it = 0; it < maxiter_outer; ++it) {
    double rn = 0.0;
    if (it == 0) {
      // x = 0: the residual is b exactly, without a pass over A
      MPDOK_HIP(hipMemcpyAsync(r, b, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
      rn = bnorm;
    } else {
      MPDOK_TRY(residual_f64(s, A64, lda, order, N, x, b, r, resid_mode, W.resid.p, W.resid.n));
      MPDOK_TRY(host_norm(s, r, N, W.dots, &rn));
    }
    res.hist.push_back(rn / bnorm);
    res.outer = it;
    if (rn / bnorm < tol) {
      res.converged = 1;
      return OK;
    }
    MPDOK_TRY(lu_solve(M, s, r, d));
    MPDOK_TRY(axpby(s, N, 1.0, x, F64, 1.0, d, F64, x, F64));
  }
  MPDOK_TRY(residual_f64(s, A64, lda, order, N, x, b, r, resid_mode, W.resid.p, W.resid.n));
  double rn = 0.0;
  MPDOK_TRY(host_norm(s, r, N, W.dots, &rn));
  res.hist.push_back(rn / bnorm);
  res.outer = maxiter_outer;
  res.converged = rn / bnorm < tol;
  return OK;
}

// LU-IR for nrhs right-hand sides (column-major B, X with ldb, ldx >= N), in
// blocks of up to 8: one residual pass over A (mpdok_residual_multi_f64) and
// one multi-right-hand-side solve per outer step serve the whole block. Each
// column runs the single loop above (x = 0, first residual = b, stop when its
// relres < tol, no update after that), so its iterates, relres and result are
// bit for bit those of lu_ir_impl. A row-major A takes the single loop per
// column. converged / relres: host arrays of nrhs (the last relres checked).
static int lu_ir_multi_impl(hipStream_t s, const double* A64, int64_t lda, int order, LUHandle* M,
                            int N, const double* B, int64_t ldb, double* X, int64_t ldx, int nrhs,
                            double tol, int maxiter_outer, int resid_mode, int* converged,
                            double* relres) {
  if (order != COL_MAJOR) {
    for (int q = 0; q < nrhs; ++q) {
      IRResult res;
      MPDOK_TRY(lu_ir_impl(s, A64, lda, order, M, N, B + (int64_t)q * ldb, X + (int64_t)q * ldx,
                           tol, maxiter_outer, resid_mode, res));
      converged[q] = res.converged;
      relres[q] = res.hist.empty() ? NAN : res.hist.back();
    }
    return OK;
  }
  constexpr int NB = 8;
  static thread_local DevBuf Rb, Db, Wr, dots;
  MPDOK_TRY(Rb.get(sizeof(double) * (size_t)N * NB));
  MPDOK_TRY(Db.get(sizeof(double) * (size_t)N * NB));
  MPDOK_TRY(Wr.get(resid_multi_workspace_bytes(N) + 64));
  double* R = static_cast<double*>(Rb.p);
  double* Dm = static_cast<double*>(Db.p);
  for (int c0 = 0; c0 < nrhs; c0 += NB) {
    const int nr = std::min(NB, nrhs - c0);
    const double* Bc = B + (int64_t)c0 * ldb;
    double* Xc = X + (int64_t)c0 * ldx;
    double bn[NB], rn[NB];
    bool active[NB];
    int nact = 0;
    for (int q = 0; q < nr; ++q) {
      MPDOK_TRY(host_norm(s, Bc + (int64_t)q * ldb, N, dots, &bn[q]));
      MPDOK_HIP(hipMemsetAsync(Xc + (int64_t)q * ldx, 0, sizeof(double) * N, s));
      active[q] = bn[q] != 0.0;
      converged[c0 + q] = active[q] ? 0 : 1;
      relres[c0 + q] = 0.0;
      nact += active[q];
    }
    for (int it = 0; it < maxiter_outer && nact > 0; ++it) {
      if (it == 0) {
        // x = 0: the residual is b exactly, without a pass over A
        MPDOK_TRY((copy2d<double, double>(s, Bc, ldb, R, N, N, nr)));
        for (int q = 0; q < nr; ++q) rn[q] = bn[q];
      } else {
        MPDOK_TRY(residual_multi_f64(s, A64, lda, N, Xc, ldx, Bc, ldb, R, N, nr, resid_mode, Wr.p,
                                     Wr.n));
        for (int q = 0; q < nr; ++q)
          if (active[q]) MPDOK_TRY(host_norm(s, R + (size_t)q * N, N, dots, &rn[q]));
      }
      for (int q = 0; q < nr; ++q) {
        if (!active[q]) continue;
        relres[c0 + q] = rn[q] / bn[q];
        if (rn[q] / bn[q] < tol) {
          converged[c0 + q] = 1;
          active[q] = false;
          --nact;
        }
      }
      if (nact == 0) break;
      MPDOK_TRY(lu_solve_multi(M, s, R, N, Dm, N, nr));
      for (int q = 0; q < nr; ++q)
        if (active[q])
          MPDOK_TRY(axpby(s, N, 1.0, Xc + (int64_t)q * ldx, F64, 1.0, Dm + (size_t)q * N, F64,
                          Xc + (int64_t)q * ldx, F64));
    }
    if (nact > 0) {
      // out of outer steps: the last update's residual decides, as in lu_ir_impl
      MPDOK_TRY(residual_multi_f64(s, A64, lda, N, Xc, ldx, Bc, ldb, R, N, nr, resid_mode, Wr.p,
                                   Wr.n));
      for (int q = 0; q < nr; ++q) {
        if (!active[q]) continue;
        MPDOK_TRY(host_norm(s, R + (size_t)q * N, N, dots, &rn[q]));
        relres[c0 + q] = rn[q] / bn[q];
        converged[c0 + q] = rn[q] / bn[q] < tol;
      }
    }
  }
  MPDOK_HIP(hipStreamSynchronize(s));
  return OK;
}

// ---------------------------------------------------------------------------
// State of the reference-compatible symbols (one slot each, like the natives).
// ---------------------------------------------------------------------------
static std::mutex g_mu;
static LUHandle* g_lu = nullptr;
static void* g_managed_gmres = nullptr;
static void* g_managed_lu = nullptr;

static int fill_hist(const IRResult& r, double* hist, int hist_cap, int* n_hist) {
  if (n_hist) *n_hist = (int)r.hist.size();
  if (hist)
    for (int i = 0; i < (int)r.hist.size() && i < hist_cap; ++i) hist[i] = r.hist[i];
  return OK;
}

}  // namespace mpdok

using namespace mpdok;

MPDOK_API int mpdok_set_stream(hipStream_t s) {
  g_stream = s;
  return 0;
}

// GMRES-IR with an optional LU preconditioner handle (NULL = none).
// order64/order32: storage order of A64 / A32 (A32 may be NULL for body = 1).
// hist: host array receiving relres per outer step (and the final check).
MPDOK_API int mpdok_gmres_ir(hipStream_t s, const double* A64, int64_t lda64, int order64,
                             const float* A32, int64_t lda32, int order32, void* lu, int N,
                             const double* b, double* x, double tol, int maxiter_outer,
                             int restart, double inner_tol, int passes, int resid_mode, int body,
                             int* converged, double* hist, int hist_cap, int* n_hist,
                             int* inner_total) {
  IROpts o;
  o.tol = tol;
  o.maxiter_outer = maxiter_outer;
  o.restart = restart;
  o.inner_tol = inner_tol;
  o.passes = passes;
  o.resid_mode = resid_mode;
  o.body = body;
  IRResult r;
  MPDOK_TRY(gmres_ir_impl(s, A64, lda64, order64, A32, lda32, order32,
                          static_cast<LUHandle*>(lu), N, b, x, o, r));
  if (converged) *converged = r.converged;
  if (inner_total) *inner_total = r.inner_total;
  return fill_hist(r, hist, hist_cap, n_hist);
}

MPDOK_API int mpdok_lu_ir(hipStream_t s, const double* A64, int64_t lda, int order, void* lu, int N,
                          const double* b, double* x, double tol, int maxiter_outer,
                          int resid_mode, int* converged, double* hist, int hist_cap,
                          int* n_hist) {
  MPDOK_REQUIRE(lu != nullptr, "lu_ir: null factorization handle");
  IRResult r;
  MPDOK_TRY(lu_ir_impl(s, A64, lda, order, static_cast<LUHandle*>(lu), N, b, x, tol,
                       maxiter_outer, resid_mode, r));
  if (converged) *converged = r.converged;
  return fill_hist(r, hist, hist_cap, n_hist);
}

MPDOK_API int mpdok_lu_ir_multi(hipStream_t s, const double* A64, int64_t lda, int order, void* lu,
                                int N, const double* B, int64_t ldb, double* X, int64_t ldx,
                                int nrhs, double tol, int maxiter_outer, int resid_mode,
                                int* converged, double* relres) {
  MPDOK_REQUIRE(lu != nullptr, "lu_ir_multi: null factorization handle");
  MPDOK_REQUIRE(A64 && B && X && converged && relres && nrhs >= 0 && N >= 1,
                "lu_ir_multi: bad arguments");
  MPDOK_REQUIRE(ldb >= N && ldx >= N, "lu_ir_multi: leading dimension < N");
  MPDOK_REQUIRE(lu_n(static_cast<LUHandle*>(lu)) == N, "lu_ir_multi: handle holds another N");
  return lu_ir_multi_impl(s, A64, lda, order, static_cast<LUHandle*>(lu), N, B, ldb, X, ldx, nrhs,
                          tol, maxiter_outer, resid_mode, converged, relres);
}

// ---- reference-compatible symbols (void, out-pointers, synchronous) -------
MPDOK_API void py_gmres_ir(void* A64, void* b64, void* x64, void* A32, int N, double tol,
                           int maxiter_outer, int restart, int* converged_out) {
  std::lock_guard<std::mutex> lk(g_mu);
  hipStream_t s = g_stream;
  int conv = 0;
  if (cast_matrix(s, A64, F64, N, COL_MAJOR, A32, F32, N, COL_MAJOR, N, N) == OK) {
    IROpts o;
    o.tol = tol;
    o.maxiter_outer = maxiter_outer;
    o.restart = restart;
    o.inner_tol = 1e-6;
    o.passes = 2;
    o.resid_mode = 0;
    o.body = 0;
    IRResult r;
    if (gmres_ir_impl(s, (const double*)A64, N, COL_MAJOR, (const float*)A32, N, COL_MAJOR,
                      nullptr, N, (const double*)b64, (double*)x64, o, r) == OK)
      conv = r.converged;
  }
  (void)hipStreamSynchronize(s);
  if (converged_out) *converged_out = conv;
}

static void alloc_managed(void** slot, int N, void** ptr_out) {
  if (*slot) {
    (void)hipFree(*slot);
    *slot = nullptr;
  }
  void* p = nullptr;
  if (N > 0 && hipMallocManaged(&p, (size_t)N * N * sizeof(double), hipMemAttachGlobal) != hipSuccess)
    p = nullptr;
  *slot = p;
  if (ptr_out) *ptr_out = p;
}

MPDOK_API void py_alloc_managed(int N, void** ptr_out) {
  std::lock_guard<std::mutex> lk(g_mu);
  alloc_managed(&g_managed_gmres, N, ptr_out);
}
MPDOK_API void py_free_managed(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_managed_gmres) (void)hipFree(g_managed_gmres);
  g_managed_gmres = nullptr;
}
MPDOK_API void py_lu_alloc_managed(int N, void** ptr_out) {
  std::lock_guard<std::mutex> lk(g_mu);
  alloc_managed(&g_managed_lu, N, ptr_out);
}
MPDOK_API void py_lu_free_managed(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_managed_lu) (void)hipFree(g_managed_lu);
  g_managed_lu = nullptr;
}

// FP32 LU with partial pivoting (the reference's Sgetrf), kept in the slot.
MPDOK_API void py_lu_factor(void* A64, int N, int* info_out) {
  std::lock_guard<std::mutex> lk(g_mu);
  int info = -1;
  if (g_lu && lu_n(g_lu) != N) {
    lu_destroy(g_lu);
    g_lu = nullptr;
  }
  if (!g_lu) lu_create(N, LU_FP32, PIV_PARTIAL, 256, 0, &g_lu);
  if (g_lu && lu_factor(g_lu, g_stream, (const double*)A64, N, COL_MAJOR, &info) != OK) info = -1;
  (void)hipStreamSynchronize(g_stream);
  if (info_out) *info_out = info;
}

MPDOK_API void py_lu_solve_factored(void* A64, void* b64, void* x64, int N, double tol,
                                    int maxiter_outer, int* converged_out) {
  std::lock_guard<std::mutex> lk(g_mu);
  int conv = 0;
  if (!g_lu || lu_n(g_lu) != N) {
    conv = -1;
  } else {
    IRResult r;
    int st = lu_ir_impl(g_stream, (const double*)A64, N, COL_MAJOR, g_lu, N, (const double*)b64,
                        (double*)x64, tol, maxiter_outer, 0, r);
    conv = st == ENOFACTOR ? -1 : (st == OK ? r.converged : 0);
  }
  (void)hipStreamSynchronize(g_stream);
  if (converged_out) *converged_out = conv;
}

MPDOK_API void py_lu_ir(void* A64, void* b64, void* x64, int N, double tol, int maxiter_outer,
                        int* converged_out) {
  int info = -1;
  py_lu_factor(A64, N, &info);
  if (info != 0) {
    if (converged_out) *converged_out = 0;
    return;
  }
  py_lu_solve_factored(A64, b64, x64, N, tol, maxiter_outer, converged_out);
}

MPDOK_API void py_lu_free_factored(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  lu_destroy(g_lu);
  g_lu = nullptr;
}

// ---------------------------------------------------------------------------
// Batched uneven kernel systems (north star's batched configuration; no
// reference counterpart). System i: N[i] points coords[i] (N x D, row-major),
// kernel (kind, param[i]) + reg I, nrhs right-hand sides B[i] (N x nrhs,
// column-major) -> X[i]. Each system: assemble K in FP64, LU (precision: FP64
// or int8-sliced; no pivoting, K is SPD), then LU-IR with the Dot2 residual per
// right-hand side until relres <= tol.
//
// Scheduling: systems go to nstreams workers, largest N first, each to the
// worker with the least N^3 so far (longest-processing-time rule). A worker is
// a host thread with its own stream, one FP64 matrix buffer and one LU handle
// sized for its largest system (lu_resize for the others), so the
// host-synchronous convergence tests of one system never stall another
// worker's GPU work and nothing is allocated between systems.
// ---------------------------------------------------------------------------
MPDOK_API int mpdok_batched_rbf_solve(int nsys, const int* N, const double* const* coords, int D,
                                      int kind, const double* param, double reg, int nrhs,
                                      const double* const* B, double* const* X, double tol,
                                      int maxiter_outer, int precision, int slices, int nb,
                                      int nstreams, int* converged, double* relres) {
  MPDOK_REQUIRE(nsys >= 0 && nrhs >= 1 && D >= 1, "batched: bad sizes");
  if (nsys == 0) return OK;
  MPDOK_REQUIRE(N && coords && param && B && X && converged && relres, "batched: null pointer");
  MPDOK_REQUIRE(nstreams >= 1 && nstreams <= 64, "batched: nstreams must be in [1, 64]");
  MPDOK_REQUIRE(precision == LU_FP64 || precision == LU_OZAKI, "batched: precision %d", precision);
  for (int i = 0; i < nsys; ++i) MPDOK_REQUIRE(N[i] >= 1, "batched: N[%d] = %d", i, N[i]);
  int dev = 0;
  MPDOK_HIP(hipGetDevice(&dev));
  std::vector<int> order(nsys);
  for (int i = 0; i < nsys; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return N[a] > N[b]; });
  const int nw = std::min(nstreams, nsys);
  std::vector<std::vector<int>> work(nw);
  std::vector<double> load(nw, 0.0);
  for (int i : order) {
    const int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    work[w].push_back(i);
    load[w] += (double)N[i] * N[i] * N[i];
  }
  std::vector<int> status(nw, OK);
  std::vector<std::string> msg(nw);
  auto worker = [&](int w) {
    int st = OK;
    hipStream_t s = nullptr;
    double* A = nullptr;
    LUHandle* h = nullptr;
    const int cap = N[work[w][0]];  // largest first
    auto run = [&]() -> int {
      MPDOK_HIP(hipSetDevice(dev));
      MPDOK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      MPDOK_HIP(hipMalloc(&A, (size_t)cap * cap * sizeof(double)));
      MPDOK_TRY(lu_create(cap, precision, PIV_NONE, nb, slices, &h));
      for (int i : work[w]) {
        const int n = N[i];
        MPDOK_TRY(rbf_assemble(s, coords[i], n, D, kind, param[i], reg, 0, 0, n, 0, n, A, F64, n,
                               COL_MAJOR));
        MPDOK_TRY(lu_resize(h, n));
        int info = 0;
        MPDOK_TRY(lu_factor(h, s, A, n, COL_MAJOR, &info));
        if (info != 0) {
          for (int r = 0; r < nrhs; ++r) {
            converged[i * nrhs + r] = 0;
            relres[i * nrhs + r] = NAN;
          }
          continue;
        }
        // all right-hand sides of the system in one block: the matrix and
        // the factors are read once per outer step for all of them
        MPDOK_TRY(lu_ir_multi_impl(s, A, n, COL_MAJOR, h, n, B[i], n, X[i], n, nrhs, tol,
                                   maxiter_outer, 1, converged + (size_t)i * nrhs,
                                   relres + (size_t)i * nrhs));
      }
      MPDOK_HIP(hipStreamSynchronize(s));
      return OK;
    };
    st = run();
    if (st != OK) msg[w] = error_string();
    if (s) (void)hipStreamSynchronize(s);
    lu_destroy(h);
    if (A) (void)hipFree(A);
    if (s) (void)hipStreamDestroy(s);
    status[w] = st;
  };
  std::vector<std::thread> threads;
  for (int w = 0; w < nw; ++w) threads.emplace_back(worker, w);
  for (auto& t : threads) t.join();
  for (int w = 0; w < nw; ++w) {
    if (status[w] != OK) {
      set_error("batched (worker %d): %s", w, msg[w].c_str());
      return status[w];
    }
  }
  return OK;
}
