// Batched uneven kernel systems in one captured HIP graph.
//
// A plan holds many independent systems K_i x_i = b_i (Gaussian or Matern
// kernel of N_i points plus a nugget, nrhs right-hand sides each) and solves
// them all with grouped launches: one launch per kind of work and step covers
// every system that still has work at that step. The whole solve (assembly,
// factorisation, refinement, copy-out) is enqueued once and captured into a
// hipGraph, so a replay is one graph launch with no host round trip. This
// path has no reference counterpart; it follows the north star's batched
// configuration (256 systems, N log-uniform in [256, 8192], 4 right-hand
// sides, each to relres 1e-12).
//
// Layout. System i is padded to P_i = 64 ceil(N_i / 64) with an identity
// block: K_pad = diag(K_i, I). Its LU is diag(L_i U_i, I), the padded rows of
// every vector stay exactly zero, and every step works on whole 64-blocks.
// One device arena holds, per system: A (P x P, column-major, the assembled
// matrix kept for the residual), F (its factors), the inverses of the 64 x 64
// diagonal blocks of L and U (P x 64 each), and P x nrhs blocks B, X, R, Y, E
// (right-hand sides, solution, residual, forward-solve result, correction).
// Systems are sorted by P descending, so the systems active at step k are a
// prefix of the table and each grouped launch gets a prefix count.
//
// Per system and right-hand side the refinement is the LU-IR loop of
// lu_ir_multi_impl (solvers.hip): x = 0, first residual = b, stop when
// relres < tol, then x += inv(LU) r with the Dot2 residual r = b - A x
// (Ogita-Rump-Oishi compensated sums) and no further update once converged.
// Convergence lives in device flags; converged columns and finished systems
// make the later launches return at once, so the graph's fixed sequence of
// maxiter_outer steps costs nothing after the batch has converged.
//
// Graph rules kept here: no hipMemsetAsync inside the captured region (every
// buffer is initialised by a kernel), no host reads inside it.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "common.h"
#include "lu.h"
#include "rbf.h"

namespace mpdok {

constexpr int BMAX_RHS = 8;

struct BSys {
  double* A;
  double* F;
  double* Linv;
  double* Uinv;
  double* B;
  double* X;
  double* R;
  double* Y;
  double* E;
  const double* coords;  // caller's N x D points (row-major)
  const double* Bu;      // caller's right-hand sides, N x nrhs, ld N
  double* Xu;            // caller's solution, N x nrhs, ld N
  double param;          // kernel parameter (gamma, or the Matern length scale)
  int N, P;
};

// Per (system, right-hand side) state: index s * nrhs + c (sorted order).
struct BState {
  double* bnorm;   // ||b||
  double* relres;  // last relres checked
  double* hist;    // relres at refinement step it: hist[it * S * nrhs + s * nrhs + c]
  int* conv;       // 1 once relres < tol (or b = 0)
  int* done;       // per system: every column converged
  int* info;       // per system: zero pivot (column + 1), 0 when none
  int* steps;      // per (system, column): updates applied
};

// Workgroup b of a grouped grid -> (system si, its tile t); see grouped_tile in
// lu_kernels.hip (same XCD ranges, same search).
__device__ __forceinline__ bool b_tile(int b, const int* __restrict__ pref, int ns, int& si,
                                       int& t) {
  const int T = pref[ns];
  const int per = (T + NXCD - 1) / NXCD;
  const int tt = (b % NXCD) * per + b / NXCD;
  if (b / NXCD >= per || tt >= T) return false;
  int lo = 0, hi = ns;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pref[mid] <= tt) lo = mid;
    else hi = mid;
  }
  si = lo;
  t = tt - pref[lo];
  return true;
}

// ---------------------------------------------------------------------------
// Initialisation: flags and counters of every (system, column), the padded
// right-hand sides, x = 0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void b_init_state_kernel(BState st, int S, int nrhs, int nhist) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < S * nrhs) {
    st.conv[i] = 0;
    st.relres[i] = NAN;
    st.steps[i] = 0;
    for (int h = 0; h < nhist; ++h) st.hist[(int64_t)h * S * nrhs + i] = NAN;
  }
  if (i < S) {
    st.done[i] = 0;
    st.info[i] = 0;
  }
}

// Row tiles of 256 over P: B = padded copy of the caller's B, X = 0.
__global__ __launch_bounds__(256) void b_copy_in_kernel(const BSys* __restrict__ sys,
                                                        const int* __restrict__ pref, int ns,
                                                        int nrhs) {
  int si, t;
  if (!b_tile(blockIdx.x, pref, ns, si, t)) return;
  const BSys q = sys[si];
  const int i = 256 * t + threadIdx.x;
  if (i >= q.P) return;
  for (int c = 0; c < nrhs; ++c) {
    q.B[i + (int64_t)c * q.P] = i < q.N ? q.Bu[i + (int64_t)c * q.N] : 0.0;
    q.X[i + (int64_t)c * q.P] = 0.0;
  }
}

__global__ __launch_bounds__(256) void b_copy_out_kernel(const BSys* __restrict__ sys,
                                                         const int* __restrict__ pref, int ns,
                                                         int nrhs) {
  int si, t;
  if (!b_tile(blockIdx.x, pref, ns, si, t)) return;
  const BSys q = sys[si];
  const int i = 256 * t + threadIdx.x;
  if (i >= q.N) return;
  for (int c = 0; c < nrhs; ++c) q.Xu[i + (int64_t)c * q.N] = q.X[i + (int64_t)c * q.P];
}

// ---------------------------------------------------------------------------
// Assembly: 64 x 64 tiles over P x P. Inside N x N the element is the one
// rbf_assemble writes for a column-major FP64 matrix (row point first in the
// Gram identity, nugget added in FP64); outside it the identity. Written to A
// and to F (the factorisation's copy) in the same pass.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void b_assemble_kernel(const BSys* __restrict__ sys,
                                                         const int* __restrict__ pref, int ns,
                                                         int D, int kind, double reg) {
  int si, t;
  if (!b_tile(blockIdx.x, pref, ns, si, t)) return;
  const BSys q = sys[si];
  const int nt = q.P / 64;
  const int i = 64 * (t % nt) + (threadIdx.x & 63);
  const int j0 = 64 * (t / nt) + 16 * (threadIdx.x >> 6);
  double ci[MAXD];
  const bool rin = i < q.N;
#pragma unroll
  for (int d = 0; d < MAXD; ++d) ci[d] = (rin && d < D) ? q.coords[(int64_t)i * D + d] : 0.0;
  const double sqi = sqnorm(ci, D);
  for (int jj = 0; jj < 16; ++jj) {
    const int j = j0 + jj;
    double v;
    if (rin && j < q.N) {
      double cj[MAXD];
#pragma unroll
      for (int d = 0; d < MAXD; ++d) cj[d] = d < D ? q.coords[(int64_t)j * D + d] : 0.0;
      const double d2 = gram_d2(ci, sqi, cj, sqnorm(cj, D), D);
      v = kernel_value(kind, q.param, d2);
      if (i == j) v = v + reg;
    } else {
      v = i == j ? 1.0 : 0.0;
    }
    const int64_t o = i + (int64_t)j * q.P;
    q.A[o] = v;
    q.F[o] = v;
  }
}

// ---------------------------------------------------------------------------
// Column norms, one workgroup per system: ||V[:, c]|| for every column c
// (V = B when `which` is 0, else R), in a fixed order (strided partial sums,
// then an LDS tree). which = 0 fills bnorm (and marks b = 0 columns
// converged); which = 1 records relres, the history entry `it`, convergence
// and the system's done flag.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void b_norm_kernel(const BSys* __restrict__ sys, BState st,
                                                     int S, int nrhs, int which, int it,
                                                     double tol) {
  const int si = blockIdx.x;
  if (which == 1 && st.done[si]) return;
  const BSys q = sys[si];
  const double* V = which == 0 ? q.B : q.R;
  __shared__ double red[256];
  int all = 1;
  for (int c = 0; c < nrhs; ++c) {
    const int id = si * nrhs + c;
    if (which == 1 && st.conv[id]) continue;
    double acc = 0.0;
    for (int i = threadIdx.x; i < q.P; i += 256) {
      const double v = V[i + (int64_t)c * q.P];
      acc = __builtin_fma(v, v, acc);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    const double nrm = sqrt(red[0]);
    __syncthreads();
    if (threadIdx.x == 0) {
      if (which == 0) {
        st.bnorm[id] = nrm;
        if (nrm == 0.0) {
          st.conv[id] = 1;
          st.relres[id] = 0.0;
        }
      } else {
        const double rel = nrm / st.bnorm[id];
        st.relres[id] = rel;
        st.hist[(int64_t)it * S * nrhs + id] = rel;
        if (rel < tol) st.conv[id] = 1;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int c = 0; c < nrhs; ++c) all &= st.conv[si * nrhs + c];
    // a system whose factorisation met a zero pivot stops here
    if (st.info[si] != 0) all = 1;
    st.done[si] = all;
  }
}

// ---------------------------------------------------------------------------
// Residual R = B - A X with the Dot2 sums of resid_cm_partial_kernel: 64-row
// tiles, the four wavefronts take one quarter of the columns each, and the
// quarters' double-double sums are merged in order before r = b - s.
// it0: R = B (x = 0, no pass over A).
// ---------------------------------------------------------------------------
template <int NR>
__global__ __launch_bounds__(256) void b_resid_kernel(const BSys* __restrict__ sys,
                                                      const int* __restrict__ pref, int ns,
                                                      BState st, int nrhs, int it0) {
  int si, t;
  if (!b_tile(blockIdx.x, pref, ns, si, t)) return;
  if (st.done[si]) return;
  const BSys q = sys[si];
  const int lane = threadIdx.x & 63, qw = threadIdx.x >> 6;
  const int i = 64 * t + lane;
  const int64_t P = q.P;
  if (it0) {
    if (qw == 0)
      for (int c = 0; c < nrhs; ++c) q.R[i + c * P] = q.B[i + c * P];
    return;
  }
  __shared__ double sh[4][NR][64], sl[4][NR][64];
  dd acc[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) acc[c] = dd{0.0, 0.0};
  const int quarter = q.P / 4;  // a multiple of 16
  const int j0 = qw * quarter, j1 = j0 + quarter;
  for (int j = j0; j < j1; j += 4) {
    double a[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] = nt_load(q.A + i + (int64_t)(j + u) * P);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < NR; ++c)
        if (c < nrhs) dd_add_prod(acc[c], a[u], q.X[j + u + c * P]);
  }
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    sh[qw][c][lane] = acc[c].hi;
    sl[qw][c][lane] = acc[c].lo;
  }
  __syncthreads();
  if (qw != 0) return;
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    if (c >= nrhs || st.conv[si * nrhs + c]) continue;
    dd s{sh[0][c][lane], sl[0][c][lane]};
    for (int w = 1; w < 4; ++w) s = dd_merge(s, dd{sh[w][c][lane], sl[w][c][lane]});
    double h, e;
    two_sum(q.B[i + c * P], -s.hi, h, e);
    q.R[i + c * P] = h + (e - s.lo);
  }
}

// ---------------------------------------------------------------------------
// Triangular solves with the 64 x 64 diagonal-block inverses. One launch per
// block step, every system with that block in the grid.
//
// Forward (unit lower), step j: y_j = inv(L_jj) r_j, written to Y by the
// system's tile 0; then r_i -= L_ij y_j for the rows below, 64 per tile (the
// four wavefronts take 16 of the block's 64 columns each; partial sums added
// in order). Every tile forms y_j itself (64 x 64 from L2), so the step is
// one launch. Backward (upper), step t: block j = P / 64 - 1 - t of each
// system, x_j = inv(U_jj) y_j into E, then y_i -= U_ij x_j for the rows above.
// ---------------------------------------------------------------------------
template <int NR, bool LOWER>
__global__ __launch_bounds__(256) void b_trsv_kernel(const BSys* __restrict__ sys,
                                                     const int* __restrict__ pref, int ns,
                                                     BState st, int nrhs, int step) {
  int si, t;
  if (!b_tile(blockIdx.x, pref, ns, si, t)) return;
  if (st.done[si]) return;
  const BSys q = sys[si];
  const int64_t P = q.P;
  const int nblk = q.P / 64;
  const int jb = LOWER ? step : nblk - 1 - step;
  const int64_t c0 = 64 * (int64_t)jb;
  const double* Dinv = (LOWER ? q.Linv : q.Uinv) + c0;  // ld P
  double* V = LOWER ? q.R : q.Y;                        // solved in place (other rows)
  double* O = LOWER ? q.Y : q.E;                        // block j's result
  __shared__ double vj[NR][64];
  __shared__ double yj[NR][64];
  __shared__ double part[4][NR][64];
  const int lane = threadIdx.x & 63, qw = threadIdx.x >> 6;
  for (int e = threadIdx.x; e < 64 * NR; e += 256) {
    const int r = e & 63, c = e >> 6;
    vj[c][r] = c < nrhs ? V[c0 + r + c * P] : 0.0;
  }
  __syncthreads();
  // y_j = Dinv v_j: row `lane`, columns qw, qw + 4 (k ascending)
  for (int c = qw; c < NR; c += 4) {
    double acc = 0.0;
#pragma unroll 16
    for (int k = 0; k < 64; ++k) acc = __builtin_fma(Dinv[lane + k * P], vj[c][k], acc);
    yj[c][lane] = acc;
  }
  __syncthreads();
  if (t == 0 && qw == 0)
    for (int c = 0; c < nrhs; ++c) O[c0 + lane + c * P] = yj[c][lane];
  // rows of tile t: below the block (forward) or above it (backward)
  const int64_t i0 = LOWER ? c0 + 64 + 64 * (int64_t)t : 64 * (int64_t)t;
  if (LOWER ? i0 >= P : i0 >= c0) return;
  const int64_t i = i0 + lane;
  const double* Mc = q.F + i + c0 * P;  // L (rows below) or U (rows above), block column j
  double acc[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) acc[c] = 0.0;
#pragma unroll
  for (int kk = 0; kk < 16; ++kk) {
    const int k = 16 * qw + kk;
    const double m = Mc[(int64_t)k * P];
#pragma unroll
    for (int c = 0; c < NR; ++c) acc[c] = __builtin_fma(m, yj[c][k], acc[c]);
  }
#pragma unroll
  for (int c = 0; c < NR; ++c) part[qw][c][lane] = acc[c];
  __syncthreads();
  if (qw != 0) return;
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    if (c >= nrhs) continue;
    const double s = ((part[0][c][lane] + part[1][c][lane]) + part[2][c][lane]) + part[3][c][lane];
    V[i + c * P] -= s;
  }
}

// x += e for the columns not yet converged; counts the updates.
__global__ __launch_bounds__(256) void b_update_kernel(const BSys* __restrict__ sys,
                                                       const int* __restrict__ pref, int ns,
                                                       BState st, int nrhs) {
  int si, t;
  if (!b_tile(blockIdx.x, pref, ns, si, t)) return;
  if (st.done[si]) return;
  const BSys q = sys[si];
  const int i = 256 * t + threadIdx.x;
  for (int c = 0; c < nrhs; ++c) {
    if (st.conv[si * nrhs + c]) continue;
    if (i < q.P) q.X[i + (int64_t)c * q.P] += q.E[i + (int64_t)c * q.P];
    if (i == 0) st.steps[si * nrhs + c] += 1;
  }
}

// ---------------------------------------------------------------------------
// The plan
// ---------------------------------------------------------------------------
struct BPlan {
  int S = 0, D = 0, kind = 0, nrhs = 0, maxit = 0, K = 0, nhist = 0;
  int upd_thr = 0;  // trailing updates of at least this many rows use 128 x 128 tiles
  double reg = 0.0, tol = 0.0;
  std::vector<int> order;  // sorted position -> caller's index
  std::vector<int> P, N;   // sorted
  void* arena = nullptr;   // device
  BSys* sys = nullptr;     // device table (sorted)
  BLUSys* lus = nullptr;   // device table (sorted)
  int* tables = nullptr;   // device prefix tables
  BState st{};
  // host copies of the launches: prefix-table offsets (into `tables`) and counts
  struct Grid {
    int off = -1, ns = 0, ntiles = 0;
  };
  Grid asm_g, rows256, rows64;
  std::vector<int> lu_cnt;                  // systems active at LU step k
  std::vector<Grid> lu_panel, lu_upd128, lu_upd64;  // lu_upd64: systems [s0, s0 + ns)
  std::vector<int> lu_upd64_s0;
  std::vector<Grid> fwd, bwd;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~BPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
    if (arena) (void)hipFree(arena);
  }
};

static inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Launch every step of the solve on stream s (eagerly, or into a capture).
static int bplan_enqueue(BPlan& p, hipStream_t s) {
  const int S = p.S, nr = p.nrhs;
  auto grid = [](int ntiles) { return (ntiles + NXCD - 1) / NXCD * NXCD; };
  auto tab = [&](const BPlan::Grid& g) { return p.tables + g.off; };
  b_init_state_kernel<<<(std::max(S * nr, S) + 255) / 256, 256, 0, s>>>(p.st, S, nr, p.nhist);
  b_copy_in_kernel<<<grid(p.rows256.ntiles), 256, 0, s>>>(p.sys, tab(p.rows256), S, nr);
  b_assemble_kernel<<<grid(p.asm_g.ntiles), 256, 0, s>>>(p.sys, tab(p.asm_g), S, p.D, p.kind,
                                                          p.reg);
  b_norm_kernel<<<S, 256, 0, s>>>(p.sys, p.st, S, nr, 0, 0, p.tol);
  MPDOK_HIP(hipGetLastError());
  for (int k = 0; k < p.K; ++k) {
    const int ns = p.lu_cnt[k];
    MPDOK_TRY(blu_diag(s, p.lus, ns, k));
    const BPlan::Grid& pg = p.lu_panel[k];
    MPDOK_TRY(blu_panel(s, p.lus, tab(pg), pg.ns, pg.ntiles, k));
    const BPlan::Grid& u1 = p.lu_upd128[k];
    if (u1.ntiles > 0) MPDOK_TRY(blu_update(s, p.lus, tab(u1), u1.ns, u1.ntiles, k, 128));
    const BPlan::Grid& u2 = p.lu_upd64[k];
    if (u2.ntiles > 0)
      MPDOK_TRY(blu_update(s, p.lus + p.lu_upd64_s0[k], tab(u2), u2.ns, u2.ntiles, k, 64));
  }
#define B_DISPATCH(KERNEL, ...)                                              \
  do {                                                                       \
    if (nr == 1) KERNEL<1><<<__VA_ARGS__>>>(ARGS);                           \
    else if (nr == 2) KERNEL<2><<<__VA_ARGS__>>>(ARGS);                      \
    else if (nr <= 4) KERNEL<4><<<__VA_ARGS__>>>(ARGS);                      \
    else KERNEL<8><<<__VA_ARGS__>>>(ARGS);                                   \
  } while (0)
  for (int it = 0; it <= p.maxit; ++it) {
#define ARGS p.sys, tab(p.rows64), S, p.st, nr, it == 0 ? 1 : 0
    B_DISPATCH(b_resid_kernel, grid(p.rows64.ntiles), 256, 0, s);
#undef ARGS
    b_norm_kernel<<<S, 256, 0, s>>>(p.sys, p.st, S, nr, 1, it, p.tol);
    MPDOK_HIP(hipGetLastError());
    if (it == p.maxit) break;  // the last residual decides
    for (int j = 0; j < p.K; ++j) {
      const BPlan::Grid& g = p.fwd[j];
#define ARGS p.sys, tab(g), g.ns, p.st, nr, j
      if (nr == 1) b_trsv_kernel<1, true><<<grid(g.ntiles), 256, 0, s>>>(ARGS);
      else if (nr == 2) b_trsv_kernel<2, true><<<grid(g.ntiles), 256, 0, s>>>(ARGS);
      else if (nr <= 4) b_trsv_kernel<4, true><<<grid(g.ntiles), 256, 0, s>>>(ARGS);
      else b_trsv_kernel<8, true><<<grid(g.ntiles), 256, 0, s>>>(ARGS);
#undef ARGS
    }
    for (int j = 0; j < p.K; ++j) {
      const BPlan::Grid& g = p.bwd[j];
#define ARGS p.sys, tab(g), g.ns, p.st, nr, j
      if (nr == 1) b_trsv_kernel<1, false><<<grid(g.ntiles), 256, 0, s>>>(ARGS);
      else if (nr == 2) b_trsv_kernel<2, false><<<grid(g.ntiles), 256, 0, s>>>(ARGS);
      else if (nr <= 4) b_trsv_kernel<4, false><<<grid(g.ntiles), 256, 0, s>>>(ARGS);
      else b_trsv_kernel<8, false><<<grid(g.ntiles), 256, 0, s>>>(ARGS);
#undef ARGS
    }
    b_update_kernel<<<grid(p.rows256.ntiles), 256, 0, s>>>(p.sys, tab(p.rows256), S, p.st, nr);
    MPDOK_HIP(hipGetLastError());
  }
#undef B_DISPATCH
  b_copy_out_kernel<<<grid(p.rows256.ntiles), 256, 0, s>>>(p.sys, tab(p.rows256), S, nr);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

static int bplan_build(BPlan& p, int nsys, const int* N, const double* const* coords,
                       const double* param, const double* const* B, double* const* X) {
  const int S = nsys, nr = p.nrhs;
  p.S = S;
  p.order.resize(S);
  std::iota(p.order.begin(), p.order.end(), 0);
  std::stable_sort(p.order.begin(), p.order.end(), [&](int a, int b) { return N[a] > N[b]; });
  p.N.resize(S);
  p.P.resize(S);
  for (int s = 0; s < S; ++s) {
    p.N[s] = N[p.order[s]];
    p.P[s] = (p.N[s] + 63) / 64 * 64;
  }
  p.K = p.P[0] / 64;
  p.nhist = p.maxit + 1;

  // prefix tables (host), then one upload
  std::vector<int> tabs;
  auto add = [&](const std::vector<int>& counts) {
    BPlan::Grid g;
    g.off = (int)tabs.size();
    g.ns = (int)counts.size();
    int acc = 0;
    for (int c : counts) {
      tabs.push_back(acc);
      acc += c;
    }
    tabs.push_back(acc);
    g.ntiles = acc;
    return g;
  };
  std::vector<int> cnt;
  for (int s = 0; s < S; ++s) cnt.push_back((p.P[s] / 64) * (p.P[s] / 64));
  p.asm_g = add(cnt);
  cnt.clear();
  for (int s = 0; s < S; ++s) cnt.push_back((p.P[s] + 255) / 256);
  p.rows256 = add(cnt);
  cnt.clear();
  for (int s = 0; s < S; ++s) cnt.push_back(p.P[s] / 64);
  p.rows64 = add(cnt);
  p.lu_cnt.assign(p.K, 0);
  p.lu_panel.resize(p.K);
  p.lu_upd128.resize(p.K);
  p.lu_upd64.resize(p.K);
  p.lu_upd64_s0.assign(p.K, 0);
  p.fwd.resize(p.K);
  p.bwd.resize(p.K);
  for (int k = 0; k < p.K; ++k) {
    int ns = 0;
    while (ns < S && p.P[ns] > 64 * k) ++ns;
    p.lu_cnt[k] = ns;
    std::vector<int> pc, u1, u2;
    int s64 = ns;  // first system whose update uses 64 x 64 tiles
    for (int s = 0; s < ns; ++s) {
      const int m = p.P[s] - 64 * (k + 1);
      pc.push_back(2 * ((m + 63) / 64));
      if (m >= p.upd_thr && m > 0) {
        u1.push_back(((m + 127) / 128) * ((m + 127) / 128));
      } else if (s64 == ns) {
        s64 = s;
      }
    }
    for (int s = s64; s < ns; ++s) {
      const int m = p.P[s] - 64 * (k + 1);
      u2.push_back(((m + 63) / 64) * ((m + 63) / 64));
    }
    // systems are sorted by P, so the 128-tile ones are the prefix [0, s64)
    u1.resize(s64);
    p.lu_panel[k] = add(pc);
    p.lu_upd128[k] = add(u1);
    p.lu_upd64[k] = add(u2);
    p.lu_upd64_s0[k] = s64;
    std::vector<int> f, b;
    for (int s = 0; s < ns; ++s) {
      const int nblk = p.P[s] / 64;
      f.push_back(std::max(1, nblk - 1 - k));  // 64-row tiles below block k
      b.push_back(std::max(1, nblk - 1 - k));  // 64-row tiles above block nblk - 1 - k
    }
    p.fwd[k] = add(f);
    p.bwd[k] = add(b);
  }

  // device arena
  size_t off = 0;
  std::vector<size_t> oA(S), oF(S), oL(S), oU(S), oV(S);
  for (int s = 0; s < S; ++s) {
    const size_t P = p.P[s];
    oA[s] = off;
    off += align256(P * P * 8);
    oF[s] = off;
    off += align256(P * P * 8);
    oL[s] = off;
    off += align256(P * 64 * 8);
    oU[s] = off;
    off += align256(P * 64 * 8);
    oV[s] = off;
    off += align256(5 * P * nr * 8);  // B, X, R, Y, E
  }
  const size_t o_sys = off;
  off += align256(sizeof(BSys) * S);
  const size_t o_lus = off;
  off += align256(sizeof(BLUSys) * S);
  const size_t o_tab = off;
  off += align256(sizeof(int) * tabs.size());
  const size_t SN = (size_t)S * nr;
  const size_t o_bn = off;
  off += align256(8 * SN);
  const size_t o_rel = off;
  off += align256(8 * SN);
  const size_t o_hist = off;
  off += align256(8 * SN * p.nhist);
  const size_t o_conv = off;
  off += align256(4 * SN);
  const size_t o_steps = off;
  off += align256(4 * SN);
  const size_t o_done = off;
  off += align256(4 * (size_t)S);
  const size_t o_info = off;
  off += align256(4 * (size_t)S);
  size_t free_b = 0, total_b = 0;
  MPDOK_HIP(hipMemGetInfo(&free_b, &total_b));
  MPDOK_REQUIRE(off < free_b, "batched plan: %zu bytes of device memory needed, %zu free", off,
                free_b);
  MPDOK_HIP(hipMalloc(&p.arena, off));
  char* base = static_cast<char*>(p.arena);
  p.sys = reinterpret_cast<BSys*>(base + o_sys);
  p.lus = reinterpret_cast<BLUSys*>(base + o_lus);
  p.tables = reinterpret_cast<int*>(base + o_tab);
  p.st.bnorm = reinterpret_cast<double*>(base + o_bn);
  p.st.relres = reinterpret_cast<double*>(base + o_rel);
  p.st.hist = reinterpret_cast<double*>(base + o_hist);
  p.st.conv = reinterpret_cast<int*>(base + o_conv);
  p.st.steps = reinterpret_cast<int*>(base + o_steps);
  p.st.done = reinterpret_cast<int*>(base + o_done);
  p.st.info = reinterpret_cast<int*>(base + o_info);
  std::vector<BSys> hs(S);
  std::vector<BLUSys> hl(S);
  for (int s = 0; s < S; ++s) {
    const int c = p.order[s];
    const size_t P = p.P[s];
    BSys& q = hs[s];
    q.A = reinterpret_cast<double*>(base + oA[s]);
    q.F = reinterpret_cast<double*>(base + oF[s]);
    q.Linv = reinterpret_cast<double*>(base + oL[s]);
    q.Uinv = reinterpret_cast<double*>(base + oU[s]);
    double* v = reinterpret_cast<double*>(base + oV[s]);
    q.B = v;
    q.X = v + P * nr;
    q.R = v + 2 * P * nr;
    q.Y = v + 3 * P * nr;
    q.E = v + 4 * P * nr;
    q.coords = coords[c];
    q.Bu = B[c];
    q.Xu = X[c];
    q.param = param[c];
    q.N = p.N[s];
    q.P = p.P[s];
    BLUSys& l = hl[s];
    l.F = q.F;
    l.Linv = q.Linv;
    l.Uinv = q.Uinv;
    l.info = p.st.info + s;
    l.P = q.P;
    l.pad_ = 0;
  }
  MPDOK_HIP(hipMemcpy(p.sys, hs.data(), sizeof(BSys) * S, hipMemcpyHostToDevice));
  MPDOK_HIP(hipMemcpy(p.lus, hl.data(), sizeof(BLUSys) * S, hipMemcpyHostToDevice));
  MPDOK_HIP(hipMemcpy(p.tables, tabs.data(), sizeof(int) * tabs.size(), hipMemcpyHostToDevice));
  return OK;
}

}  // namespace mpdok

using namespace mpdok;

// Create a plan for nsys systems (see the file comment). coords[i]: N[i] x D
// points (row-major FP64, device); param[i]: gamma (kind 0) or length scale
// (kind 1); B[i], X[i]: N[i] x nrhs column-major device blocks (ld N[i]). The
// plan keeps these pointers: a run reads coords and B and writes X, so the
// caller keeps them alive and may change their contents between runs.
// update_rows: trailing updates with at least this many rows use 128 x 128
// tiles, smaller ones 64 x 64 (the tuner's choice). use_graph: capture the
// solve into a hipGraph (1) or launch it eagerly on each run (0).
MPDOK_API int mpdok_bplan_create(int nsys, const int* N, const double* const* coords, int D,
                                 int kind, const double* param, double reg, int nrhs,
                                 const double* const* B, double* const* X, double tol,
                                 int maxiter_outer, int update_rows, int use_graph,
                                 void** plan_out) {
  MPDOK_REQUIRE(plan_out, "bplan_create: null plan_out");
  *plan_out = nullptr;
  MPDOK_REQUIRE(nsys >= 1, "bplan_create: nsys must be >= 1, got %d", nsys);
  MPDOK_REQUIRE(nrhs >= 1 && nrhs <= BMAX_RHS, "bplan_create: nrhs must be in [1, %d], got %d",
                BMAX_RHS, nrhs);
  MPDOK_REQUIRE(D >= 1 && D <= MAXD, "bplan_create: D must be in [1, %d], got %d", MAXD, D);
  MPDOK_REQUIRE(kind == 0 || kind == 1, "bplan_create: unknown kernel kind %d", kind);
  MPDOK_REQUIRE(maxiter_outer >= 1 && maxiter_outer <= 64, "bplan_create: maxiter_outer %d",
                maxiter_outer);
  MPDOK_REQUIRE(N && coords && param && B && X, "bplan_create: null pointer");
  for (int i = 0; i < nsys; ++i) {
    MPDOK_REQUIRE(N[i] >= 1 && N[i] <= 65536, "bplan_create: N[%d] = %d outside [1, 65536]", i,
                  N[i]);
    MPDOK_REQUIRE(coords[i] && B[i] && X[i], "bplan_create: null pointer for system %d", i);
    MPDOK_REQUIRE(kind == 0 || param[i] > 0.0, "bplan_create: Matern length scale of system %d", i);
  }
  BPlan* p = new (std::nothrow) BPlan();
  MPDOK_REQUIRE(p, "bplan_create: out of host memory");
  p->D = D;
  p->kind = kind;
  p->nrhs = nrhs;
  p->maxit = maxiter_outer;
  p->reg = reg;
  p->tol = tol;
  p->upd_thr = std::max(0, update_rows);
  int st = bplan_build(*p, nsys, N, coords, param, B, X);
  if (st == OK && use_graph) {
    auto capture = [&]() -> int {
      MPDOK_HIP(hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking));
      MPDOK_HIP(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal));
      const int e = bplan_enqueue(*p, p->cap);
      hipGraph_t g = nullptr;
      const hipError_t ce = hipStreamEndCapture(p->cap, &g);
      if (e != OK) {
        if (g) (void)hipGraphDestroy(g);
        return e;
      }
      MPDOK_HIP(ce);
      p->graph = g;
      MPDOK_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
      return OK;
    };
    st = capture();
  }
  if (st != OK) {
    delete p;
    return st;
  }
  *plan_out = p;
  return OK;
}

// Run the plan on stream s (one graph launch, or the eager sequence); returns
// without waiting.
MPDOK_API int mpdok_bplan_run(void* plan, hipStream_t s) {
  MPDOK_REQUIRE(plan, "bplan_run: null plan");
  BPlan* p = static_cast<BPlan*>(plan);
  if (p->exec) {
    MPDOK_HIP(hipGraphLaunch(p->exec, s));
    return OK;
  }
  return bplan_enqueue(*p, s);
}

// Results of the last run, in the caller's system order (waits for the
// device): converged / relres / steps per (system, right-hand side),
// nsys x nrhs row-major; info per system (zero pivot column + 1, else 0);
// hist (optional) the relres checked at each refinement step,
// (maxiter_outer + 1) x nsys x nrhs (NaN where a column had stopped).
MPDOK_API int mpdok_bplan_results(void* plan, int* converged, double* relres, int* steps,
                                  int* info, double* hist) {
  MPDOK_REQUIRE(plan, "bplan_results: null plan");
  BPlan* p = static_cast<BPlan*>(plan);
  MPDOK_HIP(hipDeviceSynchronize());
  const int S = p->S, nr = p->nrhs;
  std::vector<int> conv(S * nr), stp(S * nr), inf(S);
  std::vector<double> rel(S * nr), h((size_t)S * nr * p->nhist);
  MPDOK_HIP(hipMemcpy(conv.data(), p->st.conv, 4 * conv.size(), hipMemcpyDeviceToHost));
  MPDOK_HIP(hipMemcpy(stp.data(), p->st.steps, 4 * stp.size(), hipMemcpyDeviceToHost));
  MPDOK_HIP(hipMemcpy(inf.data(), p->st.info, 4 * inf.size(), hipMemcpyDeviceToHost));
  MPDOK_HIP(hipMemcpy(rel.data(), p->st.relres, 8 * rel.size(), hipMemcpyDeviceToHost));
  MPDOK_HIP(hipMemcpy(h.data(), p->st.hist, 8 * h.size(), hipMemcpyDeviceToHost));
  for (int s = 0; s < S; ++s) {
    const int c = p->order[s];
    if (info) info[c] = inf[s];
    for (int r = 0; r < nr; ++r) {
      if (converged) converged[c * nr + r] = conv[s * nr + r];
      if (relres) relres[c * nr + r] = rel[s * nr + r];
      if (steps) steps[c * nr + r] = stp[s * nr + r];
      if (hist)
        for (int it = 0; it < p->nhist; ++it)
          hist[((size_t)it * S + c) * nr + r] = h[((size_t)it * S + s) * nr + r];
    }
  }
  return OK;
}

// Device bytes the plan holds, and the number of kernel launches one run
// makes (the nodes of its graph).
MPDOK_API int mpdok_bplan_info(void* plan, size_t* bytes, int* launches) {
  MPDOK_REQUIRE(plan, "bplan_info: null plan");
  BPlan* p = static_cast<BPlan*>(plan);
  if (bytes) {
    size_t b = 0;
    (void)hipMemPtrGetInfo(p->arena, &b);
    *bytes = b;
  }
  if (launches) {
    int n = 4;
    for (int k = 0; k < p->K; ++k)
      n += 1 + (p->lu_panel[k].ntiles > 0) + (p->lu_upd128[k].ntiles > 0) +
           (p->lu_upd64[k].ntiles > 0);
    n += (p->maxit + 1) * 2 + p->maxit * (2 * p->K + 1) + 1;
    *launches = n;
  }
  return OK;
}

MPDOK_API void mpdok_bplan_destroy(void* plan) {
  if (!plan) return;
  BPlan* p = static_cast<BPlan*>(plan);
  (void)hipDeviceSynchronize();
  delete p;
}
