// Dense matrix-vector products and residuals for gfx950.
//
//   y = A x          FP32 or FP64, A row-major (C order) or column-major (F order)
//   r = b - A x      FP64, plain or Dot2 (double-double) accumulation
//
// Row-major: one wavefront owns R consecutive rows; lanes stride across the
// row with 16-byte non-temporal loads (A is streamed once, x stays in L2/MALL
// and is shared by the R rows); each row is finished with a butterfly
// reduction. Column-major: each lane owns 16 bytes of consecutive rows, the
// columns are split into chunks across blockIdx.y, partial sums go to a
// workspace and a second kernel sums the chunks in a fixed order (results are
// bit-reproducible from run to run).
//
// Reference call sites replaced: dense_krylov.py:61-63 (cp.matmul(A_fp32, x)),
// dense_krylov.py:65-79 and :391,:416 (FP64 residual), mpdok_ooc.py:158-201
// (tile SGEMV), the cublasDgemv residual inside lu_ir.so/mpdok.so
// (benchmark_mpdok.ipynb:991).
#include <mutex>

#include "common.h"
#include "gemv.h"

namespace mpdok {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

template <typename T> struct Vec;
template <> struct Vec<float> { typedef f4 type; static constexpr int W = 4; };
template <> struct Vec<double> { typedef d2 type; static constexpr int W = 2; };

template <typename T, int W>
__device__ __forceinline__ T vdot(typename Vec<T>::type a, typename Vec<T>::type x, T acc) {
#pragma unroll
  for (int i = 0; i < W; ++i) acc = __builtin_fma(a[i], x[i], acc);
  return acc;
}

// ---------------------------------------------------------------------------
// Row-major GEMV: y[i] = alpha * sum_j A[i*lda + j] * x[j]
// ---------------------------------------------------------------------------
template <typename T, int R, bool VEC>
__global__ __launch_bounds__(256) void gemv_rm_kernel(const T* __restrict__ A, int64_t lda, int M,
                                                      int N, const T* __restrict__ x,
                                                      T* __restrict__ y, T alpha,
                                                      const int* __restrict__ skip) {
  if (skip != nullptr && *skip) return;
  typedef typename Vec<T>::type V;
  constexpr int W = Vec<T>::W;
  const int lane = threadIdx.x & 63;
  const int row0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * R;
  if (row0 >= M) return;
  T acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = T(0);
  const T* rowp[R];
#pragma unroll
  for (int r = 0; r < R; ++r) rowp[r] = A + (int64_t)min(row0 + r, M - 1) * lda;

  int jt = 0;  // first column handled by the scalar tail
  if (VEC) {
    const int nv = N / W;
    const V* xv = reinterpret_cast<const V*>(x);
    int c = lane;
    for (; c + 64 < nv; c += 128) {
      V x0 = xv[c], x1 = xv[c + 64];
      V a0[R], a1[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        a0[r] = nt_load(reinterpret_cast<const V*>(rowp[r]) + c);
        a1[r] = nt_load(reinterpret_cast<const V*>(rowp[r]) + c + 64);
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        acc[r] = vdot<T, W>(a0[r], x0, acc[r]);
        acc[r] = vdot<T, W>(a1[r], x1, acc[r]);
      }
    }
    for (; c < nv; c += 64) {
      V x0 = xv[c];
#pragma unroll
      for (int r = 0; r < R; ++r)
        acc[r] = vdot<T, W>(nt_load(reinterpret_cast<const V*>(rowp[r]) + c), x0, acc[r]);
    }
    jt = nv * W;
  }
  for (int j = jt + lane; j < N; j += 64) {
    T xj = x[j];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = __builtin_fma(rowp[r][j], xj, acc[r]);
  }
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = wave_sum(acc[r]);
  if (lane == 0) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (row0 + r < M) y[row0 + r] = alpha * acc[r];
  }
}

// ---------------------------------------------------------------------------
// Column-major GEMV, pass 1: part[c*M + i] = sum_{j in chunk c} A[i + j*lda] x[j]
// ---------------------------------------------------------------------------
template <typename T, bool VEC>
__global__ __launch_bounds__(256) void gemv_cm_partial_kernel(const T* __restrict__ A, int64_t lda,
                                                              int M, int N, const T* __restrict__ x,
                                                              T* __restrict__ part, int cpc,
                                                              const int* __restrict__ skip) {
  if (skip != nullptr && *skip) return;
  typedef typename Vec<T>::type V;
  constexpr int W = Vec<T>::W;
  const int i0 = (blockIdx.x * 256 + threadIdx.x) * W;
  const int c = blockIdx.y;
  const int j0 = c * cpc, j1 = min(N, j0 + cpc);
  if (i0 >= M) return;
  T* out = part + (int64_t)c * M;
  if (VEC && i0 + W <= M) {
    // U columns per iteration: U 16-byte loads in flight per lane
    constexpr int U = 4;
    V acc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) acc[u] = V(0);
    int j = j0;
    for (; j + U - 1 < j1; j += U) {
      V a[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        a[u] = nt_load(reinterpret_cast<const V*>(A + (int64_t)(j + u) * lda + i0));
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const T xu = x[j + u];
#pragma unroll
        for (int w = 0; w < W; ++w) acc[u][w] = __builtin_fma(a[u][w], xu, acc[u][w]);
      }
    }
    for (int u = 0; j < j1; ++j, ++u) {
      V a0 = nt_load(reinterpret_cast<const V*>(A + (int64_t)j * lda + i0));
      T x0 = x[j];
#pragma unroll
      for (int w = 0; w < W; ++w) acc[u][w] = __builtin_fma(a0[w], x0, acc[u][w]);
    }
    // the U partial sums in a fixed pairwise order
#pragma unroll
    for (int h = U / 2; h > 0; h /= 2)
#pragma unroll
      for (int u = 0; u < h; ++u) acc[u] = acc[u] + acc[u + h];
#pragma unroll
    for (int w = 0; w < W; ++w) out[i0 + w] = acc[0][w];
  } else {
    for (int w = 0; w < W && i0 + w < M; ++w) {
      T acc = T(0);
      for (int j = j0; j < j1; ++j) acc = __builtin_fma(A[(int64_t)j * lda + i0 + w], x[j], acc);
      out[i0 + w] = acc;
    }
  }
}

// y = alpha * (sum of the column-chunk partial sums), 64 rows per workgroup:
// wavefront q sums chunks q, q + 4, q + 8, ... of its lane's row, and the four
// sums are added in the fixed order (s0 + s1) + (s2 + s3), so every row uses
// the same summation order on every call.
template <typename T>
__global__ __launch_bounds__(256) void gemv_cm_reduce_kernel(const T* __restrict__ part, int M,
                                                             int nchunks, T* __restrict__ y,
                                                             T alpha, const int* __restrict__ skip) {
  if (skip != nullptr && *skip) return;
  __shared__ T sh[4][64];
  const int r = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + r;
  T s = T(0);
  if (i < M)
    for (int c = q; c < nchunks; c += 4) s += part[(int64_t)c * M + i];
  sh[q][r] = s;
  __syncthreads();
  if (q == 0 && i < M) y[i] = alpha * ((sh[0][r] + sh[1][r]) + (sh[2][r] + sh[3][r]));
}

// ---------------------------------------------------------------------------
// FP64 residual r = b - A x. MODE 0: plain FP64 FMA accumulation.
// MODE 1: Dot2 (compensated) accumulation, result rounded once to FP64.
// ---------------------------------------------------------------------------
template <int MODE, int R>
__global__ __launch_bounds__(256) void resid_rm_kernel(const double* __restrict__ A, int64_t lda,
                                                       int N, const double* __restrict__ x,
                                                       const double* __restrict__ b,
                                                       double* __restrict__ r) {
  const int lane = threadIdx.x & 63;
  const int row0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * R;
  if (row0 >= N) return;
  const double* rowp[R];
#pragma unroll
  for (int q = 0; q < R; ++q) rowp[q] = A + (int64_t)min(row0 + q, N - 1) * lda;
  dd acc[R];
#pragma unroll
  for (int q = 0; q < R; ++q) acc[q] = dd{0.0, 0.0};
  const bool vec = ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(x)) & 15) == 0 &&
                   (lda & 1) == 0;
  int jt = 0;
  if (vec) {
    const int nv = N / 2;
    const d2* xv = reinterpret_cast<const d2*>(x);
    for (int c = lane; c < nv; c += 64) {
      d2 xx = xv[c];
      d2 a[R];
#pragma unroll
      for (int q = 0; q < R; ++q) a[q] = nt_load(reinterpret_cast<const d2*>(rowp[q]) + c);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        if (MODE == 0) {
          acc[q].hi = __builtin_fma(a[q][0], xx[0], acc[q].hi);
          acc[q].hi = __builtin_fma(a[q][1], xx[1], acc[q].hi);
        } else {
          dd_add_prod(acc[q], a[q][0], xx[0]);
          dd_add_prod(acc[q], a[q][1], xx[1]);
        }
      }
    }
    jt = nv * 2;
  }
  for (int j = jt + lane; j < N; j += 64) {
    double xj = x[j];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (MODE == 0)
        acc[q].hi = __builtin_fma(rowp[q][j], xj, acc[q].hi);
      else
        dd_add_prod(acc[q], rowp[q][j], xj);
    }
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    if (MODE == 0)
      acc[q].hi = wave_sum(acc[q].hi);
    else
      acc[q] = wave_sum_dd(acc[q]);
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (row0 + q >= N) break;
      if (MODE == 0) {
        r[row0 + q] = b[row0 + q] - acc[q].hi;
      } else {
        // b - (hi + lo) with one final rounding of the compensated sum.
        double s, t;
        two_sum(b[row0 + q], -acc[q].hi, s, t);
        r[row0 + q] = s + (t - acc[q].lo);
      }
    }
  }
}

template <int MODE>
__global__ __launch_bounds__(256) void resid_cm_partial_kernel(const double* __restrict__ A,
                                                               int64_t lda, int N,
                                                               const double* __restrict__ x,
                                                               double* __restrict__ part_hi,
                                                               double* __restrict__ part_lo,
                                                               int cpc) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;
  const int j0 = c * cpc, j1 = min(N, j0 + cpc);
  if (i >= N) return;
  dd acc{0.0, 0.0};
  int j = j0;
  for (; j + 3 < j1; j += 4) {
    double a0 = nt_load(A + (int64_t)j * lda + i);
    double a1 = nt_load(A + (int64_t)(j + 1) * lda + i);
    double a2 = nt_load(A + (int64_t)(j + 2) * lda + i);
    double a3 = nt_load(A + (int64_t)(j + 3) * lda + i);
    if (MODE == 0) {
      acc.hi = __builtin_fma(a0, x[j], acc.hi);
      acc.hi = __builtin_fma(a1, x[j + 1], acc.hi);
      acc.hi = __builtin_fma(a2, x[j + 2], acc.hi);
      acc.hi = __builtin_fma(a3, x[j + 3], acc.hi);
    } else {
      dd_add_prod(acc, a0, x[j]);
      dd_add_prod(acc, a1, x[j + 1]);
      dd_add_prod(acc, a2, x[j + 2]);
      dd_add_prod(acc, a3, x[j + 3]);
    }
  }
  for (; j < j1; ++j) {
    double a0 = A[(int64_t)j * lda + i];
    if (MODE == 0)
      acc.hi = __builtin_fma(a0, x[j], acc.hi);
    else
      dd_add_prod(acc, a0, x[j]);
  }
  part_hi[(int64_t)c * N + i] = acc.hi;
  if (MODE == 1) part_lo[(int64_t)c * N + i] = acc.lo;
}

template <int MODE>
__global__ __launch_bounds__(256) void resid_cm_reduce_kernel(const double* __restrict__ part_hi,
                                                              const double* __restrict__ part_lo,
                                                              int N, int nchunks,
                                                              const double* __restrict__ b,
                                                              double* __restrict__ r) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  if (MODE == 0) {
    double s = 0.0;
    for (int c = 0; c < nchunks; ++c) s += part_hi[(int64_t)c * N + i];
    r[i] = b[i] - s;
  } else {
    dd s{0.0, 0.0};
    for (int c = 0; c < nchunks; ++c)
      s = dd_merge(s, dd{part_hi[(int64_t)c * N + i], part_lo[(int64_t)c * N + i]});
    double h, t;
    two_sum(b[i], -s.hi, h, t);
    r[i] = h + (t - s.lo);
  }
}

// R = B - A X for NR right-hand sides (column-major A, X, B, R): each element
// of A is read once for NR products. For every right-hand side the partial
// sums run in the order of resid_cm_partial_kernel and the chunks are reduced
// as resid_cm_reduce_kernel does, so each column of R equals the single
// residual's bit for bit. Partials of right-hand side q, chunk c:
// part[(q nch + c) N + i].
template <int MODE, int NR>
__global__ __launch_bounds__(256) void resid_cm_multi_partial_kernel(
    const double* __restrict__ A, int64_t lda, int N, const double* __restrict__ X, int64_t ldx,
    double* __restrict__ part_hi, double* __restrict__ part_lo, int cpc) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y, nch = gridDim.y;
  const int j0 = c * cpc, j1 = min(N, j0 + cpc);
  if (i >= N) return;
  dd acc[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) acc[q] = dd{0.0, 0.0};
  int j = j0;
  for (; j + 3 < j1; j += 4) {
    const double a0 = nt_load(A + (int64_t)j * lda + i);
    const double a1 = nt_load(A + (int64_t)(j + 1) * lda + i);
    const double a2 = nt_load(A + (int64_t)(j + 2) * lda + i);
    const double a3 = nt_load(A + (int64_t)(j + 3) * lda + i);
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const double* x = X + (int64_t)q * ldx;
      if (MODE == 0) {
        acc[q].hi = __builtin_fma(a0, x[j], acc[q].hi);
        acc[q].hi = __builtin_fma(a1, x[j + 1], acc[q].hi);
        acc[q].hi = __builtin_fma(a2, x[j + 2], acc[q].hi);
        acc[q].hi = __builtin_fma(a3, x[j + 3], acc[q].hi);
      } else {
        dd_add_prod(acc[q], a0, x[j]);
        dd_add_prod(acc[q], a1, x[j + 1]);
        dd_add_prod(acc[q], a2, x[j + 2]);
        dd_add_prod(acc[q], a3, x[j + 3]);
      }
    }
  }
  for (; j < j1; ++j) {
    const double a0 = A[(int64_t)j * lda + i];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      if (MODE == 0)
        acc[q].hi = __builtin_fma(a0, X[(int64_t)q * ldx + j], acc[q].hi);
      else
        dd_add_prod(acc[q], a0, X[(int64_t)q * ldx + j]);
    }
  }
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    part_hi[((int64_t)q * nch + c) * N + i] = acc[q].hi;
    if (MODE == 1) part_lo[((int64_t)q * nch + c) * N + i] = acc[q].lo;
  }
}

template <int MODE>
__global__ __launch_bounds__(256) void resid_cm_multi_reduce_kernel(
    const double* __restrict__ part_hi, const double* __restrict__ part_lo, int N, int nchunks,
    const double* __restrict__ B, int64_t ldb, double* __restrict__ R, int64_t ldr) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int q = blockIdx.y;
  if (i >= N) return;
  const double* ph = part_hi + (int64_t)q * nchunks * N;
  const double* pl = part_lo + (int64_t)q * nchunks * N;
  const double b = B[(int64_t)q * ldb + i];
  if (MODE == 0) {
    double s = 0.0;
    for (int c = 0; c < nchunks; ++c) s += ph[(int64_t)c * N + i];
    R[(int64_t)q * ldr + i] = b - s;
  } else {
    dd s{0.0, 0.0};
    for (int c = 0; c < nchunks; ++c) s = dd_merge(s, dd{ph[(int64_t)c * N + i], pl[(int64_t)c * N + i]});
    double h, t;
    two_sum(b, -s.hi, h, t);
    R[(int64_t)q * ldr + i] = h + (t - s.lo);
  }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static int cm_chunks(int M, int rows_per_block, int N) {
  const int nrb = (M + rows_per_block - 1) / rows_per_block;
  if (nrb < 1) return 1;
  int nch = (2048 + nrb - 1) / nrb;
  int maxch = (N + 63) / 64;
  if (nch > maxch) nch = maxch;
  if (nch < 1) nch = 1;
  return nch;
}

size_t gemv_workspace_bytes(int M, int N, int dtype, int order) {
  if (order == ROW_MAJOR) return 0;
  const size_t es = dtype == F64 ? 8 : 4;
  const int W = dtype == F64 ? 2 : 4;
  const int nch = cm_chunks(M, 256 * W, N);
  return (size_t)nch * M * es;
}

size_t resid_workspace_bytes(int N, int order) {
  if (order == ROW_MAJOR) return 0;
  const int nch = cm_chunks(N, 256, N);
  return (size_t)nch * N * 8 * 2;
}

template <typename T>
static int gemv_launch(hipStream_t s, const T* A, int64_t lda, int order, int M, int N, const T* x,
                       T* y, T alpha, void* work, size_t work_bytes, const int* skip) {
  MPDOK_REQUIRE(A && x && y && M >= 0 && N >= 0, "gemv: bad arguments");
  if (M == 0) return OK;
  constexpr int W = Vec<T>::W;
  if (order == ROW_MAJOR) {
    MPDOK_REQUIRE(lda >= N, "gemv: lda < N for row-major A");
    const bool vec = ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(x)) & 15) == 0 &&
                     ((lda * sizeof(T)) & 15) == 0;
    constexpr int R = 4;
    dim3 grid((M + 4 * R - 1) / (4 * R));
    if (vec)
      gemv_rm_kernel<T, R, true><<<grid, 256, 0, s>>>(A, lda, M, N, x, y, alpha, skip);
    else
      gemv_rm_kernel<T, R, false><<<grid, 256, 0, s>>>(A, lda, M, N, x, y, alpha, skip);
  } else {
    MPDOK_REQUIRE(lda >= M, "gemv: lda < M for column-major A");
    const int nch = cm_chunks(M, 256 * W, N);
    MPDOK_REQUIRE(work != nullptr && work_bytes >= (size_t)nch * M * sizeof(T),
                  "gemv: column-major A needs a workspace of %zu bytes",
                  (size_t)nch * M * sizeof(T));
    const int cpc = (N + nch - 1) / nch;
    const bool vec = (reinterpret_cast<uintptr_t>(A) & 15) == 0 && ((lda * sizeof(T)) & 15) == 0;
    dim3 grid((M + 256 * W - 1) / (256 * W), nch);
    T* part = static_cast<T*>(work);
    if (vec)
      gemv_cm_partial_kernel<T, true><<<grid, 256, 0, s>>>(A, lda, M, N, x, part, cpc, skip);
    else
      gemv_cm_partial_kernel<T, false><<<grid, 256, 0, s>>>(A, lda, M, N, x, part, cpc, skip);
    gemv_cm_reduce_kernel<T><<<(M + 63) / 64, 256, 0, s>>>(part, M, nch, y, alpha, skip);
  }
  MPDOK_HIP(hipGetLastError());
  return OK;
}

int gemv_f32(hipStream_t s, const float* A, int64_t lda, int order, int M, int N, const float* x,
             float* y, float alpha, void* work, size_t work_bytes, const int* skip) {
  return gemv_launch<float>(s, A, lda, order, M, N, x, y, alpha, work, work_bytes, skip);
}
int gemv_f64(hipStream_t s, const double* A, int64_t lda, int order, int M, int N, const double* x,
             double* y, double alpha, void* work, size_t work_bytes, const int* skip) {
  return gemv_launch<double>(s, A, lda, order, M, N, x, y, alpha, work, work_bytes, skip);
}

int residual_f64(hipStream_t s, const double* A, int64_t lda, int order, int N, const double* x,
                 const double* b, double* r, int mode, void* work, size_t work_bytes) {
  MPDOK_REQUIRE(A && x && b && r && N >= 0, "residual: bad arguments");
  MPDOK_REQUIRE(mode == 0 || mode == 1, "residual: mode must be 0 (fp64) or 1 (dot2)");
  if (N == 0) return OK;
  if (order == ROW_MAJOR) {
    MPDOK_REQUIRE(lda >= N, "residual: lda < N");
    constexpr int R = 2;
    dim3 grid((N + 4 * R - 1) / (4 * R));
    if (mode == 0)
      resid_rm_kernel<0, R><<<grid, 256, 0, s>>>(A, lda, N, x, b, r);
    else
      resid_rm_kernel<1, R><<<grid, 256, 0, s>>>(A, lda, N, x, b, r);
  } else {
    MPDOK_REQUIRE(lda >= N, "residual: lda < N");
    const int nch = cm_chunks(N, 256, N);
    const size_t need = (size_t)nch * N * 8 * 2;
    MPDOK_REQUIRE(work != nullptr && work_bytes >= need,
                  "residual: column-major A needs a workspace of %zu bytes", need);
    const int cpc = (N + nch - 1) / nch;
    double* ph = static_cast<double*>(work);
    double* pl = ph + (size_t)nch * N;
    dim3 grid((N + 255) / 256, nch);
    if (mode == 0) {
      resid_cm_partial_kernel<0><<<grid, 256, 0, s>>>(A, lda, N, x, ph, pl, cpc);
      resid_cm_reduce_kernel<0><<<(N + 255) / 256, 256, 0, s>>>(ph, pl, N, nch, b, r);
    } else {
      resid_cm_partial_kernel<1><<<grid, 256, 0, s>>>(A, lda, N, x, ph, pl, cpc);
      resid_cm_reduce_kernel<1><<<(N + 255) / 256, 256, 0, s>>>(ph, pl, N, nch, b, r);
    }
  }
  MPDOK_HIP(hipGetLastError());
  return OK;
}

// Up to RESID_MULTI right-hand sides per pass (8, then 4, 2, 1).
constexpr int RESID_MULTI = 8;

size_t resid_multi_workspace_bytes(int N) {
  if (N <= 0) return 0;
  const int nch = cm_chunks(N, 256, N);
  return (size_t)nch * N * 8 * 2 * RESID_MULTI;
}

template <int MODE, int NR>
static void resid_multi_launch(hipStream_t s, const double* A, int64_t lda, int N, const double* X,
                               int64_t ldx, const double* B, int64_t ldb, double* R, int64_t ldr,
                               double* ph, double* pl, int nch, int cpc) {
  resid_cm_multi_partial_kernel<MODE, NR><<<dim3((N + 255) / 256, nch), 256, 0, s>>>(
      A, lda, N, X, ldx, ph, pl, cpc);
  resid_cm_multi_reduce_kernel<MODE><<<dim3((N + 255) / 256, NR), 256, 0, s>>>(ph, pl, N, nch, B,
                                                                             ldb, R, ldr);
}

int residual_multi_f64(hipStream_t s, const double* A, int64_t lda, int N, const double* X,
                       int64_t ldx, const double* B, int64_t ldb, double* R, int64_t ldr,
                       int nrhs, int mode, void* work, size_t work_bytes) {
  MPDOK_REQUIRE(A && X && B && R && N >= 0 && nrhs >= 0, "residual_multi: bad arguments");
  MPDOK_REQUIRE(mode == 0 || mode == 1, "residual_multi: mode must be 0 (fp64) or 1 (dot2)");
  MPDOK_REQUIRE(lda >= N && ldx >= N && ldb >= N && ldr >= N, "residual_multi: ld < N");
  if (N == 0 || nrhs == 0) return OK;
  const int nch = cm_chunks(N, 256, N);
  const size_t need = resid_multi_workspace_bytes(N);
  MPDOK_REQUIRE(work != nullptr && work_bytes >= need,
                "residual_multi: needs a workspace of %zu bytes", need);
  const int cpc = (N + nch - 1) / nch;
  double* ph = static_cast<double*>(work);
  double* pl = ph + (size_t)nch * N * RESID_MULTI;
  for (int c0 = 0; c0 < nrhs;) {
    const int nr = nrhs - c0 >= 8 ? 8 : nrhs - c0 >= 4 ? 4 : nrhs - c0 >= 2 ? 2 : 1;
    const double* Xc = X + (int64_t)c0 * ldx;
    const double* Bc = B + (int64_t)c0 * ldb;
    double* Rc = R + (int64_t)c0 * ldr;
#define RM(MD, NRV) resid_multi_launch<MD, NRV>(s, A, lda, N, Xc, ldx, Bc, ldb, Rc, ldr, ph, pl, nch, cpc)
    if (mode == 0) {
      if (nr == 8) RM(0, 8); else if (nr == 4) RM(0, 4); else if (nr == 2) RM(0, 2); else RM(0, 1);
    } else {
      if (nr == 8) RM(1, 8); else if (nr == 4) RM(1, 4); else if (nr == 2) RM(1, 2); else RM(1, 1);
    }
#undef RM
    c0 += nr;
  }
  MPDOK_HIP(hipGetLastError());
  return OK;
}

// ---------------------------------------------------------------------------
// Read-only ceiling: the sum of `bytes` bytes read as FP64 pairs (16-byte
// non-temporal loads, grid-stride), partial sums per workgroup, then one
// workgroup sums the partials in a fixed order. It does only what a GEMV must
// (stream the matrix in once), so its rate is the ceiling the GEMV kernels
// are measured against. No reference counterpart: measurement.
// ---------------------------------------------------------------------------
constexpr int RC_BLOCKS = 2048;

__global__ __launch_bounds__(256) void read_ceiling_kernel(const d2* __restrict__ p, size_t n2,
                                                           double* __restrict__ part) {
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n2; i += (size_t)gridDim.x * 256) {
    const d2 v = __builtin_nontemporal_load(p + i);
    acc += v[0] + v[1];
  }
  acc = wave_sum(acc);
  __shared__ double sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ __launch_bounds__(256) void read_ceiling_sum_kernel(const double* __restrict__ part,
                                                               int np, double* __restrict__ out) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) acc += part[i];
  acc = wave_sum(acc);
  __shared__ double sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

size_t read_ceiling_workspace_bytes() { return RC_BLOCKS * sizeof(double); }

int read_ceiling(hipStream_t s, const void* p, size_t bytes, double* out, void* work) {
  MPDOK_REQUIRE(p && out && work, "read_ceiling: null pointer");
  MPDOK_REQUIRE((reinterpret_cast<uintptr_t>(p) & 15) == 0 && bytes % 16 == 0,
                "read_ceiling: the buffer must be 16-byte aligned and a multiple of 16 bytes");
  const size_t n2 = bytes / 16;
  int nb = (int)((n2 + 255) / 256);
  if (nb > RC_BLOCKS) nb = RC_BLOCKS;
  if (nb < 1) nb = 1;
  double* part = static_cast<double*>(work);
  read_ceiling_kernel<<<nb, 256, 0, s>>>(static_cast<const d2*>(p), n2, part);
  read_ceiling_sum_kernel<<<1, 256, 0, s>>>(part, nb, out);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

// ---------------------------------------------------------------------------
// Allocation-free residual. The column-major residual's partial sums (high
// and low parts per column chunk and row, resid_workspace_bytes) live in a
// file-scope device array instead of a caller workspace. Each GPU holds its
// own copy of the array at its own address; the host looks that address up
// once per device (hipGetSymbolAddress, with the device current) and passes
// it to the kernels. The partial kernel writes every (chunk, row) slot that
// the reduction reads, so nothing needs clearing between calls. Calls that
// overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
constexpr size_t kResidScratchBytes = size_t(16) << 20;  // 16 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kResidScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call per device, then cached).
static int scratch_address(void** out) {
  int dev = -1;
  MPDOK_HIP(hipGetDevice(&dev));
  MPDOK_REQUIRE(dev >= 0 && dev < kMaxDevices, "residual_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[dev] == nullptr) {
    void* p = nullptr;
    MPDOK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MPDOK_REQUIRE(p != nullptr, "residual_noalloc: no device address for the scratch array");
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return OK;
}

int residual_f64_noalloc(hipStream_t s, const double* A, int64_t lda, int order, int N,
                         const double* x, const double* b, double* r, int mode) {
  MPDOK_REQUIRE(N >= 0, "residual_noalloc: N < 0");
  MPDOK_REQUIRE(order == ROW_MAJOR || order == COL_MAJOR, "residual_noalloc: bad order");
  const size_t need = resid_workspace_bytes(N, order);
  MPDOK_REQUIRE(need <= kResidScratchBytes,
                "residual_noalloc: N = %d (column-major) needs %zu bytes of scratch, more than "
                "the %zu the library holds per GPU; use mpdok_residual_f64 with a workspace",
                N, need, kResidScratchBytes);
  void* work = nullptr;
  if (need > 0 && N > 0) MPDOK_TRY(scratch_address(&work));
  return residual_f64(s, A, lda, order, N, x, b, r, mode, work, need);
}

}  // namespace mpdok

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
MPDOK_API size_t mpdok_gemv_workspace_bytes(int M, int N, int dtype, int order) {
  return mpdok::gemv_workspace_bytes(M, N, dtype, order);
}
MPDOK_API size_t mpdok_residual_workspace_bytes(int N, int order) {
  return mpdok::resid_workspace_bytes(N, order);
}
MPDOK_API int mpdok_gemv_f32(hipStream_t s, const float* A, int64_t lda, int order, int M, int N,
                             const float* x, float* y, void* work, size_t work_bytes,
                             const int* skip) {
  return mpdok::gemv_f32(s, A, lda, order, M, N, x, y, 1.0f, work, work_bytes, skip);
}
MPDOK_API int mpdok_gemv_f64(hipStream_t s, const double* A, int64_t lda, int order, int M, int N,
                             const double* x, double* y, void* work, size_t work_bytes,
                             const int* skip) {
  return mpdok::gemv_f64(s, A, lda, order, M, N, x, y, 1.0, work, work_bytes, skip);
}
MPDOK_API size_t mpdok_residual_multi_workspace_bytes(int N) {
  return mpdok::resid_multi_workspace_bytes(N);
}
MPDOK_API int mpdok_residual_multi_f64(hipStream_t s, const double* A, int64_t lda, int N,
                                       const double* X, int64_t ldx, const double* B,
                                       int64_t ldb, double* R, int64_t ldr, int nrhs, int mode,
                                       void* work, size_t work_bytes) {
  return mpdok::residual_multi_f64(s, A, lda, N, X, ldx, B, ldb, R, ldr, nrhs, mode, work,
                                   work_bytes);
}
MPDOK_API size_t mpdok_read_ceiling_workspace_bytes(void) {
  return mpdok::read_ceiling_workspace_bytes();
}
MPDOK_API int mpdok_read_ceiling(hipStream_t s, const void* p, size_t bytes, double* out,
                                 void* work) {
  return mpdok::read_ceiling(s, p, bytes, out, work);
}
MPDOK_API int mpdok_residual_f64_noalloc(hipStream_t s, const double* A, int64_t lda, int order,
                                         int N, const double* x, const double* b, double* r,
                                         int mode) {
  return mpdok::residual_f64_noalloc(s, A, lda, order, N, x, b, r, mode);
}
MPDOK_API int mpdok_residual_f64(hipStream_t s, const double* A, int64_t lda, int order, int N,
                                 const double* x, const double* b, double* r, int mode, void* work,
                                 size_t work_bytes) {
  return mpdok::residual_f64(s, A, lda, order, N, x, b, r, mode, work, work_bytes);
}
