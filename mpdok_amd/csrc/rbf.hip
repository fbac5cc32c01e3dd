// Fused kernel-matrix assembly: pairwise distance -> kernel value -> nugget
// -> output cast, written once in the target layout.
//
// Reference: rbf_kernel.py:158-227 (build_rbf_kernel: D2 = sq_i + sq_j - 2 c_i.c_j
// via the Gram identity, clamp at 0, exp(-gamma D2), then A[idx, idx] += reg in
// FP64) and mpdok_ooc.py:39-46,118-123 (_build_fp32_tile: the same FP64 value
// cast to FP32, then the nugget added in FP32 after the cast). The reference
// evaluates this in 4+ elementwise passes over (rows x N) temporaries; here one
// thread computes each output element in registers.
//
// Kernel kinds:  0 = Gaussian  exp(-p * d2)             (p = gamma)
//                1 = Matern-3/2 (1 + a) exp(-a), a = sqrt(3 d2) / p  (p = length scale)
// reg_mode:      0 = nugget added in FP64 before the cast
//                1 = nugget added after the cast, in the output precision
#include "common.h"
#include "rbf.h"

namespace mpdok {

template <typename TO> __device__ __forceinline__ TO to_out(double v);
template <> __device__ __forceinline__ double to_out<double>(double v) { return v; }
template <> __device__ __forceinline__ float to_out<float>(double v) { return (float)v; }

// Output element (i, j), i in [row0, row0+rows), j in [col0, col0+cols), stored at
// out[(i-row0)*ld + (j-col0)] (ROW_MAJOR) or out[(i-row0) + (j-col0)*ld] (COL_MAJOR).
// Block: 64 fast x 64 slow indices; the fast index follows the contiguous
// output dimension, and each of the 4 wavefronts takes 16 slow indices. The
// block's 64 slow-index points (coordinates and squared norms) are staged in
// LDS once and read by broadcast, so the loop over a wavefront's slow indices
// issues no memory reads. DT: the dimension when fixed at compile time, 0: D
// at run time. Stores are nontemporal: the output is written once and far
// larger than the caches.
template <typename TO, int ORDER, int DT>
__global__ __launch_bounds__(256) void rbf_tile_kernel(const double* __restrict__ coords, int D,
                                                       int kind, double p, double reg, int reg_mode,
                                                       int row0, int rows, int col0, int cols,
                                                       TO* __restrict__ out, int64_t ld) {
  constexpr int NC = DT ? DT : MAXD;  // coordinates kept per point
  __shared__ double sp[64][NC + 1];   // coordinates, then the squared norm
  const int t = threadIdx.x;
  const int nf = ORDER == ROW_MAJOR ? cols : rows;
  const int ns = ORDER == ROW_MAJOR ? rows : cols;
  const int goff_f = ORDER == ROW_MAJOR ? col0 : row0;
  const int goff_s = ORDER == ROW_MAJOR ? row0 : col0;
  const int n = DT ? DT : D;
  const int s0 = blockIdx.y * 64;
  if (t < 64) {
    const bool in = s0 + t < ns;
    double c[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; ++d)
      c[d] = in && d < n ? coords[(int64_t)(goff_s + s0 + t) * D + d] : 0.0;
#pragma unroll
    for (int d = 0; d < NC; ++d) sp[t][d] = c[d];
    sp[t][NC] = sqnorm<DT>(c, D);
  }
  __syncthreads();
  const int f = blockIdx.x * 64 + (t & 63);
  if (f >= nf) return;
  double cf[MAXD];
#pragma unroll
  for (int d = 0; d < MAXD; ++d) cf[d] = d < n ? coords[(int64_t)(goff_f + f) * D + d] : 0.0;
  const double sqf = sqnorm<DT>(cf, D);
  const int sw = (t >> 6) * 16;  // this wavefront's first slow index in the block
  const int kn = ns - s0 - sw < 16 ? ns - s0 - sw : 16;
  TO* o = out + (int64_t)(s0 + sw) * ld + f;
  for (int k = 0; k < kn; ++k) {
    const int sl = sw + k;
    double cs[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; ++d) cs[d] = d < NC && d < n ? sp[sl][d] : 0.0;
    const double sqs = sp[sl][NC];
    // The reference forms D2[i, j] with i the row point: keep that operand order.
    const double d2 = ORDER == ROW_MAJOR ? gram_d2<DT>(cs, sqs, cf, sqf, D)
                                         : gram_d2<DT>(cf, sqf, cs, sqs, D);
    const double v = kernel_value(kind, p, d2);
    const bool diag = goff_f + f == goff_s + s0 + sl;
    TO e;
    if (diag && reg_mode == 0) {
      e = to_out<TO>(v + reg);
    } else {
      e = to_out<TO>(v);
      if (diag) e = e + (TO)reg;
    }
    __builtin_nontemporal_store(e, o + (int64_t)k * ld);
  }
}

// A principal block (row0 == col0, rows == cols) is symmetric bit for bit:
// d2 = (sq_i + sq_j) - 2 (c_i . c_j) makes the same roundings with i and j
// exchanged (sums and products commute; a fused multiply-add rounds once,
// after its exact product). Each workgroup takes one 64 x 64 tile of the
// upper block triangle (fast-index tile bf <= slow-index tile bs), evaluates
// it once and writes it twice: along the fast index straight from registers
// and, off the diagonal, its mirror through a per-wavefront LDS transpose, so
// both writes are coalesced. Half the kernel evaluations of rbf_tile_kernel,
// the same values.
template <typename TO, int DT>
__global__ __launch_bounds__(256) void rbf_sym_kernel(const double* __restrict__ coords, int D,
                                                      int kind, double p, double reg, int reg_mode,
                                                      int n0, int n, TO* __restrict__ out,
                                                      int64_t ld) {
  constexpr int NC = DT ? DT : MAXD;
  __shared__ double sp[64][NC + 1];  // slow-index points: coordinates, squared norm
  __shared__ TO tv[4][16][65];       // per wavefront: its 16 slow x 64 fast values
  // tile pair from the linear index L = bs (bs + 1) / 2 + bf, 0 <= bf <= bs
  const int64_t L = blockIdx.x;
  int64_t bs = (int64_t)((sqrt(8.0 * (double)L + 1.0) - 1.0) * 0.5);
  while ((bs + 1) * (bs + 2) / 2 <= L) ++bs;
  while (bs * (bs + 1) / 2 > L) --bs;
  const int fb = (int)(L - bs * (bs + 1) / 2) * 64, s0 = (int)bs * 64;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int nd = DT ? DT : D;
  if (t < 64) {
    const bool in = s0 + t < n;
    double c[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; ++d)
      c[d] = in && d < nd ? coords[(int64_t)(n0 + s0 + t) * D + d] : 0.0;
#pragma unroll
    for (int d = 0; d < NC; ++d) sp[t][d] = c[d];
    sp[t][NC] = sqnorm<DT>(c, D);
  }
  __syncthreads();
  const int f = fb + lane;
  const int sw = w * 16;
  const int kn = n - s0 - sw < 16 ? n - s0 - sw : 16;
  if (f < n) {
    double cf[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; ++d) cf[d] = d < nd ? coords[(int64_t)(n0 + f) * D + d] : 0.0;
    const double sqf = sqnorm<DT>(cf, D);
    TO* o = out + (int64_t)(s0 + sw) * ld + f;
    for (int k = 0; k < kn; ++k) {
      const int sl = sw + k;
      double cs[MAXD];
#pragma unroll
      for (int d = 0; d < MAXD; ++d) cs[d] = d < NC && d < nd ? sp[sl][d] : 0.0;
      const double v = kernel_value(kind, p, gram_d2<DT>(cf, sqf, cs, sp[sl][NC], D));
      const bool diag = f == s0 + sl;
      TO e;
      if (diag && reg_mode == 0) {
        e = to_out<TO>(v + reg);
      } else {
        e = to_out<TO>(v);
        if (diag) e = e + (TO)reg;
      }
      __builtin_nontemporal_store(e, o + (int64_t)k * ld);
      tv[w][k][lane] = e;
    }
  }
  if (fb == s0) return;  // a diagonal tile is its own mirror
  // The mirror: fast index s, slow index f. Each store covers 4 slow x 16
  // fast indices; the values come from lanes of this wavefront only (LDS
  // accesses of one wavefront complete in order).
  const int kk = lane & 15, q = lane >> 4;
  if (kk >= kn) return;
  TO* m = out + (int64_t)fb * ld + s0 + sw + kk;
  for (int i = 0; i < 16; ++i) {
    const int fl = 4 * i + q;
    if (fb + fl < n) __builtin_nontemporal_store(tv[w][kk][fl], m + (int64_t)fl * ld);
  }
}

int rbf_assemble(hipStream_t st, const double* coords, int N, int D, int kind, double p,
                 double reg, int reg_mode, int row0, int rows, int col0, int cols, void* out,
                 int out_dtype, int64_t ld, int order) {
  MPDOK_REQUIRE(coords && out, "rbf_assemble: null pointer");
  MPDOK_REQUIRE(D >= 1 && D <= MAXD, "rbf_assemble: D must be in [1, %d], got %d", MAXD, D);
  MPDOK_REQUIRE(kind == 0 || kind == 1, "rbf_assemble: unknown kernel kind %d", kind);
  MPDOK_REQUIRE(row0 >= 0 && rows >= 0 && row0 + rows <= N && col0 >= 0 && cols >= 0 &&
                    col0 + cols <= N,
                "rbf_assemble: tile [%d,+%d)x[%d,+%d) outside N=%d", row0, rows, col0, cols, N);
  MPDOK_REQUIRE(order == ROW_MAJOR ? ld >= cols : ld >= rows, "rbf_assemble: ld too small");
  MPDOK_REQUIRE(!(kind == 1 && !(p > 0.0)), "rbf_assemble: Matern length scale must be > 0");
  if (rows == 0 || cols == 0) return OK;
  const int nf = order == ROW_MAJOR ? cols : rows;
  const int ns = order == ROW_MAJOR ? rows : cols;
  dim3 grid((nf + 63) / 64, (ns + 63) / 64);
  MPDOK_REQUIRE(grid.y <= 65535, "rbf_assemble: tile too tall");
  MPDOK_REQUIRE(out_dtype == F64 || out_dtype == F32, "rbf_assemble: unsupported output dtype %d",
                out_dtype);
  if (row0 == col0 && rows == cols) {
    const int64_t T = (rows + 63) / 64, pairs = T * (T + 1) / 2;
    MPDOK_REQUIRE(pairs <= 0x7fffffff, "rbf_assemble: tile too large");
#define LAUNCH_S(DT)                                                                          \
  if (out_dtype == F64)                                                                        \
    rbf_sym_kernel<double, DT><<<(unsigned)pairs, 256, 0, st>>>(coords, D, kind, p, reg,       \
                                                                reg_mode, row0, rows,          \
                                                                (double*)out, ld);             \
  else                                                                                         \
    rbf_sym_kernel<float, DT><<<(unsigned)pairs, 256, 0, st>>>(coords, D, kind, p, reg,        \
                                                               reg_mode, row0, rows, (float*)out, \
                                                               ld)
    MPDOK_RBF_DIM_DISPATCH(D, LAUNCH_S);
#undef LAUNCH_S
    MPDOK_HIP(hipGetLastError());
    return OK;
  }
#define LAUNCH_T(TO, ORD, DT)                                                           \
  rbf_tile_kernel<TO, ORD, DT><<<grid, 256, 0, st>>>(coords, D, kind, p, reg, reg_mode, row0, \
                                                     rows, col0, cols, (TO*)out, ld)
#define LAUNCH(DT)                                                 \
  if (out_dtype == F64) {                                          \
    if (order == ROW_MAJOR) LAUNCH_T(double, ROW_MAJOR, DT);       \
    else LAUNCH_T(double, COL_MAJOR, DT);                          \
  } else {                                                         \
    if (order == ROW_MAJOR) LAUNCH_T(float, ROW_MAJOR, DT);        \
    else LAUNCH_T(float, COL_MAJOR, DT);                           \
  }
  MPDOK_RBF_DIM_DISPATCH(D, LAUNCH);
#undef LAUNCH
#undef LAUNCH_T
  MPDOK_HIP(hipGetLastError());
  return OK;
}

// Diagonal of the kernel matrix, element for element what the tile kernel and
// the matrix-free products use: d2 of a point with itself by the Gram identity
// (clamped at zero; rounding can leave it slightly positive), the kernel value,
// the nugget in FP64 (elem F64) or after the FP32 cast (elem F32). Written as
// `out_dtype`, optionally inverted (the Jacobi preconditioner's M^-1).
template <typename TO>
__global__ __launch_bounds__(256) void rbf_diag_kernel(const double* __restrict__ coords, int N,
                                                       int D, int kind, double p, double reg,
                                                       int elem, int invert, TO* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  double c[MAXD];
#pragma unroll
  for (int d = 0; d < MAXD; ++d) c[d] = d < D ? coords[(int64_t)i * D + d] : 0.0;
  const double sq = sqnorm(c, D);
  const double v = kernel_value(kind, p, gram_d2(c, sq, c, sq, D));
  double e;
  if (elem == F64) {
    e = v + reg;
  } else {
    float f = (float)v;
    f = f + (float)reg;
    e = (double)f;
  }
  const TO o = (TO)e;
  out[i] = invert ? (TO)1 / o : o;
}

int rbf_diag(hipStream_t st, const double* coords, int N, int D, int kind, double p, double reg,
             int elem_dtype, int invert, void* out, int out_dtype) {
  MPDOK_REQUIRE(coords && out && N >= 0, "rbf_diag: bad arguments");
  MPDOK_REQUIRE(D >= 1 && D <= MAXD, "rbf_diag: D must be in [1, %d], got %d", MAXD, D);
  MPDOK_REQUIRE(kind == 0 || kind == 1, "rbf_diag: unknown kernel kind %d", kind);
  MPDOK_REQUIRE(!(kind == 1 && !(p > 0.0)), "rbf_diag: Matern length scale must be > 0");
  MPDOK_REQUIRE(elem_dtype == F64 || elem_dtype == F32, "rbf_diag: elem dtype must be F64/F32");
  if (N == 0) return OK;
  const dim3 g((N + 255) / 256);
  if (out_dtype == F64)
    rbf_diag_kernel<double><<<g, 256, 0, st>>>(coords, N, D, kind, p, reg, elem_dtype, invert,
                                               (double*)out);
  else if (out_dtype == F32)
    rbf_diag_kernel<float><<<g, 256, 0, st>>>(coords, N, D, kind, p, reg, elem_dtype, invert,
                                              (float*)out);
  else
    MPDOK_REQUIRE(false, "rbf_diag: unsupported output dtype %d", out_dtype);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

// Mean of the clamped squared distances over the first n points, diagonal
// included (rbf_kernel.py:57-68). Single block; n <= 512 in the reference.
__global__ __launch_bounds__(256) void mean_d2_kernel(const double* __restrict__ coords, int n,
                                                      int D, double* __restrict__ out) {
  double acc = 0.0;
  for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += 256) {
    const int i = (int)(e / n), j = (int)(e % n);
    double ci[MAXD], cj[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; ++d) {
      ci[d] = d < D ? coords[(int64_t)i * D + d] : 0.0;
      cj[d] = d < D ? coords[(int64_t)j * D + d] : 0.0;
    }
    acc += gram_d2(ci, sqnorm(ci, D), cj, sqnorm(cj, D), D);
  }
  acc = wave_sum(acc);
  __shared__ double s[4];
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = ((s[0] + s[1]) + (s[2] + s[3])) / ((double)n * (double)n);
}

int rbf_mean_d2(hipStream_t st, const double* coords, int n, int D, double* out) {
  MPDOK_REQUIRE(coords && out && n >= 1 && D >= 1 && D <= MAXD, "rbf_mean_d2: bad arguments");
  mean_d2_kernel<<<1, 256, 0, st>>>(coords, n, D, out);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

// ---------------------------------------------------------------------------
// Matrix-free products with kernel values recomputed from the coordinates and
// never stored.
//
// rbf_matvec_kernel: y_i = sum_j E(k(r_i, c_j)) x_j for row points r_i
// (i in [row0, row0 + rows) of R) and all column points c_j of C.
//   ELEM = F64: E = k + reg on the diagonal, in FP64 (the FP64 outer residual
//               of mpdok_ooc.py:205-227).
//   ELEM = F32: E = FP32 rounding of k, reg added in FP32 after the cast: bit
//               for bit the element of the stored FP32 tile (mpdok_ooc.py:39-46,
//               118-123), i.e. the matrix-free form of the FP32 inner product.
//   The diagonal (and so the nugget) exists only when R and C are the same
//   point set (same = 1); a cross product K(Q, X) x has none.
// DERIV = 1 replaces k by its derivative with respect to the kernel parameter
// p (Gaussian exp(-p d2): -d2 k; Matern-3/2 with a = sqrt(3 d2) / p:
// a^2 exp(-a) / p); with no nugget. Products and sums are FP64.
//
// Block: 32 rows x 8 column phases (256 threads). Coordinates, squared norms
// and x of a 256-column chunk are staged in LDS; each half-wavefront reads one
// column per step (broadcast reads). The 8 partial sums of a row are added in
// a fixed order, so results do not depend on timing.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double kernel_dparam(int kind, double p, double d2) {
  if (kind == 0) return -d2 * exp(-p * d2);
  const double a = sqrt(3.0 * d2) / p;
  return a * a * exp(-a) / p;
}

template <int ELEM, int DERIV, int DT, typename TX>
__device__ __forceinline__ double rbf_row_sum(const double* __restrict__ ci, double sqi, int i,
                                              const double* __restrict__ C, int NC, int D,
                                              int kind, double p, double reg, int same,
                                              const TX* __restrict__ x, double (*cj)[256],
                                              double* sqj, double* xj, int t, int ph) {
  const int n = DT ? DT : D;
  double acc = 0.0;
  for (int j0 = 0; j0 < NC; j0 += 256) {
    const int j = j0 + t;
    double c[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; ++d) c[d] = (j < NC && d < n) ? C[(int64_t)j * D + d] : 0.0;
#pragma unroll
    for (int d = 0; d < MAXD; ++d)
      if (d < n) cj[d][t] = c[d];
    sqj[t] = sqnorm<DT>(c, D);
    xj[t] = j < NC ? (double)x[j] : 0.0;
    __syncthreads();
    const int jn = NC - j0 < 256 ? NC - j0 : 256;
    for (int k = ph; k < jn; k += 8) {
      double cc[MAXD];
#pragma unroll
      for (int d = 0; d < MAXD; ++d) cc[d] = d < n ? cj[d][k] : 0.0;
      const double d2 = gram_d2<DT>(ci, sqi, cc, sqj[k], D);
      double e;
      if (DERIV) {
        e = kernel_dparam(kind, p, d2);
      } else {
        const double v = kernel_value(kind, p, d2);
        const bool diag = same && i == j0 + k;
        if (ELEM == F64) {
          e = diag ? v + reg : v;
        } else {
          float f = (float)v;
          if (diag) f = f + (float)reg;
          e = (double)f;
        }
      }
      acc = __builtin_fma(e, xj[k], acc);
    }
    __syncthreads();
  }
  return acc;
}

// y[i - row0] = sum_j E_ij x_j  (y in TY). DT: the dimension when fixed at
// compile time, 0: D at run time.
template <int ELEM, int DT, typename TX, typename TY>
__global__ __launch_bounds__(256) void rbf_matvec_kernel(const double* __restrict__ R,
                                                         const double* __restrict__ C, int NC,
                                                         int D, int kind, double p, double reg,
                                                         int same, int row0, int rows,
                                                         const TX* __restrict__ x,
                                                         TY* __restrict__ y) {
  __shared__ double cj[DT ? DT : MAXD][256];
  __shared__ double sqj[256];
  __shared__ double xj[256];
  __shared__ double part[8][33];
  const int t = threadIdx.x, r = t & 31, ph = t >> 5;
  const int il = blockIdx.x * 32 + r;
  const int i = row0 + il;
  const bool live = il < rows;
  const int n = DT ? DT : D;
  double ci[MAXD];
#pragma unroll
  for (int d = 0; d < MAXD; ++d) ci[d] = (live && d < n) ? R[(int64_t)i * D + d] : 0.0;
  const double acc = rbf_row_sum<ELEM, 0, DT, TX>(ci, sqnorm<DT>(ci, D), i, C, NC, D, kind, p, reg,
                                                 same, x, cj, sqj, xj, t, ph);
  part[ph][r] = acc;
  __syncthreads();
  if (ph == 0 && live) {
    const double s = ((part[0][r] + part[1][r]) + (part[2][r] + part[3][r])) +
                     ((part[4][r] + part[5][r]) + (part[6][r] + part[7][r]));
    y[il] = (TY)s;
  }
}

// Bilinear form u^T (dK/dp) w over rows [0, NR) of R and all points of C:
// per-block partial sums into part_out[blockIdx.x], reduced in a fixed order
// by bilinear_reduce_kernel.
template <int DT>
__global__ __launch_bounds__(256) void rbf_dbilinear_kernel(const double* __restrict__ R, int NR,
                                                            const double* __restrict__ C, int NC,
                                                            int D, int kind, double p,
                                                            const double* __restrict__ u,
                                                            const double* __restrict__ w,
                                                            double* __restrict__ part_out) {
  __shared__ double cj[DT ? DT : MAXD][256];
  __shared__ double sqj[256];
  __shared__ double xj[256];
  __shared__ double part[8][33];
  const int t = threadIdx.x, r = t & 31, ph = t >> 5;
  const int i = blockIdx.x * 32 + r;
  const bool live = i < NR;
  const int n = DT ? DT : D;
  double ci[MAXD];
#pragma unroll
  for (int d = 0; d < MAXD; ++d) ci[d] = (live && d < n) ? R[(int64_t)i * D + d] : 0.0;
  const double acc = rbf_row_sum<F64, 1, DT, double>(ci, sqnorm<DT>(ci, D), i, C, NC, D, kind, p,
                                                     0.0, 0, w, cj, sqj, xj, t, ph);
  part[ph][r] = acc;
  __syncthreads();
  if (ph == 0) {
    double s = ((part[0][r] + part[1][r]) + (part[2][r] + part[3][r])) +
               ((part[4][r] + part[5][r]) + (part[6][r] + part[7][r]));
    s = live ? s * u[i] : 0.0;
    // the 32 rows of this block, in a fixed order
    for (int off = 16; off > 0; off >>= 1) s += __shfl_down(s, off, 32);
    if (r == 0) part_out[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(256) void bilinear_reduce_kernel(const double* __restrict__ part,
                                                              int n, double* __restrict__ out) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int k = threadIdx.x; k < n; k += 256) acc += part[k];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

// ---------------------------------------------------------------------------
// Matrix-free y = E x over a whole point set (all rows, the same points on
// both sides). E is symmetric bit for bit (see rbf_sym_kernel), so each
// workgroup evaluates one pair of 256-point tiles I <= J once and uses each
// element twice: in the row sums of I (E_IJ x_J) and in the column sums of J
// (E_IJ^T x_I = E_JI x_I). The 256 partial sums of each go to P[I T + J] and
// P[J T + I] (T tiles); rbf_symv_reduce_kernel then adds P[I T + 0 .. T) in
// that order, so results do not depend on timing. Half the kernel evaluations
// of rbf_matvec_kernel; the sums are ordered differently.
// Thread t holds row point I*256 + t. Off the diagonal, lane l of a wavefront
// takes column (l + s) mod 256 at step s: every lane sees every column, the
// lanes of a step see different columns (conflict-free LDS reads), and the
// column sums are kept per wavefront in LDS (no two lanes of a step touch the
// same one; wavefronts have their own), added in a fixed order at the end.
// ---------------------------------------------------------------------------
template <int ELEM>
__device__ __forceinline__ double elem_value(int kind, double p, double reg, double d2, bool diag) {
  const double v = kernel_value(kind, p, d2);
  if (ELEM == F64) return diag ? v + reg : v;
  float f = (float)v;
  if (diag) f = f + (float)reg;
  return (double)f;
}

template <int ELEM, int DT, typename TX>
__global__ __launch_bounds__(256) void rbf_symv_kernel(const double* __restrict__ C, int N, int D,
                                                       int kind, double p, double reg,
                                                       const TX* __restrict__ x,
                                                       double* __restrict__ P, int T) {
  constexpr int NC = DT ? DT : MAXD;
  __shared__ double cj[NC][256];
  __shared__ double sqj[256];
  __shared__ double xj[256];
  __shared__ double cs[4][256];
  // tile pair from the linear index L = J (J + 1) / 2 + I, 0 <= I <= J
  const int64_t L = blockIdx.x;
  int64_t J = (int64_t)((sqrt(8.0 * (double)L + 1.0) - 1.0) * 0.5);
  while ((J + 1) * (J + 2) / 2 <= L) ++J;
  while (J * (J + 1) / 2 > L) --J;
  const int I = (int)(L - J * (J + 1) / 2);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int nd = DT ? DT : D;
  const int i = I * 256 + t, j = (int)J * 256 + t;
  {
    double c[MAXD];
#pragma unroll
    for (int d = 0; d < MAXD; ++d) c[d] = (j < N && d < nd) ? C[(int64_t)j * D + d] : 0.0;
#pragma unroll
    for (int d = 0; d < NC; ++d)
      if (d < nd) cj[d][t] = c[d];
    sqj[t] = sqnorm<DT>(c, D);
    xj[t] = j < N ? (double)x[j] : 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) cs[q][t] = 0.0;
  }
  double ci[MAXD];
#pragma unroll
  for (int d = 0; d < MAXD; ++d) ci[d] = (i < N && d < nd) ? C[(int64_t)i * D + d] : 0.0;
  const double sqi = sqnorm<DT>(ci, D);
  const double xi = i < N ? (double)x[i] : 0.0;
  __syncthreads();
  // Columns past N carry x = 0 and rows past N x = 0: their (finite) values
  // add nothing to the sums that are kept.
  double rs = 0.0;
  if (I == J) {
    for (int k = 0; k < 256; ++k) {
      double cc[MAXD];
#pragma unroll
      for (int d = 0; d < MAXD; ++d) cc[d] = d < nd ? cj[d][k] : 0.0;
      const double e =
          elem_value<ELEM>(kind, p, reg, gram_d2<DT>(ci, sqi, cc, sqj[k], D), k == t && i < N);
      rs = __builtin_fma(e, xj[k], rs);
    }
  } else {
    double* csw = cs[w];
    for (int st = 0; st < 256; ++st) {
      const int k = (lane + st) & 255;
      double cc[MAXD];
#pragma unroll
      for (int d = 0; d < MAXD; ++d) cc[d] = d < nd ? cj[d][k] : 0.0;
      const double e = elem_value<ELEM>(kind, p, reg, gram_d2<DT>(ci, sqi, cc, sqj[k], D), false);
      rs = __builtin_fma(e, xj[k], rs);
      csw[k] = __builtin_fma(e, xi, csw[k]);
    }
  }
  if (i < N) P[((int64_t)I * T + J) * 256 + t] = rs;
  if (I == J) return;
  __syncthreads();
  if (j < N) P[((int64_t)J * T + I) * 256 + t] = (cs[0][t] + cs[1][t]) + (cs[2][t] + cs[3][t]);
}

template <typename TY>
__global__ __launch_bounds__(256) void rbf_symv_reduce_kernel(const double* __restrict__ P, int N,
                                                              int T, TY* __restrict__ y) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const double* q = P + (int64_t)blockIdx.x * T * 256 + threadIdx.x;
  double s = 0.0;
  for (int J = 0; J < T; ++J) s += q[(int64_t)J * 256];
  y[i] = (TY)s;
}

static int rbf_check(const double* R, const double* C, int D, int kind, double p) {
  MPDOK_REQUIRE(R && C, "rbf: null coordinates");
  MPDOK_REQUIRE(D >= 1 && D <= MAXD, "rbf: D must be in [1, %d], got %d", MAXD, D);
  MPDOK_REQUIRE(kind == 0 || kind == 1, "rbf: unknown kernel kind %d", kind);
  MPDOK_REQUIRE(!(kind == 1 && !(p > 0.0)), "rbf: Matern length scale must be > 0");
  return OK;
}

int rbf_cross_matvec(hipStream_t st, const double* R, int NR, const double* C, int NC, int D,
                     int kind, double p, double reg, int same, int elem_dtype, const void* x,
                     void* y, int xy_dtype, int row0, int rows) {
  MPDOK_TRY(rbf_check(R, C, D, kind, p));
  MPDOK_REQUIRE(x && y, "rbf_matvec: null vector");
  MPDOK_REQUIRE(NR >= 0 && NC >= 0 && row0 >= 0 && rows >= 0 && row0 + rows <= NR,
                "rbf_matvec: rows outside the row set");
  MPDOK_REQUIRE(!same || NR == NC, "rbf_matvec: same point set needs NR == NC");
  MPDOK_REQUIRE(elem_dtype == F64 || elem_dtype == F32, "rbf_matvec: element dtype %d", elem_dtype);
  MPDOK_REQUIRE(xy_dtype == F64 || xy_dtype == F32, "rbf_matvec: vector dtype %d", xy_dtype);
  if (rows == 0) return OK;
  const int grid = (rows + 31) / 32;
#define LAUNCH_T(E, DT, T)                                                                \
  rbf_matvec_kernel<E, DT, T, T><<<grid, 256, 0, st>>>(R, C, NC, D, kind, p, reg, same, row0, \
                                                       rows, (const T*)x, (T*)y)
#define LAUNCH(DT)                                                       \
  if (elem_dtype == F64) {                                               \
    if (xy_dtype == F64) LAUNCH_T(F64, DT, double);                      \
    else LAUNCH_T(F64, DT, float);                                       \
  } else {                                                               \
    if (xy_dtype == F64) LAUNCH_T(F32, DT, double);                      \
    else LAUNCH_T(F32, DT, float);                                       \
  }
  MPDOK_RBF_DIM_DISPATCH(D, LAUNCH);
#undef LAUNCH
#undef LAUNCH_T
  MPDOK_HIP(hipGetLastError());
  return OK;
}

int rbf_matvec(hipStream_t st, const double* coords, int N, int D, int kind, double p, double reg,
               int elem_dtype, const void* x, void* y, int xy_dtype, int row0, int rows) {
  return rbf_cross_matvec(st, coords, N, coords, N, D, kind, p, reg, 1, elem_dtype, x, y,
                          xy_dtype, row0, rows);
}

size_t rbf_symv_workspace_bytes(int N) {
  const size_t T = N > 0 ? ((size_t)N + 255) / 256 : 0;
  return T * T * 256 * sizeof(double);
}

int rbf_symv(hipStream_t st, const double* coords, int N, int D, int kind, double p, double reg,
             int elem_dtype, const void* x, void* y, int xy_dtype, void* work, size_t work_bytes) {
  MPDOK_TRY(rbf_check(coords, coords, D, kind, p));
  MPDOK_REQUIRE(x && y, "rbf_symv: null vector");
  MPDOK_REQUIRE(N >= 0, "rbf_symv: N < 0");
  MPDOK_REQUIRE(elem_dtype == F64 || elem_dtype == F32, "rbf_symv: element dtype %d", elem_dtype);
  MPDOK_REQUIRE(xy_dtype == F64 || xy_dtype == F32, "rbf_symv: vector dtype %d", xy_dtype);
  if (N == 0) return OK;
  MPDOK_REQUIRE(work && work_bytes >= rbf_symv_workspace_bytes(N),
                "rbf_symv: workspace of %zu bytes, %zu needed", work_bytes,
                rbf_symv_workspace_bytes(N));
  const int T = (N + 255) / 256;
  const int64_t pairs = (int64_t)T * (T + 1) / 2;
  MPDOK_REQUIRE(pairs <= 0x7fffffff, "rbf_symv: N too large");
  double* P = static_cast<double*>(work);
#define LAUNCH_T(E, DT, TX)                                                               \
  rbf_symv_kernel<E, DT, TX><<<(unsigned)pairs, 256, 0, st>>>(coords, N, D, kind, p, reg, \
                                                              (const TX*)x, P, T)
#define LAUNCH(DT)                                       \
  if (elem_dtype == F64) {                               \
    if (xy_dtype == F64) LAUNCH_T(F64, DT, double);      \
    else LAUNCH_T(F64, DT, float);                       \
  } else {                                               \
    if (xy_dtype == F64) LAUNCH_T(F32, DT, double);      \
    else LAUNCH_T(F32, DT, float);                       \
  }
  MPDOK_RBF_DIM_DISPATCH(D, LAUNCH);
#undef LAUNCH
#undef LAUNCH_T
  if (xy_dtype == F64)
    rbf_symv_reduce_kernel<double><<<T, 256, 0, st>>>(P, N, T, (double*)y);
  else
    rbf_symv_reduce_kernel<float><<<T, 256, 0, st>>>(P, N, T, (float*)y);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

size_t rbf_dbilinear_workspace_bytes(int NR) { return (size_t)((NR + 31) / 32 + 1) * 8; }

int rbf_dbilinear(hipStream_t st, const double* R, int NR, const double* C, int NC, int D,
                  int kind, double p, const double* u, const double* w, double* out, void* work,
                  size_t work_bytes) {
  MPDOK_TRY(rbf_check(R, C, D, kind, p));
  MPDOK_REQUIRE(u && w && out && work, "rbf_dbilinear: null pointer");
  MPDOK_REQUIRE(NR >= 1 && NC >= 1, "rbf_dbilinear: empty point set");
  MPDOK_REQUIRE(work_bytes >= rbf_dbilinear_workspace_bytes(NR), "rbf_dbilinear: workspace");
  const int nb = (NR + 31) / 32;
  double* part = static_cast<double*>(work);
#define LAUNCH(DT) rbf_dbilinear_kernel<DT><<<nb, 256, 0, st>>>(R, NR, C, NC, D, kind, p, u, w, part)
  MPDOK_RBF_DIM_DISPATCH(D, LAUNCH);
#undef LAUNCH
  bilinear_reduce_kernel<<<1, 256, 0, st>>>(part, nb, out);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

}  // namespace mpdok

MPDOK_API int mpdok_rbf_matvec(hipStream_t st, const double* coords, int N, int D, int kind,
                               double p, double reg, int elem_dtype, const void* x, void* y,
                               int xy_dtype, int row0, int rows) {
  return mpdok::rbf_matvec(st, coords, N, D, kind, p, reg, elem_dtype, x, y, xy_dtype, row0, rows);
}
MPDOK_API size_t mpdok_rbf_symv_workspace_bytes(int N) { return mpdok::rbf_symv_workspace_bytes(N); }
MPDOK_API int mpdok_rbf_symv(hipStream_t st, const double* coords, int N, int D, int kind,
                             double p, double reg, int elem_dtype, const void* x, void* y,
                             int xy_dtype, void* work, size_t work_bytes) {
  return mpdok::rbf_symv(st, coords, N, D, kind, p, reg, elem_dtype, x, y, xy_dtype, work,
                         work_bytes);
}
MPDOK_API int mpdok_rbf_cross_matvec(hipStream_t st, const double* R, int NR, const double* C,
                                     int NC, int D, int kind, double p, const double* x,
                                     double* y) {
  return mpdok::rbf_cross_matvec(st, R, NR, C, NC, D, kind, p, 0.0, 0, mpdok::F64, x, y,
                                 mpdok::F64, 0, NR);
}
MPDOK_API size_t mpdok_rbf_dbilinear_workspace_bytes(int NR) {
  return mpdok::rbf_dbilinear_workspace_bytes(NR);
}
MPDOK_API int mpdok_rbf_dbilinear(hipStream_t st, const double* R, int NR, const double* C, int NC,
                                  int D, int kind, double p, const double* u, const double* w,
                                  double* out, void* work, size_t work_bytes) {
  return mpdok::rbf_dbilinear(st, R, NR, C, NC, D, kind, p, u, w, out, work, work_bytes);
}

MPDOK_API int mpdok_rbf_assemble(hipStream_t st, const double* coords, int N, int D, int kind,
                                 double p, double reg, int reg_mode, int row0, int rows, int col0,
                                 int cols, void* out, int out_dtype, int64_t ld, int order) {
  return mpdok::rbf_assemble(st, coords, N, D, kind, p, reg, reg_mode, row0, rows, col0, cols, out,
                             out_dtype, ld, order);
}
MPDOK_API int mpdok_rbf_mean_d2(hipStream_t st, const double* coords, int n, int D, double* out) {
  return mpdok::rbf_mean_d2(st, coords, n, D, out);
}
MPDOK_API int mpdok_rbf_diag(hipStream_t st, const double* coords, int N, int D, int kind,
                             double p, double reg, int elem_dtype, int invert, void* out,
                             int out_dtype) {
  return mpdok::rbf_diag(st, coords, N, D, kind, p, reg, elem_dtype, invert, out, out_dtype);
}
