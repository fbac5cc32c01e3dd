// Trailing update C -= A B of the blocked LU on the int8 matrix cores, to
// about 7*S bits, with FP64 storage of C (Ozaki scheme; Ozaki, Ogita, Oishi,
// Rump, Numer. Algorithms 59, 2012).
//
// Splitting: each row of A (m x K) and each column of B (K x n) is scaled by a
// power of two 2^e with e = frexp exponent of its max magnitude, then cut into
// S signed 7-bit digits d_1..d_S (|d| <= 127) with
//     a = 2^e * sum_q d_q 2^(-7 q)      (truncation < 2^(e - 7 S)).
// Product: for each level l = i + j in 2..S+1 the exact integer
//     acc_l = sum_{i+j=l} A_i B_j       (int32; |acc_l| <= S K 127^2)
// comes from v_mfma_i32_16x16x64_i8; pairs with i + j > S + 1 are dropped. The
// epilogue folds c = sum_l acc_l 2^(-7 l) in FP64 for l ascending from 2 and
// subtracts 2^(eA_i + eB_j) c from C. Slicing, integer sums, the fold order and
// the power-of-two scalings are all deterministic, so the result is bit for bit
// the one of the oracle's ozaki_matmul.
//
// Kernel shape: 256 threads = 2x2 wavefronts, wavefront tile 64 x 32 (4 x 2
// MFMA tiles of 16 x 16), block tile 128 x 64, K step 64. For one K step the S
// slices of both operands are staged in LDS once and all S(S+1)/2 slice pairs
// are multiplied from there, with one int32 accumulator set per level.
#include "common.h"
#include "lu.h"

namespace mpdok {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int OZ_BM = 128, OZ_BN = 64, OZ_BK = 64;

// Digits of v (|v| < 2^e) into d[0..S).
template <int S>
__device__ __forceinline__ void make_digits(double v, int e, signed char* d) {
  double r = ldexp(v, -e);
#pragma unroll
  for (int q = 0; q < S; ++q) {
    const double t = r * 128.0;
    double dq = rint(t);
    dq = fmin(fmax(dq, -127.0), 127.0);
    r = t - dq;
    d[q] = (signed char)(int)dq;
  }
}

// Rows of A (m x K, column-major, lda) -> As[q][row][k] (row stride Kp),
// rows padded to mp, k padded to Kp with zeros; ea[row] = exponent.
template <int S>
__global__ __launch_bounds__(256) void slice_rows_kernel(const double* __restrict__ A, int64_t lda,
                                                         int m, int K, int mp, int Kp,
                                                         signed char* __restrict__ As,
                                                         int* __restrict__ ea) {
  __shared__ double tile[64][65];
  __shared__ double red[4][64];
  __shared__ int se[64];
  const int t = threadIdx.x;
  const int r = t & 63, p = t >> 6;
  const int row0 = blockIdx.x * 64;
  const int i = row0 + r;
  double mx = 0.0;
  if (i < m)
    for (int k = p; k < K; k += 4) mx = fmax(mx, fabs(A[i + (int64_t)k * lda]));
  red[p][r] = mx;
  __syncthreads();
  if (t < 64) {
    const double v = fmax(fmax(red[0][t], red[1][t]), fmax(red[2][t], red[3][t]));
    int e = 0;
    if (v > 0.0) frexp(v, &e);
    se[t] = e;
    if (row0 + t < mp) ea[row0 + t] = e;
  }
  __syncthreads();
  const size_t slab = (size_t)mp * Kp;
  for (int k0 = 0; k0 < Kp; k0 += 64) {
    for (int e = t; e < 64 * 64; e += 256) {
      const int rr = e & 63, kk = e >> 6;
      const int ii = row0 + rr, k = k0 + kk;
      tile[rr][kk] = (ii < m && k < K) ? A[ii + (int64_t)k * lda] : 0.0;
    }
    __syncthreads();
    // thread -> (row rr, 16 consecutive k)
    const int rr = t >> 2, kq = (t & 3) * 16;
    const int ii = row0 + rr;
    if (ii < mp) {
      signed char dig[S][16];
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        signed char d[S];
        make_digits<S>(tile[rr][kq + c], se[rr], d);
#pragma unroll
        for (int q = 0; q < S; ++q) dig[q][c] = d[q];
      }
#pragma unroll
      for (int q = 0; q < S; ++q) {
        v4i w;
        __builtin_memcpy(&w, dig[q], 16);
        *reinterpret_cast<v4i*>(As + q * slab + (size_t)ii * Kp + k0 + kq) = w;
      }
    }
    __syncthreads();
  }
}

// Columns of B (K x n, column-major, ldb) -> Bs[q][col][k]; eb[col].
template <int S>
__global__ __launch_bounds__(256) void slice_cols_kernel(const double* __restrict__ B, int64_t ldb,
                                                         int n, int K, int np, int Kp,
                                                         signed char* __restrict__ Bs,
                                                         int* __restrict__ eb) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= np) return;
  const double* col = B + (int64_t)j * ldb;
  double mx = 0.0;
  if (j < n)
    for (int k = lane; k < K; k += 64) mx = fmax(mx, fabs(col[k]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  int e = 0;
  if (mx > 0.0) frexp(mx, &e);
  if (lane == 0) eb[j] = e;
  const size_t slab = (size_t)np * Kp;
  for (int k0 = lane * 16; k0 < Kp; k0 += 64 * 16) {
    signed char dig[S][16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const int k = k0 + c;
      const double v = (j < n && k < K) ? col[k] : 0.0;
      signed char d[S];
      make_digits<S>(v, e, d);
#pragma unroll
      for (int q = 0; q < S; ++q) dig[q][c] = d[q];
    }
#pragma unroll
    for (int q = 0; q < S; ++q) {
      v4i w;
      __builtin_memcpy(&w, dig[q], 16);
      *reinterpret_cast<v4i*>(Bs + q * slab + (size_t)j * Kp + k0) = w;
    }
  }
}

// C (m x n, column-major, ldc) -= sliced product. mp, np: padded row/column
// counts of the slice arrays (multiples of the block tile), Kp multiple of 64.
template <int S>
__global__ __launch_bounds__(256) void ozaki_gemm_kernel(int m, int n, int mp, int np, int Kp,
                                                         const signed char* __restrict__ As,
                                                         const signed char* __restrict__ Bs,
                                                         const int* __restrict__ ea,
                                                         const int* __restrict__ eb,
                                                         double* __restrict__ C, int64_t ldc) {
  constexpr int PADB = 16;  // bytes of padding per LDS row
  __shared__ __attribute__((aligned(16))) signed char sA[S][OZ_BM][OZ_BK + PADB];
  __shared__ __attribute__((aligned(16))) signed char sB[S][OZ_BN][OZ_BK + PADB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 32;
  int ti, tj;
  if (!tile_of(blockIdx.x, mp / OZ_BM, np / OZ_BN, ti, tj)) return;
  const int bm = ti * OZ_BM, bn = tj * OZ_BN;
  const size_t slabA = (size_t)mp * Kp, slabB = (size_t)np * Kp;

  v4i acc[S][4][2];
#pragma unroll
  for (int l = 0; l < S; ++l)
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[l][a][b] = v4i(0);

  // register prefetch of one K step: per thread 2 A chunks and 1 B chunk per slice
  v4i pa[S][2], pb[S];
  const int arow0 = tid >> 2, akq = (tid & 3) * 16;  // A chunks: rows arow0 and arow0 + 64
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const signed char* ap = As + q * slabA + (size_t)(bm + arow0) * Kp + k0 + akq;
      pa[q][0] = *reinterpret_cast<const v4i*>(ap);
      pa[q][1] = *reinterpret_cast<const v4i*>(ap + (size_t)64 * Kp);
      pb[q] = *reinterpret_cast<const v4i*>(Bs + q * slabB + (size_t)(bn + arow0) * Kp + k0 + akq);
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int q = 0; q < S; ++q) {
      *reinterpret_cast<v4i*>(&sA[q][arow0][akq]) = pa[q][0];
      *reinterpret_cast<v4i*>(&sA[q][arow0 + 64][akq]) = pa[q][1];
      *reinterpret_cast<v4i*>(&sB[q][arow0][akq]) = pb[q];
    }
  };
  const int fr = lane & 15, fk = (lane >> 4) * 16;
  gload(0);
  for (int k0 = 0; k0 < Kp; k0 += OZ_BK) {
    sstore();
    __syncthreads();
    if (k0 + OZ_BK < Kp) gload(k0 + OZ_BK);
    // all B-slice fragments of this K step stay in registers across the i loop
    v4i b[S][2];
#pragma unroll
    for (int j = 0; j < S; ++j)
#pragma unroll
      for (int tb = 0; tb < 2; ++tb) b[j][tb] = *reinterpret_cast<const v4i*>(&sB[j][wn + 16 * tb + fr][fk]);
#pragma unroll
    for (int i = 0; i < S; ++i) {
      v4i a[4];
#pragma unroll
      for (int ta = 0; ta < 4; ++ta) a[ta] = *reinterpret_cast<const v4i*>(&sA[i][wm + 16 * ta + fr][fk]);
#pragma unroll
      for (int j = 0; j < S - i; ++j)  // 0-based: level index i + j = l - 2
#pragma unroll
        for (int ta = 0; ta < 4; ++ta)
#pragma unroll
          for (int tb = 0; tb < 2; ++tb)
            acc[i + j][ta][tb] =
                __builtin_amdgcn_mfma_i32_16x16x64_i8(a[ta], b[j][tb], acc[i + j][ta][tb], 0, 0, 0);
    }
    __syncthreads();
  }
  // epilogue: D[i = 4*(lane/16) + r][j = lane%16] per 16 x 16 tile
#pragma unroll
  for (int ta = 0; ta < 4; ++ta)
#pragma unroll
    for (int tb = 0; tb < 2; ++tb) {
      const int j = bn + wn + 16 * tb + (lane & 15);
      if (j >= n) continue;
      const int ej = eb[j];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = bm + wm + 16 * ta + 4 * (lane >> 4) + r;
        if (i >= m) continue;
        double c = 0.0;
#pragma unroll
        for (int l = 0; l < S; ++l) c += ldexp((double)acc[l][ta][tb][r], -7 * (l + 2));
        double* cp = C + i + (int64_t)j * ldc;
        *cp = *cp - ldexp(c, ea[i] + ej);
      }
    }
}

size_t ozaki_workspace_bytes(int m, int n, int K, int S) {
  const int mp = (m + OZ_BM - 1) / OZ_BM * OZ_BM;
  const int np = (n + OZ_BN - 1) / OZ_BN * OZ_BN;
  const int Kp = (K + OZ_BK - 1) / OZ_BK * OZ_BK;
  return (size_t)S * Kp * ((size_t)mp + np) + sizeof(int) * ((size_t)mp + np) + 256;
}

template <int S>
static int ozaki_gemm_s(hipStream_t s, int m, int n, int K, const double* A, int64_t lda,
                        const double* B, int64_t ldb, double* C, int64_t ldc, void* work) {
  const int mp = (m + OZ_BM - 1) / OZ_BM * OZ_BM;
  const int np = (n + OZ_BN - 1) / OZ_BN * OZ_BN;
  const int Kp = (K + OZ_BK - 1) / OZ_BK * OZ_BK;
  signed char* As = static_cast<signed char*>(work);
  signed char* Bs = As + (size_t)S * mp * Kp;
  int* ea = reinterpret_cast<int*>(Bs + (size_t)S * np * Kp);
  int* eb = ea + mp;
  slice_rows_kernel<S><<<mp / 64, 256, 0, s>>>(A, lda, m, K, mp, Kp, As, ea);
  slice_cols_kernel<S><<<np / 4, 256, 0, s>>>(B, ldb, n, K, np, Kp, Bs, eb);
  ozaki_gemm_kernel<S><<<tiles_launch(mp / OZ_BM, np / OZ_BN), 256, 0, s>>>(m, n, mp, np, Kp, As,
                                                                              Bs, ea, eb, C, ldc);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

int ozaki_gemm(hipStream_t s, int S, int m, int n, int K, const double* A, int64_t lda,
               const double* B, int64_t ldb, double* C, int64_t ldc, void* work) {
  if (m <= 0 || n <= 0 || K <= 0) return OK;
  switch (S) {
    case 3: return ozaki_gemm_s<3>(s, m, n, K, A, lda, B, ldb, C, ldc, work);
    case 4: return ozaki_gemm_s<4>(s, m, n, K, A, lda, B, ldb, C, ldc, work);
    case 5: return ozaki_gemm_s<5>(s, m, n, K, A, lda, B, ldb, C, ldc, work);
    case 6: return ozaki_gemm_s<6>(s, m, n, K, A, lda, B, ldb, C, ldc, work);
    case 7: return ozaki_gemm_s<7>(s, m, n, K, A, lda, B, ldb, C, ldc, work);
    default: MPDOK_REQUIRE(false, "ozaki_gemm: slices must be in [3, 7], got %d", S);
  }
}

}  // namespace mpdok

// C -= sliced(A) sliced(B): exposed for parity tests of the trailing update.
MPDOK_API size_t mpdok_ozaki_workspace_bytes(int m, int n, int K, int S) {
  return mpdok::ozaki_workspace_bytes(m, n, K, S);
}
MPDOK_API int mpdok_ozaki_gemm_sub(hipStream_t s, int S, int m, int n, int K, const double* A,
                                   int64_t lda, const double* B, int64_t ldb, double* C,
                                   int64_t ldc, void* work) {
  return mpdok::ozaki_gemm(s, S, m, n, K, A, lda, B, ldb, C, ldc, work);
}
