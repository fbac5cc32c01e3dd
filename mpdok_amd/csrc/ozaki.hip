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
// counts of the slice arrays (multiples of 128), Kp multiple of 64.
// 4 wavefronts (2 x 2), block tile 128 x 64. PREFETCH: stage the next K step
// through registers while the current one is multiplied (more registers,
// one workgroup per CU); without it the kernel is lean enough for two
// workgroups per CU at S <= 5, which overlap each other's staging.
// LDS rows hold the 64 bytes of one K step of one row (column) of a slice.
// SWZ: 64-byte rows with the four 16-byte chunks permuted by
// q ^ (((row >> 1) ^ (row >> 2)) & 3), so that the 8 (or 16) lanes served
// together by the LDS, which read or write consecutive rows, always hit
// distinct bank groups; otherwise 80-byte rows (16 bytes of padding).
template <bool SWZ>
__device__ __forceinline__ int lds_off(int row, int q) {
  return SWZ ? row * 64 + 16 * (q ^ (((row >> 1) ^ (row >> 2)) & 3)) : row * 80 + 16 * q;
}

// Columns [col0, col0 + ncols) of the product (col0 a multiple of 64).
template <int S, bool PREFETCH, bool SWZ>
__global__ __launch_bounds__(256, (PREFETCH || S > 5) ? 1 : 2) void ozaki_gemm_kernel(int m, int n, int mp, int np, int Kp,
                                                             int col0, int ncols,
                                                             const signed char* __restrict__ As,
                                                             const signed char* __restrict__ Bs,
                                                             const int* __restrict__ ea,
                                                             const int* __restrict__ eb,
                                                             double* __restrict__ C, int64_t ldc) {
  constexpr int WN = 2, NT = 128 * WN, BN = 32 * WN;
  constexpr int ROWB = SWZ ? 64 : 80;
  constexpr int NCA = OZ_BM * 4 / NT, NCB = BN * 4 / NT;  // 16-byte chunks per thread
  __shared__ __attribute__((aligned(16))) signed char sA[S][OZ_BM * ROWB];
  __shared__ __attribute__((aligned(16))) signed char sB[S][BN * ROWB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 32;
  int ti, tj;
  if (!tile_of(blockIdx.x, mp / OZ_BM, (ncols + BN - 1) / BN, ti, tj)) return;
  const int bm = ti * OZ_BM, bn = col0 + tj * BN;
  const size_t slabA = (size_t)mp * Kp, slabB = (size_t)np * Kp;

  v4i acc[S][4][2];
#pragma unroll
  for (int l = 0; l < S; ++l)
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[l][a][b] = v4i(0);

  // register prefetch of one K step (16-byte chunks: row = chunk / 4, k = 16 (chunk % 4))
  v4i pa[PREFETCH ? S : 1][NCA], pb[PREFETCH ? S : 1][NCB];
  auto gload = [&](int k0) {
    if (!PREFETCH) return;
#pragma unroll
    for (int q = 0; q < S; ++q) {
#pragma unroll
      for (int u = 0; u < NCA; ++u) {
        const int c = tid + u * NT;
        pa[PREFETCH ? q : 0][u] = *reinterpret_cast<const v4i*>(As + q * slabA + (size_t)(bm + (c >> 2)) * Kp + k0 +
                                                 (c & 3) * 16);
      }
#pragma unroll
      for (int u = 0; u < NCB; ++u) {
        const int c = tid + u * NT;
        pb[PREFETCH ? q : 0][u] = *reinterpret_cast<const v4i*>(Bs + q * slabB + (size_t)(bn + (c >> 2)) * Kp + k0 +
                                                 (c & 3) * 16);
      }
    }
  };
  auto sstore = [&](int k0) {
#pragma unroll
    for (int q = 0; q < S; ++q) {
#pragma unroll
      for (int u = 0; u < NCA; ++u) {
        const int c = tid + u * NT;
        *reinterpret_cast<v4i*>(&sA[q][lds_off<SWZ>(c >> 2, c & 3)]) =
            PREFETCH ? pa[PREFETCH ? q : 0][u]
                     : *reinterpret_cast<const v4i*>(As + q * slabA + (size_t)(bm + (c >> 2)) * Kp +
                                                     k0 + (c & 3) * 16);
      }
#pragma unroll
      for (int u = 0; u < NCB; ++u) {
        const int c = tid + u * NT;
        *reinterpret_cast<v4i*>(&sB[q][lds_off<SWZ>(c >> 2, c & 3)]) =
            PREFETCH ? pb[PREFETCH ? q : 0][u]
                     : *reinterpret_cast<const v4i*>(Bs + q * slabB + (size_t)(bn + (c >> 2)) * Kp +
                                                     k0 + (c & 3) * 16);
      }
    }
  };
  const int fr = lane & 15, fq = lane >> 4;
  gload(0);
  for (int k0 = 0; k0 < Kp; k0 += OZ_BK) {
    sstore(k0);
    __syncthreads();
    if (k0 + OZ_BK < Kp) gload(k0 + OZ_BK);
    // all B-slice fragments of this K step stay in registers across the i loop
    v4i b[S][2];
#pragma unroll
    for (int j = 0; j < S; ++j)
#pragma unroll
      for (int tb = 0; tb < 2; ++tb)
        b[j][tb] = *reinterpret_cast<const v4i*>(&sB[j][lds_off<SWZ>(wn + 16 * tb + fr, fq)]);
#pragma unroll
    for (int i = 0; i < S; ++i) {
      v4i a[4];
#pragma unroll
      for (int ta = 0; ta < 4; ++ta)
        a[ta] = *reinterpret_cast<const v4i*>(&sA[i][lds_off<SWZ>(wm + 16 * ta + fr, fq)]);
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

constexpr int OZ_PAD = 128;  // row and column padding of the slice arrays

size_t ozaki_workspace_bytes(int m, int n, int K, int S) {
  const int mp = (m + OZ_PAD - 1) / OZ_PAD * OZ_PAD;
  const int np = (n + OZ_PAD - 1) / OZ_PAD * OZ_PAD;
  const int Kp = (K + OZ_BK - 1) / OZ_BK * OZ_BK;
  return (size_t)S * Kp * ((size_t)mp + np) + sizeof(int) * ((size_t)mp + np) + 256;
}

struct OzLayout {
  int mp, np, Kp;
  signed char *As, *Bs;
  int *ea, *eb;
};
static OzLayout oz_layout(int S, int m, int n, int K, void* work) {
  OzLayout L;
  L.mp = (m + OZ_PAD - 1) / OZ_PAD * OZ_PAD;
  L.np = (n + OZ_PAD - 1) / OZ_PAD * OZ_PAD;
  L.Kp = (K + OZ_BK - 1) / OZ_BK * OZ_BK;
  L.As = static_cast<signed char*>(work);
  L.Bs = L.As + (size_t)S * L.mp * L.Kp;
  L.ea = reinterpret_cast<int*>(L.Bs + (size_t)S * L.np * L.Kp);
  L.eb = L.ea + L.mp;
  return L;
}

template <int S>
static int ozaki_slice_s(hipStream_t s, int m, int n, int K, const double* A, int64_t lda,
                         const double* B, int64_t ldb, void* work) {
  const OzLayout L = oz_layout(S, m, n, K, work);
  slice_rows_kernel<S><<<L.mp / 64, 256, 0, s>>>(A, lda, m, K, L.mp, L.Kp, L.As, L.ea);
  slice_cols_kernel<S><<<L.np / 4, 256, 0, s>>>(B, ldb, n, K, L.np, L.Kp, L.Bs, L.eb);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

template <int S>
static int ozaki_apply_s(hipStream_t s, int m, int n, int K, int col0, int ncols, double* C,
                         int64_t ldc, void* work) {
  const OzLayout L = oz_layout(S, m, n, K, work);
  const int grid = tiles_launch(L.mp / OZ_BM, (ncols + 63) / 64);
  const int pf = tuning_get(TUNE_OZAKI_PREFETCH), swz = tuning_get(TUNE_OZAKI_SWIZZLE);
#define OZ_LAUNCH(PF, SW)                                                                  \
  ozaki_gemm_kernel<S, PF, SW><<<grid, 256, 0, s>>>(m, n, L.mp, L.np, L.Kp, col0, ncols, L.As, \
                                                    L.Bs, L.ea, L.eb, C, ldc)
  if (pf && swz) OZ_LAUNCH(true, true);
  else if (pf) OZ_LAUNCH(true, false);
  else if (swz) OZ_LAUNCH(false, true);
  else OZ_LAUNCH(false, false);
#undef OZ_LAUNCH
  MPDOK_HIP(hipGetLastError());
  return OK;
}

// Slice A (m x K) by rows and B (K x n) by columns into the workspace.
int ozaki_slice(hipStream_t s, int S, int m, int n, int K, const double* A, int64_t lda,
                const double* B, int64_t ldb, void* work) {
  if (m <= 0 || n <= 0 || K <= 0) return OK;
  switch (S) {
    case 3: return ozaki_slice_s<3>(s, m, n, K, A, lda, B, ldb, work);
    case 4: return ozaki_slice_s<4>(s, m, n, K, A, lda, B, ldb, work);
    case 5: return ozaki_slice_s<5>(s, m, n, K, A, lda, B, ldb, work);
    case 6: return ozaki_slice_s<6>(s, m, n, K, A, lda, B, ldb, work);
    case 7: return ozaki_slice_s<7>(s, m, n, K, A, lda, B, ldb, work);
    default: MPDOK_REQUIRE(false, "ozaki_slice: slices must be in [3, 7], got %d", S);
  }
}

// C[:, col0 : col0 + ncols] -= the sliced product of the last ozaki_slice call
// with the same (S, m, n, K, work); col0 a multiple of 64, C the base of the
// full m x n block.
int ozaki_apply(hipStream_t s, int S, int m, int n, int K, int col0, int ncols, double* C,
                int64_t ldc, void* work) {
  if (m <= 0 || ncols <= 0 || K <= 0) return OK;
  MPDOK_REQUIRE(col0 % 64 == 0 && col0 + ncols <= n, "ozaki_apply: bad column range");
  switch (S) {
    case 3: return ozaki_apply_s<3>(s, m, n, K, col0, ncols, C, ldc, work);
    case 4: return ozaki_apply_s<4>(s, m, n, K, col0, ncols, C, ldc, work);
    case 5: return ozaki_apply_s<5>(s, m, n, K, col0, ncols, C, ldc, work);
    case 6: return ozaki_apply_s<6>(s, m, n, K, col0, ncols, C, ldc, work);
    case 7: return ozaki_apply_s<7>(s, m, n, K, col0, ncols, C, ldc, work);
    default: MPDOK_REQUIRE(false, "ozaki_apply: slices must be in [3, 7], got %d", S);
  }
}

int ozaki_gemm(hipStream_t s, int S, int m, int n, int K, const double* A, int64_t lda,
               const double* B, int64_t ldb, double* C, int64_t ldc, void* work) {
  MPDOK_TRY(ozaki_slice(s, S, m, n, K, A, lda, B, ldb, work));
  return ozaki_apply(s, S, m, n, K, 0, n, C, ldc, work);
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
