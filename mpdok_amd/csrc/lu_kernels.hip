// Dense kernels of the blocked LU factorisation and triangular solves.
//
//  * gemm_kernel<T>: C = alpha A B + beta C, column-major, on the matrix cores
//    (v_mfma_f64_16x16x4f64 for FP64, v_mfma_f32_16x16x4f32 for FP32). 128x128
//    (or 64x64) block tile, 2x2 wavefronts, BK = 16, LDS double buffer with
//    register prefetch of the next tile. ROUND = 1/2 rounds both operands to
//    BF16/FP16 on their way into LDS (FP32 accumulation): the precision of a
//    BF16/FP16 matrix-core product, used by the precision ladder.
//  * diag_factor_kernel: LU without pivoting of one nb x nb (nb <= 64)
//    diagonal block in LDS, plus the explicit inverses of its unit-lower and
//    upper triangles (used by the panel and by the blocked triangular solves),
//    all in one 64 x 64 array.
//  * panel kernels for partial pivoting: pivot search with a last-block
//    reduction, row swap, column scaling and rank-1 update of the sub-panel.
//  * blocked triangular solves (forward unit-lower, backward upper) with 64-row
//    diagonal blocks applied through the stored inverses.
//
// Replaces cusolverDnSgetrf / cusolverDnSgetrs inside lu_ir.so
// (benchmark_mpdok.ipynb:989-992, mpdok_ops.py:72) and cp.linalg.solve's FP64
// getrf/getrs for the plain comparator (dense_krylov.py:526,608,681).
#include <type_traits>

#include "common.h"
#include "lu.h"

namespace mpdok {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Result-row of accumulator element r held by a lane, inside a 16 x 16 tile:
// the FP64 instruction interleaves the rows (row = lane/16 + 4 r), the FP32
// one stacks them (row = 4 (lane/16) + r); the column is lane % 16 for both.
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef v4d acc_t;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
  typedef v4f acc_t;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <int ROUND>
__device__ __forceinline__ float round_low(float v) {
  if (ROUND == 1) return (float)(__bf16)v;
  if (ROUND == 2) return (float)(_Float16)v;
  return v;
}
template <int ROUND>
__device__ __forceinline__ double round_low(double v) {
  return v;
}

constexpr int GBK = 16, GPAD = 16;

// Lane layout of the 16x16x4 MFMAs (one value per lane for A and B, four for
// the 16x16 result): A[i = l%16][k = l/16], B[k = l/16][j = l%16],
// D[i = Mfma<T>::row(l, r)][j = l%16], r = 0..3.
//
// TB x TB block tile (TB = 128 or 64), 2 x 2 wavefronts of (TB/2) x (TB/2).
// TB = 64 is for the skinny products of the panel and the block-row solve: it
// makes twice the workgroups per row of tiles and needs 38 KB of LDS (FP64)
// instead of 72 KB, so it fits next to a trailing update's workgroup on a CU.
//
// C may alias A when n <= TB (each workgroup then reads only the rows of A it
// writes) or alias B when m <= TB (only its own columns): every read of the
// operand happens in the K loop, before the workgroup's stores, and beta = 0
// leaves C unread. The panel uses this for its in-place products with the
// diagonal-block inverses.
//
// gemm_tile: the tile of C that workgroup b of a launch computes (the body of
// both kernels below).
// gemm_tile_at: the same for tile (ti, tj) given directly (grouped launches).
template <typename T, int ROUND, int TB>
__device__ __forceinline__ void gemm_tile_at(int ti, int tj, int m, int n, int k, T alpha,
                                             const T* A, int64_t lda, const T* B, int64_t ldb,
                                             T beta, T* C, int64_t ldc) {
  constexpr int KPH = 256 / TB;      // k phases of the A loads
  constexpr int NA = GBK / KPH;      // A elements per thread
  constexpr int NBL = TB / 16;       // B elements per thread
  constexpr int QT = TB / 32;        // 16 x 16 MFMA tiles per wavefront side
  __shared__ T sA[2][GBK][TB + GPAD];
  // B tile stored k-contiguous per column (+1 padding): the 16 lanes that
  // store 16 consecutive k of one column, and the 16 lanes that read one k of
  // 16 consecutive columns, both fall into distinct banks.
  __shared__ T sB[2][TB][GBK + 1];
  typedef typename Mfma<T>::acc_t acc_t;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * (TB / 2), wn = (wave >> 1) * (TB / 2);
  const int bm = ti * TB, bn = tj * TB;
  acc_t acc[QT][QT];
#pragma unroll
  for (int a = 0; a < QT; ++a)
#pragma unroll
    for (int b = 0; b < QT; ++b) acc[a][b] = acc_t(0);

  T ra[NA], rb[NBL];
  const int ai = bm + (tid % TB), ak = tid / TB;         // A: row, k phase
  const int bk = tid & 15, bj = bn + (tid >> 4);         // B: k, column
  const bool mn_inside = bm + TB <= m && bn + TB <= n;
  const T* pa = A + ai + (int64_t)ak * lda;
  const T* pb = B + bk + (int64_t)bj * ldb;
  auto gload = [&](int k0) {
    if (mn_inside && k0 + GBK <= k) {  // interior tile: no bounds checks
#pragma unroll
      for (int r = 0; r < NA; ++r) ra[r] = pa[(int64_t)(k0 + KPH * r) * lda];
#pragma unroll
      for (int r = 0; r < NBL; ++r) rb[r] = pb[k0 + (int64_t)(16 * r) * ldb];
      return;
    }
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      const int kk = k0 + ak + KPH * r;
      ra[r] = (ai < m && kk < k) ? A[ai + (int64_t)kk * lda] : T(0);
    }
#pragma unroll
    for (int r = 0; r < NBL; ++r) {
      const int kk = k0 + bk, j = bj + 16 * r;
      rb[r] = (kk < k && j < n) ? B[kk + (int64_t)j * ldb] : T(0);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int r = 0; r < NA; ++r) sA[buf][ak + KPH * r][tid % TB] = round_low<ROUND>(ra[r]);
#pragma unroll
    for (int r = 0; r < NBL; ++r) sB[buf][(tid >> 4) + 16 * r][bk] = round_low<ROUND>(rb[r]);
  };

  const int nk = (k + GBK - 1) / GBK;
  gload(0);
  sstore(0);
  __syncthreads();
  for (int t = 0; t < nk; ++t) {
    const int buf = t & 1;
    if (t + 1 < nk) gload((t + 1) * GBK);
#pragma unroll
    for (int kk = 0; kk < GBK; kk += 4) {
      T a[QT], b[QT];
#pragma unroll
      for (int q = 0; q < QT; ++q) {
        a[q] = sA[buf][kk + (lane >> 4)][wm + 16 * q + (lane & 15)];
        b[q] = sB[buf][wn + 16 * q + (lane & 15)][kk + (lane >> 4)];
      }
#pragma unroll
      for (int qa = 0; qa < QT; ++qa)
#pragma unroll
        for (int qb = 0; qb < QT; ++qb) acc[qa][qb] = Mfma<T>::mma(a[qa], b[qb], acc[qa][qb]);
    }
    if (t + 1 < nk) sstore(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int qa = 0; qa < QT; ++qa)
#pragma unroll
    for (int qb = 0; qb < QT; ++qb) {
      const int j = bn + wn + 16 * qb + (lane & 15);
      if (j >= n) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = bm + wm + 16 * qa + Mfma<T>::row(lane, r);
        if (i < m) {
          T* c = C + i + (int64_t)j * ldc;
          T v = alpha * acc[qa][qb][r];
          if (beta != T(0)) v = __builtin_fma(beta, *c, v);
          *c = v;
        }
      }
    }
}

template <typename T, int ROUND, int TB>
__device__ __forceinline__ void gemm_tile(int b, int m, int n, int k, T alpha, const T* A,
                                          int64_t lda, const T* B, int64_t ldb, T beta, T* C,
                                          int64_t ldc) {
  int ti, tj;
  if (!tile_of(b, (m + TB - 1) / TB, (n + TB - 1) / TB, ti, tj)) return;
  gemm_tile_at<T, ROUND, TB>(ti, tj, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// Batched use: blockIdx.y selects the problem (operands at strides sA_, sB_,
// sC_); blockIdx.x the tile.
template <typename T, int ROUND, int TB>
__global__ __launch_bounds__(256, 2) void gemm_kernel(int m, int n, int k, T alpha, const T* A,
                                                   int64_t lda, const T* B, int64_t ldb, T beta,
                                                   T* C, int64_t ldc, int64_t sA_, int64_t sB_,
                                                   int64_t sC_) {
  gemm_tile<T, ROUND, TB>(blockIdx.x, m, n, k, alpha, A + blockIdx.y * sA_, lda,
                          B + blockIdx.y * sB_, ldb, beta, C + blockIdx.y * sC_, ldc);
}

// Two unrelated products in one launch (blockIdx.y picks p or q), 64 x 64
// tiles: the panel's L21 = A21 inv(U11) and U12 = inv(L11) A12 of a
// sub-panel, which both wait for its diagonal block and for nothing else.
template <typename T>
__global__ __launch_bounds__(256, 2) void gemm_pair_kernel(GemmArgs<T> p, GemmArgs<T> q) {
  const GemmArgs<T> g = blockIdx.y == 0 ? p : q;
  gemm_tile<T, 0, 64>(blockIdx.x, g.m, g.n, g.k, g.alpha, g.A, g.lda, g.B, g.ldb, g.beta, g.C,
                      g.ldc);
}

template <typename T, int TB>
static void gemm_launch(hipStream_t s, dim3 grid, int round, int m, int n, int k, T alpha,
                        const T* A, int64_t lda, int64_t sA, const T* B, int64_t ldb, int64_t sB,
                        T beta, T* C, int64_t ldc, int64_t sC) {
  if (round == 0)
    gemm_kernel<T, 0, TB><<<grid, 256, 0, s>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, sA, sB, sC);
  else if (round == 1)
    gemm_kernel<T, 1, TB><<<grid, 256, 0, s>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, sA, sB, sC);
  else
    gemm_kernel<T, 2, TB><<<grid, 256, 0, s>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, sA, sB, sC);
}

// ---------------------------------------------------------------------------
// Low-precision operands on the matrix cores: C = alpha A B + beta C for FP32
// column-major operands whose values are rounded on their way into LDS to
// ROUND 1 = BF16, 2 = FP16 (v_mfma_f32_32x32x16_bf16 / _f16, FP32
// accumulation), or 3 = FP8 E4M3 after a power-of-two scale per operand
// (scales[0] for A, scales[1] for B; one scale per panel), multiplied on the
// BF16 matrix cores: an E4M3 value is exact in BF16, so the products are those
// of FP8 matrix cores, which would run them at twice the rate. The result is
// scaled back by 1 / (scales[0] scales[1]) (exact, powers of two).
//
// 128 x 128 block tile, 2 x 2 wavefronts of 64 x 64 (four 32 x 32 MFMA tiles
// each), K step 32, LDS double buffer. The MFMA takes the B tile first, so
// C's rows run along the lanes: lane l, register r of MFMA tile (ta, tb)
// holds row bm + wm + 32 ta + l % 32, column bn + wn + 32 tb + 8 (r / 4) +
// 4 (l / 32) + r % 4, and the read-modify-write of C is coalesced. Both
// operands give lane l the same eight k of its half (k group l / 32), which is
// all the product needs.
// ---------------------------------------------------------------------------
typedef __bf16 v8bf16 __attribute__((ext_vector_type(8)));
typedef _Float16 v8f16 __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

// Nearest E4M3 value (3 fraction bits, smallest normal 2^-6, subnormal step
// 2^-9, largest 448; larger magnitudes saturate), ties to even.
__device__ __forceinline__ float round_e4m3(float v) {
  const float a = fabsf(v);
  if (!(a > 0.f)) return v;
  int e;
  frexpf(a, &e);                       // a = f 2^e, f in [0.5, 1)
  const int q = e - 4 > -9 ? e - 4 : -9;  // quantum 2^(e-1-3), or the subnormal step
  float r = ldexpf(rintf(ldexpf(a, -q)), q);
  if (r > 448.f) r = 448.f;
  return copysignf(r, v);
}

template <int ROUND>
__device__ __forceinline__ auto to_low(float v, float sc) {
  if constexpr (ROUND == 2) return (_Float16)v;
  else if constexpr (ROUND == 3) return (__bf16)round_e4m3(v * sc);
  else return (__bf16)v;
}

template <int ROUND>
__global__ __launch_bounds__(256, 2) void gemm_lowp_kernel(int m, int n, int k, float alpha,
                                                           const float* __restrict__ A,
                                                           int64_t lda,
                                                           const float* __restrict__ B,
                                                           int64_t ldb, float beta,
                                                           float* __restrict__ C, int64_t ldc,
                                                           const float* __restrict__ scales) {
  constexpr int TB = 128, BK = 32, AS = BK + 8;  // AS: row stride in elements (80 bytes)
  typedef typename std::conditional<ROUND == 2, _Float16, __bf16>::type E;
  typedef typename std::conditional<ROUND == 2, v8f16, v8bf16>::type V8;
  __shared__ __attribute__((aligned(16))) E sA[2][TB][AS];
  __shared__ __attribute__((aligned(16))) E sB[2][TB][AS];
  int ti, tj;
  if (!tile_of(blockIdx.x, (m + TB - 1) / TB, (n + TB - 1) / TB, ti, tj)) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const int bm = ti * TB, bn = tj * TB;
  const float sa = ROUND == 3 ? scales[0] : 1.f, sb = ROUND == 3 ? scales[1] : 1.f;
  v16f acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = v16f(0);
  // A tile: thread t loads row t % 128 at k = 16 (t / 128) .. + 15 (each load
  // instruction: 64 consecutive rows of one column); B tile: column t / 2 at
  // k = 16 (t % 2) .. + 15 (contiguous). Both go to LDS as 16 consecutive k
  // of one row / column: two 16-byte stores.
  const int ar = tid & 127, akh = 16 * (tid >> 7);
  const int bc = tid >> 1, bkh = 16 * (tid & 1);
  float ra[16], rb[16];
  const bool inside = bm + TB <= m && bn + TB <= n;
  auto gload = [&](int k0) {
    const int i = bm + ar, j = bn + bc;
    if (inside && k0 + BK <= k) {
#pragma unroll
      for (int e = 0; e < 16; ++e) ra[e] = A[i + (int64_t)(k0 + akh + e) * lda];
#pragma unroll
      for (int e = 0; e < 16; ++e) rb[e] = B[(k0 + bkh + e) + (int64_t)j * ldb];
      return;
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int kk = k0 + akh + e;
      ra[e] = (i < m && kk < k) ? A[i + (int64_t)kk * lda] : 0.f;
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int kk = k0 + bkh + e;
      rb[e] = (j < n && kk < k) ? B[kk + (int64_t)j * ldb] : 0.f;
    }
  };
  auto sstore = [&](int buf) {
    V8 a0, a1, b0, b1;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      a0[e] = to_low<ROUND>(ra[e], sa);
      a1[e] = to_low<ROUND>(ra[8 + e], sa);
      b0[e] = to_low<ROUND>(rb[e], sb);
      b1[e] = to_low<ROUND>(rb[8 + e], sb);
    }
    *reinterpret_cast<V8*>(&sA[buf][ar][akh]) = a0;
    *reinterpret_cast<V8*>(&sA[buf][ar][akh + 8]) = a1;
    *reinterpret_cast<V8*>(&sB[buf][bc][bkh]) = b0;
    *reinterpret_cast<V8*>(&sB[buf][bc][bkh + 8]) = b1;
  };
  const int nk = (k + BK - 1) / BK;
  gload(0);
  sstore(0);
  __syncthreads();
  const int r32 = lane & 31, u = lane >> 5;
  for (int t = 0; t < nk; ++t) {
    const int buf = t & 1;
    if (t + 1 < nk) gload((t + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 16) {
      V8 fa[2], fb[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        fa[q] = *reinterpret_cast<const V8*>(&sA[buf][wm + 32 * q + r32][ks + 8 * u]);
        fb[q] = *reinterpret_cast<const V8*>(&sB[buf][wn + 32 * q + r32][ks + 8 * u]);
      }
#pragma unroll
      for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb) {
          if constexpr (ROUND == 2)
            acc[ta][tb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fb[tb], fa[ta], acc[ta][tb], 0, 0, 0);
          else
            acc[ta][tb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fb[tb], fa[ta], acc[ta][tb], 0, 0, 0);
        }
    }
    if (t + 1 < nk) sstore(buf ^ 1);
    __syncthreads();
  }
  const float back = alpha / (sa * sb);
#pragma unroll
  for (int ta = 0; ta < 2; ++ta) {
    const int i = bm + wm + 32 * ta + r32;
    if (i >= m) continue;
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int j = bn + wn + 32 * tb + 8 * (r >> 2) + 4 * u + (r & 3);
        if (j >= n) continue;
        float* c = C + i + (int64_t)j * ldc;
        float v = back * acc[ta][tb][r];
        if (beta != 0.f) v = __builtin_fma(beta, *c, v);
        *c = v;
      }
  }
}

// max |.| of an m x n column-major FP32 block into a power-of-two scale 2^e
// with max 2^e <= 448 (the E4M3 range; 1 for a zero block): partial maxima of
// E4M3_PARTS workgroups, then one workgroup takes the maximum and the scale.
constexpr int E4M3_PARTS = 256;

__global__ __launch_bounds__(256) void e4m3_max_part_kernel(const float* __restrict__ X,
                                                            int64_t ldx, int m, int n,
                                                            float* __restrict__ part) {
  float mx = 0.f;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < (int64_t)m * n;
       e += (int64_t)gridDim.x * 256)
    mx = fmaxf(mx, fabsf(X[e % m + (e / m) * ldx]));
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

__global__ __launch_bounds__(256) void e4m3_scale_kernel(const float* __restrict__ part, int np,
                                                         float* __restrict__ out) {
  float mx = 0.f;
  for (int i = threadIdx.x; i < np; i += 256) mx = fmaxf(mx, part[i]);
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    float sc = 1.f;
    if (mx > 0.f) {
      int e;
      frexpf(448.f / mx, &e);  // 448 / mx = f 2^e, f in [0.5, 1): 2^(e-1) <= 448 / mx
      sc = ldexpf(1.f, e - 1);
    }
    out[0] = sc;
  }
}

size_t gemm_lowp_scale_floats() { return 2 + 2 * E4M3_PARTS; }

int gemm_lowp(hipStream_t s, int m, int n, int k, float alpha, const float* A, int64_t lda,
              const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int round,
              float* scales) {
  if (m <= 0 || n <= 0) return OK;
  MPDOK_REQUIRE(round >= 1 && round <= 3, "gemm_lowp: round %d", round);
  MPDOK_REQUIRE(round != 3 || scales != nullptr, "gemm_lowp: FP8 needs a scale buffer");
  if (round == 3) {
    // scales: [0] A's, [1] B's, then the partial maxima of each
    float* pa = scales + 2;
    float* pb = pa + E4M3_PARTS;
    e4m3_max_part_kernel<<<E4M3_PARTS, 256, 0, s>>>(A, lda, m, k, pa);
    e4m3_max_part_kernel<<<E4M3_PARTS, 256, 0, s>>>(B, ldb, k, n, pb);
    e4m3_scale_kernel<<<1, 256, 0, s>>>(pa, E4M3_PARTS, scales);
    e4m3_scale_kernel<<<1, 256, 0, s>>>(pb, E4M3_PARTS, scales + 1);
  }
  const dim3 grid(tiles_launch((m + 127) / 128, (n + 127) / 128));
  if (round == 1)
    gemm_lowp_kernel<1><<<grid, 256, 0, s>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, scales);
  else if (round == 2)
    gemm_lowp_kernel<2><<<grid, 256, 0, s>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, scales);
  else
    gemm_lowp_kernel<3><<<grid, 256, 0, s>>>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, scales);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

// Tile choice: 64 x 64 tiles for the skinny products of the panel and of the
// block-row solve (a side or the inner dimension at most 64), or when
// 128 x 128 tiles would not give every CU two workgroups; 128 x 128 otherwise.
template <typename T>
int gemm_batched(hipStream_t s, int m, int n, int k, T alpha, const T* A, int64_t lda,
                 int64_t sA, const T* B, int64_t ldb, int64_t sB, T beta, T* C, int64_t ldc,
                 int64_t sC, int batch, int round) {
  if (m <= 0 || n <= 0 || batch <= 0) return OK;
  const int64_t t128 = (int64_t)((m + 127) / 128) * ((n + 127) / 128);
  const bool small = m <= 64 || n <= 64 || k <= 64 || t128 * batch < 512;
  const int TB = small ? 64 : 128;
  const int64_t tiles = (int64_t)((m + TB - 1) / TB) * ((n + TB - 1) / TB);
  MPDOK_REQUIRE(batch <= 65535 && tiles < (1 << 30), "gemm: grid too large (m=%d n=%d batch=%d)",
                m, n, batch);
  dim3 grid(tiles_launch((m + TB - 1) / TB, (n + TB - 1) / TB), batch);
  if (small)
    gemm_launch<T, 64>(s, grid, round, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC);
  else
    gemm_launch<T, 128>(s, grid, round, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

template <typename T>
int gemm(hipStream_t s, int m, int n, int k, T alpha, const T* A, int64_t lda, const T* B,
         int64_t ldb, T beta, T* C, int64_t ldc, int round) {
  return gemm_batched<T>(s, m, n, k, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, 1, round);
}
template int gemm<double>(hipStream_t, int, int, int, double, const double*, int64_t,
                          const double*, int64_t, double, double*, int64_t, int);

template <typename T>
int gemm_pair(hipStream_t s, const GemmArgs<T>& p, const GemmArgs<T>& q) {
  const bool ep = p.m <= 0 || p.n <= 0, eq = q.m <= 0 || q.n <= 0;
  if (ep || eq) {
    const GemmArgs<T>& g = ep ? q : p;
    if (ep && eq) return OK;
    return gemm<T>(s, g.m, g.n, g.k, g.alpha, g.A, g.lda, g.B, g.ldb, g.beta, g.C, g.ldc, 0);
  }
  const int tp = tiles_launch((p.m + 63) / 64, (p.n + 63) / 64);
  const int tq = tiles_launch((q.m + 63) / 64, (q.n + 63) / 64);
  gemm_pair_kernel<T><<<dim3(tp > tq ? tp : tq, 2), 256, 0, s>>>(p, q);
  MPDOK_HIP(hipGetLastError());
  return OK;
}
template int gemm_pair<double>(hipStream_t, const GemmArgs<double>&, const GemmArgs<double>&);
template int gemm_pair<float>(hipStream_t, const GemmArgs<float>&, const GemmArgs<float>&);
template int gemm<float>(hipStream_t, int, int, int, float, const float*, int64_t, const float*,
                         int64_t, float, float*, int64_t, int);
template int gemm_batched<double>(hipStream_t, int, int, int, double, const double*, int64_t,
                                  int64_t, const double*, int64_t, int64_t, double, double*,
                                  int64_t, int64_t, int, int);
template int gemm_batched<float>(hipStream_t, int, int, int, float, const float*, int64_t, int64_t,
                                 const float*, int64_t, int64_t, float, float*, int64_t, int64_t,
                                 int, int);

// ---------------------------------------------------------------------------
// Diagonal block: LU without pivoting, then inv(unit L) and inv(U), in place
// in one 64 x 64 LDS array (34 KB with the two staging columns). The small
// footprint matters: this kernel runs on the lookahead stream next to the
// trailing update, and it can take a CU as soon as one workgroup of the
// update leaves it instead of waiting for the whole CU's LDS.
//
// Elimination: thread t owns row r = t % 64 and the columns c = t / 64
// (mod 4). At step j every thread of row r forms the multiplier
// l = a[r][j] / a[j][j] itself and updates its columns c > j; the multiplier
// is stored into a[r][j] one step later, when no thread reads column j any
// more, so one barrier per step suffices. The arithmetic (one division per
// multiplier, one fused multiply-add per update) is the textbook order.
//
// Inversion (LAPACK's unblocked in-place algorithm): U column by column,
// ascending, u_jj <- 1 / u_jj, u_ij <- -u_jj * sum_{k=i}^{j-1} inv(U)_ik u_kj;
// unit-lower L column by column, descending, l_ij <- -(l_ij + sum_{k=j+1}^{i-1}
// inv(L)_ik l_kj). Threads 0..127 do U, 128..255 L, side by side, two lanes
// per row; each column's old values are staged first, so the in-place update
// reads no value it writes.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void diag_factor_kernel(T* __restrict__ A, int64_t lda, int nb,
                                                          T* __restrict__ Linv,
                                                          T* __restrict__ Uinv, int ldi,
                                                          int* __restrict__ info, int col0,
                                                          int factor) {
  __shared__ T s[64][65];
  __shared__ T tcol[2][65];  // staged old column: [0] U (rows <= j), [1] L (rows > j)
  const int t = threadIdx.x, r = t & 63, q = t >> 6;
  for (int c = q; c < nb; c += 4)
    if (r < nb) s[r][c] = A[r + (int64_t)c * lda];
  __syncthreads();
  if (factor) {
    T lprev = T(0);
    for (int j = 0; j < nb; ++j) {
      // multiplier of step j - 1 into column j - 1: every read of it is done
      if (j > 0 && q == 0 && r > j - 1 && r < nb) s[r][j - 1] = lprev;
      const T piv = s[j][j];
      if (t == 0 && piv == T(0)) atomicCAS(info, 0, col0 + j + 1);
      T l = T(0);
      if (r > j && r < nb) {
        l = s[r][j] / piv;
#pragma unroll 4
        for (int c = j + 1 + q; c < nb; c += 4) s[r][c] = __builtin_fma(-l, s[j][c], s[r][c]);
      }
      lprev = l;
      __syncthreads();
    }
    for (int c = q; c < nb; c += 4)
      if (r < nb) A[r + (int64_t)c * lda] = s[r][c];
    __syncthreads();
  }
  // Inversion: threads 0..127 on U, 128..255 on L; within a group, wavefront
  // w & 1 holds rows 32 (w & 1) + (lane % 32), and lanes l, l ^ 32 share a row,
  // each summing every other term (four accumulators each), then combine.
  const int tt = t & 127, half = (tt >> 5) & 1, row = (tt & 31) + 32 * (tt >> 6);
  const bool onU = t < 128;
  for (int st = 0; st < nb; ++st) {
    const int ju = st, jl = nb - 1 - st;
    if (half == 0) {
      if (onU && row <= ju) tcol[0][row] = s[row][ju];
      if (!onU && row > jl && row < nb) tcol[1][row] = s[row][jl];
    }
    __syncthreads();
    if (onU) {
      if (row <= ju) {  // both lanes of a row take the same branch
        T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
        int k = row + half;
        for (; k + 6 < ju; k += 8) {
          a0 = __builtin_fma(s[row][k], tcol[0][k], a0);
          a1 = __builtin_fma(s[row][k + 2], tcol[0][k + 2], a1);
          a2 = __builtin_fma(s[row][k + 4], tcol[0][k + 4], a2);
          a3 = __builtin_fma(s[row][k + 6], tcol[0][k + 6], a3);
        }
        for (; k < ju; k += 2) a0 = __builtin_fma(s[row][k], tcol[0][k], a0);
        T acc = (a0 + a1) + (a2 + a3);
        acc += __shfl_xor(acc, 32, 64);
        const T d = T(1) / tcol[0][ju];
        if (half == 0) s[row][ju] = row == ju ? d : -d * acc;
      }
    } else if (row > jl && row < nb) {
      T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
      int k = jl + 1 + half;
      for (; k + 6 < row; k += 8) {
        a0 = __builtin_fma(s[row][k], tcol[1][k], a0);
        a1 = __builtin_fma(s[row][k + 2], tcol[1][k + 2], a1);
        a2 = __builtin_fma(s[row][k + 4], tcol[1][k + 4], a2);
        a3 = __builtin_fma(s[row][k + 6], tcol[1][k + 6], a3);
      }
      for (; k < row; k += 2) a0 = __builtin_fma(s[row][k], tcol[1][k], a0);
      T acc = (a0 + a1) + (a2 + a3);
      acc += __shfl_xor(acc, 32, 64);
      if (half == 0) s[row][jl] = -(tcol[1][row] + acc);
    }
    __syncthreads();
  }
  for (int c = q; c < nb; c += 4) {
    if (r < nb) {
      Uinv[r + c * ldi] = r <= c ? s[r][c] : T(0);
      Linv[r + c * ldi] = r > c ? s[r][c] : (r == c ? T(1) : T(0));
    }
  }
}

// Broadcast of lane `src`'s value to the whole wavefront (through SGPRs).
__device__ __forceinline__ double lane_bcast(double x, int src) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, src);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float lane_bcast(float x, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), src));
}

// Blocked variant (the default), same contract and the same 34 KB of LDS:
//  1. LU in four 16-column sub-panels p: wavefront 0 eliminates the
//     sub-panel with lane r holding row r's 16 entries in registers and the
//     pivot row broadcast from its lane (no LDS round trips inside the
//     elimination; the operations and their order are those of the
//     unblocked kernel's elimination); the U12 rows of the sub-panel are
//     solved against its unit-lower 16 x 16 triangle (one thread per
//     column); the rest of the block takes a rank-16 update. Three
//     workgroup barriers per sub-panel.
//  2. Inverses in place: the four 16 x 16 diagonal blocks of U (wavefront 0)
//     and of the unit-lower L (wavefront 1), one lane per column, by
//     substitution in registers; then two levels of recursive doubling,
//       inv([A B; 0 D]) = [inv(A) -inv(A) B inv(D); 0 inv(D)]
//       inv([A 0; C D]) = [inv(A) 0; -inv(D) C inv(A) inv(D)],
//     16 -> 32 -> 64, each off-diagonal block formed as two products on the
//     matrix cores (16 x 16 output tiles, triangular operands masked on
//     their way out of LDS), staged in registers and written back in place.
// The rounding differs from the unblocked kernel's (other summation orders).
// Sub-panel P of the blocked diagonal kernel after its elimination: U12 of
// rows c0 .. c0 + 15 over columns c0 + 16 .. 63 (one column per thread,
// substituted in registers, k ascending per row), then the rank-16 update of
// rows and columns c0 + 16 .. 63 (thread (tr, tc) takes rows
// c0 + 16 + tr + 16 a and columns c0 + 16 + tc + 16 b, a, b < NB, so each
// operand read from LDS serves NB products; per element the even and odd k
// run in two chains, added at the end).
template <int P, typename T>
__device__ __forceinline__ void diag_blk_trailing(T (*s)[65], int t) {
  constexpr int c0 = 16 * P, ncol = 48 - c0, NB = ncol / 16;
  if (t < ncol) {
    const int c = c0 + 16 + t;
    T u[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) u[i] = s[c0 + i][c];
#pragma unroll
    for (int i = 1; i < 16; ++i) {
      T acc = u[i];
#pragma unroll
      for (int k = 0; k < i; ++k) acc = __builtin_fma(-s[c0 + i][c0 + k], u[k], acc);
      u[i] = acc;
    }
#pragma unroll
    for (int i = 1; i < 16; ++i) s[c0 + i][c] = u[i];
  }
  __syncthreads();
  {
    const int tr = t & 15, tc = t >> 4;
    T lr[NB][16], uc[NB][16];
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int k = 0; k < 16; ++k) lr[a][k] = s[c0 + 16 + tr + 16 * a][c0 + k];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int k = 0; k < 16; ++k) uc[b][k] = s[c0 + k][c0 + 16 + tc + 16 * b];
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int r = c0 + 16 + tr + 16 * a, c = c0 + 16 + tc + 16 * b;
        T a0 = s[r][c], a1 = T(0);
#pragma unroll
        for (int k = 0; k < 16; k += 2) {
          a0 = __builtin_fma(-lr[a][k], uc[b][k], a0);
          a1 = __builtin_fma(-lr[a][k + 1], uc[b][k + 1], a1);
        }
        s[r][c] = a0 + a1;
      }
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void diag_factor_blk_body(T* __restrict__ A, int64_t lda, int nb,
                                                     T* __restrict__ Linv, T* __restrict__ Uinv,
                                                     int64_t ldi, int* __restrict__ info,
                                                     int col0, int factor) {
  __shared__ T s[64][65];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  auto wave_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  // identity outside nb x nb, so the blocked steps need no bounds; the 16
  // loads of a thread go out together
  {
    T v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int e = t + 256 * q, r = e & 63, c = e >> 6;
      v[q] = (r < nb && c < nb) ? A[r + (int64_t)c * lda] : (r == c ? T(1) : T(0));
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int e = t + 256 * q;
      s[e & 63][e >> 6] = v[q];
    }
  }
  __syncthreads();
  if (factor) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int c0 = 16 * p;
      if (w == 0) {
        const int r = lane;
        T a[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) a[c] = s[r][c0 + c];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          T u[16];  // pivot row c0 + j, columns j .. 15 of the sub-panel
#pragma unroll
          for (int c = j; c < 16; ++c) u[c] = lane_bcast(a[c], c0 + j);
          if (r == 0 && c0 + j < nb && u[j] == T(0)) atomicCAS(info, 0, col0 + c0 + j + 1);
          if (r > c0 + j) {
            const T l = a[j] / u[j];
            a[j] = l;
#pragma unroll
            for (int c = j + 1; c < 16; ++c) a[c] = __builtin_fma(-l, u[c], a[c]);
          }
        }
#pragma unroll
        for (int c = 0; c < 16; ++c) s[r][c0 + c] = a[c];
      }
      __syncthreads();
      if (p == 3) break;
      if (p == 0)
        diag_blk_trailing<0, T>(s, t);
      else if (p == 1)
        diag_blk_trailing<1, T>(s, t);
      else
        diag_blk_trailing<2, T>(s, t);
    }
    for (int e = t; e < 64 * 64; e += 256) {
      const int r = e & 63, c = e >> 6;
      if (r < nb && c < nb) A[r + (int64_t)c * lda] = s[r][c];
    }
  }
  // ---- inverses of the 16 x 16 diagonal blocks: wavefront 0 U, 1 L ----
  // One lane per column, substitution in registers, right-looking (each
  // finished x_i is added into the sums of the rows it feeds at once), so
  // the dependent chain per row is one product; U's diagonal reciprocals
  // are formed up front.
  if (w < 2) {
    const int b = lane >> 4, c = lane & 15, o = 16 * b;
    T x[16], acc[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = T(0);
    if (w == 0) {  // column c of inv(U_bb): x_c = 1/u_cc, x_i = -(sum_{k>i} u_ik x_k)/u_ii
      T d[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) d[k] = T(1) / s[o + k][o + k];
#pragma unroll
      for (int k = 0; k < 16; ++k) x[k] = k == c ? d[k] : T(0);
#pragma unroll
      for (int i = 15; i >= 0; --i) {
        if (i < c) x[i] = -acc[i] * d[i];
#pragma unroll
        for (int k = 0; k < i; ++k) acc[k] = __builtin_fma(s[o + k][o + i], x[i], acc[k]);
      }
      wave_sync();
#pragma unroll
      for (int i = 0; i < 16; ++i)
        if (i <= c) s[o + i][o + c] = x[i];
    } else {  // column c of inv(L_bb): x_c = 1, x_i = -sum_{k<i} l_ik x_k (i > c)
#pragma unroll
      for (int k = 0; k < 16; ++k) x[k] = k == c ? T(1) : T(0);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        if (i > c) x[i] = -acc[i];
#pragma unroll
        for (int k = i + 1; k < 16; ++k) acc[k] = __builtin_fma(s[o + k][o + i], x[i], acc[k]);
      }
      wave_sync();
#pragma unroll
      for (int i = 0; i < 16; ++i)
        if (i > c) s[o + i][o + c] = x[i];
    }
  }
  __syncthreads();
  // ---- recursive doubling on the matrix cores: h = 16, 32 ----
  // Work units: (triangle, pair, 16 x 16 output tile), wavefront w takes
  // units w and w + 4. Lane layout of the 16 x 16 x 4 MFMA: A[l % 16][l / 16],
  // B[l / 16][l % 16], D[Mfma<T>::row(l, r)][l % 16]. Every operand of both
  // units is read from LDS first (masked values replaced by zeros after the
  // read), then the products run back to back.
  typedef typename Mfma<T>::acc_t acc_t;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int h = 16; h < 64; h *= 2) {
    const int nt = h / 16, npair = 32 / h, per = npair * nt * nt, units = 2 * per;
#pragma unroll
    for (int phase = 0; phase < 2; ++phase) {
      // phase 0: U: P = inv(A) B             L: P = C + C inv_s(A)
      // phase 1: U: X = -(P inv(D))          L: X = -(P + inv_s(D) P)
      // (inv_s: the strictly lower part of a unit-lower inverse)
      acc_t res[2];
      T av[2][8], bv[2][8];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int u = w + 4 * q;
        res[q] = acc_t(0);
        if (u >= units) continue;
        const int tri = u / per, pr = (u / (nt * nt)) % npair, tile = u % (nt * nt);
        const int o = 2 * h * pr, ti = 16 * (tile % nt), tj = 16 * (tile / nt);
        if (tri == 1) {  // L: the added term (C, or P)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            res[q][r] = s[o + h + ti + Mfma<T>::row(lane, r)][o + tj + li];
        }
        const int ia = ti + li, jb = tj + li;  // A row, B column of this lane
#pragma unroll
        for (int ks = 0; ks < h / 4; ++ks) {
          const int k = 4 * ks + lk;
          if (tri == 0) {
            if (phase == 0) {
              const T a = s[o + ia][o + k];
              av[q][ks] = k >= ia ? a : T(0);
              bv[q][ks] = s[o + k][o + h + jb];
            } else {
              const T bb = s[o + h + k][o + h + jb];
              av[q][ks] = s[o + ia][o + h + k];
              bv[q][ks] = k <= jb ? bb : T(0);
            }
          } else {
            if (phase == 0) {
              const T bb = s[o + k][o + jb];
              av[q][ks] = s[o + h + ia][o + k];
              bv[q][ks] = k > jb ? bb : T(0);
            } else {
              const T a = s[o + h + ia][o + h + k];
              av[q][ks] = k < ia ? a : T(0);
              bv[q][ks] = s[o + h + k][o + jb];
            }
          }
        }
      }
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        if (w + 4 * q >= units) continue;
        acc_t acc = res[q];
#pragma unroll
        for (int ks = 0; ks < h / 4; ++ks) acc = Mfma<T>::mma(av[q][ks], bv[q][ks], acc);
        res[q] = phase == 0 ? acc : -acc;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int u = w + 4 * q;
        if (u >= units) continue;
        const int tri = u / per, pr = (u / (nt * nt)) % npair, tile = u % (nt * nt);
        const int o = 2 * h * pr, ti = 16 * (tile % nt), tj = 16 * (tile / nt);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = ti + Mfma<T>::row(lane, r), j = tj + li;
          if (tri == 0) s[o + i][o + h + j] = res[q][r];
          else s[o + h + i][o + j] = res[q][r];
        }
      }
      __syncthreads();
    }
  }
  for (int e = t; e < 64 * 64; e += 256) {
    const int r = e & 63, c = e >> 6;
    if (r < nb && c < nb) {
      Uinv[r + c * ldi] = r <= c ? s[r][c] : T(0);
      Linv[r + c * ldi] = r > c ? s[r][c] : (r == c ? T(1) : T(0));
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void diag_factor_blk_kernel(T* __restrict__ A, int64_t lda,
                                                              int nb, T* __restrict__ Linv,
                                                              T* __restrict__ Uinv, int ldi,
                                                              int* __restrict__ info, int col0,
                                                              int factor) {
  diag_factor_blk_body<T>(A, lda, nb, Linv, Uinv, ldi, info, col0, factor);
}

template <typename T>
int diag_factor(hipStream_t s, T* A, int64_t lda, int nb, T* Linv, T* Uinv, int ldi, int* info,
                int col0, int factor) {
  MPDOK_REQUIRE(nb >= 1 && nb <= 64, "diag_factor: nb=%d", nb);
  if (tuning_get(TUNE_DIAG_KERNEL) == 1)
    diag_factor_kernel<T><<<1, 256, 0, s>>>(A, lda, nb, Linv, Uinv, ldi, info, col0, factor);
  else
    diag_factor_blk_kernel<T><<<1, 256, 0, s>>>(A, lda, nb, Linv, Uinv, ldi, info, col0, factor);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

template int diag_factor<double>(hipStream_t, double*, int64_t, int, double*, double*, int, int*,
                                 int, int);
template int diag_factor<float>(hipStream_t, float*, int64_t, int, float*, float*, int, int*, int,
                                int);

// ---------------------------------------------------------------------------
// 2-D copy (column-major) and row permutation
// ---------------------------------------------------------------------------
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void copy2d_kernel(const TI* __restrict__ src, int64_t lds,
                                                     TO* __restrict__ dst, int64_t ldd, int m,
                                                     int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  for (int j = blockIdx.y; j < n; j += gridDim.y) dst[i + (int64_t)j * ldd] = (TO)src[i + (int64_t)j * lds];
}

// FP64 copy with 16-byte accesses for even m, even leading dimensions and
// 16-byte aligned bases: each lane moves one row pair of four columns per
// iteration, the four loads issued before the stores.
typedef double cp_d2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void copy2d_v2_kernel(const cp_d2* __restrict__ src,
                                                        int64_t lds2, cp_d2* __restrict__ dst,
                                                        int64_t ldd2, int m2, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= m2) return;
  for (int j = blockIdx.y * 4; j < n; j += gridDim.y * 4) {
    cp_d2 v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (j + c < n) v[c] = __builtin_nontemporal_load(src + i + (int64_t)(j + c) * lds2);
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (j + c < n) dst[i + (int64_t)(j + c) * ldd2] = v[c];
  }
}

template <typename TI, typename TO>
int copy2d(hipStream_t s, const TI* src, int64_t lds, TO* dst, int64_t ldd, int m, int n) {
  if (m <= 0 || n <= 0) return OK;
  if constexpr (std::is_same<TI, double>::value && std::is_same<TO, double>::value) {
    if (m % 2 == 0 && lds % 2 == 0 && ldd % 2 == 0 && (reinterpret_cast<uintptr_t>(src) & 15) == 0 &&
        (reinterpret_cast<uintptr_t>(dst) & 15) == 0) {
      const int m2 = m / 2, nq = (n + 3) / 4;
      dim3 grid((m2 + 255) / 256, nq < 4096 ? nq : 4096);
      copy2d_v2_kernel<<<grid, 256, 0, s>>>(reinterpret_cast<const cp_d2*>(src), lds / 2,
                                            reinterpret_cast<cp_d2*>(dst), ldd / 2, m2, n);
      MPDOK_HIP(hipGetLastError());
      return OK;
    }
  }
  dim3 grid((m + 255) / 256, n < 4096 ? n : 4096);
  copy2d_kernel<TI, TO><<<grid, 256, 0, s>>>(src, lds, dst, ldd, m, n);
  MPDOK_HIP(hipGetLastError());
  return OK;
}
template int copy2d<double, double>(hipStream_t, const double*, int64_t, double*, int64_t, int, int);
template int copy2d<float, float>(hipStream_t, const float*, int64_t, float*, int64_t, int, int);
template int copy2d<double, float>(hipStream_t, const double*, int64_t, float*, int64_t, int, int);
template int copy2d<float, double>(hipStream_t, const float*, int64_t, double*, int64_t, int, int);

// Apply row interchanges ipiv[r0..r1) (0-based, LAPACK order) to columns
// [c0, c1) of A: one thread per column, swaps applied in sequence.
template <typename T>
__global__ __launch_bounds__(256) void laswp_kernel(T* __restrict__ A, int64_t lda, int c0, int c1,
                                                    const int* __restrict__ ipiv, int r0, int r1) {
  const int j = c0 + blockIdx.x * 256 + threadIdx.x;
  if (j >= c1) return;
  T* col = A + (int64_t)j * lda;
  for (int r = r0; r < r1; ++r) {
    const int p = ipiv[r];
    if (p != r) {
      T a = col[r];
      col[r] = col[p];
      col[p] = a;
    }
  }
}

template <typename T>
int laswp(hipStream_t s, T* A, int64_t lda, int c0, int c1, const int* ipiv, int r0, int r1) {
  if (c1 <= c0 || r1 <= r0) return OK;
  laswp_kernel<T><<<(c1 - c0 + 255) / 256, 256, 0, s>>>(A, lda, c0, c1, ipiv, r0, r1);
  MPDOK_HIP(hipGetLastError());
  return OK;
}
template int laswp<double>(hipStream_t, double*, int64_t, int, int, const int*, int, int);
template int laswp<float>(hipStream_t, float*, int64_t, int, int, const int*, int, int);

// ---------------------------------------------------------------------------
// Partial pivoting (LAPACK getf2 on one sub-panel) in one cooperative launch.
// Sub-panel: columns [j0, j0 + w), w <= 64, rows [j0, N). Workgroup b holds
// rows j0 + 256 b + t (thread t) with their w values in registers for the
// whole sub-panel. Pivoting is implicit: rows stay where they are, a flag
// marks the rows already chosen, and every column takes one grid barrier:
//   each workgroup finds its largest |a_ic| over the unchosen rows and
//   publishes that value, its row index and the row's w values (double
//   buffered by column parity); after the barrier every workgroup picks the
//   overall maximum (ties: the smallest row index), reads the winning row as
//   the pivot row u, and its unchosen rows take l = a_ic / u_c (unscaled if
//   u_c = 0, which sets info) and a_ic' -= l u_c' for c' > c.
// The multipliers, updates and their order are getf2's, so after the rows
// are permuted (laswp over every column with the returned ipiv) the panel
// holds getf2's L and U. Workgroup 0 finally turns the sequence of chosen
// rows into LAPACK's ipiv (interchanges applied in order) with two index
// maps over the panel's rows in the workspace.
// ---------------------------------------------------------------------------
constexpr int GP_W = 64;     // widest sub-panel
constexpr int GP_HDR = 1024;  // barrier words at the start of the workspace

struct GeppWork {
  unsigned* bar;   // barrier words (gp_grid_barrier)
  double* cval;    // [2][G] largest |a| per workgroup
  int* cidx;       // [2][G] its row
  double* crow;    // [2][G][GP_W] its values
  int* rowat;      // [N] row at a position of the panel (interchange maps)
  int* posof;      // [N] position of a row
};

// Device-coherent loads and stores (relaxed atomics at device scope): they
// reach the point of coherence directly, so the exchange of the panel
// kernel needs no cache-wide write-back or invalidation (a release or an
// acquire at device scope does one of those on every call).
template <typename V>
__device__ __forceinline__ V gp_ld(const V* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename V>
__device__ __forceinline__ void gp_st(V* p, V v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier of the panel kernel, entered by thread 0 of each workgroup
// after its wavefront's device-coherent stores have completed (the caller
// waits for them and synchronises the workgroup around the barrier).
// bar[0] counts the arrivals of the whole launch (zero at its start); the
// barrier of column c is passed once it reaches target = (c + 1) G.
// bar[2] is the spin-limit flag.
__device__ __forceinline__ bool gp_grid_barrier0(unsigned* bar, unsigned target) {
  if (__hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 < target) {
    // every workgroup of a cooperative launch is resident, so the others
    // arrive; the bound only turns a broken launch into an error
    long spins = 0;
    while (gp_ld(bar) < target) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > (1L << 23)) {
        atomicExch(bar + 2, 1u);
        return false;
      }
    }
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  return true;
}

// (value, index) maximum with ties to the smaller index
__device__ __forceinline__ void gp_better(double& v, int& i, double v2, int i2) {
  if (v2 > v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void gepp_panel_kernel(T* __restrict__ A, int64_t lda, int N,
                                                         int j0, int w, int* __restrict__ ipiv,
                                                         int* __restrict__ info, GeppWork ws) {
  const int G = gridDim.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int row = j0 + blockIdx.x * 256 + t;
  const bool valid = row < N;
  T a[GP_W];
#pragma unroll
  for (int c = 0; c < GP_W; ++c) a[c] = (valid && c < w) ? A[row + (int64_t)(j0 + c) * lda] : T(0);
  bool chosen = !valid;
  __shared__ double sv[4];
  __shared__ int si[4];
  __shared__ int s_best;  // this workgroup's candidate row / the pivot row
  __shared__ double s_max;
  __shared__ double cand[GP_W];
  __shared__ int s_ok;
  __shared__ T u[GP_W];
  __shared__ int seq[GP_W];
  if (t == 0) s_ok = 1;
  for (int c = 0; c < w; ++c) {
    T ac = T(0);
#pragma unroll
    for (int q = 0; q < GP_W; ++q)
      if (q == c) ac = a[q];
    // this workgroup's candidate (a NaN counts as the largest, so a row is
    // always chosen while unchosen rows remain)
    const double fa = fabs((double)ac);
    double v = chosen ? -1.0 : (fa == fa ? fa : HUGE_VAL);
    int vi = chosen ? 0x7fffffff : row;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
      gp_better(v, vi, __shfl_xor(v, off, 64), __shfl_xor(vi, off, 64));
    if (lane == 0) {
      sv[wv] = v;
      si[wv] = vi;
    }
    __syncthreads();
    if (t == 0) {
      double bv = sv[0];
      int bi = si[0];
      for (int q = 1; q < 4; ++q) gp_better(bv, bi, sv[q], si[q]);
      s_best = bi;
      s_max = bv;
    }
    __syncthreads();
    if (row == s_best) {
#pragma unroll
      for (int q = 0; q < GP_W; ++q) cand[q] = (double)a[q];
    }
    __syncthreads();
    // wavefront 0 publishes the candidate, takes the grid barrier and reads
    // the winner back: every global access of the exchange is
    // device-coherent and stays in one wavefront, whose stores complete
    // before its arrival is counted and whose reads follow the barrier
    const int slot = c & 1;
    if (wv == 0) {
      gp_st(ws.crow + ((size_t)slot * G + blockIdx.x) * GP_W + lane, cand[lane]);
      if (lane == 0) {
        gp_st(ws.cval + slot * G + blockIdx.x, s_max);
        gp_st(ws.cidx + slot * G + blockIdx.x, s_best);
      }
      // the wavefront's stores complete before its arrival is counted
      __builtin_amdgcn_s_waitcnt(0);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      if (lane == 0 && !gp_grid_barrier0(ws.bar, (unsigned)(c + 1) * (unsigned)G)) s_ok = 0;
      if (__shfl(s_ok, 0, 64)) {
        double bv = -1.0;
        int bi = 0x7fffffff, bb = 0;
        for (int q = lane; q < G; q += 64) {
          const double v2 = gp_ld(ws.cval + slot * G + q);
          const int i2 = gp_ld(ws.cidx + slot * G + q);
          if (v2 > bv || (v2 == bv && i2 < bi)) {
            bv = v2;
            bi = i2;
            bb = q;
          }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          const double v2 = __shfl_xor(bv, off, 64);
          const int i2 = __shfl_xor(bi, off, 64), b2 = __shfl_xor(bb, off, 64);
          if (v2 > bv || (v2 == bv && i2 < bi)) {
            bv = v2;
            bi = i2;
            bb = b2;
          }
        }
        u[lane] = (T)gp_ld(ws.crow + ((size_t)slot * G + bb) * GP_W + lane);
        if (lane == 0) {
          s_best = bi;
          seq[c] = bi;
          if (bv == 0.0 && blockIdx.x == 0) atomicCAS(info, 0, j0 + c + 1);
        }
      }
    }
    __syncthreads();
    if (!s_ok) {
      if (t == 0) atomicCAS(info, 0, -1);  // reported as a failed factorisation
      return;
    }
    if (row == s_best) chosen = true;
    if (!chosen) {
      const T uc = u[c];
      const T l = uc != T(0) ? ac / uc : ac;
#pragma unroll
      for (int q = 0; q < GP_W; ++q) {
        if (q == c) a[q] = l;
        if (q > c) a[q] = __builtin_fma(-l, u[q], a[q]);
      }
    }
  }
  if (valid) {
#pragma unroll
    for (int c = 0; c < GP_W; ++c)
      if (c < w) A[row + (int64_t)(j0 + c) * lda] = a[c];
  }
  if (blockIdx.x != 0) return;
  // ipiv: bring row seq[k] to position j0 + k by one interchange per k, in
  // order (maps over the panel's rows, relative to j0)
  const int h = N - j0;
  for (int q = t; q < h; q += 256) {
    ws.rowat[q] = q;
    ws.posof[q] = q;
  }
  __syncthreads();
  if (t == 0) {
    for (int k = 0; k < w; ++k) {
      const int r = seq[k] - j0;
      if (r < 0 || r >= h) {  // cannot happen while a row is chosen at every column
        atomicCAS(info, 0, -1);
        return;
      }
      const int qpos = ws.posof[r];
      const int rk = ws.rowat[k];
      ws.rowat[qpos] = rk;
      ws.posof[rk] = qpos;
      ws.rowat[k] = r;
      ws.posof[r] = k;
      ipiv[j0 + k] = j0 + qpos;
    }
  }
}

template <typename T>
int panel_gepp(hipStream_t s, T* A, int64_t lda, int N, int j0, int w, int* ipiv, int* info,
               void* work) {
  MPDOK_REQUIRE(w >= 1 && w <= GP_W && j0 + w <= N, "panel_gepp: bad sub-panel");
  const int G = (N - j0 + 255) / 256;
  char* p = static_cast<char*>(work);
  GeppWork ws;
  ws.bar = reinterpret_cast<unsigned*>(p);
  const int gmax = (N + 255) / 256;
  ws.cval = reinterpret_cast<double*>(p + GP_HDR);
  ws.cidx = reinterpret_cast<int*>(ws.cval + 2 * gmax);
  ws.crow = reinterpret_cast<double*>(p + GP_HDR + 2 * gmax * 16);
  ws.rowat = reinterpret_cast<int*>(ws.crow + (size_t)2 * gmax * GP_W);
  ws.posof = ws.rowat + N;
  MPDOK_HIP(hipMemsetAsync(ws.bar, 0, GP_HDR, s));
  void* args[] = {&A, &lda, &N, &j0, &w, &ipiv, &info, &ws};
  MPDOK_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&gepp_panel_kernel<T>),
                                       dim3(G), dim3(256), args, 0, s));
  return OK;
}
size_t panel_gepp_workspace_bytes(int N) {
  const size_t gmax = (N + 255) / 256;
  return GP_HDR + gmax * 2 * 16 + gmax * 2 * GP_W * sizeof(double) + (size_t)2 * N * sizeof(int) + 64;
}
template int panel_gepp<double>(hipStream_t, double*, int64_t, int, int, int, int*, int*, void*);
template int panel_gepp<float>(hipStream_t, float*, int64_t, int, int, int, int*, int*, void*);

// ---------------------------------------------------------------------------
// Blocked triangular solves with 64 x 64 diagonal blocks applied through
// their stored inverses (DI: nblk blocks of 64x64, column-major, ld 64).
//   forward : y = inv(L) b   (L unit lower)      backward : x = inv(U) y
// Each 64-row block step = one small kernel (block solve) + one GEMV update of
// the remaining rows (forward) or rows above (backward) in the vector type TV.
// ---------------------------------------------------------------------------
template <typename T, typename TV>
__global__ __launch_bounds__(64) void tri_block_solve_kernel(const T* __restrict__ DI, int r0,
                                                             int nb, TV* __restrict__ v) {
  __shared__ TV sv[64];
  const int t = threadIdx.x;
  if (t < nb) sv[t] = v[r0 + t];
  __syncthreads();
  if (t < nb) {
    TV acc = TV(0);
    for (int kk = 0; kk < nb; ++kk) acc = __builtin_fma((TV)DI[t + kk * 64], sv[kk], acc);
    v[r0 + t] = acc;
  }
}

// v[i] -= sum_{c in [c0, c0+nc)} M[i, c] v[c] for rows i in [i0, i1)
template <typename T, typename TV>
__global__ __launch_bounds__(256) void tri_update_kernel(const T* __restrict__ M, int64_t ldm,
                                                         int i0, int i1, int c0, int nc,
                                                         TV* __restrict__ v) {
  __shared__ TV sx[64];
  if (threadIdx.x < nc) sx[threadIdx.x] = v[c0 + threadIdx.x];
  __syncthreads();
  const int i = i0 + blockIdx.x * 256 + threadIdx.x;
  if (i >= i1) return;
  TV acc = TV(0);
  for (int c = 0; c < nc; ++c) acc = __builtin_fma((TV)M[i + (int64_t)(c0 + c) * ldm], sx[c], acc);
  v[i] -= acc;
}

template <typename T, typename TV>
int trsv_blocked(hipStream_t s, const T* LU, int64_t ld, int N, const T* Linv_all,
                 const T* Uinv_all, TV* v) {
  const int nblk = (N + 63) / 64;
  for (int b = 0; b < nblk; ++b) {  // forward, unit lower
    const int r0 = b * 64, nb = min(64, N - r0);
    tri_block_solve_kernel<T, TV><<<1, 64, 0, s>>>(Linv_all + (size_t)b * 64 * 64, r0, nb, v);
    if (r0 + nb < N)
      tri_update_kernel<T, TV><<<(N - r0 - nb + 255) / 256, 256, 0, s>>>(LU, ld, r0 + nb, N, r0,
                                                                         nb, v);
  }
  for (int b = nblk - 1; b >= 0; --b) {  // backward, upper
    const int r0 = b * 64, nb = min(64, N - r0);
    tri_block_solve_kernel<T, TV><<<1, 64, 0, s>>>(Uinv_all + (size_t)b * 64 * 64, r0, nb, v);
    if (r0 > 0)
      tri_update_kernel<T, TV><<<(r0 + 255) / 256, 256, 0, s>>>(LU, ld, 0, r0, r0, nb, v);
  }
  MPDOK_HIP(hipGetLastError());
  return OK;
}
template int trsv_blocked<double, double>(hipStream_t, const double*, int64_t, int, const double*,
                                          const double*, double*);
template int trsv_blocked<float, float>(hipStream_t, const float*, int64_t, int, const float*,
                                        const float*, float*);
template int trsv_blocked<float, double>(hipStream_t, const float*, int64_t, int, const float*,
                                         const float*, double*);

// ---------------------------------------------------------------------------
// Triangular solves with explicit inverses of TBW x TBW diagonal blocks
// (TBW = 512 or 1024). One step per diagonal block, two kernels per step:
//   diag_apply : x_i = inv(D_i) (v_i - sum_c acc[c, rows of i])
//   tri_acc    : acc[c, r] += sum_{j in chunk c of block i} M[r, j] x_j  for the
//                rows r still to be solved (below i forward, above i backward)
// Column chunks are 64 wide (TBW / 64 of them). Every (c, r) accumulator is
// written by one thread per step in stream order, so the result does not
// depend on scheduling. Wider blocks halve the number of dependent steps.
// ---------------------------------------------------------------------------
constexpr int TB_CW = 64;

// xpart[cg, r0 + rg*64 + r] = sum_{j < 64} inv(D)[rg*64 + r, cg*64 + j] v'[cg*64 + j]
// with v' = v - sum_c acc[c]; one workgroup per (row group, column group) tile;
// tiles of the zero triangle are skipped.
template <int TBW, typename T, typename TV>
__global__ __launch_bounds__(256) void diag_apply_kernel(const T* __restrict__ Binv, int r0,
                                                         const TV* __restrict__ v,
                                                         const TV* __restrict__ acc, int N,
                                                         TV* __restrict__ xpart, int lower) {
  constexpr int TB_CH = TBW / TB_CW;
  __shared__ TV vs[TB_CW];
  __shared__ TV red[4][64];
  const int rg = blockIdx.x, cg = blockIdx.y, t = threadIdx.x;
  TV* out = xpart + (size_t)cg * N + r0 + rg * 64;
  if (lower ? cg > rg : cg < rg) {
    if (t < 64) out[t] = TV(0);
    return;
  }
  // the inverse's entries do not depend on v': their loads go out first
  const int r = t & 63, p = t >> 6;
  const T* col = Binv + rg * 64 + r + (size_t)(cg * TB_CW) * TBW;
  T bv[TB_CW / 4];
#pragma unroll
  for (int q = 0; q < TB_CW / 4; ++q)
    bv[q] = __builtin_nontemporal_load(col + (size_t)(p + 4 * q) * TBW);
  if (t < TB_CW) {
    const int g = r0 + cg * TB_CW + t;
    TV s = v[g];
#pragma unroll
    for (int c = 0; c < TB_CH; ++c) s -= acc[(size_t)c * N + g];
    vs[t] = s;
  }
  __syncthreads();
  TV a = TV(0);
#pragma unroll
  for (int q = 0; q < TB_CW / 4; ++q) a = __builtin_fma((TV)bv[q], vs[p + 4 * q], a);
  red[p][r] = a;
  __syncthreads();
  if (t < 64) out[t] = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
}

// Completes x for columns [c0, c0 + TBW) from the column-group partials, and
// adds the products of those columns into the accumulators of rows
// [rbeg, rend). Workgroups with blockIdx.x == 0 store the finished x values.
template <int TBW, typename T, typename TV>
__global__ __launch_bounds__(256) void tri_acc_kernel(const T* __restrict__ M, int64_t ldm,
                                                      int rbeg, int rend, int c0,
                                                      const TV* __restrict__ xpart,
                                                      TV* __restrict__ x, TV* __restrict__ acc,
                                                      int N) {
  constexpr int TB_CH = TBW / TB_CW;
  __shared__ TV xs[TB_CW];
  const int c = blockIdx.y;
  const int cb = c0 + c * TB_CW;
  const int r = rbeg + blockIdx.x * 256 + threadIdx.x;
  const bool active = r < rend;
  // The factors are read once per solve: nontemporal loads (8% faster than
  // ordinary ones here). The row's first 16 entries and its accumulator do
  // not depend on x: their loads go out before the partial sums of x and the
  // barrier, the other 48 after the first products (measured against 32 +
  // 32, 24 + 40 and all 64 first); the sum keeps the order j = 0 .. 63.
  const T* mr = M + r + (int64_t)cb * ldm;
  T m[48];
  TV a0 = TV(0);
  if (active) {
#pragma unroll
    for (int j = 0; j < 16; ++j) m[j] = __builtin_nontemporal_load(mr + (int64_t)j * ldm);
    a0 = acc[(size_t)c * N + r];
  }
  if (threadIdx.x < TB_CW) {
    TV s = TV(0);
#pragma unroll
    for (int g = 0; g < TB_CH; ++g) s += xpart[(size_t)g * N + cb + threadIdx.x];
    xs[threadIdx.x] = s;
    if (blockIdx.x == 0) x[cb + threadIdx.x] = s;
  }
  __syncthreads();
  if (!active) return;
  TV a = TV(0);
#pragma unroll
  for (int j = 0; j < 16; ++j) a = __builtin_fma((TV)m[j], xs[j], a);
#pragma unroll
  for (int j = 0; j < 48; ++j) m[j] = __builtin_nontemporal_load(mr + (int64_t)(16 + j) * ldm);
#pragma unroll
  for (int j = 0; j < 48; ++j) a = __builtin_fma((TV)m[j], xs[16 + j], a);
  acc[(size_t)c * N + r] = a0 + a;
}

template <int TBW, typename T, typename TV>
static int trsv_block_inverse_w(hipStream_t s, const T* LU, int64_t ld, int N, const T* Lbinv,
                                const T* Ubinv, const TV* b, TV* x, TV* v, TV* acc) {
  constexpr int TB_CH = TBW / TB_CW;
  MPDOK_REQUIRE(N % TBW == 0, "trsv_block_inverse: N must be a multiple of %d", TBW);
  const int nb = N / TBW;
  TV* xpart = acc + (size_t)TB_CH * N;  // second half of the accumulator buffer
  const dim3 dgrid(TBW / 64, TB_CH);
  // forward: L y = b, y -> v
  MPDOK_HIP(hipMemsetAsync(acc, 0, sizeof(TV) * (size_t)TB_CH * N, s));
  for (int i = 0; i < nb; ++i) {
    const int r0 = i * TBW;
    diag_apply_kernel<TBW, T, TV><<<dgrid, 256, 0, s>>>(Lbinv + (size_t)i * TBW * TBW, r0, b, acc,
                                                        N, xpart, 1);
    tri_acc_kernel<TBW, T, TV>
        <<<dim3(r0 + TBW < N ? (N - r0 - TBW + 255) / 256 : 1, TB_CH), 256, 0, s>>>(
            LU, ld, r0 + TBW, N, r0, xpart, v, acc, N);
  }
  // backward: U x = y
  MPDOK_HIP(hipMemsetAsync(acc, 0, sizeof(TV) * (size_t)TB_CH * N, s));
  for (int i = nb - 1; i >= 0; --i) {
    const int r0 = i * TBW;
    diag_apply_kernel<TBW, T, TV><<<dgrid, 256, 0, s>>>(Ubinv + (size_t)i * TBW * TBW, r0, v, acc,
                                                        N, xpart, 0);
    tri_acc_kernel<TBW, T, TV><<<dim3(r0 > 0 ? (r0 + 255) / 256 : 1, TB_CH), 256, 0, s>>>(
        LU, ld, 0, r0, r0, xpart, x, acc, N);
  }
  MPDOK_HIP(hipGetLastError());
  return OK;
}

template <typename T, typename TV>
int trsv_block_inverse(hipStream_t s, const T* LU, int64_t ld, int N, const T* Lbinv,
                       const T* Ubinv, const TV* b, TV* x, TV* v, TV* acc, int tbw) {
  if (tbw == 2 * TRSV_BLOCK)
    return trsv_block_inverse_w<2 * TRSV_BLOCK, T, TV>(s, LU, ld, N, Lbinv, Ubinv, b, x, v, acc);
  MPDOK_REQUIRE(tbw == TRSV_BLOCK, "trsv_block_inverse: block width %d", tbw);
  return trsv_block_inverse_w<TRSV_BLOCK, T, TV>(s, LU, ld, N, Lbinv, Ubinv, b, x, v, acc);
}
template int trsv_block_inverse<double, double>(hipStream_t, const double*, int64_t, int,
                                                const double*, const double*, const double*,
                                                double*, double*, double*, int);
template int trsv_block_inverse<float, float>(hipStream_t, const float*, int64_t, int, const float*,
                                              const float*, const float*, float*, float*, float*,
                                              int);

// ---------------------------------------------------------------------------
// The same solves for NR right-hand sides at once (column-major N x NR
// blocks, ld N): every matrix element read serves NR products, so the
// factors stream from HBM once per block of right-hand sides. For each
// right-hand side the sums run in the single-vector kernels' order, so every
// column of the result is bit for bit that of a single solve. Accumulators:
// acc[(r TB_CH + c) N + row] for right-hand side r, column group c (xpart
// likewise).
// ---------------------------------------------------------------------------
template <int TBW, typename T, typename TV, int NR>
__global__ __launch_bounds__(256) void diag_apply_multi_kernel(const T* __restrict__ Binv, int r0,
                                                               const TV* __restrict__ v,
                                                               const TV* __restrict__ acc, int N,
                                                               TV* __restrict__ xpart, int lower) {
  constexpr int TB_CH = TBW / TB_CW;
  __shared__ TV vs[NR][TB_CW];
  __shared__ TV red[NR][4][64];
  const int rg = blockIdx.x, cg = blockIdx.y, t = threadIdx.x;
  if (lower ? cg > rg : cg < rg) {
    if (t < 64)
#pragma unroll
      for (int q = 0; q < NR; ++q) xpart[((size_t)q * TB_CH + cg) * N + r0 + rg * 64 + t] = TV(0);
    return;
  }
  const int r = t & 63, p = t >> 6;
  const T* col = Binv + rg * 64 + r + (size_t)(cg * TB_CW) * TBW;
  T bv[TB_CW / 4];
#pragma unroll
  for (int q = 0; q < TB_CW / 4; ++q)
    bv[q] = __builtin_nontemporal_load(col + (size_t)(p + 4 * q) * TBW);
  if (t < TB_CW) {
    const int g = r0 + cg * TB_CW + t;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      TV sv = v[(size_t)q * N + g];
#pragma unroll
      for (int c = 0; c < TB_CH; ++c) sv -= acc[((size_t)q * TB_CH + c) * N + g];
      vs[q][t] = sv;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    TV a = TV(0);
#pragma unroll
    for (int u = 0; u < TB_CW / 4; ++u) a = __builtin_fma((TV)bv[u], vs[q][p + 4 * u], a);
    red[q][p][r] = a;
  }
  __syncthreads();
  if (t < 64)
#pragma unroll
    for (int q = 0; q < NR; ++q)
      xpart[((size_t)q * TB_CH + cg) * N + r0 + rg * 64 + t] =
          (red[q][0][t] + red[q][1][t]) + (red[q][2][t] + red[q][3][t]);
}

template <int TBW, typename T, typename TV, int NR>
__global__ __launch_bounds__(256) void tri_acc_multi_kernel(const T* __restrict__ M, int64_t ldm,
                                                            int rbeg, int rend, int c0,
                                                            const TV* __restrict__ xpart,
                                                            TV* __restrict__ x,
                                                            TV* __restrict__ acc, int N) {
  constexpr int TB_CH = TBW / TB_CW;
  __shared__ TV xs[NR][TB_CW];
  const int c = blockIdx.y;
  const int cb = c0 + c * TB_CW;
  const int row = rbeg + blockIdx.x * 256 + threadIdx.x;
  const bool active = row < rend;
  const T* mr = M + row + (int64_t)cb * ldm;
  T m[32];
  TV a0[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) a0[q] = TV(0);
  if (active) {
#pragma unroll
    for (int j = 0; j < 32; ++j) m[j] = __builtin_nontemporal_load(mr + (int64_t)j * ldm);
#pragma unroll
    for (int q = 0; q < NR; ++q) a0[q] = acc[((size_t)q * TB_CH + c) * N + row];
  }
  if (threadIdx.x < TB_CW) {
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      TV sv = TV(0);
#pragma unroll
      for (int g = 0; g < TB_CH; ++g) sv += xpart[((size_t)q * TB_CH + g) * N + cb + threadIdx.x];
      xs[q][threadIdx.x] = sv;
      if (blockIdx.x == 0) x[(size_t)q * N + cb + threadIdx.x] = sv;
    }
  }
  __syncthreads();
  if (!active) return;
  TV a[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    a[q] = TV(0);
#pragma unroll
    for (int j = 0; j < 32; ++j) a[q] = __builtin_fma((TV)m[j], xs[q][j], a[q]);
  }
#pragma unroll
  for (int j = 0; j < 32; ++j) m[j] = __builtin_nontemporal_load(mr + (int64_t)(32 + j) * ldm);
#pragma unroll
  for (int q = 0; q < NR; ++q) {
#pragma unroll
    for (int j = 0; j < 32; ++j) a[q] = __builtin_fma((TV)m[j], xs[q][32 + j], a[q]);
    acc[((size_t)q * TB_CH + c) * N + row] = a0[q] + a[q];
  }
}

template <int TBW, typename T, typename TV, int NR>
static int trsv_multi_w(hipStream_t s, const T* LU, int64_t ld, int N, const T* Lbinv,
                        const T* Ubinv, const TV* b, TV* x, TV* v, TV* acc) {
  constexpr int TB_CH = TBW / TB_CW;
  MPDOK_REQUIRE(N % TBW == 0, "trsv_multi: N must be a multiple of %d", TBW);
  const int nb = N / TBW;
  TV* xpart = acc + (size_t)TB_CH * N * NR;
  const dim3 dgrid(TBW / 64, TB_CH);
  MPDOK_HIP(hipMemsetAsync(acc, 0, sizeof(TV) * (size_t)TB_CH * N * NR, s));
  for (int i = 0; i < nb; ++i) {
    const int r0 = i * TBW;
    diag_apply_multi_kernel<TBW, T, TV, NR><<<dgrid, 256, 0, s>>>(
        Lbinv + (size_t)i * TBW * TBW, r0, b, acc, N, xpart, 1);
    tri_acc_multi_kernel<TBW, T, TV, NR>
        <<<dim3(r0 + TBW < N ? (N - r0 - TBW + 255) / 256 : 1, TB_CH), 256, 0, s>>>(
            LU, ld, r0 + TBW, N, r0, xpart, v, acc, N);
  }
  MPDOK_HIP(hipMemsetAsync(acc, 0, sizeof(TV) * (size_t)TB_CH * N * NR, s));
  for (int i = nb - 1; i >= 0; --i) {
    const int r0 = i * TBW;
    diag_apply_multi_kernel<TBW, T, TV, NR><<<dgrid, 256, 0, s>>>(
        Ubinv + (size_t)i * TBW * TBW, r0, v, acc, N, xpart, 0);
    tri_acc_multi_kernel<TBW, T, TV, NR><<<dim3(r0 > 0 ? (r0 + 255) / 256 : 1, TB_CH), 256, 0, s>>>(
        LU, ld, 0, r0, r0, xpart, x, acc, N);
  }
  MPDOK_HIP(hipGetLastError());
  return OK;
}

template <typename T, typename TV>
int trsv_block_inverse_multi(hipStream_t s, const T* LU, int64_t ld, int N, const T* Lbinv,
                             const T* Ubinv, const TV* b, TV* x, TV* v, TV* acc, int tbw,
                             int nr) {
#define TRSV_M(W, R) return trsv_multi_w<W, T, TV, R>(s, LU, ld, N, Lbinv, Ubinv, b, x, v, acc)
  MPDOK_REQUIRE(tbw == TRSV_BLOCK || tbw == 2 * TRSV_BLOCK, "trsv_multi: block width %d", tbw);
  MPDOK_REQUIRE(nr == 1 || nr == 2 || nr == 4 || nr == 8, "trsv_multi: %d right-hand sides", nr);
  if (tbw == 2 * TRSV_BLOCK) {
    switch (nr) {
      case 8: TRSV_M(2 * TRSV_BLOCK, 8);
      case 4: TRSV_M(2 * TRSV_BLOCK, 4);
      case 2: TRSV_M(2 * TRSV_BLOCK, 2);
      default: TRSV_M(2 * TRSV_BLOCK, 1);
    }
  }
  switch (nr) {
    case 8: TRSV_M(TRSV_BLOCK, 8);
    case 4: TRSV_M(TRSV_BLOCK, 4);
    case 2: TRSV_M(TRSV_BLOCK, 2);
    default: TRSV_M(TRSV_BLOCK, 1);
  }
#undef TRSV_M
}
template int trsv_block_inverse_multi<double, double>(hipStream_t, const double*, int64_t, int,
                                                      const double*, const double*, const double*,
                                                      double*, double*, double*, int, int);
template int trsv_block_inverse_multi<float, float>(hipStream_t, const float*, int64_t, int,
                                                    const float*, const float*, const float*,
                                                    float*, float*, float*, int, int);

// 1024-wide diagonal blocks of the inverses from the 512-wide ones: block b
// takes inv(A) = in[2b] top-left and inv(D) = in[2b + 1] bottom-right; the
// off-diagonal block of the zero triangle (top-right for a lower inverse,
// bottom-left for an upper one) is zero, and the other off-diagonal block is
// left for the doubling products. One thread per element, nbig blocks.
template <typename T>
__global__ __launch_bounds__(256) void widen_block_inverse_kernel(const T* __restrict__ in,
                                                                  T* __restrict__ out, int nbig,
                                                                  int lower) {
  constexpr int B1 = TRSV_BLOCK, B2 = 2 * TRSV_BLOCK;
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)nbig * B2 * B2) return;
  const int b = (int)(e / ((size_t)B2 * B2));
  const int rc = (int)(e % ((size_t)B2 * B2));
  const int r = rc % B2, c = rc / B2;
  const int hr = r / B1, hc = c / B1;
  if (hr == hc) {
    out[e] = in[(size_t)(2 * b + hr) * B1 * B1 + (r % B1) + (size_t)(c % B1) * B1];
  } else if (lower ? hc > hr : hc < hr) {
    out[e] = T(0);
  }
}

template <typename T>
int widen_block_inverse(hipStream_t s, const T* in, T* out, int nbig, int lower) {
  if (nbig <= 0) return OK;
  const size_t n = (size_t)nbig * 4 * TRSV_BLOCK * TRSV_BLOCK;
  widen_block_inverse_kernel<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(in, out, nbig, lower);
  MPDOK_HIP(hipGetLastError());
  return OK;
}
template int widen_block_inverse<double>(hipStream_t, const double*, double*, int, int);
template int widen_block_inverse<float>(hipStream_t, const float*, float*, int, int);

// Place the 64 x 64 diagonal inverses into the TRSV_BLOCK-wide blocks (zero elsewhere).
template <typename T>
__global__ __launch_bounds__(256) void seed_block_inverse_kernel(const T* __restrict__ inv64,
                                                                 T* __restrict__ binv, int nbig) {
  // blockIdx.x: TRSV_BLOCK-wide block index; blockIdx.y: column of that block
  constexpr int TB = TRSV_BLOCK;
  const int b = blockIdx.x, col = blockIdx.y;
  T* dst = binv + (size_t)b * TB * TB + (size_t)col * TB;
  const int g = b * (TB / 64) + col / 64;  // global 64-block index
  const T* src = inv64 + (size_t)g * 4096 + (size_t)(col % 64) * 64;
  for (int r = threadIdx.x; r < TB; r += 256) {
    const int lo = (col / 64) * 64;
    dst[r] = (r >= lo && r < lo + 64) ? src[r - lo] : T(0);
  }
}

template <typename T>
int seed_block_inverse(hipStream_t s, const T* inv64, T* binv, int nbig) {
  if (nbig <= 0) return OK;
  seed_block_inverse_kernel<T><<<dim3(nbig, TRSV_BLOCK), 256, 0, s>>>(inv64, binv, nbig);
  MPDOK_HIP(hipGetLastError());
  return OK;
}
template int seed_block_inverse<double>(hipStream_t, const double*, double*, int);
template int seed_block_inverse<float>(hipStream_t, const float*, float*, int);

// Gather with a row permutation: out[i] = in[perm[i]]
template <typename T>
__global__ __launch_bounds__(256) void permute_kernel(const T* __restrict__ in, T* __restrict__ out,
                                                      const int* __restrict__ perm, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = in[perm[i]];
}
template <typename T>
int permute(hipStream_t s, const T* in, T* out, const int* perm, int n) {
  if (n <= 0) return OK;
  permute_kernel<T><<<(n + 255) / 256, 256, 0, s>>>(in, out, perm, n);
  MPDOK_HIP(hipGetLastError());
  return OK;
}
template int permute<double>(hipStream_t, const double*, double*, const int*, int);
template int permute<float>(hipStream_t, const float*, float*, const int*, int);

// ipiv (sequential interchanges) -> permutation vector perm with
// (P b)[i] = b[perm[i]]. Single thread; run once per factorisation.
__global__ void ipiv_to_perm_kernel(const int* __restrict__ ipiv, int* __restrict__ perm, int n) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int i = 0; i < n; ++i) {
    const int p = ipiv[i];
    if (p != i) {
      const int t = perm[i];
      perm[i] = perm[p];
      perm[p] = t;
    }
  }
}
int ipiv_to_perm(hipStream_t s, const int* ipiv, int* perm, int n) {
  ipiv_to_perm_kernel<<<1, 64, 0, s>>>(ipiv, perm, n);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

// ---------------------------------------------------------------------------
// Grouped launches of a batched no-pivot LU (many systems, one launch per
// kind of work and step). Step k of every system in sys[0 .. ns) (P > 64 k):
// blu_diag factors the 64 x 64 diagonal block at (64 k, 64 k) and stores the
// inverses of its unit-lower and upper triangles; blu_panel forms
// L21 = A21 inv(U11) and U12 = inv(L11) A12 in place (64 x 64 tiles, the
// aliasing rules of gemm_tile); blu_update does A22 -= L21 U12 (TB x TB
// tiles). The tiles of all systems form one grid: pref[s] is the first tile
// of system s (pref[ns] the total), and workgroups are dealt to the XCDs in
// contiguous ranges as in tile_of, so a system's tiles share an XCD's L2.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool grouped_tile(int b, const int* __restrict__ pref, int ns, int& si,
                                             int& t) {
  const int T = pref[ns];
  const int per = (T + NXCD - 1) / NXCD;
  const int tt = (b % NXCD) * per + b / NXCD;
  if (b / NXCD >= per || tt >= T) return false;
  // the largest s with pref[s] <= tt (systems without tiles are skipped)
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

__global__ __launch_bounds__(256) void blu_diag_kernel(const BLUSys* __restrict__ sys, int k) {
  const BLUSys q = sys[blockIdx.x];
  const int64_t P = q.P, c0 = 64 * (int64_t)k;
  diag_factor_blk_body<double>(q.F + c0 + c0 * P, P, 64, q.Linv + c0, q.Uinv + c0, P, q.info,
                               (int)c0, 1);
}

__global__ __launch_bounds__(256, 2) void blu_panel_kernel(const BLUSys* __restrict__ sys,
                                                           const int* __restrict__ pref, int ns,
                                                           int k) {
  int si, t;
  if (!grouped_tile(blockIdx.x, pref, ns, si, t)) return;
  const BLUSys q = sys[si];
  const int64_t P = q.P, c0 = 64 * (int64_t)k, c1 = c0 + 64;
  const int m = q.P - (int)c1;
  const int tm = (m + 63) / 64;
  double* A21 = q.F + c1 + c0 * P;
  double* A12 = q.F + c0 + c1 * P;
  const bool l21 = t < tm;  // L21 row tile t, else U12 column tile t - tm
  gemm_tile_at<double, 0, 64>(l21 ? t : 0, l21 ? 0 : t - tm, l21 ? m : 64, l21 ? 64 : m, 64, 1.0,
                              l21 ? A21 : q.Linv + c0, P, l21 ? q.Uinv + c0 : A12, P, 0.0,
                              l21 ? A21 : A12, P);
}

template <int TB>
__global__ __launch_bounds__(256, 2) void blu_update_kernel(const BLUSys* __restrict__ sys,
                                                            const int* __restrict__ pref, int ns,
                                                            int k) {
  int si, t;
  if (!grouped_tile(blockIdx.x, pref, ns, si, t)) return;
  const BLUSys q = sys[si];
  const int64_t P = q.P, c0 = 64 * (int64_t)k, c1 = c0 + 64;
  const int m = q.P - (int)c1;
  const int tm = (m + TB - 1) / TB;
  gemm_tile_at<double, 0, TB>(t % tm, t / tm, m, m, 64, -1.0, q.F + c1 + c0 * P, P,
                              q.F + c0 + c1 * P, P, 1.0, q.F + c1 + c1 * P, P);
}

int blu_diag(hipStream_t s, const BLUSys* sys, int ns, int k) {
  if (ns <= 0) return OK;
  blu_diag_kernel<<<ns, 256, 0, s>>>(sys, k);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

int blu_panel(hipStream_t s, const BLUSys* sys, const int* pref, int ns, int ntiles, int k) {
  if (ns <= 0 || ntiles <= 0) return OK;
  blu_panel_kernel<<<(ntiles + NXCD - 1) / NXCD * NXCD, 256, 0, s>>>(sys, pref, ns, k);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

int blu_update(hipStream_t s, const BLUSys* sys, const int* pref, int ns, int ntiles, int k,
               int tb) {
  MPDOK_REQUIRE(tb == 64 || tb == 128, "blu_update: tile %d", tb);
  if (ns <= 0 || ntiles <= 0) return OK;
  const int grid = (ntiles + NXCD - 1) / NXCD * NXCD;
  if (tb == 128) blu_update_kernel<128><<<grid, 256, 0, s>>>(sys, pref, ns, k);
  else blu_update_kernel<64><<<grid, 256, 0, s>>>(sys, pref, ns, k);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

}  // namespace mpdok
