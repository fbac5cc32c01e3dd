// Trailing update C -= A B of the blocked LU on the int8 matrix cores, to
// about 8 S - 1 bits, with FP64 storage of C (Ozaki scheme; Ozaki, Ogita,
// Oishi, Rump, Numer. Algorithms 59, 2012).
//
// Splitting: each row of A (m x K) and each column of B (K x n) gets a power
// of two 2^e, the smallest with max|.| * 2^(7 - e) <= 127. With P = 8 S - 1,
// an element v becomes the integer V = rint(v 2^(P - e)) (|V| <= 127 256^(S-1)),
// written in balanced base 256 from the least significant digit:
//     V = sum_q d_q 256^(S - q),  d_q in [-128, 127] (q >= 2), |d_1| <= 127,
// so v = 2^(e - P) V up to half a unit of the last digit.
// Product: for each level l = i + j in 2..S+1 the exact integer
//     acc_l = sum_{i+j=l} A_i B_j       (int32; |acc_l| <= S K 2^14)
// comes from v_mfma_i32_32x32x32_i8; pairs with i + j > S + 1 are dropped.
// The epilogue folds c = sum_l acc_l 2^(2 - 8 l) in FP64 for l ascending from
// 2 and subtracts 2^(eA_i + eB_j) c from C. Slicing, integer sums, the fold
// order and the power-of-two scalings are all deterministic, so the result is
// bit for bit the one of the oracle's ozaki_matmul.
//
// Layout of the slices (K-blocked): As[q][k / 32][row][k % 32], one 32-byte
// record per (row, 32-wide K block), so that the bytes one K step of one block
// tile needs are a single contiguous run per slice, copied straight into LDS.
//
// Trailing update of the LU (the runtime-count kernel, ozaki_gemmd_kernel):
// 512 threads = 2 x 4 wavefronts of 64 x 32 (two v_mfma_i32_32x32x32_i8
// tiles each), block tile 128 x 128, K step 32, one workgroup per CU. One K
// step of all S slices of both operands is a pipeline stage in LDS, filled by
// global_load_lds_dwordx4 (global memory to LDS without registers) three
// stages ahead of its use; a stage is waited for with s_waitcnt vmcnt and a
// plain s_barrier, so the copies of the next stages stay in flight while the
// matrix cores work on the current one. All S(S+1)/2 slice pairs of a stage
// are multiplied from LDS, with one int32 accumulator set per level. For
// slice counts up to 3 the products take the U-panel slices as the first
// operand, so C's rows run along the lanes, and the tile's FP64 C values are
// read into registers right after the first stages' copies: the read of C
// overlaps the K loop and the write-back only stores. Other shapes (tuning
// key 0) and the read of C after the K loop (key 5) are selectable.
// Products with a fixed slice count (ozaki_gemm) use 128 x 64 tiles, two
// workgroups per CU, C read after the K loop.
#include "common.h"
#include "lu.h"

namespace mpdok {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int OZ_BM = 128, OZ_BK = 32;  // block rows, K step (bytes of one record)

// Exponent of a row / column with maximum magnitude mx: the smallest e with
// mx 2^(7 - e) <= 127 (OZ_ZERO_EXP for mx = 0: its digits are all zero, so
// the exponent scales nothing, and it stays below every real one).
constexpr int OZ_ZERO_EXP = -1000;
__device__ __forceinline__ int oz_exponent(double mx) {
  int e = OZ_ZERO_EXP;
  if (mx > 0.0) {
    frexp(mx, &e);
    if (ldexp(mx, 7 - e) > 127.0) ++e;
  }
  return e;
}

// Digits of v (|v| <= 127 2^(e - 7)) into d[0..S): d[0] most significant.
template <int S>
__device__ __forceinline__ void make_digits(double v, int e, signed char* d) {
  long long V = (long long)rint(ldexp(v, 8 * S - 1 - e));
#pragma unroll
  for (int q = S - 1; q > 0; --q) {
    const signed char dq = (signed char)(V & 0xff);  // V mod 256 in [-128, 127]
    d[q] = dq;
    V = (V - dq) >> 8;  // exact: V - dq is a multiple of 256
  }
  d[0] = (signed char)V;
}

// Byte offset of digit (row, k) inside one slice of the K-blocked layout with
// `rows` padded rows.
__device__ __forceinline__ size_t oz_rec(int rows, int row, int k) {
  return (size_t)(k / OZ_BK) * rows * OZ_BK + (size_t)row * OZ_BK + (k % OZ_BK);
}

// Rows of A (m x K, column-major, lda) -> As[q] (K-blocked), rows padded to
// mp, k padded to Kp with zeros; ea[row] = exponent.
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
    const int e = oz_exponent(fmax(fmax(red[0][t], red[1][t]), fmax(red[2][t], red[3][t])));
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
    if (ii < mp && k0 + kq < Kp) {
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
        *reinterpret_cast<v4i*>(As + q * slab + oz_rec(mp, ii, k0 + kq)) = w;
      }
    }
    __syncthreads();
  }
}

// Row slicing over a 2-D grid, for panels too short to fill the chip with
// one workgroup per 64 rows. Pass 1: partial maxima of |A| over each 64 x 64
// tile, pm[kblock * mp + row]. Pass 2: each 64 x 64 tile takes its rows'
// exponents from the partial maxima (the same maximum, so the same digits as
// slice_rows_kernel) and writes its records; tiles of k block 0 write ea.
__global__ __launch_bounds__(256) void rowmax_part_kernel(const double* __restrict__ A,
                                                          int64_t lda, int m, int K, int mp,
                                                          double* __restrict__ pm) {
  __shared__ double red[4][64];
  const int t = threadIdx.x, r = t & 63, p = t >> 6;
  const int i = blockIdx.x * 64 + r, k0 = blockIdx.y * 64;
  const int k1 = K < k0 + 64 ? K : k0 + 64;
  double mx = 0.0;
  if (i < m)
#pragma unroll 4
    for (int k = k0 + p; k < k1; k += 4) mx = fmax(mx, fabs(A[i + (int64_t)k * lda]));
  red[p][r] = mx;
  __syncthreads();
  if (t < 64)
    pm[(size_t)blockIdx.y * mp + blockIdx.x * 64 + t] =
        fmax(fmax(red[0][t], red[1][t]), fmax(red[2][t], red[3][t]));
}

template <int S>
__global__ __launch_bounds__(256) void slice_rows_tile_kernel(const double* __restrict__ A,
                                                              int64_t lda, int m, int K, int mp,
                                                              int Kp, const double* __restrict__ pm,
                                                              signed char* __restrict__ As,
                                                              int* __restrict__ ea) {
  __shared__ double tile[64][65];
  __shared__ int se[64];
  const int t = threadIdx.x;
  const int row0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
  if (t < 64) {
    double mx = 0.0;
    for (int b = 0; b < Kp / 64; ++b) mx = fmax(mx, pm[(size_t)b * mp + row0 + t]);
    const int e = oz_exponent(mx);
    se[t] = e;
    if (blockIdx.y == 0) ea[row0 + t] = e;
  }
  for (int e = t; e < 64 * 64; e += 256) {
    const int rr = e & 63, kk = e >> 6;
    const int ii = row0 + rr, k = k0 + kk;
    tile[rr][kk] = (ii < m && k < K) ? A[ii + (int64_t)k * lda] : 0.0;
  }
  __syncthreads();
  // thread -> (row rr, 16 consecutive k); rows < mp and k < Kp by the grid
  const int rr = t >> 2, kq = (t & 3) * 16;
  const int ii = row0 + rr;
  const size_t slab = (size_t)mp * Kp;
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
    *reinterpret_cast<v4i*>(As + q * slab + oz_rec(mp, ii, k0 + kq)) = w;
  }
}

// Columns of B (K x n, column-major, ldb) -> Bs[q] (K-blocked); eb[col].
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
  const int e = oz_exponent(mx);
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
      *reinterpret_cast<v4i*>(Bs + q * slab + oz_rec(np, j, k0)) = w;
    }
  }
}

// The same for Kp <= OZ_SC_KMAX, one wavefront per column, the column read
// once: coalesced nontemporal loads (lane-strided) into LDS while the
// maximum is formed, then each lane takes 16 consecutive k from LDS (rows
// padded by 16 bytes every 16 values, so the 16-byte reads of a wavefront
// spread over the banks). Same digits and exponents as slice_cols_kernel.
constexpr int OZ_SC_KMAX = 2048;
template <int S>
__global__ __launch_bounds__(256) void slice_cols_lds_kernel(const double* __restrict__ B,
                                                             int64_t ldb, int n, int K, int np,
                                                             int Kp, signed char* __restrict__ Bs,
                                                             int* __restrict__ eb) {
  constexpr int LD = OZ_SC_KMAX + OZ_SC_KMAX / 16 * 2;  // padded doubles per column
  __shared__ __attribute__((aligned(16))) double colbuf[4][LD];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = blockIdx.x * 4 + w;
  if (j >= np) return;  // wavefront-uniform; no workgroup barrier below
  const double* col = B + (int64_t)j * ldb;
  double* cb = colbuf[w];
  double mx = 0.0;
  for (int k = lane; k < Kp; k += 64) {
    const double v = (j < n && k < K) ? __builtin_nontemporal_load(col + k) : 0.0;
    cb[k + (k >> 4) * 2] = v;
    mx = fmax(mx, fabs(v));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  const int e = oz_exponent(mx);
  if (lane == 0) eb[j] = e;
  // this wavefront's LDS stores before its reads
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const size_t slab = (size_t)np * Kp;
  for (int k0 = lane * 16; k0 < Kp; k0 += 64 * 16) {
    const double* src = cb + k0 + (k0 >> 4) * 2;
    signed char dig[S][16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      signed char d[S];
      make_digits<S>(src[c], e, d);
#pragma unroll
      for (int q = 0; q < S; ++q) dig[q][c] = d[q];
    }
#pragma unroll
    for (int q = 0; q < S; ++q) {
      v4i wv;
      __builtin_memcpy(&wv, dig[q], 16);
      *reinterpret_cast<v4i*>(Bs + q * slab + oz_rec(np, j, k0)) = wv;
    }
  }
}

// C (m x n, column-major, ldc) -= sliced product, columns [col0, col0 + ncols).
// mp, np: padded row / column counts of the slice arrays; Kp a multiple of 128.
//
// MFMA operands: lane l takes row (column) l % 32 and bytes [16 (l / 32),
// 16 (l / 32) + 16) of its 32-byte record (in LDS: half l / 32 of the record),
// identically for A and B (the order of k inside a product does not matter,
// only that both operands agree);
// D[i][j] sits in lane j + 32 u, register r with i = 8 (r / 4) + 4 u + r % 4.
constexpr int oz_vmcnt(int n) {  // s_waitcnt immediate: vmcnt = n, lgkmcnt / expcnt not waited
  return (n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8);
}
constexpr int OZ_LGKM0 = 15 | (3 << 14) | (7 << 4);  // s_waitcnt immediate: lgkmcnt = 0 only

// Block shape: WM x WN wavefronts, each owning 64 rows x 32 columns of the
// block tile (two 32 x 32 MFMA tiles), so BM = 64 WM, BN = 32 WN; NST
// pipeline stages; OCC workgroups per CU (launch bound).
template <int S, int WM, int WN, int NST, int OCC>
struct OzCfg {
  static constexpr int NT = 64 * WM * WN;  // threads
  static constexpr int BM = 64 * WM, BN = 32 * WN;
  // Copies of an operand with R records per slice: 2 R chunks of 16 bytes per
  // slice, NT / (2 R) slices per round, every thread busy in every round (a
  // round past the last slice copies that slice once more, see oz_issue_op).
  static constexpr int SPR_A = NT / (2 * BM), SPR_B = NT / (2 * BN);
  static constexpr int RND_A = (S + SPR_A - 1) / SPR_A, RND_B = (S + SPR_B - 1) / SPR_B;
  static constexpr int ABYTES = S * BM * OZ_BK;
  static constexpr int BBYTES = S * BN * OZ_BK;
  static constexpr int STAGE = ABYTES + BBYTES;
  static constexpr int COPIES = RND_A + RND_B;  // per thread and stage
  // epilogue staging: per wavefront 32 x 32 FP64 (column stride 33) and 32 exponents
  static constexpr int WAVES_PER_BUF = (WM * WN + NST - 1) / NST;
  static constexpr int EPI = WAVES_PER_BUF * (32 * 33 + 32) * 8;
  static constexpr int BUF = STAGE > EPI ? STAGE : EPI;
  static constexpr bool FITS = NST * BUF <= 160 * 1024;  // LDS of one workgroup
  static_assert(NT % (2 * BM) == 0 && NT % (2 * BN) == 0, "copy split");
};

// One operand of one K step: R records (32 bytes each) of each of S slices.
// In LDS a slice is stored as [k half][record][16 bytes]: the 16-byte chunk
// at position c (c = 0 .. 2R - 1) is half c / R of record c % R. The copy
// writes position = 16 * lane (global_load_lds), so each lane fetches the
// chunk that belongs there; the fragment reads then cover contiguous 512-byte
// runs per half wavefront (no bank conflicts).
template <int R, int SPR, int RND, int S>
__device__ __forceinline__ void oz_issue_op(signed char* __restrict__ dst, const signed char* src,
                                            size_t slab, int wave) {
  const int sub = (64 * wave) / (2 * R);         // slice within the round (wave-uniform)
  const int off = ((64 * wave) % (2 * R)) * 16;  // byte offset of this wavefront's 64 chunks
#pragma unroll
  for (int rd = 0; rd < RND; ++rd) {
    // past the last slice (odd S with two slices per round) the copy repeats
    // slice S - 1 into its own slot: the same bytes twice, so whichever lands
    // last leaves them unchanged, and both land before the stage is read
    const int q = rd * SPR + sub < S ? rd * SPR + sub : S - 1;
    __builtin_amdgcn_global_load_lds(src + q * slab, dst + q * R * OZ_BK + off, 16, 0, 0);
  }
}
// Source offset (bytes, within one slice's K step of the tile) of the chunk
// that thread `tid` copies for an operand of R records.
template <int R>
__device__ __forceinline__ int oz_src_off(int tid) {
  const int c = tid % (2 * R);
  return (c % R) * OZ_BK + (c / R) * 16;
}

// One K step (32 bytes of every row / column of all S slices) into a stage.
template <class Cfg, int S>
__device__ __forceinline__ void oz_issue(signed char* __restrict__ st, const signed char* a,
                                         const signed char* b, size_t slabA, size_t slabB,
                                         int wave) {
  oz_issue_op<Cfg::BM, Cfg::SPR_A, Cfg::RND_A, S>(st, a, slabA, wave);
  oz_issue_op<Cfg::BN, Cfg::SPR_B, Cfg::RND_B, S>(st + Cfg::ABYTES, b, slabB, wave);
}

typedef int v16i __attribute__((ext_vector_type(16)));

// 2^(2 - 8 (l + 2)): weight of level l (0-based) in the fold; exact power of two.
__device__ __forceinline__ constexpr double oz_level_scale(int l) {
  double w = 1.0;
  for (int i = 0; i < 8 * (l + 2) - 2; ++i) w *= 0.5;
  return w;
}

template <class Cfg, int S>
__device__ __forceinline__ void oz_compute(const signed char* __restrict__ st, v16i (&acc)[S][2],
                                           int wm, int wn, int lane) {
  const int r32 = lane & 31, hf = lane >> 5;  // record within a 32-tile, k half
  v4i b[S];
#pragma unroll
  for (int j = 0; j < S; ++j)
    b[j] = *reinterpret_cast<const v4i*>(st + Cfg::ABYTES + j * Cfg::BN * OZ_BK +
                                         hf * Cfg::BN * 16 + (wn + r32) * 16);
#pragma unroll
  for (int i = 0; i < S; ++i) {
    v4i a[2];
#pragma unroll
    for (int ta = 0; ta < 2; ++ta)
      a[ta] = *reinterpret_cast<const v4i*>(st + i * Cfg::BM * OZ_BK + hf * Cfg::BM * 16 +
                                            (wm + 32 * ta + r32) * 16);
#pragma unroll
    for (int j = 0; j < S - i; ++j)  // level index i + j (0-based)
#pragma unroll
      for (int ta = 0; ta < 2; ++ta)
        acc[i + j][ta] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[ta], b[j], acc[i + j][ta], 0, 0, 0);
  }
}

// One pipeline step on stage `st`: wait for its copies (those of the NST - 1
// later stages stay in flight), multiply, free it, refill it.
template <class Cfg, int S, int NST>
__device__ __forceinline__ void oz_step(signed char* __restrict__ st, v16i (&acc)[S][2],
                                        const signed char* a, const signed char* b, size_t slabA,
                                        size_t slabB, int wm, int wn, int lane, int wave) {
  __builtin_amdgcn_s_waitcnt(oz_vmcnt((NST - 1) * Cfg::COPIES));  // this thread's copies landed
  __builtin_amdgcn_s_barrier();                                   // ... and every thread's
  oz_compute<Cfg, S>(st, acc, wm, wn, lane);
  __builtin_amdgcn_s_waitcnt(OZ_LGKM0);  // this thread's reads of the stage are done
  __builtin_amdgcn_s_barrier();          // stage free
  oz_issue<Cfg, S>(st, a, b, slabA, slabB, wave);
}

// Kp: a multiple of 32 NST, so that the K steps come in whole rounds of stages.
template <int S, int WM, int WN, int NST, int OCC>
__global__ __launch_bounds__(64 * WM * WN, OCC) void ozaki_gemm_kernel(
    int m, int n, int mp, int np, int Kp, int col0, int ncols, const signed char* __restrict__ As,
    const signed char* __restrict__ Bs, const int* __restrict__ ea, const int* __restrict__ eb,
    double* __restrict__ C, int64_t ldc, int group) {
  typedef OzCfg<S, WM, WN, NST, OCC> Cfg;
  // one array per pipeline stage: the compiler then tells the stages apart
  // and waits only for the copies into the stage being read
  __shared__ __attribute__((aligned(16))) signed char st0[Cfg::BUF];
  __shared__ __attribute__((aligned(16))) signed char st1[Cfg::BUF];
  __shared__ __attribute__((aligned(16))) signed char st2[NST > 2 ? Cfg::BUF : 16];
  __shared__ __attribute__((aligned(16))) signed char st3[NST > 3 ? Cfg::BUF : 16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave % WM) * 64, wn = (wave / WM) * 32;
  int ti, tj;
  if (!tile_of(blockIdx.x, mp / Cfg::BM, (ncols + Cfg::BN - 1) / Cfg::BN, ti, tj, group)) return;
  const int bm = ti * Cfg::BM, bn = col0 + tj * Cfg::BN;
  const size_t slabA = (size_t)mp * Kp, slabB = (size_t)np * Kp;
  const size_t kstepA = (size_t)mp * OZ_BK, kstepB = (size_t)np * OZ_BK;
  const signed char* Ab = As + (size_t)bm * OZ_BK + oz_src_off<Cfg::BM>(tid);
  const signed char* Bb = Bs + (size_t)bn * OZ_BK + oz_src_off<Cfg::BN>(tid);

  v16i acc[S][2];
#pragma unroll
  for (int l = 0; l < S; ++l)
#pragma unroll
    for (int a = 0; a < 2; ++a) acc[l][a] = v16i(0);

  // Stage pipeline without branches in the loop body (the accumulators then
  // stay in place across iterations): every thread has Cfg::COPIES copies per
  // stage in flight; past the last K step the refills repeat the last step
  // (harmless, never read) so that the wait counts stay the same.
  const int nk = Kp / OZ_BK;
  oz_issue<Cfg, S>(st0, Ab, Bb, slabA, slabB, wave);
  oz_issue<Cfg, S>(st1, Ab + kstepA, Bb + kstepB, slabA, slabB, wave);
  if (NST > 2) oz_issue<Cfg, S>(st2, Ab + 2 * kstepA, Bb + 2 * kstepB, slabA, slabB, wave);
  if (NST > 3) oz_issue<Cfg, S>(st3, Ab + 3 * kstepA, Bb + 3 * kstepB, slabA, slabB, wave);
  for (int t = 0; t < nk; t += NST) {
    int k = min(t + NST, nk - 1);
    oz_step<Cfg, S, NST>(st0, acc, Ab + k * kstepA, Bb + k * kstepB, slabA, slabB, wm, wn, lane, wave);
    k = min(t + NST + 1, nk - 1);
    oz_step<Cfg, S, NST>(st1, acc, Ab + k * kstepA, Bb + k * kstepB, slabA, slabB, wm, wn, lane, wave);
    if (NST > 2) {
      k = min(t + NST + 2, nk - 1);
      oz_step<Cfg, S, NST>(st2, acc, Ab + k * kstepA, Bb + k * kstepB, slabA, slabB, wm, wn, lane, wave);
    }
    if (NST > 3) {
      k = min(t + NST + 3, nk - 1);
      oz_step<Cfg, S, NST>(st3, acc, Ab + k * kstepA, Bb + k * kstepB, slabA, slabB, wm, wn, lane, wave);
    }
  }
  __builtin_amdgcn_s_waitcnt(oz_vmcnt(0));
  __builtin_amdgcn_s_barrier();

  // Epilogue, per 32-row half `ta` of the wavefront's 64 x 32 C tile: the 16
  // C values of a lane are loaded together (nothing stored in between, so
  // they are all in flight at once; a load after a store to the same array
  // would wait for it), the next half's loads are issued as soon as this
  // half's accumulators are folded, then the stores follow. The levels are
  // folded (FP64, l ascending) in the MFMA layout and staged through LDS
  // unscaled, with each column's exponent beside them, so that after the
  // transpose lane (row r, columns c) scales by 2^(ea_i + eb_j) in one ldexp
  // and stores C -= value over 32 consecutive rows of a column.
  const int wb = wave % NST, ws = wave / NST;
  signed char* ebuf = wb == 0 ? st0 : wb == 1 ? st1 : wb == 2 ? st2 : st3;
  double* epi = reinterpret_cast<double*>(ebuf) + ws * (32 * 33 + 32);
  int* ecol = reinterpret_cast<int*>(epi + 32 * 33);
  const int rl = lane & 31, cph = lane >> 5;
  if (cph == 0) ecol[rl] = eb[bn + wn + rl];  // bn + wn + rl < np
  double cv[2][16];
  auto load_c = [&](int ta) {
    const int i = bm + wm + 32 * ta + rl;
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const int j = bn + wn + cph + 2 * c;
      cv[ta][c] = (i < m && j < col0 + ncols) ? C[i + (int64_t)j * ldc] : 0.0;
    }
  };
  auto fold = [&](int ta) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int il = 8 * (r >> 2) + 4 * cph + (r & 3);
      double c = 0.0;
#pragma unroll
      for (int l = 0; l < S; ++l) c += (double)acc[l][ta][r] * oz_level_scale(l);
      epi[rl * 33 + il] = c;
    }
  };
  auto store_c = [&](int ta) {
    const int i = bm + wm + 32 * ta + rl;
    const int ei = ea[i];  // i < mp
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const int jc = cph + 2 * c, j = bn + wn + jc;
      const double v = ldexp(epi[jc * 33 + rl], ei + ecol[jc]);
      if (i < m && j < col0 + ncols) C[i + (int64_t)j * ldc] = cv[ta][c] - v;
    }
  };
  auto lds_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  load_c(0);
  fold(0);
  load_c(1);
  lds_sync();
  store_c(0);
  lds_sync();
  fold(1);
  lds_sync();
  store_c(1);
}

// ---------------------------------------------------------------------------
// Single-barrier pipelined variant (the default). WM x WN wavefronts, each
// owning TA x TB MFMA tiles of 32 x 32 (wavefront tile 32 TA x 32 TB,
// 16 S TA TB accumulator registers per lane, held in the accumulation
// registers across the loop); block tile (32 TA WM) x (32 TB WN); NST LDS
// stages (2 or 4) and OCC workgroups per CU (the default shape: 2 x 2
// wavefronts of 64 x 32, two stages, two workgroups). Per K step (32 bytes of
// every row and column of all S slices) the wavefronts synchronise once:
//   multiply the first slice pair of step k (fragments in registers); wait
//   for the copies of step k + 1 and for this wavefront's reads of step k's
//   fragments; s_barrier; refill the stage of step k with step k + NST;
//   multiply the remaining pairs while reading step k + 1's fragments from
//   LDS into the registers each pair frees.
// The pair order (level i ascending, j descending) frees B fragment j after
// its pair with i = S - 1 - j and A fragment i after its group, so the reads
// of the next step go out as early as the registers allow and their latency
// hides under the remaining matrix-core work.
template <int S, int WM, int WN, int TA, int TB, int NST_, int OCC>
struct OzPCfg {
  static constexpr int NT = 64 * WM * WN, BM = 32 * TA * WM, BN = 32 * TB * WN, NST = NST_;
  static constexpr int ABYTES = S * BM * OZ_BK, BBYTES = S * BN * OZ_BK;
  static constexpr int STAGE = ABYTES + BBYTES;
  // copy rounds of an operand with R records per slice (2 R chunks of 16 B)
  static constexpr int rounds(int R) {
    return NT >= 2 * R ? (S + NT / (2 * R) - 1) / (NT / (2 * R)) : S * (2 * R / NT);
  }
  static constexpr int COPIES = rounds(BM) + rounds(BN);  // per thread and stage
  static constexpr int EPI_WAVE = (32 * 33 + 32) * 8;  // one 32 x 32 FP64 tile + its exponents
  static_assert(NST == 2 || NST == 4, "stages");
  static_assert(OCC * NST * STAGE <= 160 * 1024, "LDS");
  static_assert(NT % (2 * BM) == 0 || (2 * BM) % NT == 0, "copy split (A)");
  static_assert(NT % (2 * BN) == 0 || (2 * BN) % NT == 0, "copy split (B)");
  static_assert(2 * COPIES <= 63, "vmcnt");
};

// One K step of one operand (R records per slice) into a stage: thread tid
// copies 16-byte chunks; in LDS a slice is [k half][record][16 bytes] and
// the chunk at position c is half c / R of record c % R, so each wavefront's
// 64 copies land in 1 KB of consecutive LDS (global_load_lds writes lane l at
// 16 l) and the fragment reads cover contiguous 512-byte runs per half
// wavefront (no bank conflicts). src: the tile's first record of slice 0.
// src and slab are workgroup-uniform and the per-lane part is a 32-bit
// offset, so the copies use a scalar base address.
template <int S, int NT, int R>
__device__ __forceinline__ void ozp_issue_op(signed char* __restrict__ dst, const signed char* src,
                                             size_t slab, int tid, int wave) {
  if constexpr (NT >= 2 * R) {
    constexpr int SPR = NT / (2 * R), RND = (S + SPR - 1) / SPR;
    const int sub = wave / (2 * R / 64);  // slice within the round (wavefront-uniform)
    const int c = tid % (2 * R);
    const unsigned so = (c % R) * OZ_BK + (c / R) * 16;
    const int lo = ((64 * wave) % (2 * R)) * 16;
#pragma unroll
    for (int rd = 0; rd < RND; ++rd) {
      // a round past the last slice (S not a multiple of SPR) copies slice
      // S - 1 once more into its own slot: the same bytes twice
      const int q = rd * SPR + sub < S ? rd * SPR + sub : S - 1;
      __builtin_amdgcn_global_load_lds(src + q * slab + so, dst + q * R * OZ_BK + lo, 16, 0, 0);
    }
  } else {
    constexpr int RPS = 2 * R / NT;  // rounds per slice
#pragma unroll
    for (int rd = 0; rd < S * RPS; ++rd) {
      const int q = rd / RPS;
      const unsigned c = (rd % RPS) * NT + tid;
      __builtin_amdgcn_global_load_lds(src + q * slab + ((c % R) * OZ_BK + (c / R) * 16),
                                       dst + q * R * OZ_BK + ((rd % RPS) * NT + 64 * wave) * 16,
                                       16, 0, 0);
    }
  }
}

template <class Cfg, int S>
__device__ __forceinline__ void ozp_issue(signed char* __restrict__ st, const signed char* a,
                                          const signed char* b, size_t slabA, size_t slabB,
                                          int tid) {
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  ozp_issue_op<S, Cfg::NT, Cfg::BM>(st, a, slabA, tid, wave);
  ozp_issue_op<S, Cfg::NT, Cfg::BN>(st + Cfg::ABYTES, b, slabB, tid, wave);
}

// LDS address (32-bit) of a pointer into a __shared__ array.
__device__ __forceinline__ uint32_t oz_lds_addr(const signed char* p) {
  return static_cast<uint32_t>(
      reinterpret_cast<uintptr_t>((const __attribute__((address_space(3))) signed char*)p));
}
// Fragment reads of one K step (6 or 9 ds_read_b128) issued together and
// waited for in the same asm statement. The compiler's own wait insertion
// would place a vmcnt(0) before any LDS read while ordinary global loads and
// LDS-DMA copies are both in flight (it cannot order the two kinds), which
// would make the tile-start C loads block the first K steps; these reads
// follow the explicit stage waits instead.
__device__ __forceinline__ void oz_ds_read6(v4i& r0, v4i& r1, v4i& r2, v4i& r3, v4i& r4, v4i& r5,
                                            uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3,
                                            uint32_t a4, uint32_t a5) {
  asm volatile(
      "ds_read_b128 %0, %6\n\t"
      "ds_read_b128 %1, %7\n\t"
      "ds_read_b128 %2, %8\n\t"
      "ds_read_b128 %3, %9\n\t"
      "ds_read_b128 %4, %10\n\t"
      "ds_read_b128 %5, %11\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(r0), "=&v"(r1), "=&v"(r2), "=&v"(r3), "=&v"(r4), "=&v"(r5)
      : "v"(a0), "v"(a1), "v"(a2), "v"(a3), "v"(a4), "v"(a5)
      : "memory");
}
__device__ __forceinline__ void oz_ds_read9(v4i& r0, v4i& r1, v4i& r2, v4i& r3, v4i& r4, v4i& r5,
                                            v4i& r6, v4i& r7, v4i& r8, uint32_t a0, uint32_t a1,
                                            uint32_t a2, uint32_t a3, uint32_t a4, uint32_t a5,
                                            uint32_t a6, uint32_t a7, uint32_t a8) {
  asm volatile(
      "ds_read_b128 %0, %9\n\t"
      "ds_read_b128 %1, %10\n\t"
      "ds_read_b128 %2, %11\n\t"
      "ds_read_b128 %3, %12\n\t"
      "ds_read_b128 %4, %13\n\t"
      "ds_read_b128 %5, %14\n\t"
      "ds_read_b128 %6, %15\n\t"
      "ds_read_b128 %7, %16\n\t"
      "ds_read_b128 %8, %17\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(r0), "=&v"(r1), "=&v"(r2), "=&v"(r3), "=&v"(r4), "=&v"(r5), "=&v"(r6),
        "=&v"(r7), "=&v"(r8)
      : "v"(a0), "v"(a1), "v"(a2), "v"(a3), "v"(a4), "v"(a5), "v"(a6), "v"(a7), "v"(a8)
      : "memory");
}
// All fragments of a K step from stage st (TA = 2, TB = 1, S = 2 or 3).
template <class Cfg, int S, int TA, int TB>
__device__ __forceinline__ void oz_read_step_asm(const signed char* st, v4i (&fa)[S][TA],
                                                 v4i (&fb)[S][TB], int offA, int offB) {
  static_assert(TA == 2 && TB == 1 && (S == 2 || S == 3), "asm fragment reads");
  const uint32_t base = oz_lds_addr(st);
  auto A = [&](int i, int ta) { return base + i * Cfg::BM * OZ_BK + offA + ta * 32 * 16; };
  auto B = [&](int j) { return base + Cfg::ABYTES + j * Cfg::BN * OZ_BK + offB; };
  if constexpr (S == 2)
    oz_ds_read6(fb[0][0], fb[1][0], fa[0][0], fa[0][1], fa[1][0], fa[1][1], B(0), B(1), A(0, 0),
                A(0, 1), A(1, 0), A(1, 1));
  else
    oz_ds_read9(fb[0][0], fb[1][0], fb[2][0], fa[0][0], fa[0][1], fa[1][0], fa[1][1], fa[2][0],
                fa[2][1], B(0), B(1), B(2), A(0, 0), A(0, 1), A(1, 0), A(1, 1), A(2, 0), A(2, 1));
}

// One K step: st_cur's stage is refilled with step k + NST (tile pointers a, b
// at that step); the fragments of step k are in (fa, fb); those of step k + 1
// are read from st_nxt. sched_barrier keeps the compiler from moving the
// next step's reads back next to their uses or the barrier ahead of the
// first pair.
// VMN: the wait count before the barrier (-1: the copies of the NST - 2 later
// steps stay in flight; more when other loads were issued after step k + 1's
// copies).
// ASMRD: the next step's fragments are read after all of this step's
// products, by oz_read_step_asm (used while the tile-start C loads may be in
// flight).
template <class Cfg, int S, int TA, int TB, bool SWAP = false, int VMN = -1, bool ASMRD = false>
__device__ __forceinline__ void ozp_step(signed char* __restrict__ st_cur,
                                         const signed char* __restrict__ st_nxt,
                                         v16i (&acc)[S][TA][TB], v4i (&fa)[S][TA],
                                         v4i (&fb)[S][TB], const signed char* a,
                                         const signed char* b, size_t slabA, size_t slabB,
                                         int offA, int offB, int tid) {
#pragma unroll
  for (int i = 0; i < S; ++i) {
#pragma unroll
    for (int j = S - 1 - i; j >= 0; --j) {
#pragma unroll
      for (int ta = 0; ta < TA; ++ta)
#pragma unroll
        for (int tb = 0; tb < TB; ++tb)
          acc[i + j][ta][tb] =
              SWAP ? __builtin_amdgcn_mfma_i32_32x32x32_i8(fb[j][tb], fa[i][ta], acc[i + j][ta][tb],
                                                           0, 0, 0)
                   : __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[i][ta], fb[j][tb], acc[i + j][ta][tb],
                                                           0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
      if (i == 0 && j == S - 1) {
        // copies of step k + 1 landed (those of the later steps stay in
        // flight), and this wavefront's reads of step k's fragments are done;
        // after the barrier every wavefront's are, so st_cur is free
        __builtin_amdgcn_s_waitcnt(oz_vmcnt(VMN >= 0 ? VMN : (Cfg::NST - 2) * Cfg::COPIES) &
                                   OZ_LGKM0);
        __builtin_amdgcn_s_barrier();
        ozp_issue<Cfg, S>(st_cur, a, b, slabA, slabB, tid);
      }
      if (!ASMRD && j == S - 1 - i) {  // last use of B fragment j
#pragma unroll
        for (int tb = 0; tb < TB; ++tb)
          fb[j][tb] = *reinterpret_cast<const v4i*>(st_nxt + Cfg::ABYTES + j * Cfg::BN * OZ_BK +
                                                    offB + tb * 32 * 16);
      }
      if (!ASMRD && j == 0) {  // last use of A fragment i
#pragma unroll
        for (int ta = 0; ta < TA; ++ta)
          fa[i][ta] = *reinterpret_cast<const v4i*>(st_nxt + i * Cfg::BM * OZ_BK + offA +
                                                    ta * 32 * 16);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  if constexpr (ASMRD) oz_read_step_asm<Cfg, S, TA, TB>(st_nxt, fa, fb, offA, offB);
}

// Kp: a multiple of 4 K steps (128 bytes). The stage arrays (SB bytes each,
// at least Cfg::STAGE; one array per stage, so the compiler tells the stages
// apart and waits only for the copies into the stage being read) belong to
// the calling kernel, which may size them for more slices than S.
// EPI 0: after the K loop, C is read, updated through an LDS transpose and
// written (MFMA operands in the L-panel, U-panel order). EPI 1: operands
// swapped (U-panel slices first, so C's rows run along the lanes) and C read
// into registers at the start of the tile (see below).
template <int S, int WM, int WN, int TA, int TB, int NST, int OCC, int SB, int EPI = 0>
__device__ __forceinline__ void ozp_body(int m, int n, int mp, int np, int Kp, int col0, int ncols,
                                         const signed char* __restrict__ As,
                                         const signed char* __restrict__ Bs,
                                         const int* __restrict__ ea, const int* __restrict__ eb,
                                         double* __restrict__ C, int64_t ldc,
                                         const double* __restrict__ Ci, int64_t ldci, int group,
                                         signed char* __restrict__ st0,
                                         signed char* __restrict__ st1,
                                         signed char* __restrict__ st2,
                                         signed char* __restrict__ st3) {
  typedef OzPCfg<S, WM, WN, TA, TB, NST, OCC> Cfg;
  static_assert(SB >= Cfg::STAGE, "stage arrays");
  constexpr int WPB = SB / Cfg::EPI_WAVE;  // wavefronts staged per stage array in the epilogue
  static_assert(WPB >= 1 && WM * WN <= NST * WPB, "epilogue staging");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave % WM) * 32 * TA, wn = (wave / WM) * 32 * TB;
  int ti, tj;
  if (!tile_of(blockIdx.x, (m + Cfg::BM - 1) / Cfg::BM, (ncols + Cfg::BN - 1) / Cfg::BN, ti, tj, group)) return;
  const int bm = ti * Cfg::BM, bn = col0 + tj * Cfg::BN;
  const size_t slabA = (size_t)mp * Kp, slabB = (size_t)np * Kp;
  const size_t kstepA = (size_t)mp * OZ_BK, kstepB = (size_t)np * OZ_BK;
  const signed char* Ab = As + (size_t)bm * OZ_BK;
  const signed char* Bb = Bs + (size_t)bn * OZ_BK;
  // fragment reads: lane takes record r32 of its 32-tile, k half hf
  const int r32 = lane & 31, hf = lane >> 5;
  const int offA = hf * Cfg::BM * 16 + (wm + r32) * 16;
  const int offB = hf * Cfg::BN * 16 + (wn + r32) * 16;

  v16i acc[S][TA][TB];
#pragma unroll
  for (int l = 0; l < S; ++l)
#pragma unroll
    for (int ta = 0; ta < TA; ++ta)
#pragma unroll
      for (int tb = 0; tb < TB; ++tb) acc[l][ta][tb] = v16i(0);

  const int nk = Kp / OZ_BK;  // a multiple of 4
  ozp_issue<Cfg, S>(st0, Ab, Bb, slabA, slabB, tid);
  ozp_issue<Cfg, S>(st1, Ab + kstepA, Bb + kstepB, slabA, slabB, tid);
  if constexpr (NST > 2) {
    ozp_issue<Cfg, S>(st2, Ab + 2 * kstepA, Bb + 2 * kstepB, slabA, slabB, tid);
    ozp_issue<Cfg, S>(st3, Ab + 3 * kstepA, Bb + 3 * kstepB, slabA, slabB, tid);
  }
  // EPI 1: the tile's C values are read now, behind the first stages'
  // copies, into the registers they are updated in (swapped operands: lane
  // = row), so the FP64 read of C overlaps the whole K loop and the
  // write-back only stores. The copies of stage 0 stay the oldest loads, so
  // the first wait below still covers them.
  constexpr bool PRE = EPI == 1;
  double cpre[PRE ? TA : 1][PRE ? TB : 1][16];
  if constexpr (PRE) {
    const int rl = lane & 31, u = lane >> 5;
    // Every lane of every wavefront issues every load: the first wait below
    // counts them, and a load skipped by a whole wavefront (an edge tile
    // whose rows or columns it does not reach) would let that wait pass
    // before stage 0's copies land. Interior tiles load unconditionally;
    // edge tiles load from the block's first element where a value lies
    // outside the block and take 0 for it (never stored).
    const bool interior = bm + Cfg::BM <= m && bn + Cfg::BN <= col0 + ncols;  // workgroup-uniform
    if (interior) {
#pragma unroll
      for (int ta = 0; ta < TA; ++ta) {
        const int i = bm + wm + 32 * ta + rl;
#pragma unroll
        for (int tb = 0; tb < TB; ++tb) {
          const int jb = bn + wn + 32 * tb + 4 * u;
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int jo = 8 * (r >> 2) + (r & 3);
            cpre[ta][tb][r] = __builtin_nontemporal_load(Ci + i + (int64_t)(jb + jo) * ldci);
          }
        }
      }
    } else {
#pragma unroll
      for (int ta = 0; ta < TA; ++ta) {
        const int i = bm + wm + 32 * ta + rl;
#pragma unroll
        for (int tb = 0; tb < TB; ++tb) {
          const int jb = bn + wn + 32 * tb + 4 * u;
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int jo = 8 * (r >> 2) + (r & 3);
            const bool ok = i < m && jb + jo < col0 + ncols;
            const double v =
                __builtin_nontemporal_load(ok ? Ci + i + (int64_t)(jb + jo) * ldci : Ci);
            cpre[ta][tb][r] = ok ? v : 0.0;
          }
        }
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  // this thread's copies of stage 0 landed: the loads issued after them
  // (the later stages' copies and, for EPI 1, the C values) may be in flight
  __builtin_amdgcn_s_waitcnt(oz_vmcnt((NST - 1) * Cfg::COPIES + (PRE ? TA * TB * 16 : 0)));
  __builtin_amdgcn_s_barrier();
  v4i fa[S][TA], fb[S][TB];
  constexpr bool PEEL = PRE && NST == 4 && TA == 2 && TB == 1 && S <= 3;
  if constexpr (PEEL) oz_read_step_asm<Cfg, S, TA, TB>(st0, fa, fb, offA, offB);
#pragma unroll
  for (int q = 0; q < (PEEL ? 0 : S); ++q) {
#pragma unroll
    for (int ta = 0; ta < TA; ++ta)
      fa[q][ta] = *reinterpret_cast<const v4i*>(st0 + q * Cfg::BM * OZ_BK + offA + ta * 32 * 16);
#pragma unroll
    for (int tb = 0; tb < TB; ++tb)
      fb[q][tb] = *reinterpret_cast<const v4i*>(st0 + Cfg::ABYTES + q * Cfg::BN * OZ_BK + offB +
                                                tb * 32 * 16);
  }
  // past the last K step the refills repeat the last step (never read) and
  // the last step's next-fragment reads fetch a stage that is never used, so
  // the loop body is the same every time. With C read at the start of the
  // tile (EPI 1), the first NST - 1 steps leave those loads in flight too
  // (their copies were issued before them), so C streams in under the first
  // stages' products; the wait of step NST - 1 (for copies issued after the
  // C loads) takes what is left of it. nk >= 4 (Kp is a multiple of 128).
  int t0 = 0;
  if constexpr (PEEL) {
    constexpr int VP = (NST - 2) * Cfg::COPIES + TA * TB * 16;
    static_assert(VP <= 63, "vmcnt");
    ozp_step<Cfg, S, TA, TB, true, VP, true>(st0, st1, acc, fa, fb, Ab + min(4, nk - 1) * kstepA,
                                       Bb + min(4, nk - 1) * kstepB, slabA, slabB, offA, offB,
                                       tid);
    ozp_step<Cfg, S, TA, TB, true, VP, true>(st1, st2, acc, fa, fb, Ab + min(5, nk - 1) * kstepA,
                                       Bb + min(5, nk - 1) * kstepB, slabA, slabB, offA, offB,
                                       tid);
    ozp_step<Cfg, S, TA, TB, true, VP, true>(st2, st3, acc, fa, fb, Ab + min(6, nk - 1) * kstepA,
                                       Bb + min(6, nk - 1) * kstepB, slabA, slabB, offA, offB,
                                       tid);
    ozp_step<Cfg, S, TA, TB, true>(st3, st0, acc, fa, fb, Ab + min(7, nk - 1) * kstepA,
                                   Bb + min(7, nk - 1) * kstepB, slabA, slabB, offA, offB, tid);
    t0 = NST;
  }
  for (int t = t0; t < nk; t += NST) {
    int k = min(t + NST, nk - 1);
    ozp_step<Cfg, S, TA, TB, (EPI > 0)>(st0, st1, acc, fa, fb, Ab + k * kstepA, Bb + k * kstepB, slabA, slabB,
                             offA, offB, tid);
    k = min(t + NST + 1, nk - 1);
    ozp_step<Cfg, S, TA, TB, (EPI > 0)>(st1, NST > 2 ? st2 : st0, acc, fa, fb, Ab + k * kstepA,
                             Bb + k * kstepB, slabA, slabB, offA, offB, tid);
    if constexpr (NST > 2) {
      k = min(t + NST + 2, nk - 1);
      ozp_step<Cfg, S, TA, TB, (EPI > 0)>(st2, st3, acc, fa, fb, Ab + k * kstepA, Bb + k * kstepB, slabA,
                               slabB, offA, offB, tid);
      k = min(t + NST + 3, nk - 1);
      ozp_step<Cfg, S, TA, TB, (EPI > 0)>(st3, st0, acc, fa, fb, Ab + k * kstepA, Bb + k * kstepB, slabA,
                               slabB, offA, offB, tid);
    }
  }
  __builtin_amdgcn_s_waitcnt(oz_vmcnt(0) & OZ_LGKM0);
  __builtin_amdgcn_s_barrier();

  if constexpr (PRE) {
    const int rl = lane & 31, u = lane >> 5;
#pragma unroll
    for (int ta = 0; ta < TA; ++ta) {
      const int i = bm + wm + 32 * ta + rl;
      const int ei = ea[i];  // i < mp
#pragma unroll
      for (int tb = 0; tb < TB; ++tb) {
        const int jb = bn + wn + 32 * tb + 4 * u;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int jo = 8 * (r >> 2) + (r & 3);
          double c = 0.0;
#pragma unroll
          for (int l = 0; l < S; ++l) c += (double)acc[l][ta][tb][r] * oz_level_scale(l);
          const double v = cpre[ta][tb][r] - ldexp(c, ei + eb[jb + jo]);  // jb + jo < np
          if (i < m && jb + jo < col0 + ncols)
            __builtin_nontemporal_store(v, C + i + (int64_t)(jb + jo) * ldc);
        }
      }
    }
    return;
  }
  // Epilogue, one 32 x 32 MFMA tile (ta, tb) at a time: the next tile's C
  // values are loaded before this tile's stores; the levels are folded (FP64,
  // l ascending) in the MFMA layout, staged through LDS unscaled with the
  // tile's column exponents, and after the transpose lane (row r, columns c)
  // scales by 2^(ea_i + eb_j) and stores C -= value down 32 consecutive rows.
  const int wb = wave / WPB, ws = wave % WPB;
  signed char* ebuf = wb == 0 ? st0 : wb == 1 ? st1 : wb == 2 ? st2 : st3;
  double* epi = reinterpret_cast<double*>(ebuf) + ws * (32 * 33 + 32);
  int* ecol = reinterpret_cast<int*>(epi + 32 * 33);
  const int rl = lane & 31, cph = lane >> 5;
  constexpr int NTILE = TA * TB;
  double cv[2][16];
  auto load_c = [&](int t, double(&dst)[16]) {
    const int ta = t / TB, tb = t % TB;
    const int i = bm + wm + 32 * ta + rl;
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const int j = bn + wn + 32 * tb + cph + 2 * c;
      dst[c] = (i < m && j < col0 + ncols) ? Ci[i + (int64_t)j * ldci] : 0.0;
    }
  };
  auto lds_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  load_c(0, cv[0]);
#pragma unroll
  for (int t = 0; t < NTILE; ++t) {
    const int ta = t / TB, tb = t % TB;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int il = 8 * (r >> 2) + 4 * cph + (r & 3);
      double c = 0.0;
#pragma unroll
      for (int l = 0; l < S; ++l) c += (double)acc[l][ta][tb][r] * oz_level_scale(l);
      epi[rl * 33 + il] = c;
    }
    if (cph == 0) ecol[rl] = eb[bn + wn + 32 * tb + rl];  // < np
    if (t + 1 < NTILE) load_c(t + 1, cv[(t + 1) & 1]);
    lds_sync();
    const int i = bm + wm + 32 * ta + rl;
    const int ei = ea[i];  // i < mp
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const int jc = cph + 2 * c, j = bn + wn + 32 * tb + jc;
      const double v = ldexp(epi[jc * 33 + rl], ei + ecol[jc]);
      if (i < m && j < col0 + ncols) C[i + (int64_t)j * ldc] = cv[t & 1][c] - v;
    }
    lds_sync();
  }
}

template <int S, int WM, int WN, int TA, int TB, int NST, int OCC>
__global__ __launch_bounds__(64 * WM * WN, OCC) void ozaki_gemmp_kernel(
    int m, int n, int mp, int np, int Kp, int col0, int ncols, const signed char* __restrict__ As,
    const signed char* __restrict__ Bs, const int* __restrict__ ea, const int* __restrict__ eb,
    double* __restrict__ C, int64_t ldc, int group) {
  constexpr int SB = OzPCfg<S, WM, WN, TA, TB, NST, OCC>::STAGE;
  __shared__ __attribute__((aligned(16))) signed char st0[SB];
  __shared__ __attribute__((aligned(16))) signed char st1[SB];
  __shared__ __attribute__((aligned(16))) signed char st2[NST > 2 ? SB : 16];
  __shared__ __attribute__((aligned(16))) signed char st3[NST > 2 ? SB : 16];
  ozp_body<S, WM, WN, TA, TB, NST, OCC, SB>(m, n, mp, np, Kp, col0, ncols, As, Bs, ea, eb, C, ldc,
                                            C, ldc, group, st0, st1, st2, st3);
}

// The same with the slice count read on the device (*Sdev, 2 .. SMAX): the
// slices are those of an SMAX-slice split (only the leading *Sdev digits
// are used), so one launch serves whatever count the adaptive rule picked.
template <int SMAX, int WM, int WN, int TA, int TB, int NST, int OCC, int EPI = 0>
__global__ __launch_bounds__(64 * WM * WN, OCC) void ozaki_gemmd_kernel(
    int m, int n, int mp, int np, int Kp, int col0, int ncols, const signed char* __restrict__ As,
    const signed char* __restrict__ Bs, const int* __restrict__ ea, const int* __restrict__ eb,
    double* __restrict__ C, int64_t ldc, const double* __restrict__ Ci, int64_t ldci, int group,
    const int* __restrict__ Sdev) {
  typedef OzPCfg<SMAX, WM, WN, TA, TB, NST, OCC> Cfg;
  // a stage array also stages the epilogue's FP64 tiles (EPI 0)
  constexpr int SBE = (WM * WN + NST - 1) / NST * Cfg::EPI_WAVE;
  constexpr int SB = Cfg::STAGE > SBE ? Cfg::STAGE : SBE;
  static_assert(OCC * NST * SB <= 160 * 1024, "LDS");
  __shared__ __attribute__((aligned(16))) signed char st0[SB];
  __shared__ __attribute__((aligned(16))) signed char st1[SB];
  __shared__ __attribute__((aligned(16))) signed char st2[NST > 2 ? SB : 16];
  __shared__ __attribute__((aligned(16))) signed char st3[NST > 2 ? SB : 16];
  const int S = __builtin_amdgcn_readfirstlane(*Sdev);
  // the C prefetch (EPI 1) holds 32 TA TB more registers: S <= 3 only
  if (S <= 2)
    ozp_body<2, WM, WN, TA, TB, NST, OCC, SB, EPI>(m, n, mp, np, Kp, col0, ncols, As, Bs, ea, eb, C,
                                              ldc, Ci, ldci, group, st0, st1, st2, st3);
  else if (SMAX > 3 && S == 3)
    ozp_body<3, WM, WN, TA, TB, NST, OCC, SB, EPI>(m, n, mp, np, Kp, col0, ncols, As, Bs, ea, eb, C,
                                              ldc, Ci, ldci, group, st0, st1, st2, st3);
  else if (SMAX > 4 && S == 4)
    ozp_body<(SMAX > 4 ? 4 : SMAX), WM, WN, TA, TB, NST, OCC, SB, 0>(
        m, n, mp, np, Kp, col0, ncols, As, Bs, ea, eb, C, ldc, Ci, ldci, group, st0, st1, st2, st3);
  else
    ozp_body<SMAX, WM, WN, TA, TB, NST, OCC, SB, (SMAX > 3 ? 0 : EPI)>(m, n, mp, np, Kp, col0, ncols, As, Bs, ea, eb, C,
                                                 ldc, Ci, ldci, group, st0, st1, st2, st3);
}

// Slice count of one trailing update from the exponent maxima of its panels:
// E = max_i eA_i + max_j eB_j bounds the magnitude of every product term, and
// the error of an S-slice product is ~2^(E - (8 S - 1)). The first update of
// a factorisation (first != 0) records E0 and takes SMAX slices; a later one
// takes the fewest slices S in [smin, SMAX] whose error bound stays within
// the first update's minus `delta` bits:
//     E - (8 S - 1) <= E0 - (8 SMAX - 1) - delta,
// i.e. S = SMAX + ceil((E - E0 + delta) / 8), clamped. Rows and columns that
// are all zero carry the exponent OZ_ZERO_EXP and never decide the maximum
// unless everything is zero.
__global__ __launch_bounds__(1024) void oz_decide_kernel(const int* __restrict__ ea, int m,
                                                         const int* __restrict__ eb, int n,
                                                         int smax, int smin, int delta, int first,
                                                         int* __restrict__ e0,
                                                         int* __restrict__ s_out) {
  __shared__ int red[2][16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int a = OZ_ZERO_EXP, b = OZ_ZERO_EXP;
  for (int i = t; i < m; i += 1024) a = max(a, ea[i]);
  for (int j = t; j < n; j += 1024) b = max(b, eb[j]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a = max(a, __shfl_xor(a, off, 64));
    b = max(b, __shfl_xor(b, off, 64));
  }
  if (lane == 0) {
    red[0][w] = a;
    red[1][w] = b;
  }
  __syncthreads();
  if (t == 0) {
    for (int q = 1; q < 16; ++q) {
      a = max(a, red[0][q]);
      b = max(b, red[1][q]);
    }
    const int E = a + b;
    int S = smax;
    if (first) {
      *e0 = E;
    } else {
      const int d = E - *e0 + delta;
      S = smax + (d >= 0 ? (d + 7) / 8 : -((-d) / 8));
      S = S < smin ? smin : (S > smax ? smax : S);
    }
    *s_out = S;
  }
}

// Slice arrays: rows padded to a multiple of 128; columns to a multiple of 128
// plus 128 more, so that a column range starting at any multiple of 64 is
// covered by whole tiles of up to 128 columns.
constexpr int OZ_PAD = 128;
static int oz_mp(int m) { return (m + OZ_PAD - 1) / OZ_PAD * OZ_PAD; }
static int oz_np(int n) { return (n + OZ_PAD - 1) / OZ_PAD * OZ_PAD + OZ_PAD; }
static int oz_kp(int K) { return (K + 4 * OZ_BK - 1) / (4 * OZ_BK) * (4 * OZ_BK); }  // whole rounds of 2 or 4 stages

size_t ozaki_workspace_bytes(int m, int n, int K, int S) {
  const size_t mp = oz_mp(m), np = oz_np(n), Kp = oz_kp(K);
  return (size_t)S * Kp * (mp + np) + sizeof(int) * (mp + np) + sizeof(double) * (Kp / 64) * mp +
         256;
}

struct OzLayout {
  int mp, np, Kp;
  signed char *As, *Bs;
  int *ea, *eb;
  double* pm;  // partial row maxima of the 2-D row slicing, (Kp / 64) x mp
};
static OzLayout oz_layout(int S, int m, int n, int K, void* work) {
  OzLayout L;
  L.mp = oz_mp(m);
  L.np = oz_np(n);
  L.Kp = oz_kp(K);
  L.As = static_cast<signed char*>(work);
  L.Bs = L.As + (size_t)S * L.mp * L.Kp;
  L.ea = reinterpret_cast<int*>(L.Bs + (size_t)S * L.np * L.Kp);
  L.eb = L.ea + L.mp;
  L.pm = reinterpret_cast<double*>(L.eb + L.np);  // (mp + np) ints: a multiple of 8 bytes
  return L;
}

template <int S>
static int ozaki_slice_s(hipStream_t s, int m, int n, int K, const double* A, int64_t lda,
                         const double* B, int64_t ldb, void* work, int parts) {
  const OzLayout L = oz_layout(S, m, n, K, work);
  if (parts & 1) {
    if (tuning_get(TUNE_SLICE_ROWS) == 1) {
      slice_rows_kernel<S><<<L.mp / 64, 256, 0, s>>>(A, lda, m, K, L.mp, L.Kp, L.As, L.ea);
    } else {
      const dim3 g(L.mp / 64, L.Kp / 64);
      rowmax_part_kernel<<<g, 256, 0, s>>>(A, lda, m, K, L.mp, L.pm);
      slice_rows_tile_kernel<S><<<g, 256, 0, s>>>(A, lda, m, K, L.mp, L.Kp, L.pm, L.As, L.ea);
    }
  }
  if (parts & 2) {
    if (L.Kp <= OZ_SC_KMAX)
      slice_cols_lds_kernel<S><<<L.np / 4, 256, 0, s>>>(B, ldb, n, K, L.np, L.Kp, L.Bs, L.eb);
    else
      slice_cols_kernel<S><<<L.np / 4, 256, 0, s>>>(B, ldb, n, K, L.np, L.Kp, L.Bs, L.eb);
  }
  MPDOK_HIP(hipGetLastError());
  return OK;
}

template <int S, int WM, int WN, int NST, int OCC>
static int ozaki_launch(hipStream_t s, const OzLayout& L, int m, int n, int col0, int ncols,
                        double* C, int64_t ldc) {
  typedef OzCfg<S, WM, WN, NST, OCC> Cfg;
  if constexpr (!Cfg::FITS) {
    return ozaki_launch<S, 2, 2, 2, 1>(s, L, m, n, col0, ncols, C, ldc);  // stages too large: smallest shape
  } else {
    const int grid = tiles_launch(L.mp / Cfg::BM, (ncols + Cfg::BN - 1) / Cfg::BN);
    const int g = tuning_get(TUNE_OZAKI_GROUP);
    ozaki_gemm_kernel<S, WM, WN, NST, OCC><<<grid, Cfg::NT, 0, s>>>(
        m, n, L.mp, L.np, L.Kp, col0, ncols, L.As, L.Bs, L.ea, L.eb, C, ldc, g > 0 ? g : GROUP_M);
    MPDOK_HIP(hipGetLastError());
    return OK;
  }
}

template <int S, int WM, int WN, int TA, int TB, int NST, int OCC>
static int ozaki_launchp(hipStream_t s, const OzLayout& L, int m, int n, int col0, int ncols,
                         double* C, int64_t ldc) {
  typedef OzPCfg<S, WM, WN, TA, TB, NST, OCC> Cfg;
  const int grid = tiles_launch(L.mp / Cfg::BM, (ncols + Cfg::BN - 1) / Cfg::BN);
  const int g = tuning_get(TUNE_OZAKI_GROUP);
  ozaki_gemmp_kernel<S, WM, WN, TA, TB, NST, OCC><<<grid, Cfg::NT, 0, s>>>(
      m, n, L.mp, L.np, L.Kp, col0, ncols, L.As, L.Bs, L.ea, L.eb, C, ldc, g > 0 ? g : GROUP_M);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

template <int S>
static int ozaki_apply_s(hipStream_t s, int m, int n, int K, int col0, int ncols, double* C,
                         int64_t ldc, void* work) {
  const OzLayout L = oz_layout(S, m, n, K, work);
  // the single-barrier pipelined kernel with 128 x 64 block tiles (4
  // wavefronts of 64 x 32), two stages, two workgroups per CU; S = 6, 7 (too
  // many accumulators for it) use the two-barrier kernel with two workgroups
  // per CU (S = 6) or one (S = 7)
  if constexpr (S <= 5) return ozaki_launchp<S, 2, 2, 2, 1, 2, 2>(s, L, m, n, col0, ncols, C, ldc);
  else if constexpr (S == 6) return ozaki_launch<S, 2, 2, 2, 2>(s, L, m, n, col0, ncols, C, ldc);
  else return ozaki_launch<S, 2, 2, 2, 1>(s, L, m, n, col0, ncols, C, ldc);  // 7 S accumulators
}

template <int SMAX, int WM, int WN, int TA, int TB, int NST, int OCC, int EPI>
static int oz_dyn_launch(hipStream_t s, const OzLayout& L, int m, int n, int col0, int ncols,
                         double* C, int64_t ldc, const double* Ci, int64_t ldci, int group,
                         const int* Sdev) {
  typedef OzPCfg<SMAX, WM, WN, TA, TB, NST, OCC> Cfg;
  const int grid = tiles_launch((m + Cfg::BM - 1) / Cfg::BM, (ncols + Cfg::BN - 1) / Cfg::BN);
  ozaki_gemmd_kernel<SMAX, WM, WN, TA, TB, NST, OCC, EPI><<<grid, Cfg::NT, 0, s>>>(
      m, n, L.mp, L.np, L.Kp, col0, ncols, L.As, L.Bs, L.ea, L.eb, C, ldc, Ci, ldci, group, Sdev);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

// sel = (smax - 3) * 2 + epi
template <int WM, int WN, int TA, int TB, int NST, int OCC>
static int oz_dyn_pick(hipStream_t s, int sel, const OzLayout& L, int m, int n, int col0, int ncols,
                       double* C, int64_t ldc, const double* Ci, int64_t ldci, int group,
                       const int* Sdev) {
#define OZ_DYN(SM, EP)                                                                         \
  oz_dyn_launch<SM, WM, WN, TA, TB, NST, OCC, EP>(s, L, m, n, col0, ncols, C, ldc, Ci, ldci, group, \
                                                  Sdev)
  switch (sel) {
    case 0: return OZ_DYN(3, 0);
    case 1: return OZ_DYN(3, 1);
    case 2: return OZ_DYN(4, 0);
    case 3: return OZ_DYN(4, 1);
    case 5: return OZ_DYN(5, 1);
    default: return OZ_DYN(5, 0);
  }
#undef OZ_DYN
}

// C[:, col0 : col0 + ncols] -= the product of the leading *Sdev slices of the
// last ozaki_slice call with smax slices (smax <= 5; *Sdev read on the device).
// With Ci not null the values subtracted from are read from Ci (leading
// dimension ldci, the same block) and the results written to C: the first
// update of a factorisation reads the input matrix in place of a copy of it.
int ozaki_apply_dyn(hipStream_t s, int smax, int m, int n, int K, int row0, int nrows, int col0,
                    int ncols, double* C, int64_t ldc, void* work, const int* Sdev,
                    const double* Ci, int64_t ldci) {
  if (nrows <= 0 || ncols <= 0 || K <= 0) return OK;
  if (Ci == nullptr) {
    Ci = C;
    ldci = ldc;
  }
  MPDOK_REQUIRE(ldci >= m, "ozaki_apply_dyn: ld of the input C %lld < m = %d", (long long)ldci, m);
  MPDOK_REQUIRE(col0 % 64 == 0 && col0 + ncols <= n, "ozaki_apply_dyn: bad column range");
  MPDOK_REQUIRE(row0 % 128 == 0 && row0 >= 0 && row0 + nrows <= m, "ozaki_apply_dyn: bad row range");
  MPDOK_REQUIRE(smax >= 3 && smax <= 5, "ozaki_apply_dyn: smax must be in [3, 5], got %d", smax);
  OzLayout L = oz_layout(smax, m, n, K, work);
  // rows [row0, row0 + nrows): the K-blocked slices keep a K step's rows
  // contiguous (stride mp), so the block starts row0 records in
  L.As += (size_t)row0 * OZ_BK;
  L.ea += row0;
  C += row0;
  Ci += row0;
  m = nrows;
  const int g = tuning_get(TUNE_OZAKI_GROUP);
  const int group = g > 0 ? g : GROUP_M;
  // Block shapes (tuning key 0): 0 (default) 128 x 128 tiles, 8 wavefronts
  // of 64 x 32, four stages, one workgroup per CU; 1: 128 x 64 tiles, 4
  // wavefronts, two stages, two workgroups per CU; 2: 128 x 128, two stages;
  // 3: 256 x 64, 8 wavefronts, two stages. C handling (key 5): 0 (default)
  // C read into registers at the start of each tile (S <= 3; larger counts
  // read it after the K loop), 1: read after the K loop.
  const int shape = tuning_get(TUNE_OZAKI_SHAPE);
  const int epi = tuning_get(TUNE_OZAKI_EPI) == 0 ? 1 : 0;
  const int sel = (smax - 3) * 2 + epi;  // 0 .. 5
  switch (shape) {
    case 1: return oz_dyn_pick<2, 2, 2, 1, 2, 2>(s, sel, L, m, n, col0, ncols, C, ldc, Ci, ldci, group, Sdev);
    case 2: return oz_dyn_pick<2, 4, 2, 1, 2, 1>(s, sel, L, m, n, col0, ncols, C, ldc, Ci, ldci, group, Sdev);
    case 3: return oz_dyn_pick<4, 2, 2, 1, 2, 1>(s, sel, L, m, n, col0, ncols, C, ldc, Ci, ldci, group, Sdev);
    default: return oz_dyn_pick<2, 4, 2, 1, 4, 1>(s, sel, L, m, n, col0, ncols, C, ldc, Ci, ldci, group, Sdev);
  }
}

// Slice count for the product of the last ozaki_slice call (smax slices,
// m rows, n columns) into *s_out; first != 0 records the reference exponent
// sum in *e0 (see oz_decide_kernel).
int ozaki_decide(hipStream_t s, int smax, int m, int n, int K, void* work, int smin, int delta,
                 int first, int* e0, int* s_out) {
  const OzLayout L = oz_layout(smax, m, n, K, work);
  MPDOK_REQUIRE(smin >= 2 && smin <= smax, "ozaki_decide: smin=%d outside [2, %d]", smin, smax);
  oz_decide_kernel<<<1, 1024, 0, s>>>(L.ea, m, L.eb, n, smax, smin, delta, first, e0, s_out);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

// Slice A (m x K) by rows (parts & 1) and B (K x n) by columns (parts & 2)
// into the workspace laid out for (S, m, n, K).
int ozaki_slice(hipStream_t s, int S, int m, int n, int K, const double* A, int64_t lda,
                const double* B, int64_t ldb, void* work, int parts) {
  if (m <= 0 || n <= 0 || K <= 0) return OK;
  // int32 accumulators: a level sums at most S products of K terms of |d d'| <= 2^14
  MPDOK_REQUIRE((int64_t)S * K * 16384 <= INT32_MAX, "ozaki_slice: K=%d too long for %d slices", K,
                S);
  switch (S) {
    case 3: return ozaki_slice_s<3>(s, m, n, K, A, lda, B, ldb, work, parts);
    case 4: return ozaki_slice_s<4>(s, m, n, K, A, lda, B, ldb, work, parts);
    case 5: return ozaki_slice_s<5>(s, m, n, K, A, lda, B, ldb, work, parts);
    case 6: return ozaki_slice_s<6>(s, m, n, K, A, lda, B, ldb, work, parts);
    case 7: return ozaki_slice_s<7>(s, m, n, K, A, lda, B, ldb, work, parts);
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

// C -= the product of the leading s_fixed of the smax digits the last
// mpdok_ozaki_gemm_sub_adaptive call with the same (smax, m, n, K, work) cut,
// without slicing again: times the trailing-update kernel alone.
// dev_scratch: 2 device ints (the count is written to dev_scratch[1]).
MPDOK_API int mpdok_ozaki_apply_sliced(hipStream_t s, int smax, int m, int n, int K, double* C,
                                       int64_t ldc, void* work, int s_fixed, int* dev_scratch) {
  MPDOK_REQUIRE(s_fixed >= 2 && s_fixed <= smax, "ozaki_apply_sliced: s_fixed=%d outside [2, %d]",
                s_fixed, smax);
  MPDOK_HIP(hipMemcpyAsync(dev_scratch + 1, &s_fixed, sizeof(int), hipMemcpyHostToDevice, s));
  return mpdok::ozaki_apply_dyn(s, smax, m, n, K, 0, m, 0, n, C, ldc, work, dev_scratch + 1);
}

// The adaptive pieces, for parity tests: slice with smax slices, then
// C -= the product of the leading S slices, with S given (s_fixed >= 2) or
// picked by the decision rule against a reference exponent sum e0_in
// (s_fixed = 0; e0_in = INT_MIN records this call as the first update).
// *s_used receives the count taken.
MPDOK_API int mpdok_ozaki_gemm_sub_adaptive(hipStream_t s, int smax, int m, int n, int K,
                                            const double* A, int64_t lda, const double* B,
                                            int64_t ldb, double* C, int64_t ldc, void* work,
                                            int s_fixed, int smin, int delta, int e0_in,
                                            int* dev_scratch, int* s_used, int* e_ref) {
  MPDOK_TRY(mpdok::ozaki_slice(s, smax, m, n, K, A, lda, B, ldb, work));
  if (s_fixed >= 2) {
    MPDOK_REQUIRE(s_fixed <= smax, "ozaki_gemm_sub_adaptive: s_fixed=%d > smax=%d", s_fixed, smax);
    MPDOK_HIP(hipMemcpyAsync(dev_scratch + 1, &s_fixed, sizeof(int), hipMemcpyHostToDevice, s));
  } else {
    const int first = e0_in == INT32_MIN;
    if (!first)
      MPDOK_HIP(hipMemcpyAsync(dev_scratch, &e0_in, sizeof(int), hipMemcpyHostToDevice, s));
    MPDOK_TRY(mpdok::ozaki_decide(s, smax, m, n, K, work, smin, delta, first, dev_scratch,
                                  dev_scratch + 1));
  }
  MPDOK_TRY(mpdok::ozaki_apply_dyn(s, smax, m, n, K, 0, m, 0, n, C, ldc, work, dev_scratch + 1));
  int host[2];
  MPDOK_HIP(hipMemcpyAsync(host, dev_scratch, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  MPDOK_HIP(hipStreamSynchronize(s));
  if (s_used) *s_used = host[1];
  if (e_ref) *e_ref = host[0];
  return 0;
}
