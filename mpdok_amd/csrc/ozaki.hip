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
// comes from v_mfma_i32_32x32x32_i8; pairs with i + j > S + 1 are dropped. The
// epilogue folds c = sum_l acc_l 2^(-7 l) in FP64 for l ascending from 2 and
// subtracts 2^(eA_i + eB_j) c from C. Slicing, integer sums, the fold order and
// the power-of-two scalings are all deterministic, so the result is bit for bit
// the one of the oracle's ozaki_matmul.
//
// Layout of the slices (K-blocked): As[q][k / 32][row][k % 32], one 32-byte
// record per (row, 32-wide K block), so that the bytes one K step of one block
// tile needs are a single contiguous run per slice, copied straight into LDS.
//
// Kernel shape (default): 256 threads = 2 x 2 wavefronts, two workgroups per
// CU; wavefront tile 64 x 32 (two v_mfma_i32_32x32x32_i8 tiles), block tile
// 128 x 64, K step 32. One K step of all S slices of both operands is a
// pipeline stage in LDS, filled by global_load_lds_dwordx4 (global memory to
// LDS without registers) one stage ahead of its use; a stage is waited for
// with s_waitcnt vmcnt and a plain s_barrier, so the copies of the next stage
// stay in flight while the matrix cores work on the current one. All S(S+1)/2
// slice pairs of a stage are multiplied from LDS, with one int32 accumulator
// set per level (32 S accumulator registers per lane). Other block shapes
// (more wavefronts, stages, one workgroup per CU) are selectable for tuning.
#include "common.h"
#include "lu.h"

namespace mpdok {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int OZ_BM = 128, OZ_BK = 32;  // block rows, K step (bytes of one record)

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
      *reinterpret_cast<v4i*>(Bs + q * slab + oz_rec(np, j, k0)) = w;
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

// 2^(-7 (l + 2)): weight of level l (0-based) in the fold; exact power of two.
__device__ __forceinline__ constexpr double oz_level_scale(int l) {
  double w = 1.0;
  for (int i = 0; i < 7 * (l + 2); ++i) w *= 0.5;
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

// Slice arrays: rows padded to a multiple of 128; columns to a multiple of 128
// plus 128 more, so that a column range starting at any multiple of 64 is
// covered by whole tiles of up to 128 columns.
constexpr int OZ_PAD = 128;
static int oz_mp(int m) { return (m + OZ_PAD - 1) / OZ_PAD * OZ_PAD; }
static int oz_np(int n) { return (n + OZ_PAD - 1) / OZ_PAD * OZ_PAD + OZ_PAD; }
static int oz_kp(int K) { return (K + 4 * OZ_BK - 1) / (4 * OZ_BK) * (4 * OZ_BK); }  // whole rounds of 2 or 4 stages

size_t ozaki_workspace_bytes(int m, int n, int K, int S) {
  const size_t mp = oz_mp(m), np = oz_np(n), Kp = oz_kp(K);
  return (size_t)S * Kp * (mp + np) + sizeof(int) * (mp + np) + 256;
}

struct OzLayout {
  int mp, np, Kp;
  signed char *As, *Bs;
  int *ea, *eb;
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

template <int S>
static int ozaki_apply_s(hipStream_t s, int m, int n, int K, int col0, int ncols, double* C,
                         int64_t ldc, void* work) {
  const OzLayout L = oz_layout(S, m, n, K, work);
  // shape 0 (default): 128 x 64 tiles, two stages, two workgroups per CU, which
  // overlap each other's epilogue (the FP64 read-modify-write of C) and copies;
  // measured fastest at S = 5, 6 on 32256^2 x 512 (1.48 / 1.77 int8 POP/s).
  // S = 7 needs too many registers for two workgroups per CU: one.
  switch (tuning_get(TUNE_OZAKI_SHAPE)) {
    case 1: return ozaki_launch<S, 2, 2, 2, 1>(s, L, m, n, col0, ncols, C, ldc);  // 1 per CU
    case 2: return ozaki_launch<S, 2, 4, 2, 1>(s, L, m, n, col0, ncols, C, ldc);  // 128 x 128
    case 3: return ozaki_launch<S, 2, 2, 4, 1>(s, L, m, n, col0, ncols, C, ldc);  // four stages
    case 4: return ozaki_launch<S, 2, 4, 4, 1>(s, L, m, n, col0, ncols, C, ldc);  // 128 x 128, four stages
    default:
      if constexpr (S <= 6) return ozaki_launch<S, 2, 2, 2, 2>(s, L, m, n, col0, ncols, C, ldc);
      else return ozaki_launch<S, 2, 2, 2, 1>(s, L, m, n, col0, ncols, C, ldc);  // 7 S accumulators
  }
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
