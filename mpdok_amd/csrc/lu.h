#pragma once
#include "common.h"

namespace mpdok {

// lu_kernels.hip
template <typename T>
int gemm(hipStream_t s, int m, int n, int k, T alpha, const T* A, int64_t lda, const T* B,
         int64_t ldb, T beta, T* C, int64_t ldc, int round);
// C = alpha A B + beta C with FP32 operands rounded to BF16 (round 1), FP16
// (2) or scaled E4M3 (3; scales: gemm_lowp_scale_floats() device floats,
// the two scales first) on the matrix cores
size_t gemm_lowp_scale_floats();
int gemm_lowp(hipStream_t s, int m, int n, int k, float alpha, const float* A, int64_t lda,
              const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int round,
              float* scales);
// One product C = alpha A B + beta C (column-major) of a grouped launch.
template <typename T>
struct GemmArgs {
  int m, n, k;
  T alpha;
  const T* A;
  int64_t lda;
  const T* B;
  int64_t ldb;
  T beta;
  T* C;
  int64_t ldc;
};
// Two independent products in one launch (64 x 64 tiles, so the in-place
// aliasing rules of gemm hold for a side of at most 64).
template <typename T>
int gemm_pair(hipStream_t s, const GemmArgs<T>& p, const GemmArgs<T>& q);
template <typename T>
int diag_factor(hipStream_t s, T* A, int64_t lda, int nb, T* Linv, T* Uinv, int ldi, int* info,
                int col0, int factor);
template <typename TI, typename TO>
int copy2d(hipStream_t s, const TI* src, int64_t lds, TO* dst, int64_t ldd, int m, int n);
template <typename T>
int laswp(hipStream_t s, T* A, int64_t lda, int c0, int c1, const int* ipiv, int r0, int r1);
// getf2 on the sub-panel [j0, j0 + w) x rows [j0, N), w <= 64, in one
// cooperative launch with implicit pivoting: rows stay in place, ipiv[j0 ..
// j0 + w) receives LAPACK's interchanges, which the caller applies to every
// column (laswp over [0, N)).
template <typename T>
int panel_gepp(hipStream_t s, T* A, int64_t lda, int N, int j0, int w, int* ipiv, int* info,
               void* work);
size_t panel_gepp_workspace_bytes(int N);
template <typename T, typename TV>
int trsv_blocked(hipStream_t s, const T* LU, int64_t ld, int N, const T* Linv_all,
                 const T* Uinv_all, TV* v);
template <typename T>
int permute(hipStream_t s, const T* in, T* out, const int* perm, int n);
int ipiv_to_perm(hipStream_t s, const int* ipiv, int* perm, int n);
template <typename T>
int gemm_batched(hipStream_t s, int m, int n, int k, T alpha, const T* A, int64_t lda,
                 int64_t sA, const T* B, int64_t ldb, int64_t sB, T beta, T* C, int64_t ldc,
                 int64_t sC, int batch, int round);
constexpr int TRSV_BLOCK = 512;  // diagonal block of the inverse-based triangular solves
// tbw: width of the diagonal blocks whose inverses Lbinv / Ubinv hold
// (TRSV_BLOCK or 2 TRSV_BLOCK); acc: 2 (tbw / 64) N elements.
template <typename T, typename TV>
int trsv_block_inverse(hipStream_t s, const T* LU, int64_t ld, int N, const T* Lbinv,
                       const T* Ubinv, const TV* b, TV* x, TV* v, TV* acc, int tbw);
// The same for nr (1, 2, 4 or 8) right-hand sides, column-major N x nr
// blocks with ld N; acc: 2 (tbw / 64) N nr elements.
template <typename T, typename TV>
int trsv_block_inverse_multi(hipStream_t s, const T* LU, int64_t ld, int N, const T* Lbinv,
                             const T* Ubinv, const TV* b, TV* x, TV* v, TV* acc, int tbw, int nr);
// The off-diagonal-free parts of 2 TRSV_BLOCK-wide diagonal blocks of an
// inverse from its TRSV_BLOCK-wide ones (see lu_kernels.hip).
template <typename T>
int widen_block_inverse(hipStream_t s, const T* in, T* out, int nbig, int lower);
template <typename T>
int seed_block_inverse(hipStream_t s, const T* inv64, T* binv, int nbig);

// Grouped steps of a batched no-pivot LU (lu_kernels.hip). System table
// entry: factors F (P x P, column-major, ld P, P a multiple of 64), the
// inverses of the unit-lower and upper triangles of its 64 x 64 diagonal
// blocks (P x 64 each, ld P; block k at row 64 k) and its zero-pivot slot
// (the first zero pivot's column + 1; 0 when none).
struct BLUSys {
  double* F;
  double* Linv;
  double* Uinv;
  int* info;
  int P;
  int pad_;
};
// Step k for systems sys[0 .. ns), each with P > 64 k. pref: ns + 1 tile
// offsets (the tiles of the step's panel products, 2 ceil(m / 64) per system,
// or of its trailing update, ceil(m / tb)^2, m = P - 64 (k + 1)), ntiles =
// pref[ns] (the host's copy).
int blu_diag(hipStream_t s, const BLUSys* sys, int ns, int k);
int blu_panel(hipStream_t s, const BLUSys* sys, const int* pref, int ns, int ntiles, int k);
int blu_update(hipStream_t s, const BLUSys* sys, const int* pref, int ns, int ntiles, int k,
               int tb);

// ozaki.hip
size_t ozaki_workspace_bytes(int m, int n, int K, int S);
int ozaki_gemm(hipStream_t s, int S, int m, int n, int K, const double* A, int64_t lda,
               const double* B, int64_t ldb, double* C, int64_t ldc, void* work);
int ozaki_slice(hipStream_t s, int S, int m, int n, int K, const double* A, int64_t lda,
                const double* B, int64_t ldb, void* work, int parts = 3);
// C[row0 : row0 + nrows, col0 : col0 + ncols] -= the product of the leading
// *Sdev slices of the last ozaki_slice call for (smax, m, n, K); row0 a
// multiple of 128, col0 of 64. Ci (if not null): the block's values before
// the update, read from there (leading dimension ldci) instead of from C.
int ozaki_apply_dyn(hipStream_t s, int smax, int m, int n, int K, int row0, int nrows, int col0,
                    int ncols, double* C, int64_t ldc, void* work, const int* Sdev,
                    const double* Ci = nullptr, int64_t ldci = 0);
int ozaki_decide(hipStream_t s, int smax, int m, int n, int K, void* work, int smin, int delta,
                 int first, int* e0, int* s_out);
int ozaki_apply(hipStream_t s, int S, int m, int n, int K, int col0, int ncols, double* C,
                int64_t ldc, void* work);

// lu.hip
enum LUPrecision : int {
  LU_FP64 = 0, LU_FP32 = 1, LU_BF16 = 2, LU_FP16 = 3, LU_OZAKI = 4, LU_FP8 = 5
};
enum LUPivot : int { PIV_NONE = 0, PIV_PARTIAL = 1 };

struct LUHandle;
int lu_create(int N, int precision, int pivot, int nb, int slices, LUHandle** out);
int lu_factor(LUHandle* h, hipStream_t s, const double* A, int64_t lda, int order, int* info);
// x = inv(P L U) b  (b, x FP64; x may alias b)
int lu_solve(LUHandle* h, hipStream_t s, const double* b, double* x);
// X = inv(P L U) B for k columns (column-major, ldb / ldx >= N)
int lu_solve_multi(LUHandle* h, hipStream_t s, const double* B, int64_t ldb, double* X,
                   int64_t ldx, int k);
void lu_destroy(LUHandle* h);
int lu_n(const LUHandle* h);
// Use the first n <= capacity rows / columns of the buffers for the next
// factorisation (the handle's buffers are sized for the N it was created with).
int lu_resize(LUHandle* h, int n);

}  // namespace mpdok
