/*
 * libmpdok_hip.so — C-ABI of the MI355X (gfx950) dense-operator Krylov library.
 *
 * Two groups of entry points:
 *
 *  1. The reference's native plugin symbols, with identical argument lists, so a
 *     caller that loads mpdok.so / lu_ir.so through ctypes can point lib_path at
 *     this library instead (reference bindings: mpdok_ops.py:101-141 for
 *     lu_ir.so and mpdok_ops.py:361-382 for mpdok.so). They return void, report
 *     through out-pointers, enqueue on the library stream (mpdok_set_stream,
 *     default: the null stream) and synchronise it before returning, as the
 *     reference's Python wrappers expect (mpdok_ops.py:200,234,311,498).
 *     Matrices are FP64 column-major (Fortran order) device pointers.
 *
 *  2. New entry points (prefix mpdok_): explicit hipStream_t, int status
 *     (0 = OK, nonzero = error; text from mpdok_last_error()), no
 *     synchronisation unless stated. order: 0 = row-major, 1 = column-major.
 *     dtype: 0 = FP64, 1 = FP32.
 *
 * All pointers are device pointers unless stated otherwise.
 */
#ifndef MPDOK_HIP_H
#define MPDOK_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* hipStream_t;

/* ===== 1. Reference-compatible symbols ================================== */

/* GMRES-IR: FP64->FP32 copy of A into A32 (caller-owned N*N FP32 scratch),
 * inner FP32 GMRES(restart) with two-pass CGS and Givens (inner tolerance 1e-6),
 * FP64 residual and update; *converged = 1 when relres < tol.
 * Replaces mpdok.so::py_gmres_ir (mpdok_ops.py:362-373, called at :487-497). */
void py_gmres_ir(void* A64, void* b64, void* x64, void* A32, int N, double tol,
                 int maxiter_outer, int restart, int* converged);
/* N*N FP64 managed allocation (hipMallocManaged); one live buffer per library.
 * Replaces mpdok.so::py_alloc_managed / py_free_managed (mpdok_ops.py:375-382). */
void py_alloc_managed(int N, void** ptr_out);
void py_free_managed(void);

/* LU-IR in one call: FP32 LU with partial pivoting, FP32 triangular solves,
 * FP64 residual and update. Replaces lu_ir.so::py_lu_ir (mpdok_ops.py:102-111). */
void py_lu_ir(void* A64, void* b64, void* x64, int N, double tol, int maxiter_outer,
              int* converged);
/* Factor once (*info = 0 on success, k > 0 for a zero pivot in column k).
 * Replaces lu_ir.so::py_lu_factor (mpdok_ops.py:122-127). */
void py_lu_factor(void* A64, int N, int* info);
/* Solve with the stored factors; *converged = -1 if no factorization exists.
 * Replaces lu_ir.so::py_lu_solve_factored (mpdok_ops.py:129-138). */
void py_lu_solve_factored(void* A64, void* b64, void* x64, int N, double tol,
                          int maxiter_outer, int* converged);
/* Replaces lu_ir.so::py_lu_alloc_managed / py_lu_free_managed / py_lu_free_factored
 * (mpdok_ops.py:113-120, 140-141). */
void py_lu_alloc_managed(int N, void** ptr_out);
void py_lu_free_managed(void);
void py_lu_free_factored(void);

/* ===== 2. New entry points =============================================== */

/* Library services. */
const char* mpdok_last_error(void);
/* Stream used by the reference-compatible symbols (NULL = null stream). */
int mpdok_set_stream(hipStream_t s);
/* Kernel-shape choices picked by the tuner (0 is every key's default). Key 0:
 * block shape of the int8-sliced trailing update, 0..3 (0 = 128 x 128 tiles,
 * 8 wavefronts, four stages; 1 = 128 x 64 tiles, two workgroups per CU; 2 =
 * 128 x 128, two stages; 3 = 256 x 64). Key 1: row tiles per group of its
 * tile order (0 = 8). Key 3: 64 x 64 diagonal LU kernel (0 = blocked, 1 =
 * unblocked). Key 4: row slicing (0 = two passes over a 2-D grid, 1 = one
 * workgroup per 64 rows). Key 5: the trailing update's C (0 = read into
 * registers at the start of each tile for slice counts up to 3, 1 = read
 * after the K loop). Keys 2, 6, 7 are unused. No reference counterpart. */
int mpdok_set_tuning(int key, int value);
int mpdok_get_tuning(int key);
size_t mpdok_reduce_workspace_bytes(int n);
size_t mpdok_gemv_workspace_bytes(int M, int N, int dtype, int order);
size_t mpdok_residual_workspace_bytes(int N, int order);
size_t mpdok_arnoldi_workspace_bytes(int n, int max_k);

/* out[0] = <x, y> accumulated in FP64 (deterministic two-pass reduction);
 * work: mpdok_reduce_workspace_bytes(n). Replaces the cp.dot / cp.linalg.norm
 * reductions of dense_krylov.py:156,167,192,195,245,272,295. */
int mpdok_dot_f32(hipStream_t s, const float* x, const float* y, int n, double* out, void* work);
int mpdok_dot_f64(hipStream_t s, const double* x, const double* y, int n, double* out,
                  void* work);
/* Strided 2-D copy with dtype conversion and optional order change.
 * Replaces the FP64->FP32 casts of dense_krylov.py:47 and mpdok_ops.py:366. */
int mpdok_cast_matrix(hipStream_t s, const void* in, int in_dtype, int64_t ldi, int in_order,
                      void* out, int out_dtype, int64_t ldo, int out_order, int M, int N);
/* out = alpha*x + beta*y (y may be NULL), FP64 arithmetic, one rounding.
 * Replaces the CuPy vector expressions of dense_krylov.py:165,193-197,271,412. */
int mpdok_axpby(hipStream_t s, int n, double alpha, const void* x, int xdt, double beta,
                const void* y, int ydt, void* out, int odt);

/* y = A x. Column-major A needs a workspace of mpdok_gemv_workspace_bytes.
 * skip: optional device flag; when *skip != 0 the kernels return immediately.
 * Replaces cp.matmul(A_fp32, x) (dense_krylov.py:61-63) and the tile SGEMV of
 * mpdok_ooc.py:158-201. */
int mpdok_gemv_f32(hipStream_t s, const float* A, int64_t lda, int order, int M, int N,
                   const float* x, float* y, void* work, size_t work_bytes, const int* skip);
int mpdok_gemv_f64(hipStream_t s, const double* A, int64_t lda, int order, int M, int N,
                   const double* x, double* y, void* work, size_t work_bytes, const int* skip);
/* r = b - A x in FP64; mode 0 = FP64 accumulation (as cuBLAS Dgemv), mode 1 =
 * Dot2 compensated accumulation rounded once. Replaces dense_krylov.py:76-79,
 * :391, :416 and the natives' cublasDgemv residual (benchmark_mpdok.ipynb:991). */
int mpdok_residual_f64(hipStream_t s, const double* A, int64_t lda, int order, int N,
                       const double* x, const double* b, double* r, int mode, void* work,
                       size_t work_bytes);

/* r = b - A x exactly as mpdok_residual_f64 (same kernels, same bits), with
 * no workspace argument and no allocation: the column-major partial sums go
 * to a 16 MiB array the library holds on each GPU, so a column-major N needs
 * mpdok_residual_workspace_bytes(N, 1) <= 16 MiB (N <= 1048576; any N
 * row-major). Larger N is rejected with an error before any GPU work. Calls
 * that overlap on one GPU share that array and race. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc). */
int mpdok_residual_f64_noalloc(hipStream_t s, const double* A, int64_t lda, int order, int N,
                               const double* x, const double* b, double* r, int mode);

/* R = B - A X for nrhs right-hand sides (column-major A, X, B, R; mode as
 * mpdok_residual_f64): A is read once per block of up to 8 columns, and each
 * column of R equals mpdok_residual_f64's bit for bit. work:
 * mpdok_residual_multi_workspace_bytes(N). Replaces nrhs Dgemv residuals of
 * the reference's one-at-a-time LU-IR (benchmark_mpdok.ipynb:991). */
size_t mpdok_residual_multi_workspace_bytes(int N);
int mpdok_residual_multi_f64(hipStream_t s, const double* A, int64_t lda, int N, const double* X,
                             int64_t ldx, const double* B, int64_t ldb, double* R, int64_t ldr,
                             int nrhs, int mode, void* work, size_t work_bytes);

/* out[0] = sum of the FP64 values in `bytes` bytes at p (16-byte aligned, a
 * multiple of 16), read once with 16-byte non-temporal loads: the read-only
 * ceiling the GEMV kernels are measured against. work:
 * mpdok_read_ceiling_workspace_bytes(). No reference counterpart:
 * measurement. */
size_t mpdok_read_ceiling_workspace_bytes(void);
int mpdok_read_ceiling(hipStream_t s, const void* p, size_t bytes, double* out, void* work);

/* Kernel-matrix tile: rows [row0,row0+rows) x cols [col0,col0+cols) of
 * k(c_i, c_j) + reg*delta_ij from coords (N x D, row-major FP64). kind 0 =
 * Gaussian exp(-p d2), 1 = Matern-3/2 with length scale p. reg_mode 0 = nugget in
 * FP64 before the output cast (rbf_kernel.py:219-224), 1 = nugget after the cast
 * (mpdok_ooc.py:121-122). Replaces rbf_kernel.py:215-224 and mpdok_ooc.py:39-46. */
int mpdok_rbf_assemble(hipStream_t s, const double* coords, int N, int D, int kind, double p,
                       double reg, int reg_mode, int row0, int rows, int col0, int cols,
                       void* out, int out_dtype, int64_t ld, int order);
/* Diagonal of the kernel matrix exactly as the assembly and the matrix-free
 * products form it (elem_dtype F64: nugget in FP64; F32: the FP32 tile's
 * element, nugget after the cast), written as out_dtype, inverted when
 * invert != 0 (the Jacobi preconditioner of the out-of-core CG). No
 * reference counterpart (the reference has no Jacobi-CG). */
int mpdok_rbf_diag(hipStream_t s, const double* coords, int N, int D, int kind, double p,
                   double reg, int elem_dtype, int invert, void* out, int out_dtype);
/* out[0] = mean clamped d2 over the first n points (rbf_kernel.py:57-68). */
int mpdok_rbf_mean_d2(hipStream_t s, const double* coords, int n, int D, double* out);
/* Matrix-free y = K[row0:row0+rows, :] x, kernel values recomputed, never
 * stored. elem_dtype F64: FP64 element, nugget in FP64 (replaces
 * MPDOKOOCSolver._tiled_dgemv_fp64, mpdok_ooc.py:205-227); F32: the stored FP32
 * tile's element (cast, then nugget in FP32; the matrix-free form of
 * _tiled_sgemv, mpdok_ooc.py:39-46,158-201). x, y both xy_dtype (F64 or F32). */
int mpdok_rbf_matvec(hipStream_t s, const double* coords, int N, int D, int kind, double p,
                     double reg, int elem_dtype, const void* x, void* y, int xy_dtype, int row0,
                     int rows);
/* The same product over all N rows (row0 = 0, rows = N), each kernel value
 * evaluated once and used for both E_ij and E_ji (the matrix is symmetric bit
 * for bit): half the evaluations of mpdok_rbf_matvec, sums in another fixed
 * order. work: device memory of mpdok_rbf_symv_workspace_bytes(N) bytes
 * (N^2 / 32 rounded up to 256-point tiles), contents not kept between calls.
 * Serves the same callers (mpdok_ooc.py:205-227 and the FP32 product). */
size_t mpdok_rbf_symv_workspace_bytes(int N);
int mpdok_rbf_symv(hipStream_t s, const double* coords, int N, int D, int kind, double p,
                   double reg, int elem_dtype, const void* x, void* y, int xy_dtype, void* work,
                   size_t work_bytes);
/* Cross-kernel product y = K(R, C) x (FP64, no nugget), R: NR points, C: NC
 * points, both (n, D) row-major: the interpolant at query points
 * (rbf_demo.ipynb:527, enkf_solver.py:162-200) and K_v alpha / K_v^T g of the
 * adjoint. No reference C symbol. */
int mpdok_rbf_cross_matvec(hipStream_t s, const double* R, int NR, const double* C, int NC, int D,
                           int kind, double p, const double* x, double* y);
/* out[0] = u^T (dK(R, C)/dp) w with the kernel-parameter derivative
 * recomputed per element (Gaussian: -d2 k; Matern-3/2: a^2 exp(-a) / p),
 * deterministic block reduction. The matrix-free dK/dtheta of the adjoint
 * (north star; no reference counterpart). */
size_t mpdok_rbf_dbilinear_workspace_bytes(int NR);
int mpdok_rbf_dbilinear(hipStream_t s, const double* R, int NR, const double* C, int NC, int D,
                        int kind, double p, const double* u, const double* w, double* out,
                        void* work, size_t work_bytes);

/* Out-of-core products (mpdok_ooc.py:158-201). y[0:M] = A_host x for a
 * row-major FP32 A (M x N, leading dimension lda) in host memory (pinned for
 * full overlap), streamed in tiles of tile_rows rows through two device
 * buffers in `work` (mpdok_stream_gemv_workspace_bytes), copy of tile t+1
 * overlapping the product of tile t. Replaces _tiled_sgemv_ram/_ssd's per-tile
 * cp.asarray + matmul. */
size_t mpdok_stream_gemv_workspace_bytes(int tile_rows, int N);
int mpdok_stream_gemv_f32(hipStream_t s, const float* A_host, int64_t lda, int M, int N,
                          int tile_rows, const float* x, float* y, void* work, size_t work_bytes);
/* A_host[0:rows, :] = FP32 kernel rows [row0, row0+rows) (nugget after the
 * cast), assembled on the GPU per tile and copied to host memory, the copy of
 * one tile overlapping the assembly of the next. Replaces MPDOKOOCSolver.build's
 * tile loop (mpdok_ooc.py:116-146). Same workspace as mpdok_stream_gemv_f32. */
int mpdok_rbf_build_host(hipStream_t s, const double* coords, int N, int D, int kind, double p,
                         double reg, int row0, int rows, int tile_rows, float* A_host,
                         int64_t ldh, void* work, size_t work_bytes);

/* One Arnoldi orthogonalisation of w against V_0..V_k (vectors of length n,
 * stride ldv): passes = 1 (dense_krylov.py:289-303) or 2 (DGKS,
 * mpdok_ooc.py:263-273). Adds the coefficients into Hcol[0..k], writes
 * Hcol[k+1] = ||w|| and V_{k+1} = w/||w||, or raises the breakdown flag when
 * ||w|| <= thresh. ctl (device int[4]): ctl[0] stop, ctl[1] steps done,
 * ctl[2] breakdown. */
int mpdok_arnoldi_orth_f32(hipStream_t s, float* V, int64_t ldv, int k, float* w, int n,
                           double* Hcol, int passes, double thresh, int stop_now, int* ctl,
                           void* work, size_t work_bytes);
int mpdok_arnoldi_orth_f64(hipStream_t s, double* V, int64_t ldv, int k, double* w, int n,
                           double* Hcol, int passes, double thresh, int stop_now, int* ctl,
                           void* work, size_t work_bytes);
/* Device Givens update of Hessenberg column k (mpdok_ooc.py:284-301); sets
 * ctl[0] when |g[k+1]|/beta < tol or a breakdown was flagged. */
int mpdok_givens_step(hipStream_t s, double* H, int ldh, int k, double* R, double* cs,
                      double* sn, double* g, double beta, double tol, int* ctl);
/* x (+)= sum_{i<m} y_i V_i (dense_krylov.py:311-312, mpdok_ooc.py:310). */
int mpdok_basis_update_f32(hipStream_t s, const float* V, int64_t ldv, int m, const double* y,
                           float* x, int n, int overwrite);
int mpdok_basis_update_f32_into_f64(hipStream_t s, const float* V, int64_t ldv, int m,
                                    const double* y, double* x, int n, int overwrite);
int mpdok_basis_update_f64(hipStream_t s, const double* V, int64_t ldv, int m, const double* y,
                           double* x, int n, int overwrite);
/* out = in / sqrt(nrm2[0]) (dense_krylov.py:278, mpdok_ooc.py:249). */
int mpdok_scale_f64_to_f32(hipStream_t s, const double* in, float* out, int n,
                           const double* nrm2);
int mpdok_scale_f32(hipStream_t s, const float* in, float* out, int n, const double* nrm2);
int mpdok_scale_f64(hipStream_t s, const double* in, double* out, int n, const double* nrm2);
/* CG updates with device scalars (dense_krylov.py:191-198): alpha = rs_old/pAp,
 * x += alpha p, r -= alpha Ap, rs_new = r.r, p = r + (rs_new/rs_old) p.
 * work: 4096 doubles. */
int mpdok_cg_step_updates(hipStream_t s, float* x, float* r, float* p, const float* Ap, int n,
                          const double* rs_old, const double* pAp, double* rs_new, void* work);

/* Preconditioned CG updates with device scalars (dense_krylov.py:191-198 with
 * a diagonal preconditioner), vectors FP64 (dtype 0) or FP32 (1): alpha =
 * rz_old/pAp, x += alpha p, r -= alpha Ap, z = dinv .* r, rz_new = r.z,
 * rr_new = r.r (optional), p = z + (rz_new/rz_old) p. dinv NULL: plain CG (z
 * is r; z may be NULL). alpha and beta are rounded to the vector type. work:
 * mpdok_pcg_workspace_bytes(). The inner solver of the out-of-core Jacobi-CG
 * and the FP64 body of cg(); no reference C symbol. */
size_t mpdok_pcg_workspace_bytes(void);
int mpdok_pcg_step_updates(hipStream_t s, int dtype, void* x, void* r, void* z, void* p,
                           const void* Ap, const void* dinv, int n, const double* rz_old,
                           const double* pAp, double* rz_new, double* rr_new, void* work);
/* out = d .* in (dtype 0 FP64 / 1 FP32): the Jacobi preconditioner apply. */
int mpdok_diag_scale(hipStream_t s, int dtype, const void* d, const void* in, void* out, int n);

/* ---- LU factorisation ------------------------------------------------------
 * Opaque handle replacing the global factor slot of lu_ir.so
 * (mpdok_ops.py:122-141,245-247). precision: 0 = FP64, 1 = FP32, 2 = BF16
 * trailing updates on the BF16 matrix cores (FP32 storage and accumulation),
 * 3 = the same in FP16, 4 = int8-sliced trailing updates with `slices` digits
 * and FP64 storage, 5 = FP8 (E4M3) operands with one power-of-two scale per
 * operand of each update (the precision ladder's FP8 experiment).
 * pivot: 0 = none (SPD / diagonally dominant), 1 = partial (getrf).
 * nb: outer block (multiple of 64; 0 = 512). */
int mpdok_lu_create(int N, int precision, int pivot, int nb, int slices, void** handle);
/* Copies A (FP64, row- or column-major) into the handle and factors it;
 * synchronises s; *info = 0 or the 1-based column of the first zero pivot.
 * Replaces lu_ir.so::py_lu_factor's Sgetrf (benchmark_mpdok.ipynb:989). */
int mpdok_lu_factor(void* handle, hipStream_t s, const double* A, int64_t lda, int order,
                    int* info);
/* x = inv(P L U) b (FP64 vectors; x may alias b). Replaces Sgetrs
 * (benchmark_mpdok.ipynb:990,992). */
int mpdok_lu_solve(void* handle, hipStream_t s, const double* b, double* x);
/* X[:, 0:nrhs] = inv(P L U) B[:, 0:nrhs] (FP64, column-major, ldb / ldx >= N;
 * X must not alias B). Blocks of up to 8 right-hand sides share each read of
 * the factors; every column equals mpdok_lu_solve's result bit for bit. The
 * multi-right-hand-side Sgetrs the reference's factor-once use case would call
 * once per field (run_demo.py:75-109 solves 72 right-hand sides with
 * lu_ir.so::py_lu_solve_factored, mpdok_ops.py:204-243, one at a time). */
int mpdok_lu_solve_multi(void* handle, hipStream_t s, const double* B, int64_t ldb, double* X,
                         int64_t ldx, int nrhs);
int mpdok_lu_destroy(void* handle);
/* Use the first n <= N (N as created) rows and columns of the handle's
 * buffers for the next factorisation: one handle serves systems of many
 * sizes without reallocating. No reference counterpart. */
int mpdok_lu_resize(void* handle, int n);
/* HIP-event timing of the trailing-update launches of later factorisations;
 * mpdok_lu_stats returns the summed time (ms), the operations done (int8
 * operations, 2 m n k S(S+1)/2 per launch, for the sliced variant; flops
 * otherwise), the launch count and the algorithmic flops of the updates
 * (2 m n k per launch) of the last factorisation. No reference counterpart:
 * measurement. */
int mpdok_lu_set_profiling(void* handle, int on);
/* Factor the next panel on a side stream while the trailing update of the
 * current step finishes (default on; no-pivot handles only). */
int mpdok_lu_set_lookahead(void* handle, int on);
int mpdok_lu_stats(void* handle, double* trailing_ms, double* trailing_ops, int* launches,
                   double* trailing_flops);
/* Adaptive slice counts for the int8-sliced variant (default on when
 * slices <= 5): every panel pair is cut into `slices` digits, and each
 * trailing update multiplies only the leading S of them, S the fewest in
 * [smin, slices] whose error bound 2^(E - (8 S - 1)) (E = max row exponent
 * of L21 + max column exponent of U12) stays within the first update's
 * minus `delta` bits. Decided on the device; no host synchronisation. No
 * reference counterpart (the reference factors in FP32). */
int mpdok_lu_set_adaptive(void* handle, int on, int smin, int delta);
/* Slice counts of the last factorisation, one per trailing update (*nsteps
 * of them; at most n written). */
int mpdok_lu_slice_counts(void* handle, int* out, int n, int* nsteps);
/* First columns of the last factorisation's steps, one trailing update per
 * step (a panel, or a group of panels); *nsteps of them, at most n written.
 * No reference counterpart. */
int mpdok_lu_step_starts(void* handle, int* out, int n, int* nsteps);
/* Row lookahead on the no-pivot lookahead path (the int8-sliced variant
 * needs adaptive slice counts): while the trailing matrix has at least
 * min_rows rows, a step's block-row solve, U12 slicing and slice-count
 * decision run on the lookahead stream right after its panels, and the main
 * stream updates the next step's rows before the rest of its trailing update.
 * 0 turns it off; default 20480. No reference counterpart. */
int mpdok_lu_set_row_lookahead(void* handle, int min_rows);
/* Panels per trailing update of the int8-sliced variant on the lookahead
 * path: groups of `panels` (1 or 2), and of `panels_early` (panels .. 4)
 * while the trailing matrix has at least early_rows rows. A group is
 * factored right-looking inside (FP64) and its trailing update takes
 * K = (panels in the group) nb, so the trailing matrix is read and written
 * once per group. Defaults: 2, 2, 0. No reference counterpart. */
int mpdok_lu_set_update_panels(void* handle, int panels, int panels_early, int early_rows);
/* Device pointers to the packed factors (column-major, ld N) and the row
 * permutation ((P b)[i] = b[perm[i]], NULL without pivoting). */
int mpdok_lu_factors(void* handle, void** LU, int** perm, int* elem_size);
/* C = alpha A B + beta C on the matrix cores (column-major); dtype 0 FP64,
 * 1 FP32; round (FP32 only) 0 = FP32 products, 1 = BF16 operands on the BF16
 * matrix cores, 2 = FP16, 3 = E4M3 values after one power-of-two scale per
 * operand (on the BF16 matrix cores, where E4M3 values are exact); FP32
 * accumulation. The trailing update of the factorisation (in place of
 * cuSOLVER's internal GEMM). */
int mpdok_gemm(hipStream_t s, int dtype, int m, int n, int k, double alpha, const void* A,
               int64_t lda, const void* B, int64_t ldb, double beta, void* C, int64_t ldc,
               int round);
/* C -= A B with both operands cut into `slices` int8 digits per row / column
 * (int8 matrix cores, exact int32 sums, FP64 fold). work: workspace bytes. */
size_t mpdok_ozaki_workspace_bytes(int m, int n, int K, int slices);
int mpdok_ozaki_gemm_sub(hipStream_t s, int slices, int m, int n, int K, const double* A,
                         int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc,
                         void* work);
/* The adaptive pieces of the factorisation, for parity tests: cut A and B
 * into smax (3..5) digits, then C -= the product of the leading S digits,
 * S = s_fixed (>= 2) or, for s_fixed = 0, the decision rule of
 * mpdok_lu_set_adaptive against the reference exponent sum e0_in (INT_MIN:
 * this call is the first update and records it). dev_scratch: 2 device
 * ints. *s_used and *e_ref receive S and the reference sum. Synchronises s. */
int mpdok_ozaki_gemm_sub_adaptive(hipStream_t s, int smax, int m, int n, int K, const double* A,
                                  int64_t lda, const double* B, int64_t ldb, double* C,
                                  int64_t ldc, void* work, int s_fixed, int smin, int delta,
                                  int e0_in, int* dev_scratch, int* s_used, int* e_ref);
/* C -= the product of the leading s_fixed of the smax digits that the last
 * mpdok_ozaki_gemm_sub_adaptive call with the same (smax, m, n, K, work)
 * cut, without slicing again (times the trailing-update kernel alone).
 * dev_scratch: 2 device ints. No reference counterpart: measurement. */
int mpdok_ozaki_apply_sliced(hipStream_t s, int smax, int m, int n, int K, double* C, int64_t ldc,
                             void* work, int s_fixed, int* dev_scratch);

/* ---- refinement drivers -----------------------------------------------------
 * GMRES-IR: FP64 outer refinement (residual mode 0 FP64 / 1 Dot2) around one
 * GMRES(restart) cycle per outer step with body 0 = FP32 (A32) / 1 = FP64
 * (A64), CGS `passes`, inner stop |g|/beta < inner_tol, optional right
 * preconditioner lu (NULL = none). hist: host array of hist_cap doubles
 * receiving relres per outer step plus the final check. Unpreconditioned FP32
 * body with passes = 2 is py_gmres_ir (mpdok_ops.py:362-373); the LU-
 * preconditioned form has no reference counterpart. */
int mpdok_gmres_ir(hipStream_t s, const double* A64, int64_t lda64, int order64, const float* A32,
                   int64_t lda32, int order32, void* lu, int N, const double* b, double* x,
                   double tol, int maxiter_outer, int restart, double inner_tol, int passes,
                   int resid_mode, int body, int* converged, double* hist, int hist_cap,
                   int* n_hist, int* inner_total);
/* LU-IR with a factorisation handle: x = 0; repeat { r = b - A x (FP64 /
 * Dot2); stop if relres < tol; x += inv(PLU) r }. Replaces
 * lu_ir.so::py_lu_solve_factored's loop (mpdok_ops.py:225-243). */
int mpdok_lu_ir(hipStream_t s, const double* A64, int64_t lda, int order, void* lu, int N,
                const double* b, double* x, double tol, int maxiter_outer, int resid_mode,
                int* converged, double* hist, int hist_cap, int* n_hist);

/* mpdok_lu_ir for nrhs right-hand sides (column-major B, X, ldb / ldx >= N):
 * blocks of up to 8 share one residual pass over A and one solve with the
 * factors per outer step; each column follows mpdok_lu_ir's loop exactly (the
 * same iterates and result bit for bit, no update once it has converged).
 * converged, relres: host arrays of nrhs. Synchronises s. The multi-right-hand-
 * side form of lu_ir.so::py_lu_solve_factored (mpdok_ops.py:204-243). */
int mpdok_lu_ir_multi(hipStream_t s, const double* A64, int64_t lda, int order, void* lu, int N,
                      const double* B, int64_t ldb, double* X, int64_t ldx, int nrhs, double tol,
                      int maxiter_outer, int resid_mode, int* converged, double* relres);

/* Batched uneven kernel systems (north star's batched configuration; no
 * reference counterpart). System i: N[i] points coords[i] (device, N x D
 * row-major), kernel (kind, param[i]) + reg I, nrhs right-hand sides B[i]
 * (device, N x nrhs column-major) solved into X[i] (same layout) to relres <=
 * tol by LU (precision 0 FP64 or 4 int8-sliced, no pivoting) + LU-IR with the
 * Dot2 residual, all right-hand sides of a system refined together
 * (mpdok_lu_ir_multi). nstreams workers (host threads with their own streams),
 * largest systems first, longest-processing-time assignment. converged and
 * relres (host, nsys x nrhs) receive the per-right-hand-side outcome. */
int mpdok_batched_rbf_solve(int nsys, const int* N, const double* const* coords, int D, int kind,
                            const double* param, double reg, int nrhs, const double* const* B,
                            double* const* X, double tol, int maxiter_outer, int precision,
                            int slices, int nb, int nstreams, int* converged, double* relres);

/* Batched plan: the same systems solved with grouped launches (one launch per
 * kind of work and block step covers every system with work at that step)
 * captured into one hipGraph; no reference counterpart (north star's batched
 * configuration). Systems are padded to multiples of 64 with an identity
 * block; FP64 no-pivot LU with 64-wide steps; LU-IR with the Dot2 residual and
 * device-side convergence flags (converged columns and finished systems make
 * the remaining launches return at once). coords[i], B[i], X[i]: device
 * pointers kept by the plan (each run reads coords and B and writes X).
 * update_rows: trailing updates of at least this many rows use 128 x 128
 * tiles, smaller ones 64 x 64 (chosen by the Python tuner). use_graph: 1 to
 * capture at creation and replay on each run, 0 to launch eagerly. nrhs <= 8,
 * N[i] <= 65536. */
int mpdok_bplan_create(int nsys, const int* N, const double* const* coords, int D, int kind,
                       const double* param, double reg, int nrhs, const double* const* B,
                       double* const* X, double tol, int maxiter_outer, int update_rows,
                       int use_graph, void** plan_out);
/* Enqueue one solve on s (a graph launch, or the eager sequence); no wait. */
int mpdok_bplan_run(void* plan, hipStream_t s);
/* Outcome of the last run in the caller's system order (waits for the
 * device): converged, relres, steps (updates applied): nsys x nrhs; info:
 * nsys (zero-pivot column + 1, else 0); hist: relres per refinement step,
 * (maxiter_outer + 1) x nsys x nrhs, NaN after a column stopped. Any pointer
 * may be NULL. */
int mpdok_bplan_results(void* plan, int* converged, double* relres, int* steps, int* info,
                        double* hist);
/* Device bytes held and kernel launches per run (graph nodes). */
int mpdok_bplan_info(void* plan, size_t* bytes, int* launches);
void mpdok_bplan_destroy(void* plan);

#ifdef __cplusplus
}
#endif
#endif /* MPDOK_HIP_H */
