// Blocked right-looking LU factorisation and solve behind an opaque handle.
//
// Storage: the factors live in a handle-owned N x N column-major array (FP64
// for the FP64 and int8-sliced variants, FP32 for the FP32/BF16/FP16
// variants), L unit-lower and U packed LAPACK-style, plus the inverses of all
// 64 x 64 diagonal blocks of L and U for the blocked triangular solves.
//
// Algorithm (outer block nb, inner block 64):
//   for each outer panel [k, k+kb):
//     for each 64-column sub-panel j:
//       no pivoting : factor the 64 x 64 diagonal block in LDS; L21 = A21 inv(U11)
//       partial     : column-by-column pivot search / swap / eliminate over the
//                     full height, then apply the interchanges to the other columns
//       U12 of the rest of the panel = inv(L11) A12; panel update by GEMM
//     U12 for the columns right of the panel (blocked forward substitution)
//     trailing update A22 -= L21 U12 (the O(N^3) part):
//       FP64  : FP64 matrix cores        FP32 : FP32 matrix cores
//       BF16/FP16 : operands rounded to BF16/FP16, FP32 accumulation
//       OZAKI : int8 matrix cores on S sliced digits, FP64 storage
//
// Reference: the sgetrf/sgetrs of lu_ir.so (mpdok_ops.py:69-320,
// benchmark_mpdok.ipynb:983-996); pivoting semantics follow LAPACK getrf
// (ipiv = sequential interchanges).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>
#include "common.h"
#include "gemv.h"
#include <type_traits>

#include "lu.h"

namespace mpdok {

struct LUHandle {
  int N = 0, precision = 0, pivot = 0, nb = 512, slices = 6;
  int cap = 0;  // largest N the buffers hold (lu_resize)
  bool factored = false;
  size_t es = 8;
  void* LU = nullptr;
  void* Linv = nullptr;
  void* Uinv = nullptr;
  void* W = nullptr;
  size_t w_elems = 0;  // elements W holds
  int* ipiv = nullptr;
  int* perm = nullptr;
  int* d_info = nullptr;
  void* piv_work = nullptr;
  void* oz_work = nullptr;   // int8 slices of even steps
  void* oz_work2 = nullptr;  // ... and of odd steps (the panel slices the next step's rows)
  void* v1 = nullptr;
  void* v2 = nullptr;
  // inverse-based triangular solves (N a multiple of TRSV_BLOCK)
  bool fast_solve = false;
  void* Lbinv = nullptr;
  void* Ubinv = nullptr;
  // 2 TRSV_BLOCK-wide diagonal-block inverses (allocated for N >= 2
  // TRSV_BLOCK, used when N is a multiple of it): the solves then take half
  // as many dependent steps (wide_solve)
  void* Lbinv2 = nullptr;
  void* Ubinv2 = nullptr;
  bool wide_solve = false;
  void* v3 = nullptr;
  void* acc = nullptr;
  // multi-right-hand-side solves (allocated on first use): three N x
  // MULTI_NR vector blocks and the accumulators of MULTI_NR solves
  static constexpr int MULTI_NR = 8;
  void* mv = nullptr;
  void* macc = nullptr;
  // lookahead (no pivoting): panel k+1 is factored on `side` while the main
  // stream finishes the trailing update of step k; Wp is the side's scratch
  hipStream_t side = nullptr;
  // ev_first: the first panel of the group being factored is done
  hipEvent_t ev_next_cols = nullptr, ev_panel = nullptr, ev_first = nullptr;
  void* Wp = nullptr;
  bool lookahead = true;
  // int8-sliced variant: per-update slice counts picked on the device from the
  // panels' exponents (adaptive, slices <= 5); d_S[0] is the first update's
  // exponent sum, d_S[1 + step] the count of step `step`
  bool adaptive = true;
  int smin = 2, delta = 0;
  // panels per trailing update on the lookahead path of the int8-sliced
  // variant: groups of `upd` panels (1 or 2), and of `upd_early` (up to
  // MAX_GROUP) while the trailing matrix has at least early_rows rows. A
  // group is factored right-looking inside (FP64) and its trailing update
  // takes K = panels nb, so C is read and written once per group.
  static constexpr int MAX_GROUP = 4;
  int upd = 1, upd_early = 1, early_rows = 0;
  // the running factorisation's groups: first columns (plus N at the end)
  // and the step of the group starting at panel k / nb
  std::vector<int> group_start, step_idx;
  // row lookahead: while the trailing matrix has at least rows_side_min rows
  // (0: never), a group's block-row solve, U12 slicing and slice-count
  // decision run on the lookahead stream right after its panels (the main
  // stream updates the group's rows before the rest of its trailing update)
  int rows_side_min = 0;
  // the input matrix (column-major FP64) while the first trailing update of
  // the int8-sliced variant reads its values from there instead of from a
  // copy in the handle (null otherwise)
  const double* c_in = nullptr;
  int64_t ld_in = 0;
  int* d_S = nullptr;
  float* d_sc = nullptr;  // FP8 operand scales of the running trailing update
  std::vector<int> steps_S;  // host copy after a factorisation
  // optional HIP-event timing of the trailing-update launches
  bool profile = false;
  std::vector<hipEvent_t> ev;
  std::vector<double> ev_ops;  // algorithmic operations of each timed launch (per slice pair)
  std::vector<int> ev_step;    // step of each timed launch
  double trailing_ms = 0.0, trailing_ops = 0.0, trailing_flops = 0.0;
  int trailing_launches = 0;
};

int lu_n(const LUHandle* h) { return h ? h->N : 0; }

int lu_resize(LUHandle* h, int n) {
  MPDOK_REQUIRE(h != nullptr, "lu_resize: null handle");
  MPDOK_REQUIRE(n >= 1 && n <= h->cap, "lu_resize: n=%d outside [1, %d]", n, h->cap);
  h->N = n;
  h->fast_solve = (n % TRSV_BLOCK) == 0 && h->Lbinv != nullptr;
  h->wide_solve = false;
  h->factored = false;
  return OK;
}

static bool fp64_storage(int precision) { return precision == LU_FP64 || precision == LU_OZAKI; }

void lu_destroy(LUHandle* h) {
  if (!h) return;
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  if (h->ev_next_cols) (void)hipEventDestroy(h->ev_next_cols);
  if (h->ev_panel) (void)hipEventDestroy(h->ev_panel);
  if (h->ev_first) (void)hipEventDestroy(h->ev_first);
  if (h->side) (void)hipStreamDestroy(h->side);
  if (h->Wp) (void)hipFree(h->Wp);
  void* ptrs[] = {h->LU, h->Linv, h->Uinv, h->W, h->ipiv, h->perm, h->d_info,
                  h->piv_work, h->oz_work, h->oz_work2, h->v1, h->v2, h->Lbinv, h->Ubinv, h->v3,
                  h->Lbinv2, h->Ubinv2,
                  h->acc, h->mv, h->macc,
                  h->d_S, h->d_sc};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete h;
}

int lu_create(int N, int precision, int pivot, int nb, int slices, LUHandle** out) {
  MPDOK_REQUIRE(out != nullptr, "lu_create: null out");
  *out = nullptr;
  MPDOK_REQUIRE(N >= 1, "lu_create: N must be >= 1");
  MPDOK_REQUIRE(precision >= LU_FP64 && precision <= LU_FP8, "lu_create: bad precision %d",
                precision);
  MPDOK_REQUIRE(pivot == PIV_NONE || pivot == PIV_PARTIAL, "lu_create: bad pivot %d", pivot);
  if (nb <= 0) nb = 512;
  MPDOK_REQUIRE(nb % 64 == 0, "lu_create: nb must be a multiple of 64");
  MPDOK_REQUIRE(precision != LU_OZAKI || (slices >= 3 && slices <= 7),
                "lu_create: slices must be in [3, 7]");
  LUHandle* h = new LUHandle();
  h->N = N;
  h->cap = N;
  h->precision = precision;
  h->pivot = pivot;
  h->nb = nb;
  h->slices = slices;
  h->es = fp64_storage(precision) ? 8 : 4;
  const size_t nblk = (N + 63) / 64;
  size_t wel = (size_t)N * (nb > 128 ? nb : 128);
  // the wide inverses' products: TRSV_BLOCK^2 per 2 TRSV_BLOCK-wide block
  if (N >= 2 * TRSV_BLOCK && wel < (size_t)N * (TRSV_BLOCK / 2))
    wel = (size_t)N * (TRSV_BLOCK / 2);
  h->w_elems = wel;
  int st = OK;
  auto alloc = [&](void** p, size_t bytes) {
    if (st != OK) return;
    if (hipMalloc(p, bytes) != hipSuccess) {
      set_error("lu_create: hipMalloc of %zu bytes failed", bytes);
      st = ENOMEM_;
    }
  };
  alloc(&h->LU, (size_t)N * N * h->es);
  alloc(&h->Linv, nblk * 64 * 64 * h->es);
  alloc(&h->Uinv, nblk * 64 * 64 * h->es);
  alloc(&h->W, wel * h->es);
  alloc((void**)&h->ipiv, (size_t)N * sizeof(int));
  alloc((void**)&h->perm, (size_t)N * sizeof(int));
  alloc((void**)&h->d_info, 64);
  alloc(&h->piv_work, panel_gepp_workspace_bytes(N));
  alloc(&h->v1, (size_t)N * 8);
  alloc(&h->v2, (size_t)N * 8);
  if (precision == LU_OZAKI) {
    alloc(&h->oz_work, ozaki_workspace_bytes(N, N, LUHandle::MAX_GROUP * nb, slices));
    alloc(&h->oz_work2, ozaki_workspace_bytes(N, N, LUHandle::MAX_GROUP * nb, slices));
    alloc((void**)&h->d_S, sizeof(int) * (N / nb + 3));
  }
  if (precision == LU_FP8) alloc((void**)&h->d_sc, gemm_lowp_scale_floats() * sizeof(float));
  h->adaptive = precision == LU_OZAKI && slices <= 5;
  // row lookahead while the trailing matrix has at least 20480 rows:
  // measured 175 -> 172 ms for the N = 32768 factorisation (the early
  // trailing updates saturate the GPU, so the moved work still competes for
  // it; the gain is the latency it no longer adds to the main stream)
  h->rows_side_min = 20480;
  if (precision == LU_OZAKI) {
    // pairs throughout: groups of four while the trailing matrix is large
    // shorten the trailing updates but cost more FP64 block-row work and
    // lower the later slice counts (the first group's exponents are the
    // reference); measured slower end to end at N = 32768
    h->upd = 2;
    h->upd_early = 2;
    h->early_rows = 0;
  }
  h->fast_solve = (N % TRSV_BLOCK) == 0;
  if (N >= TRSV_BLOCK) {  // also for a later lu_resize to a multiple of TRSV_BLOCK
    alloc(&h->Lbinv, (size_t)N * TRSV_BLOCK * h->es);
    alloc(&h->Ubinv, (size_t)N * TRSV_BLOCK * h->es);
    alloc(&h->v3, (size_t)N * 8);
    // accumulators + column-group partials: 2 (block width / 64) per row
    alloc(&h->acc, (size_t)N * 8 * 32);
    // whenever a size up to N can use them, so that the solve a system gets
    // depends on its own size only, not on the capacity of the handle
    if (N >= 2 * TRSV_BLOCK) {
      alloc(&h->Lbinv2, (size_t)N * 2 * TRSV_BLOCK * h->es);
      alloc(&h->Ubinv2, (size_t)N * 2 * TRSV_BLOCK * h->es);
    }
  }
  if (st == OK && hipMemset(h->piv_work, 0, panel_gepp_workspace_bytes(N)) != hipSuccess) {
    set_error("lu_create: hipMemset failed");
    st = EHIP;
  }
  if (st == OK && pivot == PIV_NONE) {
    // the panel's scratch also serves the recursive doubling of its inverse
    // block ((TRSV_BLOCK / 2)^2 elements) and the block-row solves inside a
    // group of panels (nb (MAX_GROUP - 1) nb)
    size_t wp = (size_t)N * 64 > (size_t)65536 ? (size_t)N * 64 : (size_t)65536;
    if (wp < (size_t)nb * nb * (LUHandle::MAX_GROUP - 1)) wp = (size_t)nb * nb * (LUHandle::MAX_GROUP - 1);
    if (wp < (size_t)nb * N) wp = (size_t)nb * N;  // block-row solves of the row lookahead
    alloc(&h->Wp, wp * h->es);
    // default-priority stream (no priority is ever set on any stream)
    if (st == OK && (hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) != hipSuccess ||
                     hipEventCreateWithFlags(&h->ev_next_cols, hipEventDisableTiming) != hipSuccess ||
                     hipEventCreateWithFlags(&h->ev_panel, hipEventDisableTiming) != hipSuccess ||
                     hipEventCreateWithFlags(&h->ev_first, hipEventDisableTiming) != hipSuccess)) {
      set_error("lu_create: stream/event creation failed");
      st = EHIP;
    }
  }
  if (st != OK) {
    lu_destroy(h);
    return st;
  }
  *out = h;
  return OK;
}

// Slice workspace of the trailing update of step k (alternating, so that the
// panel can slice the next step's rows while this step's update reads its own).
// Index of the trailing update that starts at column k.
static int step_of(const LUHandle* h, int k) { return h->step_idx[k / h->nb]; }
static void* oz_ws(const LUHandle* h, int k) { return (step_of(h, k) & 1) ? h->oz_work2 : h->oz_work; }

// Whether the panel [k, k + kb) is a whole TRSV_BLOCK diagonal block whose
// inverses the handle keeps (built right after the panel: the block-row
// solve of the next step multiplies by the L one).
static bool panel_block_inverse(const LUHandle* h, int k, int kb) {
  return h->pivot == PIV_NONE && h->Lbinv != nullptr && kb == TRSV_BLOCK &&
         k % TRSV_BLOCK == 0 && k + TRSV_BLOCK <= h->N;
}

// Inverses of the TRSV_BLOCK x TRSV_BLOCK diagonal blocks of L and U by
// recursive doubling from the 64 x 64 inverses:
//   inv([A 0; C D]) = [inv(A) 0; -inv(D) C inv(A)  inv(D)]
//   inv([A B; 0 D]) = [inv(A) -inv(A) B inv(D); 0  inv(D)]
// batched over all big blocks (uniform strides along the diagonal).
// Blocks b0 .. b0 + nbig - 1; W: scratch of (TRSV_BLOCK / 2)^2 nbig elements.
// parts: INV_L, INV_U or both.
constexpr int INV_L = 1, INV_U = 2, INV_LU = 3;
template <typename T>
static int build_block_inverses(LUHandle* h, hipStream_t s, int b0, int nbig, T* W,
                                int parts = INV_LU) {
  const int N = h->N, TBk = TRSV_BLOCK;
  if (nbig <= 0) return OK;
  const int64_t ld = N;
  const int64_t sLU = (int64_t)TBk * (ld + 1), sB = (int64_t)TBk * TBk;
  const T* LU = static_cast<const T*>(h->LU) + b0 * sLU;
  T* Lb = static_cast<T*>(h->Lbinv) + b0 * sB;
  T* Ub = static_cast<T*>(h->Ubinv) + b0 * sB;
  const size_t i64 = (size_t)b0 * (TBk / 64) * 4096;
  const bool doL = parts & INV_L, doU = parts & INV_U;
  if (doL) MPDOK_TRY(seed_block_inverse<T>(s, static_cast<const T*>(h->Linv) + i64, Lb, nbig));
  if (doU) MPDOK_TRY(seed_block_inverse<T>(s, static_cast<const T*>(h->Uinv) + i64, Ub, nbig));
  for (int hh = 64; hh < TBk; hh *= 2) {
    const int64_t sW = (int64_t)hh * hh;
    if (nbig == 1) {
      // one big block: the TBk / (2 hh) pairs of this level are batched
      // (uniform strides along the diagonal), four launches per level
      const int np = TBk / (2 * hh);
      const int64_t pLU = 2 * (int64_t)hh * (ld + 1), pB = 2 * (int64_t)hh * (TBk + 1);
      const int64_t dD = (int64_t)hh * (TBk + 1);
      const T* C = LU + hh;
      if (doL) {
        MPDOK_TRY(gemm_batched<T>(s, hh, hh, hh, T(1), C, ld, pLU, Lb, TBk, pB, T(0), W, hh, sW, np,
                                  0));
        MPDOK_TRY(gemm_batched<T>(s, hh, hh, hh, T(-1), Lb + dD, TBk, pB, W, hh, sW, T(0), Lb + hh,
                                  TBk, pB, np, 0));
      }
      const T* Bm = LU + (int64_t)hh * ld;
      if (doU) {
        MPDOK_TRY(gemm_batched<T>(s, hh, hh, hh, T(1), Bm, ld, pLU, Ub + dD, TBk, pB, T(0), W, hh,
                                  sW, np, 0));
        MPDOK_TRY(gemm_batched<T>(s, hh, hh, hh, T(-1), Ub, TBk, pB, W, hh, sW, T(0),
                                  Ub + (int64_t)hh * TBk, TBk, pB, np, 0));
      }
      continue;
    }
    for (int off = 0; off < TBk; off += 2 * hh) {
      const int64_t dA = (int64_t)off * (TBk + 1);         // inv(A) inside the big block
      const int64_t dD = (int64_t)(off + hh) * (TBk + 1);  // inv(D)
      // L: T = C inv(A); X21 = -inv(D) T
      const T* C = LU + (off + hh) + (int64_t)off * ld;
      if (doL) {
        MPDOK_TRY(gemm_batched<T>(s, hh, hh, hh, T(1), C, ld, sLU, Lb + dA, TBk, sB, T(0), W, hh,
                                  sW, nbig, 0));
        MPDOK_TRY(gemm_batched<T>(s, hh, hh, hh, T(-1), Lb + dD, TBk, sB, W, hh, sW, T(0),
                                  Lb + (off + hh) + (int64_t)off * TBk, TBk, sB, nbig, 0));
      }
      // U: T = B inv(D); X12 = -inv(A) T
      const T* Bm = LU + off + (int64_t)(off + hh) * ld;
      if (doU) {
        MPDOK_TRY(gemm_batched<T>(s, hh, hh, hh, T(1), Bm, ld, sLU, Ub + dD, TBk, sB, T(0), W, hh,
                                  sW, nbig, 0));
        MPDOK_TRY(gemm_batched<T>(s, hh, hh, hh, T(-1), Ub + dA, TBk, sB, W, hh, sW, T(0),
                                  Ub + off + (int64_t)(off + hh) * TBk, TBk, sB, nbig, 0));
      }
    }
  }
  return OK;
}

// The 2 TRSV_BLOCK-wide diagonal-block inverses from the TRSV_BLOCK-wide
// ones, one more level of the doubling batched over all N / (2 TRSV_BLOCK)
// blocks:
//   inv([A 0; C D]) = [inv(A) 0; -inv(D) C inv(A) inv(D)]
//   inv([A B; 0 D]) = [inv(A) -inv(A) B inv(D); 0 inv(D)]
// W: scratch of TRSV_BLOCK^2 N / (2 TRSV_BLOCK) elements.
template <typename T>
static int build_wide_inverses(LUHandle* h, hipStream_t s, T* W) {
  const int N = h->N, B1 = TRSV_BLOCK, B2 = 2 * TRSV_BLOCK, nbig = N / B2;
  const int64_t ld = N;
  const int64_t sLU = (int64_t)B2 * (ld + 1), sB = (int64_t)B2 * B2, sW = (int64_t)B1 * B1;
  const T* LU = static_cast<const T*>(h->LU);
  T* L2 = static_cast<T*>(h->Lbinv2);
  T* U2 = static_cast<T*>(h->Ubinv2);
  MPDOK_TRY(widen_block_inverse<T>(s, static_cast<const T*>(h->Lbinv), L2, nbig, 1));
  MPDOK_TRY(widen_block_inverse<T>(s, static_cast<const T*>(h->Ubinv), U2, nbig, 0));
  MPDOK_REQUIRE((size_t)nbig * sW <= h->w_elems, "lu: scratch of %zu elements below %zu",
                h->w_elems, (size_t)nbig * sW);
  const int64_t dD = (int64_t)B1 * (B2 + 1);  // inv(D) inside a wide block
  // L: T = C inv(A); X21 = -inv(D) T
  MPDOK_TRY(gemm_batched<T>(s, B1, B1, B1, T(1), LU + B1, ld, sLU, L2, B2, sB, T(0), W, B1, sW,
                            nbig, 0));
  MPDOK_TRY(gemm_batched<T>(s, B1, B1, B1, T(-1), L2 + dD, B2, sB, W, B1, sW, T(0), L2 + B1, B2,
                            sB, nbig, 0));
  // U: T = B inv(D); X12 = -inv(A) T
  MPDOK_TRY(gemm_batched<T>(s, B1, B1, B1, T(1), LU + (int64_t)B1 * ld, ld, sLU, U2 + dD, B2, sB,
                            T(0), W, B1, sW, nbig, 0));
  MPDOK_TRY(gemm_batched<T>(s, B1, B1, B1, T(-1), U2, B2, sB, W, B1, sW, T(0),
                            U2 + (int64_t)B1 * B2, B2, sB, nbig, 0));
  return OK;
}

// Factor the panel of columns [k, k + kb) over rows [k, N): 64-column
// sub-panels (diagonal block LU and the inverses of its triangles in LDS;
// L21 = A21 inv(U11) and the U12 inside the panel = inv(L11) A12 as GEMMs
// with those inverses; panel update by GEMM). Without pivoting, when the
// panel is a whole inverse block of the solves, its TRSV_BLOCK-wide inverses
// follow by recursive doubling: the block-row solve of the next step
// multiplies by the L one. W: scratch (the side stream's, under lookahead).
template <typename T>
static int panel_factor(LUHandle* h, hipStream_t s, int k, int kb, T* W, bool slice = true,
                        int inv_parts = INV_LU) {
  const int N = h->N;
  T* A = static_cast<T*>(h->LU);
  T* Linv = static_cast<T*>(h->Linv);
  T* Uinv = static_cast<T*>(h->Uinv);
  const int64_t ld = N;
  auto at = [&](int i, int j) { return A + i + (int64_t)j * ld; };
  for (int j = k; j < k + kb; j += 64) {
    const int jb = 64 < k + kb - j ? 64 : k + kb - j;
    T* Li = Linv + (size_t)(j / 64) * 4096;
    T* Ui = Uinv + (size_t)(j / 64) * 4096;
    const int m2 = N - j - jb, n2 = k + kb - j - jb;
    if (h->pivot == PIV_PARTIAL) {
      // the sub-panel is factored with implicit pivoting (rows in place);
      // its interchanges then move whole rows, panel columns included
      MPDOK_TRY(panel_gepp<T>(s, A, ld, N, j, jb, h->ipiv, h->d_info, h->piv_work));
      MPDOK_TRY(laswp<T>(s, A, ld, 0, N, h->ipiv, j, j + jb));
      MPDOK_TRY(diag_factor<T>(s, at(j, j), ld, jb, Li, Ui, 64, h->d_info, j, 0));
      // U12 = inv(L11) A12, in place (m = jb <= 64: each workgroup its own columns)
      if (n2 > 0)
        MPDOK_TRY(gemm<T>(s, jb, n2, jb, T(1), Li, 64, at(j, j + jb), ld, T(0), at(j, j + jb), ld,
                          0));
    } else {
      MPDOK_TRY(diag_factor<T>(s, at(j, j), ld, jb, Li, Ui, 64, h->d_info, j, 1));
      // in place, in one launch: L21 = A21 inv(U11) (n = jb <= 64: each
      // workgroup its own rows) and U12 = inv(L11) A12 (m = jb <= 64: its own
      // columns); neither reads what the other writes
      const GemmArgs<T> l21{m2, jb, jb, T(1), at(j + jb, j), ld, Ui, 64, T(0), at(j + jb, j), ld};
      const GemmArgs<T> u12{jb, n2, jb, T(1), Li, 64, at(j, j + jb), ld, T(0), at(j, j + jb), ld};
      MPDOK_TRY(gemm_pair<T>(s, l21, u12));
    }
    if (n2 > 0) {
      if (m2 > 0)
        MPDOK_TRY(gemm<T>(s, m2, n2, jb, T(-1), at(j + jb, j), ld, at(j, j + jb), ld, T(1),
                          at(j + jb, j + jb), ld, 0));
    }
  }
  if (panel_block_inverse(h, k, kb))
    MPDOK_TRY(build_block_inverses<T>(h, s, k / TRSV_BLOCK, 1, W, inv_parts));
  // int8-sliced variant: the rows of L21 for this step's trailing update
  const int n3 = N - k - kb;
  if (slice && h->precision == LU_OZAKI && n3 > 0)
    MPDOK_TRY(ozaki_slice(s, h->slices, n3, n3, kb, reinterpret_cast<const double*>(at(k + kb, k)),
                          ld, nullptr, ld, oz_ws(h, k), 1));
  return OK;
}

// U12 = inv(L11) A12 for the block row [k, k + kb), columns [c0, c0 + ncols)
// (ncols < 0: up to N): one GEMM with the TRSV_BLOCK-wide inverse when the
// panel built it, else blocked forward substitution with the 64 x 64
// inverses. W: scratch of kb ncols elements.
template <typename T>
static int block_row_solve(LUHandle* h, hipStream_t s, int k, int kb, int c0, T* W,
                           int ncols = -1) {
  const int N = h->N, n3 = ncols < 0 ? N - c0 : ncols;
  T* A = static_cast<T*>(h->LU);
  T* Linv = static_cast<T*>(h->Linv);
  const int64_t ld = N;
  auto at = [&](int i, int j) { return A + i + (int64_t)j * ld; };
  if (panel_block_inverse(h, k, kb)) {
    const T* Lb = static_cast<const T*>(h->Lbinv) + (size_t)(k / TRSV_BLOCK) * TRSV_BLOCK * TRSV_BLOCK;
    MPDOK_TRY(gemm<T>(s, kb, n3, kb, T(1), Lb, TRSV_BLOCK, at(k, c0), ld, T(0), W, kb, 0));
    MPDOK_TRY((copy2d<T, T>(s, W, kb, at(k, c0), ld, kb, n3)));
    return OK;
  }
  for (int j = k; j < k + kb; j += 64) {
    const int jb = 64 < k + kb - j ? 64 : k + kb - j;
    T* Li = Linv + (size_t)(j / 64) * 4096;
    MPDOK_TRY(gemm<T>(s, jb, n3, jb, T(1), Li, 64, at(j, c0), ld, T(0), W, jb, 0));
    MPDOK_TRY((copy2d<T, T>(s, W, jb, at(j, c0), ld, jb, n3)));
    const int mrem = k + kb - j - jb;
    if (mrem > 0)
      MPDOK_TRY(gemm<T>(s, mrem, n3, jb, T(-1), at(j + jb, j), ld, at(j, c0), ld, T(1),
                        at(j + jb, c0), ld, 0));
  }
  return OK;
}

static bool adaptive_on(const LUHandle* h) {
  return h->precision == LU_OZAKI && h->adaptive && h->slices <= 5 && h->d_S != nullptr;
}

// Column slices of U12 for step k (its rows of L21 were sliced with the
// panel, smax digits), and with adaptive slice counts the decision for its
// trailing update.
static int slice_step(LUHandle* h, hipStream_t s, int k, int kb, int n3, const double* U12,
                      int64_t ld) {
  MPDOK_TRY(ozaki_slice(s, h->slices, n3, n3, kb, nullptr, ld, U12, ld, oz_ws(h, k), 2));
  if (adaptive_on(h))
    MPDOK_TRY(ozaki_decide(s, h->slices, n3, n3, kb, oz_ws(h, k), h->smin, h->delta, k == 0,
                           h->d_S, h->d_S + 1 + step_of(h, k)));
  return OK;
}

// Trailing update of step k restricted to rows [k+kb+r_off, k+kb+r_off+nrows)
// and columns [k+kb+c_off, k+kb+c_off+ncols). For the int8-sliced variant the
// operands must have been sliced already (row ranges need the adaptive
// kernel; r_off a multiple of 128).
template <typename T>
static int trailing_block(LUHandle* h, hipStream_t s, int k, int kb, int r_off, int nrows,
                          int c_off, int ncols) {
  const int N = h->N, m3 = N - k - kb, n3 = N - k - kb;
  if (ncols <= 0 || nrows <= 0) return OK;
  T* A = static_cast<T*>(h->LU);
  const int64_t ld = N;
  auto at = [&](int i, int j) { return A + i + (int64_t)j * ld; };
  const int round = h->precision == LU_BF16 ? 1
                    : h->precision == LU_FP16 ? 2
                    : h->precision == LU_FP8  ? 3
                                              : 0;
  const size_t nev = h->ev_ops.size();
  if (h->profile) {
    // int8-sliced: S(S+1)/2 int8 products of 2 m n k operations each
    h->ev_ops.push_back(2.0 * nrows * ncols * kb);
    h->ev_step.push_back(step_of(h, k));
    while (h->ev.size() < 2 * h->ev_ops.size()) {
      hipEvent_t e;
      MPDOK_HIP(hipEventCreate(&e));
      h->ev.push_back(e);
    }
    MPDOK_HIP(hipEventRecord(h->ev[2 * nev], s));
  }
  if (adaptive_on(h)) {
    const double* ci = k == 0 && h->c_in ? h->c_in + kb + (int64_t)kb * h->ld_in : nullptr;
    MPDOK_TRY(ozaki_apply_dyn(s, h->slices, m3, n3, kb, r_off, nrows, c_off, ncols,
                              reinterpret_cast<double*>(at(k + kb, k + kb)), ld, oz_ws(h, k),
                              h->d_S + 1 + step_of(h, k), ci, h->ld_in));
  } else if (h->precision == LU_OZAKI) {
    MPDOK_REQUIRE(r_off == 0 && nrows == m3, "trailing update: row ranges need adaptive slicing");
    MPDOK_TRY(ozaki_apply(s, h->slices, m3, n3, kb, c_off, ncols,
                          reinterpret_cast<double*>(at(k + kb, k + kb)), ld, oz_ws(h, k)));
  } else if (round > 0) {
    // BF16 / FP16 / scaled FP8 operands on the matrix cores, FP32 accumulation
    if constexpr (std::is_same<T, float>::value)
      MPDOK_TRY(gemm_lowp(s, nrows, ncols, kb, -1.f, at(k + kb + r_off, k), ld,
                          at(k, k + kb + c_off), ld, 1.f, at(k + kb + r_off, k + kb + c_off), ld,
                          round, h->d_sc));
  } else {
    MPDOK_TRY(gemm<T>(s, nrows, ncols, kb, T(-1), at(k + kb + r_off, k), ld, at(k, k + kb + c_off),
                      ld, T(1), at(k + kb + r_off, k + kb + c_off), ld, 0));
  }
  if (h->profile) MPDOK_HIP(hipEventRecord(h->ev[2 * nev + 1], s));
  return OK;
}
template <typename T>
static int trailing_cols(LUHandle* h, hipStream_t s, int k, int kb, int c_off, int ncols) {
  return trailing_block<T>(h, s, k, kb, 0, h->N - k - kb, c_off, ncols);
}

// Factor the columns [k, k + kbs) (a group of panels of nb columns) over rows
// [k, N), right-looking inside the group: each panel, then the block-row
// solve of its rows over the group's later columns and their FP64 update.
// The int8-sliced variant then slices the rows of the group's L21 (K = kbs)
// for the group's trailing update. Of the panels' TRSV_BLOCK-wide inverses
// only the L ones are built here (the block-row solves need them); the U ones
// serve only the solves after the factorisation (group_u_inverses).
// ev_first (if not null) is recorded once the first panel is factored.
template <typename T>
static int panel_group_factor(LUHandle* h, hipStream_t s, int k, int kbs, T* W,
                              hipEvent_t ev_first = nullptr) {
  const int N = h->N, nb = h->nb;
  T* A = static_cast<T*>(h->LU);
  const int64_t ld = N;
  auto at = [&](int i, int j) { return A + i + (int64_t)j * ld; };
  for (int j = k; j < k + kbs; j += nb) {
    const int jb = nb < k + kbs - j ? nb : k + kbs - j;
    MPDOK_TRY(panel_factor<T>(h, s, j, jb, W, false, INV_L));
    if (ev_first && j == k) MPDOK_HIP(hipEventRecord(ev_first, s));
    const int rem = k + kbs - j - jb;
    if (rem > 0) {
      MPDOK_TRY(block_row_solve<T>(h, s, j, jb, j + jb, W, rem));
      MPDOK_TRY(gemm<T>(s, N - j - jb, rem, jb, T(-1), at(j + jb, j), ld, at(j, j + jb), ld, T(1),
                        at(j + jb, j + jb), ld, 0));
    }
  }
  const int n3 = N - k - kbs;
  if (h->precision == LU_OZAKI && n3 > 0)
    MPDOK_TRY(ozaki_slice(s, h->slices, n3, n3, kbs,
                          reinterpret_cast<const double*>(at(k + kbs, k)), ld, nullptr, ld,
                          oz_ws(h, k), 1));
  return OK;
}

// The TRSV_BLOCK-wide U inverses of the panels of the group [k, k + kbs).
template <typename T>
static int group_u_inverses(LUHandle* h, hipStream_t s, int k, int kbs, T* W) {
  for (int j = k; j < k + kbs; j += h->nb) {
    const int jb = h->nb < k + kbs - j ? h->nb : k + kbs - j;
    if (panel_block_inverse(h, j, jb))
      MPDOK_TRY(build_block_inverses<T>(h, s, j / TRSV_BLOCK, 1, W, INV_U));
  }
  return OK;
}

// U12 of the block rows [k, k + kbs) over columns [c0, N), panel by panel:
// each block row's solve, then the update of the group's block rows below it.
// Only the group's panels p0 .. p1 - 1 (p1 < 0: to the end); the first
// panel's part needs only that panel to be factored.
template <typename T>
static int block_rows_solve(LUHandle* h, hipStream_t s, int k, int kbs, int c0, T* W,
                            int p0 = 0, int p1 = -1) {
  const int N = h->N, nb = h->nb;
  T* A = static_cast<T*>(h->LU);
  const int64_t ld = N;
  auto at = [&](int i, int j) { return A + i + (int64_t)j * ld; };
  const int jend = p1 < 0 ? k + kbs : std::min(k + kbs, k + p1 * nb);
  for (int j = k + p0 * nb; j < jend; j += nb) {
    const int jb = nb < k + kbs - j ? nb : k + kbs - j;
    MPDOK_TRY(block_row_solve<T>(h, s, j, jb, c0, W));
    const int rem = k + kbs - j - jb;
    if (rem > 0)
      MPDOK_TRY(gemm<T>(s, rem, N - c0, jb, T(-1), at(j + jb, j), ld, at(j, c0), ld, T(1),
                        at(j + jb, c0), ld, 0));
  }
  return OK;
}

static bool lookahead_on(const LUHandle* h) {
  return h->lookahead && h->pivot == PIV_NONE && h->side != nullptr;
}

// Steps of the next factorisation: groups of panels on the lookahead path of
// the int8-sliced variant, single panels otherwise.
static void build_schedule(LUHandle* h) {
  const int N = h->N, nb = h->nb;
  const bool grouped = lookahead_on(h) && h->precision == LU_OZAKI;
  h->group_start.clear();
  for (int k = 0; k < N;) {
    h->group_start.push_back(k);
    int g = 1;
    if (grouped) g = h->upd_early > h->upd && N - k >= h->early_rows ? h->upd_early : h->upd;
    k += g * nb;
  }
  h->group_start.push_back(N);
  h->step_idx.assign(N / nb + 2, 0);
  for (size_t i = 0; i + 1 < h->group_start.size(); ++i) h->step_idx[h->group_start[i] / nb] = (int)i;
}

// The input matrix's columns [c0, c1) into the handle's storage.
static int copy_columns(LUHandle* h, hipStream_t s, const double* A, int64_t lda, int order, int c0,
                        int c1) {
  if (c1 <= c0) return OK;
  const double* src = order == COL_MAJOR ? A + (int64_t)c0 * lda : A + c0;
  char* dst = static_cast<char*>(h->LU) + (size_t)c0 * h->N * h->es;
  return cast_matrix(s, src, F64, lda, order, dst, fp64_storage(h->precision) ? F64 : F32, h->N,
                     COL_MAJOR, h->N, c1 - c0);
}

// Factor the handle's storage. On the lookahead path the caller has copied
// only the first step's columns; the rest of the input (rest_A, rest_lda,
// rest_order) is copied here on the main stream while the lookahead stream
// factors the first step.
template <typename T>
static int factor_impl(LUHandle* h, hipStream_t s, const double* rest_A, int64_t rest_lda,
                       int rest_order) {
  const int N = h->N, nb = h->nb;
  T* A = static_cast<T*>(h->LU);
  T* W = static_cast<T*>(h->W);
  const int64_t ld = N;
  auto at = [&](int i, int j) { return A + i + (int64_t)j * ld; };
  const bool look = lookahead_on(h);
  if (look) {
    // Step k: [main] U12 of block row k, slicing, trailing update of the next
    // panel's columns; [side] factor the next panel as soon as its columns are
    // final; [main] the rest of the trailing update meanwhile. The panel
    // touches only columns [k+kb, k+2kb) and its own scratch Wp; the rest of
    // the trailing update only columns >= k+2kb.
    // one step per group of panels (one trailing update each)
    T* Wp = static_cast<T*>(h->Wp);
    const std::vector<int>& gs = h->group_start;
    // the first step's columns are in place: factor them on the lookahead
    // stream while the main stream copies the rest of the input
    MPDOK_HIP(hipEventRecord(h->ev_next_cols, s));
    MPDOK_HIP(hipStreamWaitEvent(h->side, h->ev_next_cols, 0));
    MPDOK_TRY(panel_group_factor<T>(h, h->side, 0, gs[1], Wp, h->ev_first));
    MPDOK_HIP(hipEventRecord(h->ev_panel, h->side));
    // row lookahead for group g (see LUHandle::rows_side_min); the first
    // group's rows stay on the main stream (its copy runs there)
    // The int8-sliced update takes row ranges that start on a 128-row tile,
    // so a group whose width is not a multiple of 128 keeps its rows on the
    // main stream.
    auto rows_on_side = [&](size_t g) {
      return g >= 1 && g + 1 < gs.size() && gs[g + 1] < N && h->rows_side_min > 0 &&
             N - gs[g] >= h->rows_side_min && (adaptive_on(h) || h->precision != LU_OZAKI) &&
             (h->precision != LU_OZAKI || (gs[g + 1] - gs[g]) % 128 == 0);
    };
    // The U inverses of group g go to the stream that waits for the other:
    // the lookahead stream while the trailing updates are long (the first
    // group and the row-lookahead steps), the main stream, after its part of
    // the step, once it waits for each group's panels.
    auto u_on_main = [&](size_t g) { return g >= 1 && !rows_on_side(g); };
    MPDOK_TRY(group_u_inverses<T>(h, h->side, 0, gs[1], Wp));
    // The first trailing update writes every entry of rows and columns
    // [gs[1], N): with adaptive int8 slices and a column-major input it reads
    // them from the input, and only the first group's rows of the later
    // columns (its block-row solve works on them in place) are copied.
    h->c_in = nullptr;
    if (rest_A && rest_order == COL_MAJOR && adaptive_on(h) && gs.size() > 2) {
      MPDOK_TRY((copy2d<double, double>(s, rest_A + (int64_t)gs[1] * rest_lda, rest_lda,
                                        reinterpret_cast<double*>(at(0, gs[1])), ld, gs[1],
                                        N - gs[1])));
      h->c_in = rest_A;
      h->ld_in = rest_lda;
    } else if (rest_A) {
      MPDOK_TRY(copy_columns(h, s, rest_A, rest_lda, rest_order, gs[1], N));
    }
    for (size_t g = 0; g + 1 < gs.size(); ++g) {
      const int k = gs[g], kb = gs[g + 1] - k;
      const int n3 = N - k - kb;
      if (n3 > 0 && !rows_on_side(g)) {
        // the first panel's block row (and its update of the group's other
        // rows) while the lookahead stream factors the group's other panels
        MPDOK_HIP(hipStreamWaitEvent(s, h->ev_first, 0));
        MPDOK_TRY(block_rows_solve<T>(h, s, k, kb, k + kb, W, 0, 1));
      }
      MPDOK_HIP(hipStreamWaitEvent(s, h->ev_panel, 0));
      if (n3 <= 0) {
        if (u_on_main(g)) MPDOK_TRY(group_u_inverses<T>(h, s, k, kb, W));
        break;
      }
      if (!rows_on_side(g)) {
        MPDOK_TRY(block_rows_solve<T>(h, s, k, kb, k + kb, W, 1));
        if (h->precision == LU_OZAKI)
          MPDOK_TRY(slice_step(h, s, k, kb, n3, reinterpret_cast<const double*>(at(k, k + kb)), ld));
      }
      const int nk = gs[g + 2] - gs[g + 1];  // the next group
      const bool nxt = rows_on_side(g + 1);
      // the next group's columns, then (row lookahead) its rows
      MPDOK_TRY(trailing_block<T>(h, s, k, kb, 0, n3, 0, nk));
      if (nxt) MPDOK_TRY(trailing_block<T>(h, s, k, kb, 0, nk, nk, n3 - nk));
      MPDOK_HIP(hipEventRecord(h->ev_next_cols, s));
      // the rest of the update is enqueued before the panel's launches, so
      // that it never waits for the host to issue them
      MPDOK_TRY(trailing_block<T>(h, s, k, kb, nxt ? nk : 0, nxt ? n3 - nk : n3, nk, n3 - nk));
      MPDOK_HIP(hipStreamWaitEvent(h->side, h->ev_next_cols, 0));
      MPDOK_TRY(panel_group_factor<T>(h, h->side, k + kb, nk, Wp, nxt ? nullptr : h->ev_first));
      if (nxt) {
        const int k2 = k + kb, n32 = N - k2 - nk;
        MPDOK_TRY(block_rows_solve<T>(h, h->side, k2, nk, k2 + nk, Wp));
        if (h->precision == LU_OZAKI)
          MPDOK_TRY(slice_step(h, h->side, k2, nk, n32,
                               reinterpret_cast<const double*>(at(k2, k2 + nk)), ld));
      }
      MPDOK_HIP(hipEventRecord(h->ev_panel, h->side));
      if (!u_on_main(g + 1)) MPDOK_TRY(group_u_inverses<T>(h, h->side, k + kb, nk, Wp));
      if (u_on_main(g)) MPDOK_TRY(group_u_inverses<T>(h, s, k, kb, W));
    }
    h->c_in = nullptr;
    // everything the lookahead stream holds, U inverses included
    MPDOK_HIP(hipEventRecord(h->ev_panel, h->side));
    MPDOK_HIP(hipStreamWaitEvent(s, h->ev_panel, 0));
    if (h->fast_solve && nb != TRSV_BLOCK)
      MPDOK_TRY(build_block_inverses<T>(h, s, 0, N / TRSV_BLOCK, W));
    h->wide_solve = h->fast_solve && h->Lbinv2 != nullptr && N % (2 * TRSV_BLOCK) == 0;
    if (h->wide_solve) MPDOK_TRY(build_wide_inverses<T>(h, s, W));
    return OK;
  }
  for (int k = 0; k < N; k += nb) {
    const int kb = nb < N - k ? nb : N - k;
    MPDOK_TRY(panel_factor<T>(h, s, k, kb, W));
    const int n3 = N - k - kb;
    if (n3 <= 0) continue;
    MPDOK_TRY(block_row_solve<T>(h, s, k, kb, k + kb, W));
    if (h->precision == LU_OZAKI)
      MPDOK_TRY(slice_step(h, s, k, kb, n3, reinterpret_cast<const double*>(at(k, k + kb)), ld));
    MPDOK_TRY(trailing_cols<T>(h, s, k, kb, 0, n3));
  }
  if (h->pivot == PIV_PARTIAL) MPDOK_TRY(ipiv_to_perm(s, h->ipiv, h->perm, N));
  if (h->fast_solve && !(h->pivot == PIV_NONE && nb == TRSV_BLOCK))
    MPDOK_TRY(build_block_inverses<T>(h, s, 0, N / TRSV_BLOCK, W));
  h->wide_solve = h->fast_solve && h->Lbinv2 != nullptr && N % (2 * TRSV_BLOCK) == 0;
  if (h->wide_solve) MPDOK_TRY(build_wide_inverses<T>(h, s, W));
  return OK;
}

int lu_factor(LUHandle* h, hipStream_t s, const double* A, int64_t lda, int order, int* info) {
  MPDOK_REQUIRE(h && A, "lu_factor: null handle or matrix");
  MPDOK_REQUIRE(lda >= h->N, "lu_factor: lda < N");
  h->factored = false;
  const int N = h->N;
  h->ev_ops.clear();
  h->ev_step.clear();
  MPDOK_HIP(hipMemsetAsync(h->d_info, 0, sizeof(int), s));
  // copy (and convert / transpose) A into the handle's column-major storage:
  // on the lookahead path only the first step's columns before the
  // factorisation starts (it copies the rest beside the first panel)
  build_schedule(h);
  h->c_in = nullptr;
  const bool split = lookahead_on(h);
  MPDOK_TRY(copy_columns(h, s, A, lda, order, 0, split ? h->group_start[1] : N));
  const int st = fp64_storage(h->precision)
                     ? factor_impl<double>(h, s, split ? A : nullptr, lda, order)
                     : factor_impl<float>(h, s, split ? A : nullptr, lda, order);
  if (st != OK) {
    // work already enqueued on either stream may still use the handle's
    // buffers: drain both before the caller can reuse or destroy them
    if (h->side) (void)hipStreamSynchronize(h->side);
    (void)hipStreamSynchronize(s);
    return st;
  }
  int hinfo = 0;
  MPDOK_HIP(hipMemcpyAsync(&hinfo, h->d_info, sizeof(int), hipMemcpyDeviceToHost, s));
  const int nsteps = (int)h->group_start.size() - 1;
  h->steps_S.assign(nsteps, h->precision == LU_OZAKI ? h->slices : 0);
  if (adaptive_on(h) && N > h->nb)
    MPDOK_HIP(hipMemcpyAsync(h->steps_S.data(), h->d_S + 1, sizeof(int) * (nsteps - 1),
                             hipMemcpyDeviceToHost, s));
  MPDOK_HIP(hipStreamSynchronize(s));
  if (info) *info = hinfo;
  h->factored = (hinfo == 0);
  if (h->profile) {
    h->trailing_ms = 0.0;
    h->trailing_ops = 0.0;
    h->trailing_flops = 0.0;
    h->trailing_launches = (int)h->ev_ops.size();
    for (size_t i = 0; i < h->ev_ops.size(); ++i) {
      float ms = 0.f;
      MPDOK_HIP(hipEventElapsedTime(&ms, h->ev[2 * i], h->ev[2 * i + 1]));
      h->trailing_ms += ms;
      const int Sk = h->precision == LU_OZAKI ? h->steps_S[h->ev_step[i]] : 1;
      h->trailing_ops += h->ev_ops[i] * (h->precision == LU_OZAKI ? Sk * (Sk + 1) / 2.0 : 1.0);
      h->trailing_flops += h->ev_ops[i];
    }
  }
  return OK;
}

template <typename T, typename TV>
static int solve_impl(LUHandle* h, hipStream_t s, const double* b, double* x) {
  const int N = h->N;
  TV* v = static_cast<TV*>(h->v1);
  TV* v2 = static_cast<TV*>(h->v2);
  if (sizeof(TV) == 8) {
    MPDOK_TRY((copy2d<double, double>(s, b, N, reinterpret_cast<double*>(v), N, N, 1)));
  } else {
    MPDOK_TRY((copy2d<double, float>(s, b, N, reinterpret_cast<float*>(v), N, N, 1)));
  }
  if (h->pivot == PIV_PARTIAL) {
    MPDOK_TRY(permute<TV>(s, v, v2, h->perm, N));
    TV* t = v;
    v = v2;
    v2 = t;
  }
  if (h->fast_solve) {
    // v: permuted right-hand side; result into v2 (v3 holds the forward result)
    MPDOK_TRY((trsv_block_inverse<T, TV>(
        s, static_cast<const T*>(h->LU), N, N,
        static_cast<const T*>(h->wide_solve ? h->Lbinv2 : h->Lbinv),
        static_cast<const T*>(h->wide_solve ? h->Ubinv2 : h->Ubinv), v, v2,
        static_cast<TV*>(h->v3), static_cast<TV*>(h->acc),
        h->wide_solve ? 2 * TRSV_BLOCK : TRSV_BLOCK)));
    v = v2;
  } else {
    MPDOK_TRY((trsv_blocked<T, TV>(s, static_cast<const T*>(h->LU), N, N,
                                   static_cast<const T*>(h->Linv),
                                   static_cast<const T*>(h->Uinv), v)));
  }
  if (sizeof(TV) == 8) {
    MPDOK_TRY((copy2d<double, double>(s, reinterpret_cast<double*>(v), N, x, N, N, 1)));
  } else {
    MPDOK_TRY((copy2d<float, double>(s, reinterpret_cast<float*>(v), N, x, N, N, 1)));
  }
  return OK;
}

// X[:, 0:k] = inv(P L U) B[:, 0:k] (FP64, column-major, leading dimensions
// ldb / ldx >= N): blocks of up to MULTI_NR right-hand sides go through the
// triangular solves together (8, then 4, 2, 1), so the factors are read once
// per block; each column equals the single solve's bit for bit. Without the
// inverse-based solves (N not a multiple of TRSV_BLOCK) the columns are solved
// one at a time.
template <typename T, typename TV>
static int solve_multi_impl(LUHandle* h, hipStream_t s, const double* B, int64_t ldb, double* X,
                            int64_t ldx, int k) {
  const int N = h->N;
  constexpr int NR = LUHandle::MULTI_NR;
  const size_t vb = (size_t)N * NR * sizeof(TV);
  const int tbw = h->wide_solve ? 2 * TRSV_BLOCK : TRSV_BLOCK;
  if (!h->mv) MPDOK_HIP(hipMalloc(&h->mv, 3 * vb));
  if (!h->macc) MPDOK_HIP(hipMalloc(&h->macc, 2 * (size_t)(2 * TRSV_BLOCK / 64) * vb));
  TV* v = static_cast<TV*>(h->mv);
  TV* v2 = v + (size_t)N * NR;
  TV* v3 = v2 + (size_t)N * NR;
  for (int c0 = 0; c0 < k;) {
    const int nr = k - c0 >= 8 ? 8 : k - c0 >= 4 ? 4 : k - c0 >= 2 ? 2 : 1;
    if (sizeof(TV) == 8)
      MPDOK_TRY((copy2d<double, double>(s, B + (int64_t)c0 * ldb, ldb,
                                        reinterpret_cast<double*>(v), N, N, nr)));
    else
      MPDOK_TRY((copy2d<double, float>(s, B + (int64_t)c0 * ldb, ldb,
                                       reinterpret_cast<float*>(v), N, N, nr)));
    TV* in = v;
    if (h->pivot == PIV_PARTIAL) {
      for (int q = 0; q < nr; ++q)
        MPDOK_TRY(permute<TV>(s, v + (size_t)q * N, v2 + (size_t)q * N, h->perm, N));
      in = v2;
    }
    TV* out = in == v ? v2 : v;
    MPDOK_TRY((trsv_block_inverse_multi<T, TV>(
        s, static_cast<const T*>(h->LU), N, N,
        static_cast<const T*>(h->wide_solve ? h->Lbinv2 : h->Lbinv),
        static_cast<const T*>(h->wide_solve ? h->Ubinv2 : h->Ubinv), in, out, v3,
        static_cast<TV*>(h->macc), tbw, nr)));
    if (sizeof(TV) == 8)
      MPDOK_TRY((copy2d<double, double>(s, reinterpret_cast<double*>(out), N,
                                        X + (int64_t)c0 * ldx, ldx, N, nr)));
    else
      MPDOK_TRY((copy2d<float, double>(s, reinterpret_cast<float*>(out), N,
                                       X + (int64_t)c0 * ldx, ldx, N, nr)));
    c0 += nr;
  }
  return OK;
}

int lu_solve_multi(LUHandle* h, hipStream_t s, const double* B, int64_t ldb, double* X,
                   int64_t ldx, int k) {
  MPDOK_REQUIRE(h && B && X && k >= 0, "lu_solve_multi: bad arguments");
  MPDOK_REQUIRE(ldb >= h->N && ldx >= h->N, "lu_solve_multi: leading dimension < N");
  if (!h->factored) {
    set_error("lu_solve_multi: no factorization available");
    return ENOFACTOR;
  }
  if (k == 0) return OK;
  if (!h->fast_solve) {
    for (int c = 0; c < k; ++c) MPDOK_TRY(lu_solve(h, s, B + (int64_t)c * ldb, X + (int64_t)c * ldx));
    return OK;
  }
  if (fp64_storage(h->precision)) return solve_multi_impl<double, double>(h, s, B, ldb, X, ldx, k);
  return solve_multi_impl<float, float>(h, s, B, ldb, X, ldx, k);
}

int lu_solve(LUHandle* h, hipStream_t s, const double* b, double* x) {
  MPDOK_REQUIRE(h && b && x, "lu_solve: null argument");
  if (!h->factored) {
    set_error("lu_solve: no factorization available");
    return ENOFACTOR;
  }
  if (fp64_storage(h->precision)) return solve_impl<double, double>(h, s, b, x);
  return solve_impl<float, float>(h, s, b, x);
}

}  // namespace mpdok

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
MPDOK_API int mpdok_lu_create(int N, int precision, int pivot, int nb, int slices, void** out) {
  return mpdok::lu_create(N, precision, pivot, nb, slices, reinterpret_cast<mpdok::LUHandle**>(out));
}
MPDOK_API int mpdok_lu_factor(void* h, hipStream_t s, const double* A, int64_t lda, int order,
                              int* info) {
  return mpdok::lu_factor(static_cast<mpdok::LUHandle*>(h), s, A, lda, order, info);
}
MPDOK_API int mpdok_lu_solve(void* h, hipStream_t s, const double* b, double* x) {
  return mpdok::lu_solve(static_cast<mpdok::LUHandle*>(h), s, b, x);
}
MPDOK_API int mpdok_lu_solve_multi(void* h, hipStream_t s, const double* B, int64_t ldb, double* X,
                                   int64_t ldx, int nrhs) {
  return mpdok::lu_solve_multi(static_cast<mpdok::LUHandle*>(h), s, B, ldb, X, ldx, nrhs);
}
MPDOK_API int mpdok_lu_resize(void* h, int n) {
  return mpdok::lu_resize(static_cast<mpdok::LUHandle*>(h), n);
}
MPDOK_API int mpdok_lu_destroy(void* h) {
  mpdok::lu_destroy(static_cast<mpdok::LUHandle*>(h));
  return 0;
}
// Device pointers to the packed factors (column-major, ld = N, element size
// 8 for FP64/int8-sliced and 4 otherwise) and the row permutation (NULL
// without pivoting), for tests and inspection.
MPDOK_API int mpdok_lu_factors(void* hv, void** LU, int** perm, int* elem_size) {
  auto* h = static_cast<mpdok::LUHandle*>(hv);
  MPDOK_REQUIRE(h != nullptr, "lu_factors: null handle");
  if (LU) *LU = h->LU;
  if (perm) *perm = h->pivot == mpdok::PIV_PARTIAL ? h->perm : nullptr;
  if (elem_size) *elem_size = (int)h->es;
  return 0;
}
// Overlap of the next panel with the trailing update (no-pivot handles only).
MPDOK_API int mpdok_lu_set_lookahead(void* hv, int on) {
  auto* h = static_cast<mpdok::LUHandle*>(hv);
  MPDOK_REQUIRE(h != nullptr, "lu_set_lookahead: null handle");
  h->lookahead = on != 0;
  return 0;
}
// HIP-event timing of the trailing-update launches of the next factorisations.
MPDOK_API int mpdok_lu_set_profiling(void* hv, int on) {
  auto* h = static_cast<mpdok::LUHandle*>(hv);
  MPDOK_REQUIRE(h != nullptr, "lu_set_profiling: null handle");
  h->profile = on != 0;
  return 0;
}
// Totals of the last profiled factorisation: summed event time of the trailing
// updates (ms), their algorithmic operations (int8 ops for the sliced variant,
// flops otherwise) and the number of timed launches.
MPDOK_API int mpdok_lu_stats(void* hv, double* trailing_ms, double* trailing_ops, int* launches,
                             double* trailing_flops) {
  auto* h = static_cast<mpdok::LUHandle*>(hv);
  MPDOK_REQUIRE(h != nullptr, "lu_stats: null handle");
  if (trailing_ms) *trailing_ms = h->trailing_ms;
  if (trailing_ops) *trailing_ops = h->trailing_ops;
  if (launches) *launches = h->trailing_launches;
  if (trailing_flops) *trailing_flops = h->trailing_flops;
  return 0;
}
// Plain GEMM on the matrix cores (C = alpha A B + beta C, column-major), for
// parity tests of the trailing-update kernel. dtype 0 = FP64, 1 = FP32;
// round 0/1/2 = none/BF16/FP16 operand rounding (FP32 only).
MPDOK_API int mpdok_gemm(hipStream_t s, int dtype, int m, int n, int k, double alpha,
                         const void* A, int64_t lda, const void* B, int64_t ldb, double beta,
                         void* C, int64_t ldc, int round) {
  if (dtype == mpdok::F64)
    return mpdok::gemm<double>(s, m, n, k, alpha, (const double*)A, lda, (const double*)B, ldb,
                               beta, (double*)C, ldc, 0);
  if (round >= 1 && round <= 3) {
    static float* sc = nullptr;  // the FP8 scales of this entry point's launches
    if (round == 3 && sc == nullptr)
      MPDOK_HIP(hipMalloc(&sc, mpdok::gemm_lowp_scale_floats() * sizeof(float)));
    return mpdok::gemm_lowp(s, m, n, k, (float)alpha, (const float*)A, lda, (const float*)B, ldb,
                            (float)beta, (float*)C, ldc, round, sc);
  }
  return mpdok::gemm<float>(s, m, n, k, (float)alpha, (const float*)A, lda, (const float*)B, ldb,
                            (float)beta, (float*)C, ldc, 0);
}

// Adaptive slice counts of the int8-sliced variant (default on for slices
// <= 5): each trailing update takes the fewest slices in [smin, slices] whose
// error bound stays within the first update's minus `delta` bits.
MPDOK_API int mpdok_lu_set_adaptive(void* hv, int on, int smin, int delta) {
  auto* h = static_cast<mpdok::LUHandle*>(hv);
  MPDOK_REQUIRE(h != nullptr, "lu_set_adaptive: null handle");
  MPDOK_REQUIRE(!on || (smin >= 2 && smin <= h->slices),
                "lu_set_adaptive: smin=%d outside [2, %d]", smin, h->slices);
  MPDOK_REQUIRE(!on || h->slices <= 5, "lu_set_adaptive: needs slices <= 5 (got %d)", h->slices);
  h->adaptive = on != 0;
  if (on) {
    h->smin = smin;
    h->delta = delta;
  }
  return 0;
}
// Panels per trailing update on the lookahead path of the int8-sliced
// variant: groups of `panels` (1 or 2), and of `panels_early` (panels ..
// MAX_GROUP) while the trailing matrix has at least early_rows rows
// (defaults 2, 2 and 0).
MPDOK_API int mpdok_lu_set_update_panels(void* hv, int panels, int panels_early, int early_rows) {
  auto* h = static_cast<mpdok::LUHandle*>(hv);
  MPDOK_REQUIRE(h != nullptr, "lu_set_update_panels: null handle");
  MPDOK_REQUIRE(panels == 1 || panels == 2, "lu_set_update_panels: panels must be 1 or 2, got %d",
                panels);
  MPDOK_REQUIRE(panels_early >= panels && panels_early <= mpdok::LUHandle::MAX_GROUP,
                "lu_set_update_panels: panels_early=%d outside [%d, %d]", panels_early, panels,
                mpdok::LUHandle::MAX_GROUP);
  MPDOK_REQUIRE(early_rows >= 0, "lu_set_update_panels: early_rows < 0");
  MPDOK_REQUIRE(panels_early == 1 || h->precision == mpdok::LU_OZAKI,
                "lu_set_update_panels: panel groups need the int8-sliced variant");
  h->upd = panels;
  h->upd_early = panels_early;
  h->early_rows = early_rows;
  return 0;
}
// First columns of the last factorisation's steps (one trailing update per
// step: a panel, or a group of panels), n of them at most.
MPDOK_API int mpdok_lu_step_starts(void* hv, int* out, int n, int* nsteps) {
  auto* h = static_cast<mpdok::LUHandle*>(hv);
  MPDOK_REQUIRE(h != nullptr, "lu_step_starts: null handle");
  const int ns = h->group_start.empty() ? 0 : (int)h->group_start.size() - 1;
  for (int i = 0; i < ns && i < n; ++i) out[i] = h->group_start[i];
  if (nsteps) *nsteps = ns;
  return 0;
}
// Row lookahead (no-pivot lookahead path; the int8-sliced variant needs
// adaptive slice counts): while the trailing matrix has at least min_rows
// rows, a step's block-row solve, U12 slicing and slice-count decision run on
// the lookahead stream right after its panels. 0 turns it off.
MPDOK_API int mpdok_lu_set_row_lookahead(void* hv, int min_rows) {
  auto* h = static_cast<mpdok::LUHandle*>(hv);
  MPDOK_REQUIRE(h != nullptr, "lu_set_row_lookahead: null handle");
  MPDOK_REQUIRE(min_rows >= 0, "lu_set_row_lookahead: min_rows < 0");
  h->rows_side_min = min_rows;
  return 0;
}
// Slice counts the last factorisation used, one per trailing update (n of them at most).
MPDOK_API int mpdok_lu_slice_counts(void* hv, int* out, int n, int* nsteps) {
  auto* h = static_cast<mpdok::LUHandle*>(hv);
  MPDOK_REQUIRE(h != nullptr, "lu_slice_counts: null handle");
  const int ns = (int)h->steps_S.size();
  for (int i = 0; i < ns && i < n; ++i) out[i] = h->steps_S[i];
  if (nsteps) *nsteps = ns;
  return 0;
}
