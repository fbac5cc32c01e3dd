// Shared definitions for the MI355X (gfx950) dense-operator Krylov library.
//
// Every exported entry point is `extern "C"`, takes plain pointers and sizes,
// and (for the new entry points) an explicit hipStream_t plus an int status.
// Device code is written for 64-wide CDNA4 wavefronts.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

#define MPDOK_API extern "C" __attribute__((visibility("default")))

namespace mpdok {

enum Status : int {
  OK = 0,
  EINVAL_ = -1,   // bad argument (shape, null pointer, unsupported layout)
  ENOMEM_ = -2,   // device/host allocation failed
  EHIP = -3,      // HIP runtime error
  ENOFACTOR = -4, // solve requested without a factorization
  ESINGULAR = -5, // zero pivot met during factorization
};

// Thread-local last-error text, readable through mpdok_last_error().
void set_error(const char* fmt, ...);
const char* error_string();  // this thread's last error message

// Process-wide kernel-shape choices (set by the tuner through
// mpdok_set_tuning; read at launch time).
enum TuneKey : int {
  TUNE_OZAKI_SHAPE = 0,  // trailing update's block shape (0: 128 x 128, four stages; 1: 128 x 64,
                         // two workgroups per CU; 2: 128 x 128, two stages; 3: 256 x 64)
  TUNE_OZAKI_GROUP = 1,  // int8-sliced GEMM: row tiles per group of the tile order (0: 8)
  TUNE_DIAG_KERNEL = 3,    // 64 x 64 diagonal LU: 0 = blocked, 1 = unblocked
  TUNE_SLICE_ROWS = 4,     // int8 row slicing: 0 = 2-D grid (two passes), 1 = one workgroup per 64 rows
  TUNE_OZAKI_EPI = 5,      // trailing update's C: 0 = read at the tile's start (S <= 3), 1 = after the K loop
  TUNE_COUNT = 8,
};
int tuning_get(int key);

constexpr int WAVE = 64;

// Memory order of a dense matrix argument.
enum Order : int { ROW_MAJOR = 0, COL_MAJOR = 1 };

// Element type codes used by the C-ABI.
enum DType : int { F64 = 0, F32 = 1, BF16 = 2, F16 = 3 };

}  // namespace mpdok

#define MPDOK_HIP(expr)                                                          \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) {                                                      \
      mpdok::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr,              \
                       hipGetErrorString(e_));                                   \
      return e_ == hipErrorOutOfMemory ? mpdok::ENOMEM_ : mpdok::EHIP;           \
    }                                                                            \
  } while (0)

#define MPDOK_TRY(expr)                 \
  do {                                  \
    int s_ = (expr);                    \
    if (s_ != mpdok::OK) return s_;     \
  } while (0)

#define MPDOK_REQUIRE(cond, ...)                 \
  do {                                           \
    if (!(cond)) {                               \
      mpdok::set_error(__VA_ARGS__);             \
      return mpdok::EINVAL_;                     \
    }                                            \
  } while (0)

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
namespace mpdok {

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Error-free transformations (Knuth TwoSum, FMA-based TwoProd). Used by the
// Dot2 residual of Ogita, Rump and Oishi (SIAM J. Sci. Comput. 26(6), 2005).
// Contraction of a*b+c into an FMA must stay off in these functions: a fused
// product would no longer match the separately rounded p the corrections use.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
#pragma clang fp contract(off)
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void two_prod(double a, double b, double& p, double& e) {
#pragma clang fp contract(off)
  p = a * b;
  e = __builtin_fma(a, b, -p);
}
// A double-double accumulator: value = hi + lo, |lo| small.
struct dd {
  double hi, lo;
};
__device__ __forceinline__ void dd_add_prod(dd& acc, double a, double b) {
#pragma clang fp contract(off)
  double p, q, s, t;
  two_prod(a, b, p, q);
  two_sum(acc.hi, p, s, t);
  acc.hi = s;
  acc.lo += t + q;
}
__device__ __forceinline__ dd dd_merge(dd a, dd b) {
#pragma clang fp contract(off)
  double s, t;
  two_sum(a.hi, b.hi, s, t);
  return dd{s, a.lo + b.lo + t};
}
__device__ __forceinline__ dd wave_sum_dd(dd v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    dd o{__shfl_xor(v.hi, off, 64), __shfl_xor(v.lo, off, 64)};
    v = dd_merge(v, o);
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T nt_load(const T* p) {
  return __builtin_nontemporal_load(p);
}

// Tile order of the matrix-core GEMMs: consecutive workgroups are dealt
// round-robin to the 8 XCDs, so workgroup b is mapped to tile
// (b % 8) * per_xcd + b / 8, giving every XCD one contiguous range of tiles;
// inside a range, tiles run down a group of GROUP_M row blocks before moving to
// the next column block, so a band of A rows stays in that XCD's L2. Launch
// with tiles_launch(tm, tn) workgroups; surplus workgroups get `false`.
constexpr int NXCD = 8, GROUP_M = 8;
__device__ __forceinline__ bool tile_of(int b, int tm, int tn, int& ti, int& tj,
                                        int group = GROUP_M) {
  const int T = tm * tn;
  const int per = (T + NXCD - 1) / NXCD;
  const int t = (b % NXCD) * per + b / NXCD;
  if (b / NXCD >= per || t >= T) return false;
  const int width = group * tn;
  const int g = t / width, first = g * group;
  const int gsz = min(group, tm - first);
  const int r = t % width;
  ti = first + r % gsz;
  tj = r / gsz;
  return true;
}
inline int tiles_launch(int tm, int tn) { return (tm * tn + NXCD - 1) / NXCD * NXCD; }

}  // namespace mpdok
