#pragma once
#include "common.h"

namespace mpdok {
constexpr int MAXD = 8;

__device__ __forceinline__ double kernel_value(int kind, double p, double d2) {
  if (kind == 0) return exp(-p * d2);
  const double a = sqrt(3.0 * d2) / p;
  return (1.0 + a) * exp(-a);
}

// Squared distance by the Gram identity, evaluated in the reference's order:
// (sq_i + sq_j) - 2 * (c_i . c_j), clamped at zero.
// DT: the dimension when it is fixed at compile time (1 .. MAXD), 0: D at run
// time. Both forms do the same operations in the same order.
template <int DT = 0>
__device__ __forceinline__ double gram_d2(const double* ci, double sqi, const double* cj, double sqj,
                                          int D) {
  const int n = DT ? DT : D;
  double dot = ci[0] * cj[0];
#pragma unroll
  for (int d = 1; d < MAXD; ++d)
    if (d < n) dot = __builtin_fma(ci[d], cj[d], dot);
  double d2 = (sqi + sqj) - 2.0 * dot;
  return d2 > 0.0 ? d2 : 0.0;
}

template <int DT = 0>
__device__ __forceinline__ double sqnorm(const double* c, int D) {
  const int n = DT ? DT : D;
  double s = c[0] * c[0];
#pragma unroll
  for (int d = 1; d < MAXD; ++d)
    if (d < n) s += c[d] * c[d];
  return s;
}

// Launches MACRO(DT) with DT = D for the common dimensions 2 and 3, else 0.
#define MPDOK_RBF_DIM_DISPATCH(D, MACRO) \
  do {                                   \
    if ((D) == 2) {                      \
      MACRO(2);                          \
    } else if ((D) == 3) {               \
      MACRO(3);                          \
    } else {                             \
      MACRO(0);                          \
    }                                    \
  } while (0)

int rbf_assemble(hipStream_t st, const double* coords, int N, int D, int kind, double p,
                 double reg, int reg_mode, int row0, int rows, int col0, int cols, void* out,
                 int out_dtype, int64_t ld, int order);
int rbf_mean_d2(hipStream_t st, const double* coords, int n, int D, double* out);
int rbf_matvec(hipStream_t st, const double* coords, int N, int D, int kind, double p, double reg,
               int elem_dtype, const void* x, void* y, int xy_dtype, int row0, int rows);
}  // namespace mpdok
