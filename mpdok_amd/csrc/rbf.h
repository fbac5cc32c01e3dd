#pragma once
#include "common.h"

namespace mpdok {
int rbf_assemble(hipStream_t st, const double* coords, int N, int D, int kind, double p,
                 double reg, int reg_mode, int row0, int rows, int col0, int cols, void* out,
                 int out_dtype, int64_t ld, int order);
int rbf_mean_d2(hipStream_t st, const double* coords, int n, int D, double* out);
int rbf_matvec(hipStream_t st, const double* coords, int N, int D, int kind, double p, double reg,
               int elem_dtype, const void* x, void* y, int xy_dtype, int row0, int rows);
}  // namespace mpdok
