// Internal declarations shared between the translation units of libmpdok_hip.
#pragma once
#include "common.h"

namespace mpdok {

size_t gemv_workspace_bytes(int M, int N, int dtype, int order);
size_t resid_workspace_bytes(int N, int order);
int gemv_f32(hipStream_t s, const float* A, int64_t lda, int order, int M, int N, const float* x,
             float* y, float alpha, void* work, size_t work_bytes, const int* skip);
int gemv_f64(hipStream_t s, const double* A, int64_t lda, int order, int M, int N, const double* x,
             double* y, double alpha, void* work, size_t work_bytes, const int* skip);
int residual_f64(hipStream_t s, const double* A, int64_t lda, int order, int N, const double* x,
                 const double* b, double* r, int mode, void* work, size_t work_bytes);
// R = B - A X for nrhs columns, column-major A (the single residual's bits per column)
size_t resid_multi_workspace_bytes(int N);
int residual_multi_f64(hipStream_t s, const double* A, int64_t lda, int N, const double* X,
                       int64_t ldx, const double* B, int64_t ldb, double* R, int64_t ldr,
                       int nrhs, int mode, void* work, size_t work_bytes);

// blas1.hip
size_t reduce_workspace_bytes(int n);
// out[0] = sum x[i]*y[i] in FP64 (x, y FP32 or FP64), deterministic two-pass.
int dot_f32(hipStream_t s, const float* x, const float* y, int n, double* out, void* work);
int dot_f64(hipStream_t s, const double* x, const double* y, int n, double* out, void* work);
int cast_f64_to_f32(hipStream_t s, const double* in, float* out, int64_t n);
int cast_f32_to_f64(hipStream_t s, const float* in, double* out, int64_t n);
int axpby(hipStream_t s, int64_t n, double alpha, const void* x, int xdt, double beta,
          const void* y, int ydt, void* out, int odt);
int cast_matrix(hipStream_t s, const void* in, int in_dtype, int64_t ldi, int in_order, void* out,
                int out_dtype, int64_t ldo, int out_order, int M, int N);

}  // namespace mpdok
