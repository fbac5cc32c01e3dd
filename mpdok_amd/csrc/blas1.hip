// Vector reductions, casts and error reporting.
//
// Dot products and norms accumulate in FP64 whatever the input type and reduce
// in a fixed order (per-block partials, then one block), so a result never
// depends on scheduling. Matrix casts replace the FP64->FP32 copies of
// dense_krylov.py:47 and the native scratch cast of mpdok_ops.py:366,478-482.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include "common.h"
#include "gemv.h"
#include "lu.h"

namespace mpdok {

static thread_local char g_err[1024] = "";

static int g_tuning[TUNE_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0};
int tuning_get(int key) { return key >= 0 && key < TUNE_COUNT ? g_tuning[key] : 0; }
int tuning_set(int key, int value) {
  MPDOK_REQUIRE(key >= 0 && key < TUNE_COUNT, "tuning: unknown key %d", key);
  MPDOK_REQUIRE(key != TUNE_OZAKI_GROUP || (value >= 0 && value <= 256),
                "tuning: group %d out of range", value);
  MPDOK_REQUIRE(key != TUNE_OZAKI_SHAPE || (value >= 0 && value <= 3),
                "tuning: trailing-update shape must be in [0, 3], got %d", value);
  MPDOK_REQUIRE(key != 2 && key != 6 && key != 7, "tuning: key %d is not used", key);
  MPDOK_REQUIRE((key != TUNE_DIAG_KERNEL && key != TUNE_SLICE_ROWS && key != TUNE_OZAKI_EPI) ||
                    (value >= 0 && value <= 1),
                "tuning: key %d takes 0 or 1, got %d", key, value);
  g_tuning[key] = value;
  return OK;
}

const char* error_string() { return g_err; }
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

constexpr int RED_BLOCKS = 512;

size_t reduce_workspace_bytes(int n) { return RED_BLOCKS * sizeof(double); }

template <typename T>
__global__ __launch_bounds__(256) void dot_partial_kernel(const T* __restrict__ x,
                                                          const T* __restrict__ y, int n,
                                                          double* __restrict__ part) {
  double acc = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    acc = __builtin_fma((double)x[i], (double)y[i], acc);
  acc = wave_sum(acc);
  __shared__ double s[4];
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
}

__global__ __launch_bounds__(256) void sum_partials_kernel(const double* __restrict__ part,
                                                           int np, double* __restrict__ out) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) acc += part[i];
  acc = wave_sum(acc);
  __shared__ double s[4];
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = (s[0] + s[1]) + (s[2] + s[3]);
}

template <typename T>
static int dot_launch(hipStream_t s, const T* x, const T* y, int n, double* out, void* work) {
  MPDOK_REQUIRE(out && work && n >= 0 && (n == 0 || (x && y)), "dot: bad arguments");
  int nb = (n + 255) / 256;
  if (nb > RED_BLOCKS) nb = RED_BLOCKS;
  if (nb < 1) nb = 1;
  double* part = static_cast<double*>(work);
  dot_partial_kernel<T><<<nb, 256, 0, s>>>(x, y, n, part);
  sum_partials_kernel<<<1, 256, 0, s>>>(part, nb, out);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

int dot_f32(hipStream_t s, const float* x, const float* y, int n, double* out, void* work) {
  return dot_launch<float>(s, x, y, n, out, work);
}
int dot_f64(hipStream_t s, const double* x, const double* y, int n, double* out, void* work) {
  return dot_launch<double>(s, x, y, n, out, work);
}

template <typename TI, typename TO>
__global__ __launch_bounds__(256) void cast_kernel(const TI* __restrict__ in, TO* __restrict__ out,
                                                   int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    out[i] = (TO)in[i];
}

static dim3 grid_1d(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return dim3((unsigned)g);
}

int cast_f64_to_f32(hipStream_t s, const double* in, float* out, int64_t n) {
  MPDOK_REQUIRE(in && out && n >= 0, "cast: bad arguments");
  if (n == 0) return OK;
  cast_kernel<double, float><<<grid_1d(n), 256, 0, s>>>(in, out, n);
  MPDOK_HIP(hipGetLastError());
  return OK;
}
int cast_f32_to_f64(hipStream_t s, const float* in, double* out, int64_t n) {
  MPDOK_REQUIRE(in && out && n >= 0, "cast: bad arguments");
  if (n == 0) return OK;
  cast_kernel<float, double><<<grid_1d(n), 256, 0, s>>>(in, out, n);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

// Strided 2-D cast with optional transpose of the storage order, 64x64 tiles
// staged through LDS so both the read and the write are coalesced.
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void cast2d_kernel(const TI* __restrict__ in, int64_t ldi,
                                                     int in_order, TO* __restrict__ out,
                                                     int64_t ldo, int out_order, int M, int N) {
  __shared__ TO tile[64][65];
  const int bi = blockIdx.y * 64, bj = blockIdx.x * 64;
  const int t = threadIdx.x;
  // Read: the fast index follows the input storage order.
  for (int e = t; e < 64 * 64; e += 256) {
    int fast = e & 63, slow = e >> 6;
    int i, j;
    if (in_order == ROW_MAJOR) { i = bi + slow; j = bj + fast; }
    else { i = bi + fast; j = bj + slow; }
    if (i < M && j < N) {
      TI v = in_order == ROW_MAJOR ? in[(int64_t)i * ldi + j] : in[i + (int64_t)j * ldi];
      tile[i - bi][j - bj] = (TO)v;
    }
  }
  __syncthreads();
  for (int e = t; e < 64 * 64; e += 256) {
    int fast = e & 63, slow = e >> 6;
    int i, j;
    if (out_order == ROW_MAJOR) { i = bi + slow; j = bj + fast; }
    else { i = bi + fast; j = bj + slow; }
    if (i < M && j < N) {
      TO v = tile[i - bi][j - bj];
      if (out_order == ROW_MAJOR) out[(int64_t)i * ldo + j] = v;
      else out[i + (int64_t)j * ldo] = v;
    }
  }
}

int cast_matrix(hipStream_t s, const void* in, int in_dtype, int64_t ldi, int in_order, void* out,
                int out_dtype, int64_t ldo, int out_order, int M, int N) {
  MPDOK_REQUIRE(in && out && M >= 0 && N >= 0, "cast_matrix: bad arguments");
  if (M == 0 || N == 0) return OK;
  // a column-major FP64 copy: the copy kernel's 16-byte path where it applies
  if (in_dtype == F64 && out_dtype == F64 && in_order == COL_MAJOR && out_order == COL_MAJOR)
    return copy2d<double, double>(s, (const double*)in, ldi, (double*)out, ldo, M, N);
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  MPDOK_REQUIRE(grid.y <= 65535 * 64, "cast_matrix: too many rows");
  if (in_dtype == F64 && out_dtype == F32)
    cast2d_kernel<double, float><<<grid, 256, 0, s>>>((const double*)in, ldi, in_order, (float*)out, ldo, out_order, M, N);
  else if (in_dtype == F64 && out_dtype == F64)
    cast2d_kernel<double, double><<<grid, 256, 0, s>>>((const double*)in, ldi, in_order, (double*)out, ldo, out_order, M, N);
  else if (in_dtype == F32 && out_dtype == F32)
    cast2d_kernel<float, float><<<grid, 256, 0, s>>>((const float*)in, ldi, in_order, (float*)out, ldo, out_order, M, N);
  else if (in_dtype == F32 && out_dtype == F64)
    cast2d_kernel<float, double><<<grid, 256, 0, s>>>((const float*)in, ldi, in_order, (double*)out, ldo, out_order, M, N);
  else
    MPDOK_REQUIRE(false, "cast_matrix: unsupported dtype pair %d->%d", in_dtype, out_dtype);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

// out = alpha * x + beta * y, each operand FP32 or FP64 (y may be null),
// evaluated in FP64 and rounded once to the output type.
__device__ __forceinline__ double ld_any(const void* p, int dt, int64_t i) {
  return dt == F64 ? static_cast<const double*>(p)[i] : (double)static_cast<const float*>(p)[i];
}
__global__ __launch_bounds__(256) void axpby_kernel(int64_t n, double alpha, const void* x, int xdt,
                                                    double beta, const void* y, int ydt, void* out,
                                                    int odt) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double v = alpha * ld_any(x, xdt, i);
    if (y != nullptr) v = __builtin_fma(beta, ld_any(y, ydt, i), v);
    if (odt == F64)
      static_cast<double*>(out)[i] = v;
    else
      static_cast<float*>(out)[i] = (float)v;
  }
}

int axpby(hipStream_t s, int64_t n, double alpha, const void* x, int xdt, double beta,
          const void* y, int ydt, void* out, int odt) {
  MPDOK_REQUIRE(x && out && n >= 0, "axpby: bad arguments");
  MPDOK_REQUIRE((xdt == F64 || xdt == F32) && (odt == F64 || odt == F32) &&
                    (y == nullptr || ydt == F64 || ydt == F32),
                "axpby: dtypes must be F64 or F32");
  if (n == 0) return OK;
  axpby_kernel<<<grid_1d(n), 256, 0, s>>>(n, alpha, x, xdt, beta, y, ydt, out, odt);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

}  // namespace mpdok

MPDOK_API int mpdok_axpby(hipStream_t s, int n, double alpha, const void* x, int xdt, double beta,
                          const void* y, int ydt, void* out, int odt) {
  return mpdok::axpby(s, n, alpha, x, xdt, beta, y, ydt, out, odt);
}

MPDOK_API const char* mpdok_last_error(void) { return mpdok::g_err; }
MPDOK_API int mpdok_set_tuning(int key, int value) { return mpdok::tuning_set(key, value); }
MPDOK_API int mpdok_get_tuning(int key) { return mpdok::tuning_get(key); }
MPDOK_API size_t mpdok_reduce_workspace_bytes(int n) { return mpdok::reduce_workspace_bytes(n); }
MPDOK_API int mpdok_dot_f32(hipStream_t s, const float* x, const float* y, int n, double* out,
                            void* work) {
  return mpdok::dot_f32(s, x, y, n, out, work);
}
MPDOK_API int mpdok_dot_f64(hipStream_t s, const double* x, const double* y, int n, double* out,
                            void* work) {
  return mpdok::dot_f64(s, x, y, n, out, work);
}
MPDOK_API int mpdok_cast_matrix(hipStream_t s, const void* in, int in_dtype, int64_t ldi,
                                int in_order, void* out, int out_dtype, int64_t ldo, int out_order,
                                int M, int N) {
  return mpdok::cast_matrix(s, in, in_dtype, ldi, in_order, out, out_dtype, ldo, out_order, M, N);
}
