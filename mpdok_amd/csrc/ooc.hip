// Out-of-core operator: the FP32 kernel matrix lives in host memory (RAM, or
// pages read from an on-disk copy) and crosses PCIe tile by tile.
//
//  * stream_gemv_f32: y = A x for a row-major FP32 A (M x N) in host memory.
//    Double-buffered: tile t + 1 is copied on a library-owned copy stream
//    while tile t is multiplied on the caller's stream, so the product runs at
//    the host-to-device copy rate. Replaces mpdok_ooc.py:158-201, which copies
//    one tile at a time (pageable, no overlap) and multiplies it after.
//  * rbf_build_host: A[row0 : row0 + rows, :] (FP32, nugget added after the
//    cast, mpdok_ooc.py:39-46,118-123) assembled on the GPU tile by tile and
//    copied to host memory, the device-to-host copy of tile t overlapping the
//    assembly of tile t + 1.
//
// The copy stream and its events are created once per device (default
// priority, non-blocking) and kept for the life of the process. Calls on one
// device serialise their enqueueing through a mutex; the work itself runs in
// stream order on the caller's stream, which must not reuse `work` before the
// products have run (any later work on the same stream is ordered after them).
#include <mutex>

#include "common.h"
#include "gemv.h"
#include "rbf.h"

namespace mpdok {

namespace {

constexpr int MAX_DEVICES = 64;

struct CopyEngine {
  hipStream_t stream = nullptr;
  hipEvent_t ready[2] = {nullptr, nullptr};  // tile b is in (or out of) the buffer
  hipEvent_t freed[2] = {nullptr, nullptr};  // buffer b may be overwritten
};

std::mutex g_mu;
CopyEngine g_engines[MAX_DEVICES];

int copy_engine(CopyEngine** out) {
  int dev = 0;
  MPDOK_HIP(hipGetDevice(&dev));
  MPDOK_REQUIRE(dev >= 0 && dev < MAX_DEVICES, "ooc: device %d out of range", dev);
  CopyEngine& ce = g_engines[dev];
  if (!ce.stream) {
    MPDOK_HIP(hipStreamCreateWithFlags(&ce.stream, hipStreamNonBlocking));
    for (int b = 0; b < 2; ++b) {
      MPDOK_HIP(hipEventCreateWithFlags(&ce.ready[b], hipEventDisableTiming));
      MPDOK_HIP(hipEventCreateWithFlags(&ce.freed[b], hipEventDisableTiming));
    }
  }
  *out = &ce;
  return OK;
}

}  // namespace

size_t stream_gemv_workspace_bytes(int tile_rows, int N) {
  return 2 * (size_t)(tile_rows > 0 ? tile_rows : 0) * (size_t)(N > 0 ? N : 0) * sizeof(float);
}

int stream_gemv_f32(hipStream_t s, const float* A_host, int64_t lda, int M, int N, int tile_rows,
                    const float* x, float* y, void* work, size_t work_bytes) {
  MPDOK_REQUIRE(M >= 0 && N >= 0 && tile_rows >= 1, "stream_gemv: bad sizes");
  if (M == 0) return OK;
  MPDOK_REQUIRE(A_host && x && y && work, "stream_gemv: null pointer");
  MPDOK_REQUIRE(lda >= N, "stream_gemv: lda < N");
  MPDOK_REQUIRE(work_bytes >= stream_gemv_workspace_bytes(tile_rows, N),
                "stream_gemv: workspace too small (%zu < %zu)", work_bytes,
                stream_gemv_workspace_bytes(tile_rows, N));
  std::lock_guard<std::mutex> lk(g_mu);
  CopyEngine* ce = nullptr;
  MPDOK_TRY(copy_engine(&ce));
  float* buf[2] = {static_cast<float*>(work), static_cast<float*>(work) + (size_t)tile_rows * N};
  // both buffers become free once the caller's stream reaches this point
  MPDOK_HIP(hipEventRecord(ce->freed[0], s));
  MPDOK_HIP(hipEventRecord(ce->freed[1], s));
  for (int t = 0, i = 0; i < M; ++t, i += tile_rows) {
    const int b = t & 1, rows = tile_rows < M - i ? tile_rows : M - i;
    MPDOK_HIP(hipStreamWaitEvent(ce->stream, ce->freed[b], 0));
    const float* src = A_host + (int64_t)i * lda;
    if (lda == N)
      MPDOK_HIP(hipMemcpyAsync(buf[b], src, (size_t)rows * N * sizeof(float),
                               hipMemcpyHostToDevice, ce->stream));
    else
      MPDOK_HIP(hipMemcpy2DAsync(buf[b], (size_t)N * sizeof(float), src,
                                 (size_t)lda * sizeof(float), (size_t)N * sizeof(float), rows,
                                 hipMemcpyHostToDevice, ce->stream));
    MPDOK_HIP(hipEventRecord(ce->ready[b], ce->stream));
    MPDOK_HIP(hipStreamWaitEvent(s, ce->ready[b], 0));
    MPDOK_TRY(gemv_f32(s, buf[b], N, ROW_MAJOR, rows, N, x, y + i, 1.0f, nullptr, 0, nullptr));
    MPDOK_HIP(hipEventRecord(ce->freed[b], s));
  }
  return OK;
}

int rbf_build_host(hipStream_t s, const double* coords, int N, int D, int kind, double p,
                   double reg, int row0, int rows, int tile_rows, float* A_host, int64_t ldh,
                   void* work, size_t work_bytes) {
  MPDOK_REQUIRE(rows >= 0 && tile_rows >= 1 && row0 >= 0 && row0 + rows <= N,
                "rbf_build_host: bad rows");
  if (rows == 0) return OK;
  MPDOK_REQUIRE(coords && A_host && work, "rbf_build_host: null pointer");
  MPDOK_REQUIRE(ldh >= N, "rbf_build_host: ldh < N");
  MPDOK_REQUIRE(work_bytes >= stream_gemv_workspace_bytes(tile_rows, N),
                "rbf_build_host: workspace too small");
  std::lock_guard<std::mutex> lk(g_mu);
  CopyEngine* ce = nullptr;
  MPDOK_TRY(copy_engine(&ce));
  float* buf[2] = {static_cast<float*>(work), static_cast<float*>(work) + (size_t)tile_rows * N};
  MPDOK_HIP(hipEventRecord(ce->freed[0], s));
  MPDOK_HIP(hipEventRecord(ce->freed[1], s));
  for (int t = 0, i = 0; i < rows; ++t, i += tile_rows) {
    const int b = t & 1, r = tile_rows < rows - i ? tile_rows : rows - i;
    // buffer b is free once its previous tile has left for the host
    MPDOK_HIP(hipStreamWaitEvent(s, ce->freed[b], 0));
    MPDOK_TRY(rbf_assemble(s, coords, N, D, kind, p, reg, 1, row0 + i, r, 0, N, buf[b], F32, N,
                           ROW_MAJOR));
    MPDOK_HIP(hipEventRecord(ce->ready[b], s));
    MPDOK_HIP(hipStreamWaitEvent(ce->stream, ce->ready[b], 0));
    float* dst = A_host + (int64_t)i * ldh;
    if (ldh == N)
      MPDOK_HIP(hipMemcpyAsync(dst, buf[b], (size_t)r * N * sizeof(float), hipMemcpyDeviceToHost,
                               ce->stream));
    else
      MPDOK_HIP(hipMemcpy2DAsync(dst, (size_t)ldh * sizeof(float), buf[b],
                                 (size_t)N * sizeof(float), (size_t)N * sizeof(float), r,
                                 hipMemcpyDeviceToHost, ce->stream));
    MPDOK_HIP(hipEventRecord(ce->freed[b], ce->stream));
  }
  // the caller's stream (and a synchronize on it) covers the last copies
  MPDOK_HIP(hipStreamWaitEvent(s, ce->freed[0], 0));
  MPDOK_HIP(hipStreamWaitEvent(s, ce->freed[1], 0));
  return OK;
}

}  // namespace mpdok

MPDOK_API size_t mpdok_stream_gemv_workspace_bytes(int tile_rows, int N) {
  return mpdok::stream_gemv_workspace_bytes(tile_rows, N);
}
MPDOK_API int mpdok_stream_gemv_f32(hipStream_t s, const float* A_host, int64_t lda, int M, int N,
                                    int tile_rows, const float* x, float* y, void* work,
                                    size_t work_bytes) {
  return mpdok::stream_gemv_f32(s, A_host, lda, M, N, tile_rows, x, y, work, work_bytes);
}
MPDOK_API int mpdok_rbf_build_host(hipStream_t s, const double* coords, int N, int D, int kind,
                                   double p, double reg, int row0, int rows, int tile_rows,
                                   float* A_host, int64_t ldh, void* work, size_t work_bytes) {
  return mpdok::rbf_build_host(s, coords, N, D, kind, p, reg, row0, rows, tile_rows, A_host, ldh,
                               work, work_bytes);
}
