// Krylov building blocks kept on the device.
//
// Arnoldi orthogonalisation (dense_krylov.py:289-303 one-pass batched CGS;
// mpdok_ooc.py:263-282 two-pass DGKS) is split into
//   mdot   : part[i, c] = <V_i, w> over chunk c            grid (chunks, k+1)
//   update : h_i = sum_c part[i, c];  w -= sum_i h_i V_i;  H[i, k] += h_i;
//            per-block partial of ||w||^2 (last pass only)
//   finish : nrm = sqrt(sum of partials); H[k+1, k] = nrm; V_{k+1} = w / nrm,
//            or raise the breakdown flag when nrm <= 1e-12.
// A control block `ctl` lives in device memory: ctl[0] = stop (every later
// kernel of the cycle returns at once), ctl[1] = completed steps m, ctl[2] =
// breakdown seen. The host reads it once per restart cycle instead of
// synchronising on every dot product as the reference does.
//
// The device Givens step follows mpdok_ooc.py:284-301 (the port of the
// native inner_gmres): rotate column k by the previous rotations, build the
// rotation that zeroes H[k+1, k], update g, and stop when |g[k+1]|/beta < tol.
// One deliberate difference: a breakdown still rotates column k before the
// cycle stops (mpdok_ooc.py:278-282 leaves it unrotated, which makes the
// back-substitution use a stale column).
#include "common.h"
#include "krylov.h"

namespace mpdok {

constexpr int MDOT_EPT = 16;                 // elements per thread in mdot
constexpr int MDOT_CHUNK = 256 * MDOT_EPT;   // elements per mdot block
constexpr int UPD_CHUNK = 1024;              // elements per update block (4 per thread)

static int nchunks_mdot(int n) { return (n + MDOT_CHUNK - 1) / MDOT_CHUNK; }
static int nblocks_upd(int n) { return (n + UPD_CHUNK - 1) / UPD_CHUNK; }

size_t arnoldi_workspace_bytes(int n, int max_k) {
  // mdot partials + norm partials
  return ((size_t)(max_k + 1) * nchunks_mdot(n) + (size_t)nblocks_upd(n) + 64) * sizeof(double);
}

__device__ __forceinline__ double block_sum_256(double v, double* sh) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  return (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

template <typename T>
__global__ __launch_bounds__(256) void mdot_kernel(const T* __restrict__ V, int64_t ldv,
                                                   const T* __restrict__ w, int n,
                                                   double* __restrict__ part, int nch,
                                                   const int* __restrict__ ctl) {
  if (ctl[0]) return;
  const int c = blockIdx.x, i = blockIdx.y;
  const T* vi = V + (int64_t)i * ldv;
  double acc = 0.0;
  const int base = c * MDOT_CHUNK;
#pragma unroll 4
  for (int e = 0; e < MDOT_EPT; ++e) {
    const int j = base + e * 256 + threadIdx.x;
    if (j < n) acc = __builtin_fma((double)vi[j], (double)w[j], acc);
  }
  __shared__ double sh[4];
  acc = block_sum_256(acc, sh);
  if (threadIdx.x == 0) part[(int64_t)i * nch + c] = acc;
}

// w -= sum_{i<=k} h_i V_i with h reduced from the mdot partials. Block 0 adds h
// into column k of H (ldh = leading dimension, column-major (m+1) x m).
template <typename T>
__global__ __launch_bounds__(256) void cgs_update_kernel(const T* __restrict__ V, int64_t ldv,
                                                         T* __restrict__ w, int n, int kp1,
                                                         const double* __restrict__ part, int nch,
                                                         double* __restrict__ Hcol,
                                                         double* __restrict__ nrm_part,
                                                         const int* __restrict__ ctl) {
  if (ctl[0]) return;
  __shared__ double h[MAX_RESTART + 1];
  for (int i = threadIdx.x; i < kp1; i += 256) {
    double s = 0.0;
    for (int c = 0; c < nch; ++c) s += part[(int64_t)i * nch + c];
    h[i] = s;
  }
  __syncthreads();
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < kp1; i += 256) Hcol[i] += h[i];
  const int j0 = blockIdx.x * UPD_CHUNK + threadIdx.x;
  double nn = 0.0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int j = j0 + e * 256;
    if (j < n) {
      double acc = (double)w[j];
      for (int i = 0; i < kp1; ++i) acc -= h[i] * (double)V[(int64_t)i * ldv + j];
      T wn = (T)acc;
      w[j] = wn;
      nn = __builtin_fma((double)wn, (double)wn, nn);
    }
  }
  if (nrm_part != nullptr) {
    __shared__ double sh[4];
    nn = block_sum_256(nn, sh);
    if (threadIdx.x == 0) nrm_part[blockIdx.x] = nn;
  }
}

// nrm = ||w||; H[k+1,k] = nrm; V_{k+1} = w * (1/nrm) unless nrm <= thresh.
template <typename T>
__global__ __launch_bounds__(256) void arnoldi_finish_kernel(T* __restrict__ vnext,
                                                             const T* __restrict__ w, int n,
                                                             const double* __restrict__ nrm_part,
                                                             int nparts, double* __restrict__ Hsub,
                                                             int k, double thresh, int stop_now,
                                                             int* __restrict__ ctl) {
  if (ctl[0]) return;
  __shared__ double sh[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) s += nrm_part[i];
  s = block_sum_256(s, sh);
  const double nrm = sqrt(s);
  const bool brk = !(nrm > thresh);
  if (!brk) {
    const T inv = (T)(1.0 / nrm);
    for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) vnext[j] = w[j] * inv;
  }
  // Every block has the same nrm; the control words are written by block 0 of
  // the grid after all blocks have read ctl[0] at entry (all blocks of this
  // kernel already passed the check, and no block reads ctl again).
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *Hsub = nrm;
    ctl[1] = k + 1;
    if (brk) ctl[2] = 1;
    if (brk && stop_now) ctl[0] = 1;
  }
}

// Single-lane Givens update of column k (see file header).
__global__ void givens_kernel(double* __restrict__ H, int ldh, int k, double* __restrict__ R,
                              double* __restrict__ cs, double* __restrict__ sn,
                              double* __restrict__ g, double beta, double tol,
                              int* __restrict__ ctl) {
  if (ctl[0] || threadIdx.x != 0) return;
  // R keeps the rotated copy; H stays the plain Hessenberg matrix.
  double* rk = R + (int64_t)k * ldh;
  const double* hk = H + (int64_t)k * ldh;
  for (int i = 0; i <= k + 1; ++i) rk[i] = hk[i];
  for (int j = 0; j < k; ++j) {
    const double t = cs[j] * rk[j] + sn[j] * rk[j + 1];
    rk[j + 1] = -sn[j] * rk[j] + cs[j] * rk[j + 1];
    rk[j] = t;
  }
  const double rho = hypot(rk[k], rk[k + 1]);
  double c = 1.0, s = 0.0;
  if (rho > 0.0) {
    c = rk[k] / rho;
    s = rk[k + 1] / rho;
  }
  cs[k] = c;
  sn[k] = s;
  rk[k] = rho;
  rk[k + 1] = 0.0;
  g[k + 1] = -s * g[k];
  g[k] = c * g[k];
  if (ctl[2] || fabs(g[k + 1]) / beta < tol) ctl[0] = 1;
}

// x (TX) += sum_{i<m} y_i V_i   (y FP64 on the device)
template <typename T, typename TX>
__global__ __launch_bounds__(256) void basis_update_kernel(const T* __restrict__ V, int64_t ldv,
                                                           int m, const double* __restrict__ y,
                                                           TX* __restrict__ x, int n, int overwrite) {
  __shared__ double ys[MAX_RESTART + 1];
  for (int i = threadIdx.x; i < m; i += 256) ys[i] = y[i];
  __syncthreads();
  for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) {
    double acc = 0.0;
    for (int i = 0; i < m; ++i) acc = __builtin_fma(ys[i], (double)V[(int64_t)i * ldv + j], acc);
    if (overwrite)
      x[j] = (TX)acc;
    else
      x[j] = (TX)((double)x[j] + (double)(T)acc);
  }
}

// v_out = v_in * scale (scale = 1/beta computed from a device norm^2)
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void scale_by_inv_norm_kernel(const TI* __restrict__ in,
                                                                TO* __restrict__ out, int n,
                                                                const double* __restrict__ nrm2) {
  const double nrm = sqrt(nrm2[0]);
  const TO inv = nrm > 0.0 ? (TO)(1.0 / nrm) : (TO)0;
  for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) out[j] = (TO)in[j] * inv;
}

static dim3 grid_1d(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return dim3((unsigned)g);
}

template <typename T>
int arnoldi_orth(hipStream_t s, T* V, int64_t ldv, int k, T* w, int n, double* Hcol,
                 int passes, double thresh, int stop_now, int* ctl, void* work, size_t work_bytes) {
  MPDOK_REQUIRE(V && w && Hcol && ctl && work, "arnoldi_orth: null pointer");
  MPDOK_REQUIRE(k >= 0 && k < MAX_RESTART, "arnoldi_orth: k=%d out of range", k);
  MPDOK_REQUIRE(passes == 1 || passes == 2, "arnoldi_orth: passes must be 1 or 2");
  MPDOK_REQUIRE(work_bytes >= arnoldi_workspace_bytes(n, k), "arnoldi_orth: workspace too small");
  const int nch = nchunks_mdot(n), nub = nblocks_upd(n);
  double* part = static_cast<double*>(work);
  double* npart = part + (size_t)(k + 1) * nch;
  for (int p = 0; p < passes; ++p) {
    mdot_kernel<T><<<dim3(nch, k + 1), 256, 0, s>>>(V, ldv, w, n, part, nch, ctl);
    cgs_update_kernel<T><<<nub, 256, 0, s>>>(V, ldv, w, n, k + 1, part, nch, Hcol,
                                             p == passes - 1 ? npart : nullptr, ctl);
  }
  arnoldi_finish_kernel<T><<<grid_1d(n), 256, 0, s>>>(V + (int64_t)(k + 1) * ldv, w, n, npart, nub,
                                                      Hcol + k + 1, k, thresh, stop_now, ctl);
  MPDOK_HIP(hipGetLastError());
  return OK;
}
template int arnoldi_orth<float>(hipStream_t, float*, int64_t, int, float*, int, double*, int,
                                 double, int, int*, void*, size_t);
template int arnoldi_orth<double>(hipStream_t, double*, int64_t, int, double*, int, double*, int,
                                  double, int, int*, void*, size_t);

int givens_step(hipStream_t s, double* H, int ldh, int k, double* R, double* cs, double* sn,
                double* g, double beta, double tol, int* ctl) {
  givens_kernel<<<1, 64, 0, s>>>(H, ldh, k, R, cs, sn, g, beta, tol, ctl);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

template <typename T, typename TX>
int basis_update(hipStream_t s, const T* V, int64_t ldv, int m, const double* y, TX* x, int n,
                 int overwrite) {
  MPDOK_REQUIRE(m >= 0 && m <= MAX_RESTART + 1, "basis_update: m=%d out of range", m);
  if (n == 0) return OK;
  if (m == 0 && !overwrite) return OK;
  basis_update_kernel<T, TX><<<grid_1d(n), 256, 0, s>>>(V, ldv, m, y, x, n, overwrite);
  MPDOK_HIP(hipGetLastError());
  return OK;
}
template int basis_update<float, float>(hipStream_t, const float*, int64_t, int, const double*,
                                        float*, int, int);
template int basis_update<float, double>(hipStream_t, const float*, int64_t, int, const double*,
                                         double*, int, int);
template int basis_update<double, double>(hipStream_t, const double*, int64_t, int, const double*,
                                          double*, int, int);

template <typename TI, typename TO>
int scale_by_inv_norm(hipStream_t s, const TI* in, TO* out, int n, const double* nrm2) {
  scale_by_inv_norm_kernel<TI, TO><<<grid_1d(n), 256, 0, s>>>(in, out, n, nrm2);
  MPDOK_HIP(hipGetLastError());
  return OK;
}
template int scale_by_inv_norm<double, float>(hipStream_t, const double*, float*, int,
                                              const double*);
template int scale_by_inv_norm<float, float>(hipStream_t, const float*, float*, int, const double*);
template int scale_by_inv_norm<double, double>(hipStream_t, const double*, double*, int,
                                               const double*);

// ---------------------------------------------------------------------------
// Conjugate gradients (dense_krylov.py:191-198) with device scalars.
//   cg_update1: alpha = rs / pAp; x += alpha p; r -= alpha Ap; partial ||r||^2
//   cg_update2: rs_new = sum(partials); beta = rs_new / rs; p = r + beta p
// alpha and beta are rounded to FP32 before use, as the reference's CuPy
// expressions on FP32 arrays do.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void cg_update1_kernel(float* __restrict__ x, float* __restrict__ r,
                                                         const float* __restrict__ p,
                                                         const float* __restrict__ Ap, int n,
                                                         const double* __restrict__ rs,
                                                         const double* __restrict__ pAp,
                                                         double* __restrict__ part) {
  const float alpha = (float)(rs[0] / pAp[0]);
  double acc = 0.0;
  for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) {
    x[j] = x[j] + alpha * p[j];
    const float rn = r[j] - alpha * Ap[j];
    r[j] = rn;
    acc = __builtin_fma((double)rn, (double)rn, acc);
  }
  __shared__ double sh[4];
  acc = block_sum_256(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void cg_update2_kernel(const float* __restrict__ r,
                                                         float* __restrict__ p, int n,
                                                         const double* __restrict__ part, int np,
                                                         const double* __restrict__ rs_old,
                                                         double* __restrict__ rs_new) {
  __shared__ double sh[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) s += part[i];
  s = block_sum_256(s, sh);
  const float beta = (float)(s / rs_old[0]);
  for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) p[j] = r[j] + beta * p[j];
  if (blockIdx.x == 0 && threadIdx.x == 0) rs_new[0] = s;
}

int cg_step_updates(hipStream_t s, float* x, float* r, float* p, const float* Ap, int n,
                    const double* rs_old, const double* pAp, double* rs_new, void* work) {
  dim3 g = grid_1d(n);
  if (g.x > 512) g.x = 512;
  double* part = static_cast<double*>(work);
  cg_update1_kernel<<<g, 256, 0, s>>>(x, r, p, Ap, n, rs_old, pAp, part);
  cg_update2_kernel<<<g, 256, 0, s>>>(r, p, n, part, g.x, rs_old, rs_new);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

// ---------------------------------------------------------------------------
// Preconditioned conjugate gradients with a diagonal (Jacobi) preconditioner,
// FP32 or FP64 vectors, every scalar on the device (dense_krylov.py:191-198
// with z = M^-1 r in place of r in the direction update):
//   pcg_update1: alpha = rz / pAp (rounded to T); x += alpha p; r -= alpha Ap;
//                z = dinv .* r (no preconditioner: z is r); per-block
//                partials of r.z and r.r in FP64
//   pcg_update2: rz_new, rr_new = the partials summed in a fixed order by every
//                block; beta = rz_new / rz (rounded to T); p = z + beta p
// rr_new gives the recursive residual norm for the stopping test without a
// second pass over r.
// ---------------------------------------------------------------------------
constexpr int PCG_BLOCKS = 512;

template <typename T>
__global__ __launch_bounds__(256) void pcg_update1_kernel(
    T* __restrict__ x, T* __restrict__ r, T* __restrict__ z, const T* __restrict__ p,
    const T* __restrict__ Ap, const T* __restrict__ dinv, int n, const double* __restrict__ rz,
    const double* __restrict__ pAp, double* __restrict__ part) {
  const T alpha = (T)(rz[0] / pAp[0]);
  double a_rz = 0.0, a_rr = 0.0;
  for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) {
    x[j] = x[j] + alpha * p[j];
    const T rn = r[j] - alpha * Ap[j];
    r[j] = rn;
    T zn = rn;
    if (dinv) {
      zn = rn * dinv[j];
      z[j] = zn;
    }
    a_rz = __builtin_fma((double)rn, (double)zn, a_rz);
    a_rr = __builtin_fma((double)rn, (double)rn, a_rr);
  }
  __shared__ double sh[4];
  a_rz = block_sum_256(a_rz, sh);
  a_rr = block_sum_256(a_rr, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a_rz;
    part[PCG_BLOCKS + blockIdx.x] = a_rr;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void pcg_update2_kernel(const T* __restrict__ z,
                                                          T* __restrict__ p, int n,
                                                          const double* __restrict__ part, int np,
                                                          const double* __restrict__ rz_old,
                                                          double* __restrict__ rz_new,
                                                          double* __restrict__ rr_new) {
  __shared__ double sh[4];
  double srz = 0.0, srr = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) {
    srz += part[i];
    srr += part[PCG_BLOCKS + i];
  }
  srz = block_sum_256(srz, sh);
  srr = block_sum_256(srr, sh);
  const T beta = (T)(srz / rz_old[0]);
  for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256)
    p[j] = z[j] + beta * p[j];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    rz_new[0] = srz;
    if (rr_new) rr_new[0] = srr;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void diag_scale_kernel(const T* __restrict__ d,
                                                         const T* __restrict__ in,
                                                         T* __restrict__ out, int n) {
  for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) out[j] = in[j] * d[j];
}

size_t pcg_workspace_bytes() { return 2 * PCG_BLOCKS * sizeof(double); }

template <typename T>
int pcg_step_updates(hipStream_t s, T* x, T* r, T* z, T* p, const T* Ap, const T* dinv, int n,
                     const double* rz_old, const double* pAp, double* rz_new, double* rr_new,
                     void* work) {
  MPDOK_REQUIRE(x && r && p && Ap && rz_old && pAp && rz_new && work && n >= 1,
                "pcg_step_updates: bad arguments");
  MPDOK_REQUIRE(dinv == nullptr || z != nullptr, "pcg_step_updates: dinv needs a z vector");
  dim3 g = grid_1d(n);
  if (g.x > PCG_BLOCKS) g.x = PCG_BLOCKS;
  double* part = static_cast<double*>(work);
  pcg_update1_kernel<T><<<g, 256, 0, s>>>(x, r, z, p, Ap, dinv, n, rz_old, pAp, part);
  pcg_update2_kernel<T><<<g, 256, 0, s>>>(dinv ? z : r, p, n, part, g.x, rz_old, rz_new, rr_new);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

template <typename T>
int diag_scale(hipStream_t s, const T* d, const T* in, T* out, int n) {
  MPDOK_REQUIRE(d && in && out && n >= 0, "diag_scale: bad arguments");
  if (n == 0) return OK;
  diag_scale_kernel<T><<<grid_1d(n), 256, 0, s>>>(d, in, out, n);
  MPDOK_HIP(hipGetLastError());
  return OK;
}

}  // namespace mpdok

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
MPDOK_API size_t mpdok_arnoldi_workspace_bytes(int n, int max_k) {
  return mpdok::arnoldi_workspace_bytes(n, max_k);
}
MPDOK_API int mpdok_arnoldi_orth_f32(hipStream_t s, float* V, int64_t ldv, int k, float* w, int n,
                                     double* Hcol, int passes, double thresh, int stop_now,
                                     int* ctl, void* work, size_t work_bytes) {
  return mpdok::arnoldi_orth<float>(s, V, ldv, k, w, n, Hcol, passes, thresh, stop_now, ctl, work,
                                    work_bytes);
}
MPDOK_API int mpdok_arnoldi_orth_f64(hipStream_t s, double* V, int64_t ldv, int k, double* w, int n,
                                     double* Hcol, int passes, double thresh, int stop_now,
                                     int* ctl, void* work, size_t work_bytes) {
  return mpdok::arnoldi_orth<double>(s, V, ldv, k, w, n, Hcol, passes, thresh, stop_now, ctl, work,
                                     work_bytes);
}
MPDOK_API int mpdok_givens_step(hipStream_t s, double* H, int ldh, int k, double* R, double* cs,
                                double* sn, double* g, double beta, double tol, int* ctl) {
  return mpdok::givens_step(s, H, ldh, k, R, cs, sn, g, beta, tol, ctl);
}
MPDOK_API int mpdok_basis_update_f32(hipStream_t s, const float* V, int64_t ldv, int m,
                                     const double* y, float* x, int n, int overwrite) {
  return mpdok::basis_update<float, float>(s, V, ldv, m, y, x, n, overwrite);
}
MPDOK_API int mpdok_basis_update_f32_into_f64(hipStream_t s, const float* V, int64_t ldv, int m,
                                              const double* y, double* x, int n, int overwrite) {
  return mpdok::basis_update<float, double>(s, V, ldv, m, y, x, n, overwrite);
}
MPDOK_API int mpdok_basis_update_f64(hipStream_t s, const double* V, int64_t ldv, int m,
                                     const double* y, double* x, int n, int overwrite) {
  return mpdok::basis_update<double, double>(s, V, ldv, m, y, x, n, overwrite);
}
MPDOK_API int mpdok_scale_f64_to_f32(hipStream_t s, const double* in, float* out, int n,
                                     const double* nrm2) {
  return mpdok::scale_by_inv_norm<double, float>(s, in, out, n, nrm2);
}
MPDOK_API int mpdok_scale_f32(hipStream_t s, const float* in, float* out, int n,
                              const double* nrm2) {
  return mpdok::scale_by_inv_norm<float, float>(s, in, out, n, nrm2);
}
MPDOK_API int mpdok_scale_f64(hipStream_t s, const double* in, double* out, int n,
                              const double* nrm2) {
  return mpdok::scale_by_inv_norm<double, double>(s, in, out, n, nrm2);
}
MPDOK_API int mpdok_cg_step_updates(hipStream_t s, float* x, float* r, float* p, const float* Ap,
                                    int n, const double* rs_old, const double* pAp,
                                    double* rs_new, void* work) {
  return mpdok::cg_step_updates(s, x, r, p, Ap, n, rs_old, pAp, rs_new, work);
}
MPDOK_API size_t mpdok_pcg_workspace_bytes(void) { return mpdok::pcg_workspace_bytes(); }
MPDOK_API int mpdok_pcg_step_updates(hipStream_t s, int dtype, void* x, void* r, void* z, void* p,
                                     const void* Ap, const void* dinv, int n,
                                     const double* rz_old, const double* pAp, double* rz_new,
                                     double* rr_new, void* work) {
  if (dtype == mpdok::F32)
    return mpdok::pcg_step_updates<float>(s, (float*)x, (float*)r, (float*)z, (float*)p,
                                          (const float*)Ap, (const float*)dinv, n, rz_old, pAp,
                                          rz_new, rr_new, work);
  MPDOK_REQUIRE(dtype == mpdok::F64, "pcg_step_updates: dtype must be 0 (FP64) or 1 (FP32)");
  return mpdok::pcg_step_updates<double>(s, (double*)x, (double*)r, (double*)z, (double*)p,
                                         (const double*)Ap, (const double*)dinv, n, rz_old, pAp,
                                         rz_new, rr_new, work);
}
MPDOK_API int mpdok_diag_scale(hipStream_t s, int dtype, const void* d, const void* in, void* out,
                               int n) {
  if (dtype == mpdok::F32)
    return mpdok::diag_scale<float>(s, (const float*)d, (const float*)in, (float*)out, n);
  MPDOK_REQUIRE(dtype == mpdok::F64, "diag_scale: dtype must be 0 (FP64) or 1 (FP32)");
  return mpdok::diag_scale<double>(s, (const double*)d, (const double*)in, (double*)out, n);
}
