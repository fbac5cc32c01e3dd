#pragma once
#include "common.h"

namespace mpdok {

constexpr int MAX_RESTART = 256;  // largest Krylov dimension per restart cycle

size_t arnoldi_workspace_bytes(int n, int max_k);
template <typename T>
int arnoldi_orth(hipStream_t s, T* V, int64_t ldv, int k, T* w, int n, double* Hcol, int passes,
                 double thresh, int stop_now, int* ctl, void* work, size_t work_bytes);
int givens_step(hipStream_t s, double* H, int ldh, int k, double* R, double* cs, double* sn,
                double* g, double beta, double tol, int* ctl);
template <typename T, typename TX>
int basis_update(hipStream_t s, const T* V, int64_t ldv, int m, const double* y, TX* x, int n,
                 int overwrite);
template <typename TI, typename TO>
int scale_by_inv_norm(hipStream_t s, const TI* in, TO* out, int n, const double* nrm2);
int cg_step_updates(hipStream_t s, float* x, float* r, float* p, const float* Ap, int n,
                    const double* rs_old, const double* pAp, double* rs_new, void* work);

}  // namespace mpdok
