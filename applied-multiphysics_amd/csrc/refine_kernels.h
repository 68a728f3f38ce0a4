// FP64 vector kernels used by the device refinement drivers (iterative refinement and GMRES-IR).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace amp {

constexpr int kMaxRestart = 64;

void launchResidual(const int32_t* rowptr, const int32_t* colidx, const double* val, const double* x, const double* b,
                    double* r, int32_t n, hipStream_t st);
// R_k = B_k - A X_k for q right-hand sides n apart (B == nullptr: R_k = A X_k), the matrix read once per eight
void launchResidualMulti(const int32_t* rowptr, const int32_t* colidx, const double* val, const double* X,
                         const double* B, double* R, int32_t n, int32_t q, hipStream_t st);
// out[j] = V_j . w for j < nv (deterministic two-stage reduction through `partial`)
void launchMultiDot(const double* V, int64_t ldv, int32_t nv, const double* w, int32_t n, double* partial, double* out,
                    hipStream_t st);
// w -= sum_j h_j V_j, h on the device
void launchMultiAxpy(const double* V, int64_t ldv, int32_t nv, const double* h, double* w, int32_t n, hipStream_t st);
// y = a x + c y
void launchAxpby(double a, const double* x, double c, double* y, int32_t n, hipStream_t st);
// x += sum_j coef_j Z_j, coef on the device
void launchCombine(const double* Z, int64_t ldz, int32_t nz, const double* coef, double* x, int32_t n, hipStream_t st);
// batches of nb independent systems (vectors of system j at j n, values at j nnz, b_j at j bStride: 0 shares one)
void launchBatchResidual(const int32_t* rowptr, const int32_t* colidx, const double* val, int64_t nnz, const double* x,
                         const double* b, int64_t bStride, double* r, int32_t n, int32_t nb, hipStream_t st);
// out[j] = u_j . w_j (deterministic two-stage reduction through `partial`, at least partialEntries(nb) entries)
void launchBatchDot(const double* u, const double* w, int32_t n, int32_t nb, double* partial, double* out,
                    hipStream_t st);
// y_j = a_j x_j + c_j y_j (a, c on the device; c == nullptr: c_j = 1)
void launchBatchAxpby(const double* a, const double* x, const double* c, double* y, int32_t n, int32_t nb,
                      hipStream_t st);
int64_t partialEntries(int32_t maxVectors);

}  // namespace amp
