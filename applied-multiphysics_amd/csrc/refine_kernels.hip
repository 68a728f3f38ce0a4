// FP64 vector kernels of the device refinement: CSR residual, dot products and updates (gfx950).
//
// Iterative refinement (solveRefined, ldlt.mjs:253-275) and right-preconditioned restarted GMRES (gmres-real.mjs:31-104)
// keep the residual r = b - A x, the solution update and the Krylov basis in FP64 in HBM; the factor of working
// precision only enters through the preconditioner solve. Reductions are two-stage (per-block partials, then one block)
// so results are deterministic run to run.
#include "refine_kernels.h"

namespace amp {

namespace {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 512;

__device__ inline double blockSum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x < blockDim.x / 64) t = sh[threadIdx.x];
  if (wave == 0)
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  __syncthreads();
  return t;
}

}  // namespace

// r = b - A x (row order of the CSR pattern, one row per thread); b == nullptr gives r = A x
__global__ __launch_bounds__(256) void k_csr_residual(const int32_t* __restrict__ rowptr,
                                                      const int32_t* __restrict__ colidx,
                                                      const double* __restrict__ val, const double* __restrict__ x,
                                                      const double* __restrict__ b, double* __restrict__ r, int32_t n) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = b ? b[i] : 0.0;
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; p++) s -= val[p] * x[colidx[p]];
    r[i] = b ? s : -s;
  }
}

// R_k = B_k - A X_k for k < q <= Q right-hand sides (each n apart), one row per thread: the row's entries are read once
// for all of them, in the order of k_csr_residual (so each R_k equals its result bit for bit); B == nullptr gives
// R_k = A X_k
template <int Q>
__global__ __launch_bounds__(256) void k_csr_residual_multi(const int32_t* __restrict__ rowptr,
                                                            const int32_t* __restrict__ colidx,
                                                            const double* __restrict__ val,
                                                            const double* __restrict__ X,
                                                            const double* __restrict__ B, double* __restrict__ R,
                                                            int32_t n, int32_t q) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) s[k] = (B && k < q) ? B[int64_t(k) * n + i] : 0.0;
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; p++) {
      const double a = val[p];
      const int64_t c = colidx[p];
#pragma unroll
      for (int k = 0; k < Q; k++)
        if (k < q) s[k] -= a * X[int64_t(k) * n + c];
    }
#pragma unroll
    for (int k = 0; k < Q; k++)
      if (k < q) R[int64_t(k) * n + i] = B ? s[k] : -s[k];
  }
}

// partial[j * gridDim.x + block] = sum over the block's rows of V_j . w, for j < nv (V_j = V + j ldv)
__global__ __launch_bounds__(kRedThreads) void k_multi_dot(const double* __restrict__ V, int64_t ldv, int32_t nv,
                                                            const double* __restrict__ w, int32_t n,
                                                            double* __restrict__ partial) {
  __shared__ double sh[kRedThreads / 64];
  for (int32_t j = 0; j < nv; j++) {
    const double* vj = V + j * ldv;
    double s = 0;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += vj[i] * w[i];
    s = blockSum(s, sh);
    if (threadIdx.x == 0) partial[int64_t(j) * gridDim.x + blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(kRedThreads) void k_reduce_partials(const double* __restrict__ partial, int32_t nparts,
                                                                  int32_t nv, double* __restrict__ out) {
  __shared__ double sh[kRedThreads / 64];
  for (int32_t j = 0; j < nv; j++) {
    double s = 0;
    for (int32_t i = threadIdx.x; i < nparts; i += blockDim.x) s += partial[int64_t(j) * nparts + i];
    s = blockSum(s, sh);
    if (threadIdx.x == 0) out[j] = s;
  }
}

// w -= sum_j h_j V_j (h on the device)
__global__ void k_multi_axpy(const double* __restrict__ V, int64_t ldv, int32_t nv, const double* __restrict__ h,
                             double* __restrict__ w, int32_t n) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = w[i];
    for (int32_t j = 0; j < nv; j++) s -= h[j] * V[j * ldv + i];
    w[i] = s;
  }
}

// y = a x + c y (scalars from the host); with c == 0 the old y is not read (fresh storage may hold NaN patterns)
__global__ void k_axpby(double a, const double* __restrict__ x, double c, double* __restrict__ y, int32_t n) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    y[i] = c == 0.0 ? a * x[i] : a * x[i] + c * y[i];
}

// x += sum_j coef_j Z_j (coefficients from the host, at most kMaxRestart)
__global__ void k_combine(const double* __restrict__ Z, int64_t ldz, int32_t nz, const double* __restrict__ coef,
                          double* __restrict__ x, int32_t n) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = 0;
    for (int32_t j = 0; j < nz; j++) s += coef[j] * Z[j * ldz + i];
    x[i] += s;
  }
}

// ---- batches of independent systems (one right-hand side each, vectors of system j at j n) --------------------------

// r_j = b_j - A_j x_j (b == nullptr: r_j = A_j x_j); values of system j at j nnz, b_j at j bStride
__global__ __launch_bounds__(256) void k_batch_residual(const int32_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ colidx,
                                                        const double* __restrict__ val, int64_t nnz,
                                                        const double* __restrict__ x, const double* __restrict__ b,
                                                        int64_t bStride, double* __restrict__ r, int32_t n, int32_t nb) {
  const int32_t j = blockIdx.y;
  const double* vj = val + int64_t(j) * nnz;
  const double* xj = x + int64_t(j) * n;
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = b ? b[int64_t(j) * bStride + i] : 0.0;
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; p++) s -= vj[p] * xj[colidx[p]];
    r[int64_t(j) * n + i] = b ? s : -s;
  }
}

// partial[j gridDim.x + block] = block's share of u_j . w_j
__global__ __launch_bounds__(kRedThreads) void k_batch_dot(const double* __restrict__ u, const double* __restrict__ w,
                                                           int32_t n, double* __restrict__ partial) {
  __shared__ double sh[kRedThreads / 64];
  const int32_t j = blockIdx.y;
  const double* uj = u + int64_t(j) * n;
  const double* wj = w + int64_t(j) * n;
  double s = 0;
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += uj[i] * wj[i];
  s = blockSum(s, sh);
  if (threadIdx.x == 0) partial[int64_t(j) * gridDim.x + blockIdx.x] = s;
}

// y_j = a_j x_j + c_j y_j with per-system coefficients on the device; c_j == 0 does not read y_j
__global__ void k_batch_axpby(const double* __restrict__ a, const double* __restrict__ x,
                              const double* __restrict__ c, double* __restrict__ y, int32_t n) {
  const int32_t j = blockIdx.y;
  const double aj = a[j], cj = c ? c[j] : 1.0;
  const double* xj = x + int64_t(j) * n;
  double* yj = y + int64_t(j) * n;
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    yj[i] = cj == 0.0 ? aj * xj[i] : aj * xj[i] + cj * yj[i];
}

static unsigned gridFor(int32_t n) { return unsigned(std::min<int64_t>((int64_t(n) + 255) / 256, 4096)); }

void launchResidual(const int32_t* rowptr, const int32_t* colidx, const double* val, const double* x, const double* b,
                    double* r, int32_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_csr_residual, dim3(gridFor(n)), dim3(256), 0, st, rowptr, colidx, val, x, b, r, n);
}

void launchResidualMulti(const int32_t* rowptr, const int32_t* colidx, const double* val, const double* X,
                         const double* B, double* R, int32_t n, int32_t q, hipStream_t st) {
  constexpr int Q = 8;
  for (int32_t k0 = 0; k0 < q; k0 += Q) {
    const int64_t o = int64_t(k0) * n;
    hipLaunchKernelGGL(k_csr_residual_multi<Q>, dim3(gridFor(n)), dim3(256), 0, st, rowptr, colidx, val, X + o,
                       B ? B + o : nullptr, R + o, n, std::min(Q, q - k0));
  }
}

void launchMultiDot(const double* V, int64_t ldv, int32_t nv, const double* w, int32_t n, double* partial, double* out,
                    hipStream_t st) {
  const unsigned g = unsigned(std::min<int64_t>((int64_t(n) + kRedThreads - 1) / kRedThreads, kRedBlocks));
  hipLaunchKernelGGL(k_multi_dot, dim3(g), dim3(kRedThreads), 0, st, V, ldv, nv, w, n, partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kRedThreads), 0, st, partial, int32_t(g), nv, out);
}

void launchMultiAxpy(const double* V, int64_t ldv, int32_t nv, const double* h, double* w, int32_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_multi_axpy, dim3(gridFor(n)), dim3(256), 0, st, V, ldv, nv, h, w, n);
}

void launchAxpby(double a, const double* x, double c, double* y, int32_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_axpby, dim3(gridFor(n)), dim3(256), 0, st, a, x, c, y, n);
}

void launchCombine(const double* Z, int64_t ldz, int32_t nz, const double* coef, double* x, int32_t n,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_combine, dim3(gridFor(n)), dim3(256), 0, st, Z, ldz, nz, coef, x, n);
}

void launchBatchResidual(const int32_t* rowptr, const int32_t* colidx, const double* val, int64_t nnz, const double* x,
                         const double* b, int64_t bStride, double* r, int32_t n, int32_t nb, hipStream_t st) {
  const unsigned g = unsigned(std::min<int64_t>((int64_t(n) + 255) / 256, 1024));
  hipLaunchKernelGGL(k_batch_residual, dim3(g, nb), dim3(256), 0, st, rowptr, colidx, val, nnz, x, b, bStride, r, n, nb);
}

void launchBatchDot(const double* u, const double* w, int32_t n, int32_t nb, double* partial, double* out,
                    hipStream_t st) {
  const unsigned g = unsigned(std::min<int64_t>((int64_t(n) + kRedThreads - 1) / kRedThreads, kRedBlocks));
  hipLaunchKernelGGL(k_batch_dot, dim3(g, nb), dim3(kRedThreads), 0, st, u, w, n, partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kRedThreads), 0, st, partial, int32_t(g), nb, out);
}

void launchBatchAxpby(const double* a, const double* x, const double* c, double* y, int32_t n, int32_t nb,
                      hipStream_t st) {
  const unsigned g = unsigned(std::min<int64_t>((int64_t(n) + 255) / 256, 1024));
  hipLaunchKernelGGL(k_batch_axpby, dim3(g, nb), dim3(256), 0, st, a, x, c, y, n);
}

int64_t partialEntries(int32_t maxVectors) { return int64_t(kRedBlocks) * maxVectors; }

}  // namespace amp
