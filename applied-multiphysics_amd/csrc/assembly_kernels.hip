// Device kernels of the P1 assembly (assembly.cpp): parametric element coefficients and the gather of each restricted
// CSR entry's element contributions.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace amp {

// values[b nnz + z] = sum over z's contributions (element e, local entry ab), in element order, of k[b ne + e] Kgeo[e,
// ab], starting from +0 as the reference's zero-filled K (scalar.ts:110, :137-140). Products and sums round separately
// (contraction into fused multiply-adds is switched off), in the order of the reference's scatter. One thread per entry, every coefficient set in turn.
__global__ __launch_bounds__(256) void k_assemble_gather(const int32_t* __restrict__ zptr,
                                                         const int32_t* __restrict__ contrib,
                                                         const double* __restrict__ Kgeo, const double* __restrict__ k,
                                                         int32_t ne, double* __restrict__ values, int64_t nnz,
                                                         int32_t nb, int32_t kk) {
#pragma clang fp contract(off)
  for (int64_t z = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; z < nnz; z += int64_t(gridDim.x) * blockDim.x) {
    const int32_t p0 = zptr[z], p1 = zptr[z + 1];
    for (int32_t b = 0; b < nb; b++) {
      const double* kb = k + int64_t(b) * ne;
      double v = 0.0;
      for (int32_t p = p0; p < p1; p++) {
        const int32_t e = contrib[2 * p], ab = contrib[2 * p + 1];
        v = v + kb[e] * Kgeo[int64_t(e) * kk + ab];
      }
      values[int64_t(b) * nnz + z] = v;
    }
  }
}

// k[b ne + e] = exp(sum_m theta[b nm + m] Phi[e nm + m]), the sum in mode order
__global__ __launch_bounds__(256) void k_param_coefficients(const double* __restrict__ Phi, int32_t nm, int32_t ne,
                                                            const double* __restrict__ theta, double* __restrict__ k,
                                                            int32_t nb) {
#pragma clang fp contract(off)
  const int64_t total = int64_t(nb) * ne;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
    const int32_t b = int32_t(i / ne), e = int32_t(i % ne);
    double s = 0.0;
    for (int32_t m = 0; m < nm; m++) s = s + theta[int64_t(b) * nm + m] * Phi[int64_t(e) * nm + m];
    k[i] = exp(s);
  }
}

void launchAssembleGather(const int32_t* zptr, const int32_t* contrib, const double* Kgeo, const double* k, int32_t ne,
                          double* values, int64_t nnz, int32_t nb, int32_t kk, hipStream_t st) {
  const unsigned g = unsigned((nnz + 255) / 256 < 65536 ? (nnz + 255) / 256 : 65536);
  hipLaunchKernelGGL(k_assemble_gather, dim3(g), dim3(256), 0, st, zptr, contrib, Kgeo, k, ne, values, nnz, nb, kk);
}

void launchParamCoefficients(const double* Phi, int32_t nm, int32_t ne, const double* theta, double* k, int32_t nb,
                             hipStream_t st) {
  const int64_t total = int64_t(nb) * ne;
  const unsigned g = unsigned((total + 255) / 256 < 65536 ? (total + 255) / 256 : 65536);
  hipLaunchKernelGGL(k_param_coefficients, dim3(g), dim3(256), 0, st, Phi, nm, ne, theta, k, nb);
}

}  // namespace amp
