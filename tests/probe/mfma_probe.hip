// Test-only probe of the matrix-core lane layouts used by the library: one wave runs one 16x16x4 MFMA (FP64 and FP32)
// on given A (16x4) and B (4x16), each lane supplying A[l % 16][l / 16] and B[l / 16][l % 16], and writes every lane's
// four result components as they come out (raw[4 l + r]).
#include <hip/hip_runtime.h>

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__global__ void k_probe(const double* A, const double* B, double* raw64, float* raw32) {
  const int l = threadIdx.x;
  const double a = A[(l % 16) * 4 + l / 16], b = B[(l / 16) * 16 + l % 16];
  v4d d = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, v4d{0, 0, 0, 0}, 0, 0, 0);
  v4f f = __builtin_amdgcn_mfma_f32_16x16x4f32(float(a), float(b), v4f{0, 0, 0, 0}, 0, 0, 0);
  for (int r = 0; r < 4; r++) {
    raw64[4 * l + r] = d[r];
    raw32[4 * l + r] = f[r];
  }
}

extern "C" int probe_mfma(const double* A, const double* B, double* raw64, float* raw32) {
  double *dA, *dB, *d64;
  float* d32;
  if (hipMalloc(&dA, 64 * 8) || hipMalloc(&dB, 64 * 8) || hipMalloc(&d64, 256 * 8) || hipMalloc(&d32, 256 * 4)) return -1;
  (void)hipMemcpy(dA, A, 64 * 8, hipMemcpyHostToDevice);
  (void)hipMemcpy(dB, B, 64 * 8, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k_probe, dim3(1), dim3(64), 0, 0, dA, dB, d64, d32);
  hipError_t e = hipDeviceSynchronize();
  (void)hipMemcpy(raw64, d64, 256 * 8, hipMemcpyDeviceToHost);
  (void)hipMemcpy(raw32, d32, 256 * 4, hipMemcpyDeviceToHost);
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(d64);
  (void)hipFree(d32);
  return e == hipSuccess ? 0 : -2;
}
