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

// bf16 16x16x32: lane l supplies A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16], j < 8 (A 16 x 32, B 32 x 16,
// small integers, exact in bf16); raw[4 l + r] as it comes out
typedef __bf16 bf16x8 __attribute__((__vector_size__(16)));
__global__ void k_probe_bf16(const float* A, const float* B, float* raw) {
  const int l = threadIdx.x;
  bf16x8 a, b;
  for (int j = 0; j < 8; j++) {
    a[j] = (__bf16)A[(l % 16) * 32 + 8 * (l / 16) + j];
    b[j] = (__bf16)B[(8 * (l / 16) + j) * 16 + l % 16];
  }
  v4f d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, v4f{0, 0, 0, 0}, 0, 0, 0);
  for (int r = 0; r < 4; r++) raw[4 * l + r] = d[r];
}

extern "C" int probe_mfma_bf16(const float* A, const float* B, float* raw) {
  float *dA, *dB, *dr;
  if (hipMalloc(&dA, 512 * 4) || hipMalloc(&dB, 512 * 4) || hipMalloc(&dr, 256 * 4)) return -1;
  (void)hipMemcpy(dA, A, 512 * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(dB, B, 512 * 4, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k_probe_bf16, dim3(1), dim3(64), 0, 0, dA, dB, dr);
  hipError_t e = hipDeviceSynchronize();
  (void)hipMemcpy(raw, dr, 256 * 4, hipMemcpyDeviceToHost);
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dr);
  return e == hipSuccess ? 0 : -2;
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
