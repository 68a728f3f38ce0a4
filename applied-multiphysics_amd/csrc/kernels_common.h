// Device helpers shared by the factor and solve kernels (gfx950, wave64, MFMA).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "plan.h"

namespace amp {

// Kernels of the factor's panel chain (main stream) raise their waves' issue priority: the side stream's bulk updates
// keep every CU occupied with long-running workgroups, and a chain wave placed beside them would otherwise take a
// round-robin share of its SIMD's issue slots while every later super-panel waits on it.
__device__ inline void chainPriority() { __builtin_amdgcn_s_setprio(3); }

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
struct Vec4;
template <>
struct Vec4<double> {
  using type = v4d;
};
template <>
struct Vec4<float> {
  using type = v4f;
};

// 16x16x4 matrix-core step D = A B + C. Lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]. The result rows
// differ by type: FP64 component r of lane l is D[4 r + l / 16][l % 16]; FP32 component r is D[4 (l / 16) + r][l % 16].
__device__ inline v4d mfma16x16x4(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline v4f mfma16x16x4(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <typename T>
__device__ inline int mfmaRow(int lane, int r);
template <>
__device__ inline int mfmaRow<double>(int lane, int r) {
  return 4 * r + (lane >> 4);
}
template <>
__device__ inline int mfmaRow<float>(int lane, int r) {
  return 4 * (lane >> 4) + r;
}

// Broadcast lane `l`'s value to the whole wave through a scalar register (l must be the same in every lane).
__device__ inline double laneValue(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}
__device__ inline float laneValue(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Lane l's value within each segment of G lanes (lanes G g .. G g + G - 1 of the wave read lane G g + l), through
// ds_bpermute, which reads nothing of LDS memory. l may differ between segments; the source lane must be active (an
// inactive one gives 0).
template <int G>
__device__ inline float segValue(float v, int l) {
  static_assert(G > 0 && G < 64 && (G & (G - 1)) == 0, "segments are a power of two below the wave");
  const int src = ((int(threadIdx.x & 63) & ~(G - 1)) + l) << 2;
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(v)));
}
template <int G>
__device__ inline double segValue(double v, int l) {
  static_assert(G > 0 && G < 64 && (G & (G - 1)) == 0, "segments are a power of two below the wave");
  const int src = ((int(threadIdx.x & 63) & ~(G - 1)) + l) << 2;
  const int lo = __builtin_amdgcn_ds_bpermute(src, __double2loint(v));
  const int hi = __builtin_amdgcn_ds_bpermute(src, __double2hiint(v));
  return __hiloint2double(hi, lo);
}

// Whether any lane of the wave has `p` (a wave-uniform answer, for loop exits of segment-varying bounds).
__device__ inline bool waveAny(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

// Kernel-side view of the factor storage and of the plan's per-supernode arrays.
template <typename T>
struct FactorArgs {
  T* Lx;
  T* U;
  T* W;
  T* D;
  uint8_t* pert;
  int32_t* failCol;
  const int32_t* superStart;
  const int32_t* Sp;
  const int64_t* Lxp;
  const int64_t* Uoff;
  const int64_t* Woff;
  const int32_t* childHead;
  const int32_t* childNext;
  const int32_t* rel;
  const int32_t* Si;
  const int32_t* lowerP;
  const int32_t* lowerSrc;
  const int32_t* amap;
  int32_t ldlt;
  double tau;
  int32_t bulkBf16;  // FP32 factor: bulk trailing updates as three bf16 products (k_gemm_bf16x3)
  double growthLimit;  // FP32 factor: diagonal blocks whose largest elimination multiplier exceeds it are redone in
                       // FP64 arithmetic (0: never)
  int32_t* fallbacks;  // count of such blocks in the current factorization
  // bf16 bulk products: the split panel of the current bulk launch, four arrays of splitPitch entries (A heads, A
  // tails, B heads, B tails; for Cholesky the B arrays are the A arrays)
  __bf16* split;
  int64_t splitPitch;
  // fused panel steps: one counter per kPanel launch (workgroups whose inputs are read), and the flag a wait on one
  // sets when it gives up
  int32_t* panelCtr;
  int32_t* panelErr;
};

}  // namespace amp
