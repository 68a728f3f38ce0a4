// Pivot rule shared by the factor kernels (device).
#pragma once
#include "kernels_common.h"

namespace amp {

// Pivot of column k (value d already updated by every earlier column). Cholesky: d -> sqrt(d), failure unless d > 0.
// LDL^T: |d| < tau is replaced by +-tau (counted), a non-finite small pivot fails; the column keeps a unit diagonal.
// Writes the scaling factor to *inv and the weight multiplier (1 or d) to *mult. Failures only record the column:
// the factorization runs to the end with the IEEE result and the host reports the first failing column.
template <typename T>
__device__ inline T pivotValue(T d, int32_t col, const FactorArgs<T>& a, T* inv, T* mult) {
  if (!a.ldlt) {
    if (!(d > T(0))) atomicMin(a.failCol, col);
    const T sq = sqrt(d);
    *inv = T(1) / sq;
    *mult = T(1);
    return sq;
  }
  const T tau = T(a.tau);
  uint8_t perturbed = 0;
  if (!(fabs(d) >= tau)) {
    if (!isfinite(d))
      atomicMin(a.failCol, col);
    else {
      d = d < T(0) ? -tau : tau;
      perturbed = 1;
    }
  }
  a.pert[col] = perturbed;
  a.D[col] = d;
  *inv = T(1) / d;
  *mult = d;
  return T(1);
}

// 1 / sqrt(d) and 1 / d from the hardware estimates plus Newton steps (y <- y (3/2 - d y^2 / 2), y <- y (2 - d y)):
// within an ulp or two of the correctly rounded values at a fraction of the latency of the IEEE square root and
// division, which sit on the serial chain of every elimination step of a diagonal block. Non-positive d gives NaN,
// as sqrt does.
__device__ inline double invSqrt(double d) {
  double y = __builtin_amdgcn_rsq(d);
  const double h = 0.5 * d;
  y = y * (1.5 - h * y * y);
  return y * (1.5 - h * y * y);
}
// FP32: the hardware estimates are accurate to about an ulp already (no Newton step on the serial chain).
__device__ inline float invSqrt(float d) { return __builtin_amdgcn_rsqf(d); }
__device__ inline double recip(double d) {
  double y = __builtin_amdgcn_rcp(d);
  y = y * (2.0 - d * y);
  return y * (2.0 - d * y);
}
__device__ inline float recip(float d) { return __builtin_amdgcn_rcpf(d); }

// The rule of pivotValue evaluated by every lane of a wave on a broadcast value (scaling from invSqrt / recip); only
// the lane with `record` writes D, the perturbation flag and the failure column (so the stores happen once).
template <typename T>
__device__ inline T pivotRule(T d, int32_t col, const FactorArgs<T>& a, T* inv, T* mult, bool record) {
  if (!a.ldlt) {
    if (record && !(d > T(0))) atomicMin(a.failCol, col);
    const T r = invSqrt(d);
    *inv = r;
    *mult = T(1);
    return d * r;
  }
  const T tau = T(a.tau);
  uint8_t perturbed = 0;
  if (!(fabs(d) >= tau)) {
    if (!isfinite(d)) {
      if (record) atomicMin(a.failCol, col);
    } else {
      d = d < T(0) ? -tau : tau;
      perturbed = 1;
    }
  }
  if (record) {
    a.pert[col] = perturbed;
    a.D[col] = d;
  }
  *inv = recip(d);
  *mult = d;
  return T(1);
}

// Branch-free form of pivotRule for a kernel built for one factorization kind (LDLT). Every lane evaluates the rule on
// the broadcast pivot d; the lane with `record` keeps the column's results in its own registers (rec_d: D, rec_p:
// perturbation flag, rec_fail: the column if the pivot fails, else unchanged) for a single store at the end of the
// kernel, so the elimination step has no branch and no memory access. Cholesky records no failure here: a pivot
// d <= 0 (or NaN) gives the factored diagonal d * d^-1/2 = NaN, which the store checks once per column.
template <typename T, bool LDLT>
__device__ inline T pivotRuleNB(T d, int32_t col, T tau, T* inv, T* mult, bool record, T& rec_d, uint8_t& rec_p,
                                int32_t& rec_fail) {
  if constexpr (!LDLT) {
    const T r = invSqrt(d);
    *inv = r;
    *mult = T(1);
    return d * r;
  } else {
    const bool small = !(fabs(d) >= tau);
    const bool bad = small && !isfinite(d);
    const bool pert = small && !bad;
    d = pert ? (d < T(0) ? -tau : tau) : d;
    rec_fail = (record && bad) ? min(rec_fail, col) : rec_fail;
    rec_d = record ? d : rec_d;
    rec_p = record ? uint8_t(pert) : rec_p;
    *inv = recip(d);
    *mult = d;
    return T(1);
  }
}

template <typename T>
void launchDiagBlocks(const DiagTask* tasks, unsigned count, const FactorArgs<T>& a, hipStream_t st);

}  // namespace amp
