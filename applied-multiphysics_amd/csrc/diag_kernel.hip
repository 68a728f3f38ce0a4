// Diagonal blocks of tiled fronts (gfx950): factor a 64 x 64 block and build the inverse of its lower factor.
#include "pivot.h"

namespace amp {

namespace {

// Broadcast lane `l`'s value to the whole wave through a scalar register (l is a compile-time constant here).
__device__ inline double laneValue(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}
__device__ inline float laneValue(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

}  // namespace

// One wave per block, no barriers. Lane i keeps row i of the block in registers (right-looking elimination: at step k
// every lane reads the pivot from lane k, lane k records it, every lane scales its entry of column k, and the column's
// entries are broadcast lane by lane through scalar registers for the rank-1 update of each row). The inverse
// W = L^-1 is built in the same steps by Gauss-Jordan elimination with lane c keeping column c of W:
// W[k][c] /= l_kk, then W[i][c] -= l_ik W[k][c] for i > k, with the same broadcast column. All loops over the 64
// columns are unrolled, so every operand is a register.
template <typename T>
__global__ __launch_bounds__(64) void k_diag_block(const DiagTask* __restrict__ tasks, FactorArgs<T> a) {
  constexpr int E = kTileEdge, P = kTileEdge + 1;
  __shared__ T Ws[E * P];
  const DiagTask t = tasks[blockIdx.x];
  const int lane = threadIdx.x, nb = t.nb;
  T* F = a.Lx + t.f_off;
  T r[E], w[E];
#pragma unroll
  for (int j = 0; j < E; j++) {
    r[j] = (j < nb && lane < nb && lane >= j) ? F[lane + int64_t(j) * t.ld] : T(0);
    w[j] = j == lane ? T(1) : T(0);
  }
#pragma unroll
  for (int k = 0; k < E; k++) {
    if (k < nb) {
      // every lane evaluates the pivot rule on the same value; only lane k records it (D, flags, failure)
      T inv, mult;
      const T dk = laneValue(r[k], k);
      const T piv = pivotRule<T>(dk, t.col0 + k, a, &inv, &mult, lane == k);
      if (lane == k) r[k] = piv;
      if (lane > k) r[k] *= inv;
      const T lik = r[k];
      const T lkk = laneValue(r[k], k);
      const T wk = a.ldlt ? w[k] : w[k] / lkk;
      w[k] = wk;
#pragma unroll
      for (int j = k + 1; j < E; j++) {
        const T ljk = laneValue(lik, j);
        if (lane >= j) r[j] -= lik * (ljk * mult);
        w[j] -= ljk * wk;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < E; j++)
    if (j < nb && lane < nb && lane >= j) F[lane + int64_t(j) * t.ld] = r[j];
  // W column `lane` -> LDS (element (i, lane)), then stores along columns of W
#pragma unroll
  for (int i = 0; i < E; i++) Ws[i * P + lane] = (i < nb && lane < nb) ? w[i] : T(0);
  __syncthreads();
  T* Wt = a.W + t.w_off;
  for (int j = 0; j < E; j++) Wt[lane + j * E] = Ws[lane * P + j];
}

template <typename T>
void launchDiagBlocks(const DiagTask* tasks, unsigned count, const FactorArgs<T>& a, hipStream_t st) {
  hipLaunchKernelGGL(k_diag_block<T>, dim3(count), dim3(64), 0, st, tasks, a);
}

template void launchDiagBlocks<double>(const DiagTask*, unsigned, const FactorArgs<double>&, hipStream_t);
template void launchDiagBlocks<float>(const DiagTask*, unsigned, const FactorArgs<float>&, hipStream_t);

}  // namespace amp
