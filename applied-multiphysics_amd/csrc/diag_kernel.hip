// Diagonal blocks of tiled fronts (gfx950): factor a 64 x 64 block and build the inverse of its lower factor.
#include "pivot.h"

namespace amp {

// One wave per block. Lane i keeps row i of the block in registers (right-looking elimination: at step k lane k takes
// the pivot, every lane scales its entry of column k, the column is broadcast through LDS and each lane updates its
// row). The inverse W = L^-1 is built in the same steps by Gauss-Jordan elimination with lane c keeping column c of W:
// W[k][c] /= l_kk, then W[i][c] -= l_ik W[k][c] for i > k, reading l_ik from the same broadcast column. All loops
// over the 64 columns are unrolled, so every operand is a register or a broadcast LDS read.
template <typename T>
__global__ __launch_bounds__(64) void k_diag_block(const DiagTask* __restrict__ tasks, FactorArgs<T> a) {
  constexpr int E = kTileEdge, P = kTileEdge + 1;
  __shared__ T colL[E], colW[E];
  __shared__ T Ws[E * P];
  __shared__ T s_inv, s_mult;
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
      if (lane == k) {
        T inv, mult;
        r[k] = pivotValue<T>(r[k], t.col0 + k, a, &inv, &mult);
        s_inv = inv;
        s_mult = mult;
      }
      __syncthreads();
      const T inv = s_inv, mult = s_mult;
      if (lane > k) r[k] *= inv;
      colL[lane] = r[k];
      colW[lane] = r[k] * mult;
      __syncthreads();
      const T lik = r[k];
#pragma unroll
      for (int j = k + 1; j < E; j++)
        if (lane >= j) r[j] -= lik * colW[j];
      const T wk = a.ldlt ? w[k] : w[k] / colL[k];
      w[k] = wk;
#pragma unroll
      for (int i = k + 1; i < E; i++) w[i] -= colL[i] * wk;
      __syncthreads();
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
