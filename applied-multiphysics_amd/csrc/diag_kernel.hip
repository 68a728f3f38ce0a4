// Diagonal blocks of tiled fronts (gfx950): factor a 64 x 64 block and build the inverse of its lower factor.
#include <type_traits>

#include "diag_core.h"

namespace amp {

// Eight waves per block; the elimination runs in steps of kDiagStep columns (a column group), one barrier per step.
// Wave w owns the groups G = NW jj + w (lane i keeps A[i][j] of their columns, and row j of the inverse W = L^-1), so
// at every step the columns and rows still active are spread evenly over the waves. The pivot step of group G (by its
// owner wave, no barrier inside): pivot each column (read from its pivot lane; every lane evaluates the pivot rule on
// the broadcast value, the pivot lane records it), scale it, finish its row of W (Gauss-Jordan: W[k][:] *= 1 / l_kk),
// apply its rank-1 update to the group's later columns and W rows; the group's columns are published through LDS as
// l and as l d (zero at and above their pivot rows), with their W rows (double-buffered by step parity). The update of
// step G: every wave applies the rank-B update to its active columns (A[i][j] -= (l_ik d_k) l_jk, l_jk read as LDS
// broadcasts; rows above the diagonal of column j are never read or stored) and W rows (W[j][:] -= l_jk W[k][:]).
// Lookahead: the owner of group G + 1 updates its columns first and runs its pivot step while the other waves finish
// step G. The step loop runs at run time; within a step every register index is a constant.
// GROWTH (FP32 factor with a growth limit, a.growthLimit > 0): the block is eliminated in FP32; when its largest
// elimination multiplier exceeds the limit, the block (still unmodified in memory) is eliminated again in FP64
// arithmetic, with FP64 pivots, and stored rounded to FP32; a.fallbacks counts such blocks.
template <typename T, bool LDLT, bool GROWTH>
__global__ __launch_bounds__(64 * kDiagWaves) void k_diag_block(const DiagTask* __restrict__ tasks, FactorArgs<T> a) {
  constexpr int J = kDiagSlots;
  __shared__ __attribute__((aligned(32))) double ldsRaw[3 * 2 * kDiagStep * kTileEdge];
  __shared__ double gred[kDiagWaves];
  chainPriority();
  const DiagTask t = tasks[blockIdx.x];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const T* F = a.Lx + t.f_off;
  T recD = T(0), gmax = T(0);
  uint8_t recP = 0;
  int32_t recFail = INT32_MAX;
  T rA[J], rW[J];
  diagCore<T, LDLT>(t, F, T(a.tau), w, lane, rA, rW, reinterpret_cast<T*>(ldsRaw), recD, recP, recFail, gmax);
  if constexpr (GROWTH && std::is_same<T, float>::value) {
    {
      float g = gmax;
      for (int o = 32; o > 0; o >>= 1) g = fmaxf(g, __shfl_xor(g, o, 64));
      if (lane == 0) gred[w] = double(g);
      __syncthreads();
      double gb = 0;
#pragma unroll
      for (int i = 0; i < kDiagWaves; i++) gb = fmax(gb, gred[i]);
      // NaN multipliers (a failed pivot) keep the FP32 result, whose failure is recorded the same way
      if (gb > a.growthLimit) {
        double dA[J], dW[J], dD = 0, dg = 0;
        uint8_t dP = 0;
        int32_t dFail = INT32_MAX;
        diagCore<double, LDLT>(t, F, a.tau, w, lane, dA, dW, ldsRaw, dD, dP, dFail, dg);
        if (threadIdx.x == 0) atomicAdd(a.fallbacks, 1);
        diagStore<double, LDLT>(t, a, w, lane, dA, dW, dD, dP, dFail);
        return;
      }
    }
  }
  diagStore<T, LDLT>(t, a, w, lane, rA, rW, recD, recP, recFail);
}

template <typename T>
void launchDiagBlocks(const DiagTask* tasks, unsigned count, const FactorArgs<T>& a, hipStream_t st) {
  // the instance with the growth check (and its FP64 path) only when an FP32 factor asks for it: it holds more
  // registers than the plain FP32 elimination
  const bool growth = std::is_same<T, float>::value && a.growthLimit > 0;
  if (a.ldlt) {
    if (growth)
      hipLaunchKernelGGL((k_diag_block<T, true, true>), dim3(count), dim3(64 * kDiagWaves), 0, st, tasks, a);
    else
      hipLaunchKernelGGL((k_diag_block<T, true, false>), dim3(count), dim3(64 * kDiagWaves), 0, st, tasks, a);
  } else if (growth) {
    hipLaunchKernelGGL((k_diag_block<T, false, true>), dim3(count), dim3(64 * kDiagWaves), 0, st, tasks, a);
  } else {
    hipLaunchKernelGGL((k_diag_block<T, false, false>), dim3(count), dim3(64 * kDiagWaves), 0, st, tasks, a);
  }
}

template void launchDiagBlocks<double>(const DiagTask*, unsigned, const FactorArgs<double>&, hipStream_t);
template void launchDiagBlocks<float>(const DiagTask*, unsigned, const FactorArgs<float>&, hipStream_t);

}  // namespace amp
