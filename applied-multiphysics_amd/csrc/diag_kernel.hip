// Diagonal blocks of tiled fronts (gfx950): factor a 64 x 64 block and build the inverse of its lower factor.
#include "pivot.h"

namespace amp {

// Eight waves per block. Wave w owns the columns j = 8 jj + w of the block (lane i keeps A[i][j] in rA[jj]) and the
// rows i = 8 ii + w of the inverse W = L^-1 (lane c keeps W[i][c] in rW[ii]), so at every elimination step the
// columns and rows still active are spread evenly over the waves and each wave has at most eight to update. The pivot
// step of column k (by its owner wave): read the pivot from lane k (every lane evaluates the pivot rule on the
// broadcast value, lane k records it), scale column k, finish row k of W (Gauss-Jordan: W[k][:] *= 1 / l_kk) and
// publish both through LDS (double-buffered by step parity, one barrier per step); the column is published twice, as
// l and as l d, both zero at and above the pivot row. The update of step k: every wave applies the rank-1 update to
// its active columns (A[i][j] -= (l_ik d) l_jk, with l_jk read as an LDS broadcast of entry j of the published column;
// rows at or above the pivot receive zero, and rows above the diagonal of column j are never read or stored) and rows
// of W (W[i][:] -= l_jk W[k][:]). Lookahead: the owner of column k + 1 updates that column first and runs its pivot
// step while the other waves finish the update of step k. The step loop is unrolled, so every register index is a
// constant.
constexpr int kDiagWaves = 8;

template <typename T>
__global__ __launch_bounds__(64 * kDiagWaves) void k_diag_block(const DiagTask* __restrict__ tasks, FactorArgs<T> a) {
  constexpr int E = kTileEdge, NW = kDiagWaves, J = kTileEdge / NW;
  __shared__ T colL[2][E], colD[2][E], rowW[2][E];
  const DiagTask t = tasks[blockIdx.x];
  const int lane = threadIdx.x & 63, nb = t.nb;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  T* F = a.Lx + t.f_off;
  T rA[J], rW[J];
#pragma unroll
  for (int jj = 0; jj < J; jj++) {
    const int j = NW * jj + w;
    rA[jj] = (j < nb && lane < nb && lane >= j) ? F[lane + int64_t(j) * t.ld] : T(0);
    rW[jj] = j == lane ? T(1) : T(0);
  }
  auto pivotStep = [&](const int k) {
    const int kk = k / NW, p = k & 1;
    T inv, mult;
    const T dk = laneValue(rA[kk], k);
    const T piv = pivotRule<T>(dk, t.col0 + k, a, &inv, &mult, lane == k);
    const T l = lane > k ? rA[kk] * inv : T(0);
    if (lane == k) rA[kk] = piv;
    if (lane > k) rA[kk] = l;
    const T wk = a.ldlt ? rW[kk] : rW[kk] * inv;
    rW[kk] = wk;
    colL[p][lane] = l;
    colD[p][lane] = l * mult;
    rowW[p][lane] = wk;
  };
  if (nb > 0 && w == 0) pivotStep(0);
  __syncthreads();
#pragma clang loop unroll(full)
  for (int k = 0; k < E; k++) {
    if (k >= nb) continue;
    const int p = k & 1, k1 = k + 1;
    const T cd = colD[p][lane], rw = rowW[p][lane];
    const bool ahead = k1 < nb && w == (k1 % NW);
    if (k1 < E && ahead) {
      const int jj = k1 / NW;
      const T ljk = colL[p][k1];
      rA[jj] -= cd * ljk;
      rW[jj] -= ljk * rw;
      pivotStep(k1);
    }
#pragma unroll
    for (int jj = 0; jj < J; jj++) {
      const int j = NW * jj + w;
      if (NW * jj + NW - 1 > k && j > k && !(ahead && j == k1)) {
        const T ljk = colL[p][j];
        rA[jj] -= cd * ljk;
        rW[jj] -= ljk * rw;
      }
    }
    __syncthreads();
  }
  // W (column-major, edge 64) straight from the registers: each store writes row j of W across the 64 columns, the
  // L2 merges the rows into whole lines; no LDS transpose, so the kernel keeps a small LDS footprint and fits on a CU
  // beside the trailing-update workgroups it runs with
  T* Wt = a.W + t.w_off;
#pragma unroll
  for (int jj = 0; jj < J; jj++) {
    const int j = NW * jj + w;
    if (j < nb && lane < nb && lane >= j) F[lane + int64_t(j) * t.ld] = rA[jj];
    Wt[j + E * lane] = (j < nb && lane < nb) ? rW[jj] : T(0);
  }
}

template <typename T>
void launchDiagBlocks(const DiagTask* tasks, unsigned count, const FactorArgs<T>& a, hipStream_t st) {
  hipLaunchKernelGGL(k_diag_block<T>, dim3(count), dim3(64 * kDiagWaves), 0, st, tasks, a);
}

template void launchDiagBlocks<double>(const DiagTask*, unsigned, const FactorArgs<double>&, hipStream_t);
template void launchDiagBlocks<float>(const DiagTask*, unsigned, const FactorArgs<float>&, hipStream_t);

}  // namespace amp
