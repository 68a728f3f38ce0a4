// Diagonal blocks of tiled fronts (gfx950): factor a 64 x 64 block and build the inverse of its lower factor.
#include "pivot.h"

namespace amp {

// Eight waves per block; the elimination runs in steps of two columns (a column pair), one barrier per step. Wave w
// owns the pairs P = 8 jj + w, i.e. columns 2 P and 2 P + 1 (lane i keeps A[i][j] in rA[2 jj + c] for j = 2 P + c),
// and the rows i = 2 P + c of the inverse W = L^-1 (lane c' keeps W[i][c'] in rW[2 jj + c]), so at every step the
// columns and rows still active are spread evenly over the waves. The pivot step of pair P (by its owner wave, no
// barrier inside): pivot column 2 P (read from lane 2 P; every lane evaluates the pivot rule on the broadcast value,
// the pivot lane records it), scale it, finish row 2 P of W (Gauss-Jordan: W[k][:] *= 1 / l_kk), apply its rank-1
// update to column 2 P + 1 and row 2 P + 1 of W (the multiplier of row 2 P + 1 read from its own lane), then the same
// for column 2 P + 1; both columns are published through LDS as l and as l d (zero at and above their pivot rows), with
// their W rows (double-buffered by step parity). The update of step P: every wave applies the rank-2 update to its
// active columns (A[i][j] -= (l_ik d_k) l_jk over the two columns k of the pair, l_jk read as LDS broadcasts; rows above
// the diagonal of column j are never read or stored) and W rows (W[j][:] -= l_jk W[k][:]). Lookahead: the owner of
// pair P + 1 updates its two columns first and runs its pivot step while the other waves finish step P. The step
// loop is unrolled, so every register index is a constant.
constexpr int kDiagWaves = 8;

template <typename T>
__global__ __launch_bounds__(64 * kDiagWaves) void k_diag_block(const DiagTask* __restrict__ tasks, FactorArgs<T> a) {
  constexpr int E = kTileEdge, NW = kDiagWaves, NP = E / 2, JP = NP / NW, J = 2 * JP;
  __shared__ T colL[2][2][E], colD[2][2][E], rowW[2][2][E];
  const DiagTask t = tasks[blockIdx.x];
  const int lane = threadIdx.x & 63, nb = t.nb;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  T* F = a.Lx + t.f_off;
  T rA[J], rW[J];
#pragma unroll
  for (int s = 0; s < J; s++) {
    const int j = 2 * (NW * (s >> 1) + w) + (s & 1);
    rA[s] = (j < nb && lane < nb && lane >= j) ? F[lane + int64_t(j) * t.ld] : T(0);
    rW[s] = j == lane ? T(1) : T(0);
  }
  // pivot and scale column k held in slot s; publish it as entry c of buffer p. Sets l (zero at and above row k), its
  // update weight l d and the finished W row.
  auto pivotColumn = [&](const int s, const int k, const int p, const int c, T& lo, T& cdo, T& wko) {
    T inv, mult;
    const T dk = laneValue(rA[s], k);
    const T piv = pivotRule<T>(dk, t.col0 + k, a, &inv, &mult, lane == k);
    const T l = lane > k ? rA[s] * inv : T(0);
    if (lane == k) rA[s] = piv;
    if (lane > k) rA[s] = l;
    const T wk = a.ldlt ? rW[s] : rW[s] * inv;
    rW[s] = wk;
    colL[p][c][lane] = l;
    colD[p][c][lane] = l * mult;
    rowW[p][c][lane] = wk;
    lo = l;
    cdo = l * mult;
    wko = wk;
  };
  // pivot step of pair P (owner wave, slots s0 = 2 jj, s0 + 1): columns 2 P and 2 P + 1 (the second only if < nb)
  auto pivotPair = [&](const int P) {
    const int s0 = 2 * (P / NW), k = 2 * P, p = P & 1;
    T l, cd, wk;
    pivotColumn(s0, k, p, 0, l, cd, wk);
    if (k + 1 < nb) {
      const T lk1 = laneValue(l, k + 1);  // l of row k + 1
      rA[s0 + 1] -= cd * lk1;
      rW[s0 + 1] -= lk1 * wk;
      pivotColumn(s0 + 1, k + 1, p, 1, l, cd, wk);
    } else {
      colL[p][1][lane] = T(0);
      colD[p][1][lane] = T(0);
      rowW[p][1][lane] = T(0);
    }
  };
  if (nb > 0 && w == 0) pivotPair(0);
  __syncthreads();
#pragma clang loop unroll(full)
  for (int P = 0; P < NP; P++) {
    if (2 * P >= nb) continue;
    const int p = P & 1, P1 = P + 1;
    const T cd0 = colD[p][0][lane], cd1 = colD[p][1][lane], rw0 = rowW[p][0][lane], rw1 = rowW[p][1][lane];
    const bool ahead = 2 * P1 < nb && w == (P1 % NW);
    if (P1 < NP && ahead) {
      const int s0 = 2 * (P1 / NW);
#pragma unroll
      for (int c = 0; c < 2; c++) {
        const int j = 2 * P1 + c;
        const T l0 = colL[p][0][j], l1 = colL[p][1][j];
        rA[s0 + c] -= cd0 * l0 + cd1 * l1;
        rW[s0 + c] -= l0 * rw0 + l1 * rw1;
      }
      pivotPair(P1);
    }
#pragma unroll
    for (int s = 0; s < J; s++) {
      const int Ps = NW * (s >> 1) + w, j = 2 * Ps + (s & 1);
      if (NW * (s >> 1) + NW - 1 > P && Ps > P && !(ahead && Ps == P1)) {
        const T l0 = colL[p][0][j], l1 = colL[p][1][j];
        rA[s] -= cd0 * l0 + cd1 * l1;
        rW[s] -= l0 * rw0 + l1 * rw1;
      }
    }
    __syncthreads();
  }
  // W (column-major, edge 64) straight from the registers: each store writes row j of W across the 64 columns, the
  // L2 merges the rows into whole lines; no LDS transpose, so the kernel keeps a small LDS footprint and fits on a CU
  // beside the trailing-update workgroups it runs with
  T* Wt = a.W + t.w_off;
#pragma unroll
  for (int s = 0; s < J; s++) {
    const int j = 2 * (NW * (s >> 1) + w) + (s & 1);
    if (j < nb && lane < nb && lane >= j) F[lane + int64_t(j) * t.ld] = rA[s];
    Wt[j + E * lane] = (j < nb && lane < nb) ? rW[s] : T(0);
  }
}

template <typename T>
void launchDiagBlocks(const DiagTask* tasks, unsigned count, const FactorArgs<T>& a, hipStream_t st) {
  hipLaunchKernelGGL(k_diag_block<T>, dim3(count), dim3(64 * kDiagWaves), 0, st, tasks, a);
}

template void launchDiagBlocks<double>(const DiagTask*, unsigned, const FactorArgs<double>&, hipStream_t);
template void launchDiagBlocks<float>(const DiagTask*, unsigned, const FactorArgs<float>&, hipStream_t);

}  // namespace amp
