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
constexpr int kDiagStep = 4;  // columns eliminated per barrier

template <typename T, bool LDLT>
__global__ __launch_bounds__(64 * kDiagWaves) void k_diag_block(const DiagTask* __restrict__ tasks, FactorArgs<T> a) {
  constexpr int E = kTileEdge, NW = kDiagWaves, B = kDiagStep, NG = E / B, JG = NG / NW, J = B * JG;
  __shared__ T colL[2][B][E], colD[2][B][E], rowW[2][B][E];
  const DiagTask t = tasks[blockIdx.x];
  const int lane = threadIdx.x & 63, nb = t.nb;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  T* F = a.Lx + t.f_off;
  const T tau = T(a.tau);
  // pivot results of column col0 + lane, recorded by this lane (owner waves only) and stored at the end
  T recD = T(0);
  uint8_t recP = 0;
  int32_t recFail = INT32_MAX;
  T rA[J], rW[J];  // slot s: column / W row B (NW (s / B) + w) + s % B
#pragma unroll
  for (int s = 0; s < J; s++) {
    const int j = B * (NW * (s / B) + w) + s % B;
    rA[s] = (j < nb && lane < nb && lane >= j) ? F[lane + int64_t(j) * t.ld] : T(0);
    rW[s] = j == lane ? T(1) : T(0);
  }
  // pivot step of group G (owner wave, slots s0 .. s0 + B - 1): its columns one after another, each pivoted, scaled
  // and applied to the group's later columns and W rows (multipliers read from their own lanes); published as entry
  // c of buffer p. Columns at or beyond nb publish zeros.
  auto pivotGroup = [&](const int G) {
    const int s0 = B * (G / NW), p = G & 1;
#pragma unroll
    for (int c = 0; c < B; c++) {
      const int k = B * G + c;
      if (k < nb) {
        T inv, mult;
        const T dk = laneValue(rA[s0 + c], k);
        const T piv = pivotRuleNB<T, LDLT>(dk, t.col0 + k, tau, &inv, &mult, lane == k, recD, recP, recFail);
        const T l = lane > k ? rA[s0 + c] * inv : T(0);
        if (lane == k) rA[s0 + c] = piv;
        if (lane > k) rA[s0 + c] = l;
        const T wk = LDLT ? rW[s0 + c] : rW[s0 + c] * inv;
        rW[s0 + c] = wk;
        const T cd = l * mult;
        colL[p][c][lane] = l;
        colD[p][c][lane] = cd;
        rowW[p][c][lane] = wk;
#pragma unroll
        for (int c2 = c + 1; c2 < B; c2++) {
          const T lk = laneValue(l, B * G + c2);  // l of the group's row B G + c2
          rA[s0 + c2] -= cd * lk;
          rW[s0 + c2] -= lk * wk;
        }
      } else {
        colL[p][c][lane] = T(0);
        colD[p][c][lane] = T(0);
        rowW[p][c][lane] = T(0);
      }
    }
  };
  if (nb > 0 && w == 0) pivotGroup(0);
  __syncthreads();
  // rank-B update of slot s from the group published in buffer p (cd, rw: this lane's weights and W row entries)
  auto update = [&](const int s, const int j, const int p, const T (&cd)[B], const T (&rw)[B]) {
    T da = T(0), dw = T(0);
#pragma unroll
    for (int c = 0; c < B; c++) {
      const T lj = colL[p][c][j];
      da += cd[c] * lj;
      dw += lj * rw[c];
    }
    rA[s] -= da;
    rW[s] -= dw;
  };
#pragma clang loop unroll(full)
  for (int G = 0; G < NG; G++) {
    if (B * G >= nb) continue;
    const int p = G & 1, G1 = G + 1;
    T cd[B], rw[B];
#pragma unroll
    for (int c = 0; c < B; c++) {
      cd[c] = colD[p][c][lane];
      rw[c] = rowW[p][c][lane];
    }
    const bool ahead = B * G1 < nb && w == (G1 % NW);
    if (G1 < NG && ahead) {
      const int s0 = B * (G1 / NW);
#pragma unroll
      for (int c = 0; c < B; c++) update(s0 + c, B * G1 + c, p, cd, rw);
      pivotGroup(G1);
    }
#pragma unroll
    for (int s = 0; s < J; s++) {
      const int Gs = NW * (s / B) + w, j = B * Gs + s % B;
      if (NW * (s / B) + NW - 1 > G && Gs > G && !(ahead && Gs == G1)) update(s, j, p, cd, rw);
    }
    __syncthreads();
  }
  // W (column-major, edge 64) straight from the registers: each store writes row j of W across the 64 columns, the
  // L2 merges the rows into whole lines; no LDS transpose, so the kernel keeps a small LDS footprint and fits on a CU
  // beside the trailing-update workgroups it runs with
  // the pivot lane of every column belongs to the column's owner wave
  const bool owner = lane < nb && w == (lane / B) % NW;
  if (owner && recFail != INT32_MAX) atomicMin(a.failCol, recFail);
  if (LDLT && owner) {
    a.D[t.col0 + lane] = recD;
    a.pert[t.col0 + lane] = recP;
  }
  T* Wt = a.W + t.w_off;
#pragma unroll
  for (int s = 0; s < J; s++) {
    const int j = B * (NW * (s / B) + w) + s % B;
    if (j < nb && lane < nb && lane >= j) F[lane + int64_t(j) * t.ld] = rA[s];
    Wt[j + E * lane] = (j < nb && lane < nb) ? rW[s] : T(0);
  }
}

template <typename T>
void launchDiagBlocks(const DiagTask* tasks, unsigned count, const FactorArgs<T>& a, hipStream_t st) {
  if (a.ldlt)
    hipLaunchKernelGGL((k_diag_block<T, true>), dim3(count), dim3(64 * kDiagWaves), 0, st, tasks, a);
  else
    hipLaunchKernelGGL((k_diag_block<T, false>), dim3(count), dim3(64 * kDiagWaves), 0, st, tasks, a);
}

template void launchDiagBlocks<double>(const DiagTask*, unsigned, const FactorArgs<double>&, hipStream_t);
template void launchDiagBlocks<float>(const DiagTask*, unsigned, const FactorArgs<float>&, hipStream_t);

}  // namespace amp
