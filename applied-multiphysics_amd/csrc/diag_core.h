// Elimination of one 64 x 64 diagonal block of a tiled front (device helpers of k_diag_block and k_panel_step).
#pragma once
#include "pivot.h"

namespace amp {

constexpr int kDiagWaves = 8;
constexpr int kDiagStep = 4;  // columns eliminated per barrier
constexpr int kDiagSlots = kDiagStep * (kTileEdge / kDiagStep / kDiagWaves);  // columns per lane

// The elimination of one block in compute precision TC from storage of precision TS (F: the block, leading dimension
// t.ld). Leaves the factored columns in rA, the inverse's rows in rW, the owner lanes' pivot records, and in gmax the
// lane's largest elimination multiplier |l_ik / l_kk| (Cholesky) or |l_ik| (LDL^T). lds: 3 x 2 x B x 64 entries of TC.
template <typename TC, bool LDLT, typename TS>
__device__ __forceinline__ void diagCore(const DiagTask& t, const TS* F, TC tau, int w, int lane, TC (&rA)[kDiagSlots],
                                         TC (&rW)[kDiagSlots], TC* lds, TC& recD, uint8_t& recP, int32_t& recFail,
                                         TC& gmax) {
  constexpr int E = kTileEdge, NW = kDiagWaves, B = kDiagStep, NG = E / B, J = kDiagSlots;
  TC* colL = lds;                // [2][B][E]
  TC* colD = lds + 2 * B * E;    // [2][B][E]
  TC* rowW = lds + 4 * B * E;    // [2][B][E]
  const int nb = t.nb;
#pragma unroll
  for (int s = 0; s < J; s++) {
    const int j = B * (NW * (s / B) + w) + s % B;
    rA[s] = (j < nb && lane < nb && lane >= j) ? TC(F[lane + int64_t(j) * t.ld]) : TC(0);
    rW[s] = j == lane ? TC(1) : TC(0);
  }
  // pivot step of group G (owner wave, slots s0 .. s0 + B - 1): its columns one after another, each pivoted, scaled
  // and applied to the group's later columns and W rows (multipliers read from their own lanes); published as entry
  // c of buffer p. Columns at or beyond nb publish zeros.
  auto pivotGroup = [&](const int G) {
    const int s0 = B * (G / NW), p = G & 1;
#pragma unroll
    for (int c = 0; c < B; c++) {
      const int k = B * G + c;
      const int at = (p * B + c) * E + lane;
      if (k < nb) {
        TC inv, mult;
        const TC dk = laneValue(rA[s0 + c], k);
        const TC piv = pivotRuleNB<TC, LDLT>(dk, t.col0 + k, tau, &inv, &mult, lane == k, recD, recP, recFail);
        const TC l = lane > k ? rA[s0 + c] * inv : TC(0);
        gmax = fmax(gmax, fabs(LDLT ? l : l * inv));
        if (lane == k) rA[s0 + c] = piv;
        if (lane > k) rA[s0 + c] = l;
        const TC wk = LDLT ? rW[s0 + c] : rW[s0 + c] * inv;
        rW[s0 + c] = wk;
        const TC cd = l * mult;
        colL[at] = l;
        colD[at] = cd;
        rowW[at] = wk;
#pragma unroll
        for (int c2 = c + 1; c2 < B; c2++) {
          const TC lk = laneValue(l, B * G + c2);  // l of the group's row B G + c2
          rA[s0 + c2] -= cd * lk;
          rW[s0 + c2] -= lk * wk;
        }
      } else {
        colL[at] = TC(0);
        colD[at] = TC(0);
        rowW[at] = TC(0);
      }
    }
  };
  if (nb > 0 && w == 0) pivotGroup(0);
  __syncthreads();
  // rank-B update of slot s from the group published in buffer p (cd, rw: this lane's weights and W row entries)
  auto update = [&](const int s, const int j, const int p, const TC (&cd)[B], const TC (&rw)[B]) {
    TC da = TC(0), dw = TC(0);
#pragma unroll
    for (int c = 0; c < B; c++) {
      const TC lj = colL[(p * B + c) * E + j];
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
    TC cd[B], rw[B];
#pragma unroll
    for (int c = 0; c < B; c++) {
      cd[c] = colD[(p * B + c) * E + lane];
      rw[c] = rowW[(p * B + c) * E + lane];
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
}

// Store the block's factor (in place), the inverse W (column-major, edge 64, straight from the registers: each store
// writes row j of W across the 64 columns, the L2 merges the rows into whole lines; no LDS transpose, so the kernel
// keeps a small LDS footprint and fits on a CU beside the trailing-update workgroups it runs with) and, from the owner
// lanes, the pivot records.
template <typename TC, bool LDLT, typename TS>
__device__ __forceinline__ void diagStore(const DiagTask& t, const FactorArgs<TS>& a, int w, int lane,
                                          const TC (&rA)[kDiagSlots], const TC (&rW)[kDiagSlots], TC recD,
                                          uint8_t recP, int32_t recFail) {
  constexpr int E = kTileEdge, NW = kDiagWaves, B = kDiagStep, J = kDiagSlots;
  const int nb = t.nb;
  TS* F = a.Lx + t.f_off;
  // the pivot lane of every column belongs to the column's owner wave
  const bool owner = lane < nb && w == (lane / B) % NW;
  if (owner && recFail != INT32_MAX) atomicMin(a.failCol, recFail);
  if (LDLT && owner) {
    a.D[t.col0 + lane] = TS(recD);
    a.pert[t.col0 + lane] = recP;
  }
  TS* Wt = a.W + t.w_off;
#pragma unroll
  for (int s = 0; s < J; s++) {
    const int j = B * (NW * (s / B) + w) + s % B;
    if (j < nb && lane < nb && lane >= j) F[lane + int64_t(j) * t.ld] = TS(rA[s]);
    Wt[j + E * lane] = (j < nb && lane < nb) ? TS(rW[s]) : TS(0);
  }
}

}  // namespace amp
