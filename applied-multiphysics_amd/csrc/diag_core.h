// Elimination of one 64 x 64 diagonal block of a tiled front (device helpers of k_diag_block and k_panel_step).
#pragma once
#include <type_traits>

#include "pivot.h"

namespace amp {

constexpr int kDiagWaves = 8;
constexpr int kDiagStep = 4;  // columns eliminated per barrier
constexpr int kDiagSlots = kDiagStep * (kTileEdge / kDiagStep / kDiagWaves);  // columns per lane

// The block's lower part (F: the block, leading dimension t.ld, storage precision TS) into the slots rA in compute
// precision TC, and the identity into rW: the elimination's inputs.
template <typename TC, typename TS>
__device__ __forceinline__ void diagLoad(const DiagTask& t, const TS* F, int w, int lane, TC (&rA)[kDiagSlots],
                                         TC (&rW)[kDiagSlots]) {
  constexpr int E = kTileEdge, NW = kDiagWaves, B = kDiagStep, J = kDiagSlots;
  static_assert(E == B * NW * (J / B), "slots cover the block");
  const int nb = t.nb;
#pragma unroll
  for (int s = 0; s < J; s++) {
    const int j = B * (NW * (s / B) + w) + s % B;
    rA[s] = (j < nb && lane < nb && lane >= j) ? TC(F[lane + int64_t(j) * t.ld]) : TC(0);
    rW[s] = j == lane ? TC(1) : TC(0);
  }
}

// The elimination of the loaded slots in compute precision TC. Leaves the factored columns in rA, the inverse's rows
// in rW, the owner lanes' pivot records, and in gmax the lane's largest elimination multiplier |l_ik / l_kk|
// (Cholesky) or |l_ik| (LDL^T). lds: 3 x 2 x B x 64 entries of TC, aligned to B entries.
template <typename TC, bool LDLT>
__device__ __forceinline__ void diagEliminate(const DiagTask& t, TC tau, int w, int lane, TC (&rA)[kDiagSlots],
                                              TC (&rW)[kDiagSlots], TC* lds, TC& recD, uint8_t& recP,
                                              int32_t& recFail, TC& gmax) {
  constexpr int E = kTileEdge, NW = kDiagWaves, B = kDiagStep, NG = E / B, J = kDiagSlots;
  using TV = typename Vec4<TC>::type;
  static_assert(B == 4, "a group's published entries of one row are one vector");
  // published groups, row-major ([2][E] vectors of the group's B entries): one vector access per row and array
  TV* colL = reinterpret_cast<TV*>(lds);              // multipliers l
  TV* colD = reinterpret_cast<TV*>(lds + 2 * B * E);  // weights l * mult
  TV* rowW = reinterpret_cast<TV*>(lds + 4 * B * E);  // W row entries
  const int nb = t.nb;
  // pivot step of group G (owner wave, slots s0 .. s0 + B - 1): its columns one after another, each pivoted, scaled
  // and applied to the group's later columns and W rows (multipliers read from their own lanes); published as entry
  // c of buffer p. Columns at or beyond nb publish zeros.
  auto pivotGroup = [&](const int G, auto s1) {
    constexpr int s0 = decltype(s1)::value;  // B (G / NW)
    const int p = G & 1;
    TV lv, cdv, wv;
#pragma unroll
    for (int c = 0; c < B; c++) {
      const int k = B * G + c;
      if (k < nb) {
        TC inv, mult;
        const TC dk = laneValue(rA[s0 + c], k);
        const TC piv = pivotRuleNB<TC, LDLT>(dk, t.col0 + k, tau, &inv, &mult, lane == k, recD, recP, recFail);
        // the rows above the diagonal (lanes < k) hold values that are never stored or read again (pivots read the
        // diagonal, updates read the rows of later columns), so the multipliers are not masked there
        const TC l = rA[s0 + c] * inv;
        gmax = fmax(gmax, lane > k ? fabs(LDLT ? l : l * inv) : TC(0));
        rA[s0 + c] = lane > k ? l : piv;
        const TC wk = LDLT ? rW[s0 + c] : rW[s0 + c] * inv;
        rW[s0 + c] = wk;
        const TC cd = l * mult;
        lv[c] = l;
        cdv[c] = cd;
        wv[c] = wk;
#pragma unroll
        for (int c2 = c + 1; c2 < B; c2++) {
          const TC lk = laneValue(l, B * G + c2);  // l of the group's row B G + c2
          rA[s0 + c2] -= cd * lk;
          rW[s0 + c2] -= lk * wk;
        }
      } else {
        lv[c] = TC(0);
        cdv[c] = TC(0);
        wv[c] = TC(0);
      }
    }
    colL[p * E + lane] = lv;
    colD[p * E + lane] = cdv;
    rowW[p * E + lane] = wv;
  };
  if (nb > 0 && w == 0) pivotGroup(0, std::integral_constant<int, 0>{});
  __syncthreads();
  // rank-B update of slot s from the group published in buffer p (cd, rw: this lane's weights and W row entries)
  auto update = [&](const int s, const int j, const int p, const TV& cd, const TV& rw) {
    const TV lj = colL[p * E + j];
    TC da = TC(0), dw = TC(0);
#pragma unroll
    for (int c = 0; c < B; c++) {
      da += cd[c] * lj[c];
      dw += lj[c] * rw[c];
    }
    rA[s] -= da;
    rW[s] -= dw;
  };
  // group G's step: the owner of group G + 1 updates that group's slots and pivots it, every wave applies group G to
  // its later slots. A runtime loop (the straight-line form is too large for the instruction cache, and a workgroup
  // runs it once); `s1`: the first slot of group G + 1 in its owner wave, B (G + 1) / (B NW) (compile-time, so that the
  // slots stay in registers)
  auto groupStep = [&](const int G, auto s1) {
    const int p = G & 1, G1 = G + 1;
    const TV cd = colD[p * E + lane], rw = rowW[p * E + lane];
    const bool ahead = B * G1 < nb && w == (G1 % NW);
    if (G1 < NG && ahead) {
      constexpr int s0 = decltype(s1)::value;
#pragma unroll
      for (int c = 0; c < B; c++) update(s0 + c, B * G1 + c, p, cd, rw);
      pivotGroup(G1, s1);
    }
#pragma unroll
    for (int s = 0; s < J; s++) {
      const int Gs = NW * (s / B) + w, j = B * Gs + s % B;
      if (NW * (s / B) + NW - 1 > G && Gs > G && !(ahead && Gs == G1)) update(s, j, p, cd, rw);
    }
    __syncthreads();
  };
  static_assert(J == 2 * B, "two groups per wave: groups 0 .. NW - 1 in slots 0 .. B - 1, the rest in B .. 2 B - 1");
  const int ng = (nb + B - 1) / B;  // groups with columns
#pragma unroll 1
  for (int G = 0; G < min(ng, NW - 1); G++) groupStep(G, std::integral_constant<int, 0>{});
#pragma unroll 1
  for (int G = NW - 1; G < ng; G++) groupStep(G, std::integral_constant<int, B>{});
}

// diagLoad, then diagEliminate.
template <typename TC, bool LDLT, typename TS>
__device__ __forceinline__ void diagCore(const DiagTask& t, const TS* F, TC tau, int w, int lane, TC (&rA)[kDiagSlots],
                                         TC (&rW)[kDiagSlots], TC* lds, TC& recD, uint8_t& recP, int32_t& recFail,
                                         TC& gmax) {
  diagLoad(t, F, w, lane, rA, rW);
  diagEliminate<TC, LDLT>(t, tau, w, lane, rA, rW, lds, recD, recP, recFail, gmax);
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
  if constexpr (LDLT) {
    if (owner && recFail != INT32_MAX) atomicMin(a.failCol, recFail);
  } else {
    // Cholesky: the factored diagonal of column `lane` (its pivot lane's slot) is NaN where the pivot was not positive
    TC piv = TC(1);
#pragma unroll
    for (int s = 0; s < J; s++)
      if (B * (NW * (s / B) + w) + s % B == lane) piv = rA[s];
    if (owner && !(piv > TC(0))) atomicMin(a.failCol, t.col0 + lane);
  }
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
