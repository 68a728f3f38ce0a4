// Diagonal blocks of tiled fronts (gfx950): factor a 64 x 64 block and build the inverse of its lower factor.
#include <type_traits>

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
// step G. The step loop is unrolled, so every register index is a constant.
// GROWTH (FP32 factor with a growth limit, a.growthLimit > 0): the block is eliminated in FP32; when its largest
// elimination multiplier exceeds the limit, the block (still unmodified in memory) is eliminated again in FP64
// arithmetic, with FP64 pivots, and stored rounded to FP32; a.fallbacks counts such blocks.
template <typename T, bool LDLT, bool GROWTH>
__global__ __launch_bounds__(64 * kDiagWaves) void k_diag_block(const DiagTask* __restrict__ tasks, FactorArgs<T> a) {
  constexpr int J = kDiagSlots;
  __shared__ double ldsRaw[3 * 2 * kDiagStep * kTileEdge];
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
